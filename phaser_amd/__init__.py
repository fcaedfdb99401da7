"""phaser_amd: MI355X (gfx950) backend for phaser's reconstruction engines.

HIP kernels (phaser_amd/csrc, C ABI in include/phaser_mi355x.h) driven from Python through
ctypes, with PyTorch-ROCm tensors providing device memory and streams.

    from phaser_amd.engines.conventional import run_engine
    state = run_engine({'data': patterns, 'state': state, 'dtype': numpy.float32,
                        'seed': 0, 'observer': None, 'recons_name': 'x'}, plan)
"""
BACKEND_NAME = "mi355x"

__all__ = ["BACKEND_NAME"]
