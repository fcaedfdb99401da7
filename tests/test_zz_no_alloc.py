"""The no-allocation probe step (phx_lsqml_probe_apply_strips_static: its work plane is a
fixed device array of the library) against phx_lsqml_probe_apply_strips with a work plane
passed in, through the C ABI: the same kernels on the same inputs must give the same
bits.  Two calls in a row, the second on a smaller problem (the array keeps the first
call's contents), and a size beyond the array rejected before any launch."""
import ctypes

import numpy
import pytest

pytestmark = pytest.mark.gpu


def _after_backward_sweep(n, M, S, scan, seed):
    """A workspace whose group 0 has run its forward and backward sweeps and the
    probe-spectrum reduction: the inputs of the probe step."""
    from phaser_amd import _lib
    from phaser_amd.device import GroupWorkspace, stream_handle
    from phaser_amd.groups import GroupManager
    from phaser_amd.optics import make_propagators
    from phaser_amd.synthetic import SyntheticConfig, make_dataset
    from tests.helpers import oracle_simulate
    cfg = SyntheticConfig(n=n, n_modes=M, n_slices=S, scan_shape=scan, scan_step_px=6.0, jitter_px=0.5, seed=seed)
    state, pats, _truth = make_dataset(cfg, simulate=oracle_simulate)
    state = state.to_device()
    props = make_propagators(state.probe.sampling, state.wavelength, state.object.thicknesses, 2 / 3)
    groups = list(GroupManager(state.scan, 8, False, seed=3).iter(state.scan))
    ws = GroupWorkspace(state, pats.patterns, pats.pattern_mask, props, noise_offset=0.0, noise_eps=1e-3,
                        n_groups=len(groups), bmax=max(numpy.asarray(g).shape[-1] for g in groups), keep_chi=True)
    ws.set_positions(state.object.sampling, state.scan)
    ws.set_groups(groups)
    ws.set_params(beta_object=1.0, beta_probe=1.0, reg_object=1e-2, reg_probe=1e-2, gamma=1e-4)
    d, st = ws.desc_refs[0], stream_handle()
    for name, *args in (("phx_group_begin",), ("phx_probe_spectra",), ("phx_group_forward",),
                        ("phx_group_obj_mag",), ("phx_lsqml_backward_sweep", 1, 1, 0), ("phx_probe_spectrum_reduce",)):
        _lib.call(name, d, *args, st)
    return ws


def _probe_step(ws, static):
    """Run one form of the probe step from the current inputs; returns (probe, probe_k,
    d_probe) after it."""
    import torch
    from phaser_amd import _lib
    from phaser_amd.device import stream_handle
    d, st = ws.desc_refs[0], stream_handle()
    if static:
        _lib.call("phx_lsqml_probe_apply_strips_static", d, st)
    else:
        _lib.call("phx_lsqml_probe_apply_strips", d, ctypes.c_void_p(ws.probe_tmp.data_ptr()), st)
    torch.cuda.synchronize()
    return ws.probe.clone(), ws.probe_k.clone(), ws.d_probe.clone()


def test_static_work_plane_matches_work_plane_argument():
    import torch
    from phaser_amd import _lib
    cap = int(_lib.load().phx_probe_strip_scratch_elems())
    assert cap == 131072
    # the second problem is smaller than the first (fewer modes, smaller planes)
    for n, M, S, scan, seed in ((128, 4, 3, (4, 4), 8), (64, 2, 2, (4, 4), 9)):
        print(f"no-alloc probe step: n={n} modes={M} slices={S} scan={scan}", flush=True)
        ws = _after_backward_sweep(n, M, S, scan, seed)
        snap = (ws.probe.clone(), ws.probe_k.clone(), ws.d_probe.clone())
        ref = _probe_step(ws, static=False)
        ws.probe.copy_(snap[0])
        ws.probe_k.copy_(snap[1])
        ws.d_probe.copy_(snap[2])
        torch.cuda.synchronize()
        got = _probe_step(ws, static=True)
        assert not torch.equal(ref[0], snap[0])          # the step did move the probe
        assert torch.count_nonzero(ref[2]) == 0          # and consumed the spectrum
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
        assert ws.finite_group() is None


def test_static_work_plane_rejects_sizes_beyond_the_array():
    """n_modes n^2 above the array's capacity is refused before anything is launched."""
    from phaser_amd import _lib
    from phaser_amd.device import stream_handle
    ws = _after_backward_sweep(32, 1, 1, (3, 3), 4)
    big = _lib.GroupDesc.from_buffer_copy(ws.descs[0])
    big.n = 128
    big.n_modes = 9                      # 9 * 128^2 = 147456 > 131072
    with pytest.raises(ValueError, match="fixed device work array"):
        _lib.call("phx_lsqml_probe_apply_strips_static", ctypes.byref(big), stream_handle())
