"""ctypes binding of libphaser_mi355x.so (declared in include/phaser_mi355x.h).

The library is loaded from this package's directory only; if it is missing the backend
refuses to run (there is no CPU fallback)."""
import ctypes
import os
import threading

from .build import LIB_PATH

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u32 = ctypes.c_uint32
c_f32 = ctypes.c_float
c_vp = ctypes.c_void_p

PHX_EINVAL = -1000
PHX_EUNSUPPORTED = -1001
PHX_ETOOLARGE = -1002


class GroupDesc(ctypes.Structure):
    """Mirror of `phx_group_desc` (include/phaser_mi355x.h)."""
    _fields_ = [
        ("n", c_i32), ("n_modes", c_i32), ("n_slices", c_i32), ("obj_h", c_i32), ("obj_w", c_i32),
        ("bmax", c_i32), ("group_index", c_i32), ("n_pos", c_i32),
        ("group_pos", c_vp), ("group_size", c_vp), ("start", c_vp), ("shift", c_vp),
        ("ky_s", c_vp), ("kx_s", c_vp), ("twiddle", c_vp), ("object", c_vp), ("probe", c_vp),
        ("props", c_vp), ("patterns_s", c_vp), ("mask_s", c_vp), ("params", c_vp),
        ("psi", c_vp), ("chi_a", c_vp), ("chi_b", c_vp), ("farfield", c_vp), ("intensity", c_vp),
        ("gpm", c_vp), ("d_obj", c_vp), ("d_probe", c_vp), ("probe_k", c_vp), ("gom", c_vp),
        ("red", c_vp), ("errors", c_vp), ("aux", c_vp), ("finite_flag", c_vp),
        ("psi_slots", c_i32), ("prop_zero_waves", c_i32), ("grad_obj", c_vp), ("grad_probe", c_vp),
        ("chi_slices", c_vp), ("tilt", c_vp), ("slice_dz", c_vp), ("tprops", c_vp),
    ]


_P = ctypes.POINTER(GroupDesc)

# name -> (restype, argtypes)
SIGNATURES = {
    "phx_get_cutouts_f32": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_f32, c_vp]),
    "phx_get_cutouts_c64": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_f32, c_f32, c_vp]),
    "phx_add_cutouts_f32": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp]),
    "phx_add_cutouts_c64": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp]),
    "phx_set_cutouts_f32": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp]),
    "phx_set_cutouts_c64": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp]),
    "phx_fill_u32": (c_int, [c_vp, c_u32, c_i64, c_vp]),
    "phx_fft2_c64": (c_int, [c_vp, c_vp, c_i64, c_i64, c_int, c_vp, c_vp]),
    "phx_scramble": (c_int, [c_vp, c_vp, c_i64, c_i64, c_int, c_int, c_vp]),
    "phx_scale_c64": (c_int, [c_vp, c_i64, c_f32, c_vp]),
    "phx_check_finite_f32": (c_int, [c_vp, c_i64, c_vp, c_i32, c_vp]),
    "phx_clamp_amplitude_c64": (c_int, [c_vp, c_i64, c_f32, c_vp, c_i32, c_vp]),
    "phx_probe_limit_support": (c_int, [_P, c_vp, c_vp]),
    "phx_regularize_layers_c64": (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_f32, c_f32, c_vp]),
    "phx_cost_regularizers_c64": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_int, c_vp]),
    "phx_object_filter_c64": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp,
                                      c_vp, c_i64, c_vp]),
    "phx_red_doubles": (c_i64, [c_i32, c_i32, c_i32]),
    "phx_group_begin": (c_int, [_P, c_vp]),
    "phx_tilt_propagators": (c_int, [_P, c_vp]),
    "phx_probe_spectra": (c_int, [_P, c_vp]),
    "phx_group_obj_mag": (c_int, [_P, c_vp]),
    "phx_shift_probes": (c_int, [_P, c_int, c_int, c_vp]),
    "phx_forward_slice": (c_int, [_P, c_int, c_int, c_vp]),
    "phx_farfield": (c_int, [_P, c_int, c_int, c_vp]),
    "phx_lsqml_backward": (c_int, [_P, c_int, c_int, c_int, c_vp]),
    "phx_group_forward": (c_int, [_P, c_vp]),
    "phx_farfield_update": (c_int, [_P, c_int, c_vp]),
    "phx_lsqml_backward_sweep": (c_int, [_P, c_int, c_int, c_int, c_vp]),
    "phx_lsqml_object_update": (c_int, [_P, c_int, c_vp]),
    "phx_farfield_gradient_update": (c_int, [_P, c_int, c_vp]),
    "phx_gradient_sweep": (c_int, [_P, c_int, c_int, c_int, c_vp]),
    "phx_dry_run_stats_intensity": (c_int, [_P, c_vp]),
    "phx_gradient_sweep_farfield": (c_int, [_P, c_int, c_int, c_int, c_int, c_vp]),
    "phx_gradient_group_sweeps": (c_int, [_P, c_int, c_int, c_int, c_int, c_vp]),
    "phx_plane_task_plan": (c_int, [c_int, c_int, c_int, c_vp, c_int]),
    "phx_gradient_object_gather": (c_int, [_P, c_int, c_int, c_vp]),
    "phx_gradient_object_gather_begin": (c_int, [_P, c_int, c_int, _P, c_vp]),
    "phx_gemm_bf16": (c_int, [c_vp, c_i64, c_int, c_vp, c_i64, c_int, c_i64, c_i64, c_i64, c_vp, c_vp, c_i64, c_vp,
                              c_int, c_vp, c_i64, c_int, c_vp]),
    "phx_gemm_bf16_nt": (c_int, [c_vp, c_i64, c_vp, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp, c_i64, c_vp, c_i64,
                                 c_vp, c_int, c_vp, c_i64, c_int, c_vp]),
    "phx_transpose_bf16": (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp]),
    "phx_mse_grad_t": (c_int, [c_vp, c_vp, c_i64, c_i64, c_f32, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "phx_colsum_f32": (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp]),
    "phx_adam_f32_bf16": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_f32, c_f32, c_f32,
                                  c_f32, c_vp]),
    "phx_f32_to_bf16": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "phx_mse_grad": (c_int, [c_vp, c_vp, c_i64, c_f32, c_vp, c_vp, c_vp, c_vp]),
    "phx_pattern_features": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_f32, c_vp, c_vp]),
    "phx_object_patches": (c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_i64, c_f32, c_vp, c_vp]),
    "phx_stitch_patches": (c_int, [c_vp, c_i64, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp,
                                   c_vp]),
    "phx_lsqml_object_update_slices": (c_int, [_P, c_int, c_int, c_vp]),
    "phx_lsqml_object_update_slices_begin": (c_int, [_P, c_int, c_int, _P, c_vp]),
    "phx_probe_spectrum_reduce": (c_int, [_P, c_vp]),
    "phx_lsqml_probe_apply": (c_int, [_P, c_vp]),
    "phx_lsqml_probe_apply_strips": (c_int, [_P, c_vp, c_vp]),
    "phx_lsqml_probe_apply_strips_static": (c_int, [_P, c_vp]),
    "phx_probe_strip_scratch_elems": (c_i64, []),
    "phx_epie_backward": (c_int, [_P, c_int, c_int, c_int, c_vp]),
    "phx_epie_object_apply": (c_int, [_P, c_int, c_vp]),
    "phx_epie_probe_apply": (c_int, [_P, c_vp]),
    "phx_dry_run_stats": (c_int, [_P, c_vp]),
    "phx_position_gradient": (c_int, [_P, c_int, c_vp]),
    "phx_position_finish": (c_int, [_P, c_vp, c_vp]),
    "phx_exit_intensity": (c_int, [_P, c_vp, c_vp]),
    "phx_farfield_gradient": (c_int, [_P, c_int, c_vp]),
    "phx_gradient_backward": (c_int, [_P, c_int, c_int, c_int, c_int, c_vp]),
    "phx_gradient_probe_finish": (c_int, [_P, c_vp]),
    "phx_dry_run_stats_gradient": (c_int, [_P, c_vp]),
    "phx_sgd_step": (c_int, [c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_int, c_int, c_vp, c_i32, c_vp]),
    "phx_adam_step": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_f32, c_f32, c_f32, c_i32,
                              c_int, c_vp, c_i32, c_vp]),
    "phx_sqnorm_accumulate": (c_int, [c_vp, c_i64, c_f32, c_vp, c_vp]),
    "phx_polyak_step": (c_int, [c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_f32, c_f32, c_vp, c_vp, c_vp, c_i32,
                                c_vp]),
}

_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load the HIP library (once).  Raises RuntimeError when it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"{os.path.basename(path)} is not built; run `python -m phaser_amd.build` "
                "(the MI355X backend has no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


class HipError(RuntimeError):
    pass


def check(rc: int, what: str):
    if rc == 0:
        return
    if rc == PHX_EINVAL:
        raise ValueError(f"{what}: invalid argument")
    if rc == PHX_EUNSUPPORTED:
        raise NotImplementedError(f"{what}: size not supported by the HIP kernels")
    if rc == PHX_ETOOLARGE:
        raise ValueError(f"{what}: size exceeds the library's fixed device work array")
    raise HipError(f"{what}: HIP error {-rc}")


def call(name: str, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
