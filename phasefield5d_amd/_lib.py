"""ctypes binding of the C ABI in include/pf5d.h (libpf5d.so, built for gfx950).

There is no CPU fallback: if the library is missing or no GPU is visible,
``lib()`` raises and every device entry point fails with that error.
"""
import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpf5d.so")

F64, F32 = 0, 1
_CODES = {-1: ValueError, -2: RuntimeError, -3: NotImplementedError, -4: MemoryError}

vp, i32, i64, f64 = C.c_void_p, C.c_int, C.c_long, C.c_double
P_i64, P_f64, P_i32 = C.POINTER(C.c_long), C.POINTER(C.c_double), C.POINTER(C.c_int)


class SimDesc(C.Structure):
    """Mirror of ``pf5d_sim_desc``."""
    _fields_ = [("dim", i32), ("shape", i64 * 3), ("cell_size", f64), ("kappa", f64 * 4), ("X0", f64 * 4),
                ("l", f64 * 4), ("elastic", i32), ("ctld", f64), ("dt_cfl", f64), ("dt0", f64),
                ("lower_limit", f64), ("upper_limit", f64), ("t_end", f64), ("oob_interval", i32),
                ("tol_soft", f64), ("tol_hard", f64)]


_SIGS = {
    "pf5d_last_error": (C.c_char_p, []),
    "pf5d_version": (i32, []),
    "pf5d_device_info": (i32, [P_i32]),
    "pf5d_table_create": (i32, [i32, i32, i32, P_f64, f64, C.POINTER(vp)]),
    "pf5d_table_destroy": (i32, [vp]),
    "pf5d_interpolate": (i32, [i32, vp, i64, vp, vp, i64, vp]),
    "pf5d_laplacian": (i32, [i32, vp, i32, P_i64, f64, vp, vp]),
    "pf5d_kappa_subtract": (i32, [i32, vp, vp, i64, i64, P_f64, vp]),
    "pf5d_divergence_from_mu": (i32, [i32, vp, vp, i64, vp, i64, i32, P_i64, f64, vp, vp]),
    "pf5d_abs_max": (i32, [i32, vp, i64, P_f64, vp]),
    "pf5d_gradients_pm": (i32, [i32, vp, i64, i32, P_i64, f64, vp, vp, vp]),
    "pf5d_fluxes_pm": (i32, [i32, vp, vp, i64, i32, P_i64, vp, vp, vp, vp, vp]),
    "pf5d_field_stats": (i32, [i32, vp, i64, i64, i32, P_f64, vp]),
    "pf5d_field_dot": (i32, [i32, vp, vp, i64, i32, P_f64, vp]),
    "pf5d_update": (i32, [i32, vp, vp, i64, f64, vp]),
    "pf5d_clamp_simplex": (i32, [i32, vp, i64, vp]),
    "pf5d_bounds_stats": (i32, [i32, vp, i64, P_f64, vp]),
    "pf5d_bounds_stats_noalloc": (i32, [i32, vp, i64, P_f64, vp]),
    "pf5d_elastic_create": (i32, [i32, i32, P_i64, P_f64, P_f64, P_f64, C.POINTER(vp)]),
    "pf5d_elastic_apply": (i32, [vp, vp, i64, vp, vp]),
    "pf5d_elastic_psi": (vp, [vp]),
    "pf5d_elastic_destroy": (i32, [vp]),
    "pf5d_engine_create": (i32, [i32, vp, C.POINTER(vp)]),
    "pf5d_engine_destroy": (i32, [vp]),
    "pf5d_engine_add": (i32, [vp, C.POINTER(SimDesc), P_f64, P_f64, vp, P_i32]),
    "pf5d_engine_add_generated": (i32, [vp, C.POINTER(SimDesc), C.c_uint64, f64, P_f64, vp, P_i32]),
    "pf5d_engine_remove": (i32, [vp, i32]),
    "pf5d_engine_num_variants": (i32, [vp, P_i32, P_i32]),
    "pf5d_engine_set_tiling": (i32, [vp, i32, i32]),
    "pf5d_engine_set_fft": (i32, [vp, i32, i32]),
    "pf5d_engine_set_fft_planes": (i32, [vp, i32, i32]),
    "pf5d_engine_set_fft_wave": (i32, [vp, i32]),
    "pf5d_engine_step": (i32, [vp, i32]),
    "pf5d_engine_prepare": (i32, [vp, i32]),
    "pf5d_engine_sync": (i32, [vp]),
    "pf5d_engine_get_state": (i32, [vp, i32, P_f64]),
    "pf5d_engine_set_time": (i32, [vp, i32, f64, f64]),
    "pf5d_engine_get_all_states": (i32, [vp, i32, P_f64, P_i32]),
    "pf5d_engine_get_field": (i32, [vp, i32, i32, P_f64]),
    "pf5d_engine_set_field": (i32, [vp, i32, P_f64]),
    "pf5d_engine_field_ptr": (vp, [vp, i32, i32]),
    "pf5d_engine_stream": (vp, [vp]),
    "pf5d_engine_stats": (i32, [vp, P_i64, P_i64]),
    "pf5d_engine_launch_info": (i32, [vp, P_i64]),
    "pf5d_engine_num_groups": (i32, [vp, P_i32]),
    "pf5d_engine_time_kernel": (i32, [vp, i32, i32, P_f64]),
    "pf5d_engine_time_step_stages": (i32, [vp, i32, i32, P_f64]),
}

EXPORTED = tuple(_SIGS)

_lock = threading.Lock()
_lib = None


def load_library(path=LIB_PATH):
    """dlopen the library and attach prototypes (no GPU access happens here)."""
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                          "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    h = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    return h


def lib():
    global _lib
    with _lock:
        if _lib is None:
            _lib = load_library()
        return _lib


def check(rc):
    if rc != 0:
        msg = lib().pf5d_last_error().decode(errors="replace")
        raise _CODES.get(rc, RuntimeError)(msg)
    return rc


def require_gpu():
    """Raise unless a ROCm GPU is visible to torch (the step has no CPU path)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("phasefield5d_amd needs a ROCm GPU: torch.cuda.is_available() is False "
                           "(the explicit step has no CPU fallback)")


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def darr(vals):
    return (C.c_double * len(vals))(*[float(v) for v in vals])


def larr(vals):
    return (C.c_long * len(vals))(*[int(v) for v in vals])


def ptr_f64(a):
    """ctypes double* to a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(P_f64)
