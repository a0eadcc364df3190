"""ctypes binding of libswm_gpu.so (declarations in include/swm_gpu.h).

The library is built in-tree by cubedsphere_sbp_swm_amd/build.py.  There is no CPU
fallback: if the library is missing or the device call fails, an exception is raised.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libswm_gpu.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

# name -> (restype, argtypes); the full exported C ABI of include/swm_gpu.h
SIGNATURES = {
    "swm_last_error": (ctypes.c_char_p, []),
    "swm_ctx_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                            ctypes.c_char_p, _d, _d, _d, _i]),
    "swm_ctx_destroy": (_i, [_vp]),
    "swm_ctx_info": (_i, [_vp, _ip, _ip, ctypes.POINTER(ctypes.c_longlong), _ip]),
    "swm_ctx_set_stream": (_i, [_vp, _vp]),
    "swm_ctx_synchronize": (_i, [_vp]),
    "swm_state_alloc": (_i, [_vp, ctypes.POINTER(_vp)]),
    "swm_state_free": (_i, [_vp, _vp]),
    "swm_state_upload": (_i, [_vp, _vp, _dp, _dp, _dp, _i]),
    "swm_state_download": (_i, [_vp, _vp, _dp, _dp, _dp, _i]),
    "swm_rhs": (_i, [_vp, _vp, _vp]),
    "swm_assign": (_i, [_vp, _vp, _d, _vp, _d, _vp]),
    "swm_update": (_i, [_vp, _vp, _d, _vp, _d, _vp]),
    "swm_rk4_steps": (_i, [_vp, _vp, _d, _i]),
    "swm_diagnostics": (_i, [_vp, _vp, _dp]),
    "swm_diagnostics_noalloc": (_i, [_vp, _vp, _vp]),
    "swm_set_orography": (_i, [_vp, _dp, _i]),
    "swm_ic_spherical_harmonics": (_i, [_vp, ctypes.c_ulonglong, _i, _d, _d, _d, _dp, _dp, _dp, _i]),
    "swm_ctx_set_tile_shape": (_i, [_vp, _i, _i, _i]),
    "swm_ctx_get_tile_shape": (_i, [_vp, _ip, _ip, _ip]),
    "swm_ctx_set_tiling": (_i, [_vp, _i, _i]),
    "swm_ctx_get_tiling": (_i, [_vp, _ip, _ip]),
    "swm_ctx_autotune": (_i, [_vp, _vp, _i, _dp]),
    "swm_bench_stage": (_i, [_vp, _vp, _i, _i]),
    "swm_debug_geometry": (_i, [_vp, _i, _dp, _i]),
    "swm_debug_halo_table": (_i, [_vp, _ip]),
    "swm_rhs_vjp": (_i, [_vp, _vp, _vp, _vp]),
    "swm_adj_create": (_i, [_vp, _i, _d, _i, ctypes.POINTER(_vp)]),
    "swm_adj_destroy": (_i, [_vp]),
    "swm_adj_forward": (_i, [_vp, _vp, _vp]),
    "swm_adj_backward": (_i, [_vp, _vp, _vp]),
    "swm_adj_set_tiling": (_i, [_vp, _i]),
    "swm_laplace": (_i, [_vp, ctypes.c_char_p, _vp, _vp]),
    "swm_diff_apply": (_i, [_vp, _vp, _vp, ctypes.c_char_p, _d, ctypes.c_char_p, _d]),
    "swm_diff_step": (_i, [_vp, _vp, _d, ctypes.c_char_p, _d, ctypes.c_char_p, _d]),
    "swm_state_copy_from_device": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "swm_state_copy_to_device": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "swm_ens_create": (_i, [ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i, ctypes.POINTER(_vp)]),
    "swm_ens_destroy": (_i, [_vp]),
    "swm_ens_rk4_steps": (_i, [_vp, _dp, _i]),
    "swm_ens_set_tile_shape": (_i, [_vp, _i, _i, _i, _i]),
    "swm_ens_get_tile_shape": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "swm_ens_set_tiling": (_i, [_vp, _i, _i, _i]),
    "swm_ens_get_tiling": (_i, [_vp, _ip, _ip, _ip]),
    "swm_ens_autotune": (_i, [_vp, _i, _dp]),
    "swm_ens_bench_stage": (_i, [_vp, _i]),
    "swm_ecs_halo_xy": (_i, [_vp, _vp, _i, _vp]),
    "swm_ecs_halo_xy_host": (_i, [_vp, _vp, _i, _dp]),
    "swm_ecs_halo_o": (_i, [_vp, _vp, _i, _vp]),
    "swm_ecs_halo_o_host": (_i, [_vp, _dp, _i, _dp]),
    "swm_ecs_halo_vec_a": (_i, [_vp, _vp, _vp, _i, _vp, _vp]),
    "swm_ecs_halo_vec_a_host": (_i, [_vp, _dp, _dp, _i, _dp, _dp]),
    "swm_latlon_outputer_create": (_i, [_vp, _i, _i, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "swm_latlon_vec_outputer_create": (_i, [_vp, _i, _i, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "swm_outputer_destroy": (_i, [_vp]),
    "swm_latlon_write": (_i, [_vp, _vp, ctypes.c_char_p, _i]),
    "swm_latlon_vec_write": (_i, [_vp, _vp, _vp, ctypes.c_char_p, ctypes.c_char_p, _i]),
    "swm_paneled_write": (_i, [_vp, _vp, ctypes.c_char_p, _i]),
    "swm_c_create": (_i, [ctypes.POINTER(_vp), _i] + [ctypes.c_char_p] * 8 + [_d, _d, _d]),
    "swm_c_destroy": (_i, [_vp]),
    "swm_c_info": (_i, [_vp, _ip, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_longlong)]),
    "swm_c_synchronize": (_i, [_vp]),
    "swm_c_state_alloc": (_i, [_vp, ctypes.POINTER(_vp)]),
    "swm_c_state_free": (_i, [_vp, _vp]),
    "swm_c_state_upload": (_i, [_vp, _vp, _dp, _dp, _dp, _i]),
    "swm_c_state_download": (_i, [_vp, _vp, _dp, _dp, _dp, _i]),
    "swm_c_set_orography": (_i, [_vp, _dp, _i]),
    "swm_c_rhs": (_i, [_vp, _vp, _vp]),
    "swm_c_rk4_steps": (_i, [_vp, _vp, _d, _i]),
    "swm_c_diagnostics": (_i, [_vp, _vp, _dp]),
    "swm_c_ecs_halo_vec": (_i, [_vp, ctypes.c_char_p, _vp, _i, _vp, _vp]),
    "swm_c_grad": (_i, [_vp, ctypes.c_char_p, _vp, _vp, _vp]),
    "swm_c_massflux": (_i, [_vp, ctypes.c_char_p, _vp, _vp, _vp, _vp, _vp]),
}

# host set-up routines (geometry.h), exported by the same library; they need no device
HOST_SIGNATURES = {
    "swmh_geometry": (None, [_i, _i, _d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "swmh_sync_weights": (None, [_i, _dp, _dp, _dp, _dp]),
    "swmh_quad_weights": (_i, [_i, _i, _dp]),
    "swmh_halo_table": (None, [_i, _i, _ip]),
    "swmh_ic_sh": (_i, [_i, _i, ctypes.c_ulonglong, _i, _d, _d, _d, _dp, _dp, _dp, _i]),
    "swmh_edge_sync_matrix": (_i, [_i, _i, _ip, _ip, _dp, _i]),
    "swmh_ecs_xy_weights": (_i, [_i, _i, _dp, _ip]),
    "swmh_basis_b": (None, [_i, _i, _dp, _dp]),
    "swmh_ecs_o_weights": (None, [_i, _i, _dp, _ip]),
    "swmh_ecs_vec_a_tm": (None, [_i, _i, _dp]),
    "swmh_latlon_table": (None, [_i, _i, _i, _ip, _ip, _ip, _dp, _dp, _dp]),
    "swmh_c_geometry": (None, [_i, _i, _i, _i, _i, _d, _dp, _dp, _dp, _dp, _dp, _dp]),
    "swmh_c_halo_lists": (_i, [_i, _i, _i, _i, ctypes.c_longlong, _i, _i] +
                          [ctypes.POINTER(ctypes.c_longlong)] * 2 + [ctypes.c_void_p] +
                          [ctypes.POINTER(ctypes.c_longlong)] * 2 + [ctypes.c_void_p] +
                          [ctypes.POINTER(ctypes.c_longlong)] * 4 + [_ip]),
    "swmh_c_ecs_vec_tables": (_i, [_i, _i, _i, _dp, _ip, _dp, _ip, _dp, _ip, _dp, _dp, _dp]),
    "swmh_c_ecs_strip_sources": (_i, [_i, _i, _i, _i, ctypes.c_longlong, _i] +
                                 [ctypes.POINTER(ctypes.c_longlong), ctypes.c_void_p] * 2 +
                                 [ctypes.POINTER(ctypes.c_longlong)]),
}

_lib = None


class SwmError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SwmError(f"{LIB_PATH} is missing: build it with cubedsphere_sbp_swm_amd/build.py "
                           "(there is no CPU fallback)")
        # PyTorch ships its own HIP runtime under the same soname (libamdhip64.so.7) as ROCm's.
        # Whichever is loaded first serves the whole process; when the library loads ROCm's
        # first, PyTorch initialised afterwards finds no device.  Loading PyTorch first makes
        # both share its runtime (the Fortran host, without PyTorch, uses ROCm's).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in {**SIGNATURES, **HOST_SIGNATURES}.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().swm_last_error()
        raise SwmError(f"swm status {rc}: {msg.decode() if msg else ''}")
    return rc
