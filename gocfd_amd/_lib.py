"""ctypes binding of libgocfd_euler2d.so (C ABI in include/gocfd_euler2d.h).

The library is built in-tree by gocfd_amd.build. Loading fails loudly when it
is missing: there is no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgocfd_euler2d.so")

OK, ERR_ARG, ERR_MESH, ERR_SETUP, ERR_DEVICE, ERR_NAN = 0, -1, -2, -3, -4, -5

# Every symbol declared in include/gocfd_euler2d.h
EXPORTS = [
    "gocfd_last_error", "gocfd_mesh_read", "gocfd_mesh_from_arrays", "gocfd_mesh_dims",
    "gocfd_mesh_arrays", "gocfd_mesh_free", "gocfd_euler2d_create", "gocfd_euler2d_destroy",
    "gocfd_euler2d_dims", "gocfd_euler2d_set_state", "gocfd_euler2d_get_state",
    "gocfd_euler2d_rhs", "gocfd_euler2d_step", "gocfd_euler2d_step_async", "gocfd_euler2d_sync",
    "gocfd_euler2d_time", "gocfd_euler2d_set_time", "gocfd_euler2d_get_residual",
    "gocfd_euler2d_get_edges", "gocfd_euler2d_get_connectivity", "gocfd_euler2d_get_operators",
    "gocfd_euler2d_get_geometry", "gocfd_euler2d_get_coords", "gocfd_euler2d_kernel_time",
    "gocfd_euler2d_tiling", "gocfd_euler2d_get_shock", "gocfd_euler2d_get_element_dt",
    "gocfd_euler2d_interp_matrix",
    "gocfd_euler2d_vjp", "gocfd_euler2d_jvp",
    "gocfd_euler2d_rhs_noalloc",
    "gocfd_ensemble_create", "gocfd_ensemble_destroy", "gocfd_ensemble_case_dims", "gocfd_ensemble_step",
    "gocfd_ensemble_get_state", "gocfd_ensemble_time", "gocfd_ensemble_kernel_time",
    "gocfd_ensemble_set_launch", "gocfd_ensemble_get_launch", "gocfd_ensemble_tune",
]


class Params(C.Structure):
    """gocfd_params (include/gocfd_euler2d.h)"""
    _fields_ = [
        ("polynomial_order", C.c_int32), ("flux_type", C.c_int32), ("init_type", C.c_int32),
        ("max_iterations", C.c_int32), ("cfl", C.c_double), ("final_time", C.c_double),
        ("minf", C.c_double), ("gamma", C.c_double), ("alpha", C.c_double),
        ("local_time_stepping", C.c_int32), ("limiter", C.c_int32), ("kappa", C.c_double),
        ("precision", C.c_int32), ("device", C.c_int32),
    ]


class GocfdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p


def load() -> C.CDLL:
    """Load the solver library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m gocfd_amd.build` "
            "(the solver has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.gocfd_last_error.restype = C.c_char_p
    L.gocfd_last_error.argtypes = []
    L.gocfd_mesh_read.argtypes = [C.c_char_p, C.POINTER(_vp)]
    L.gocfd_mesh_from_arrays.argtypes = [C.c_int32, C.c_int32, _dp, _dp, _i32p, C.c_int32,
                                         C.POINTER(C.c_char_p), _i32p, _i32p, _i32p, C.POINTER(_vp)]
    L.gocfd_mesh_dims.argtypes = [_vp, _i32p, _i32p, _i32p]
    L.gocfd_mesh_arrays.argtypes = [_vp, _dp, _dp, _i32p]
    L.gocfd_mesh_free.argtypes = [_vp]
    L.gocfd_mesh_free.restype = None
    L.gocfd_euler2d_create.argtypes = [_vp, C.POINTER(Params), C.POINTER(_vp)]
    L.gocfd_euler2d_destroy.argtypes = [_vp]
    L.gocfd_euler2d_destroy.restype = None
    L.gocfd_euler2d_dims.argtypes = [_vp, _i32p, _i32p, _i32p, _i32p, _i32p]
    L.gocfd_euler2d_set_state.argtypes = [_vp, _dp]
    L.gocfd_euler2d_get_state.argtypes = [_vp, _dp]
    L.gocfd_euler2d_get_residual.argtypes = [_vp, _dp]
    L.gocfd_euler2d_rhs.argtypes = [_vp, _dp, C.c_double, _dp]
    L.gocfd_euler2d_rhs_noalloc.argtypes = [_vp, _vp, C.c_double, _vp, _vp]
    L.gocfd_euler2d_step.argtypes = [_vp, C.c_int32]
    L.gocfd_euler2d_step_async.argtypes = [_vp, C.c_int32]
    L.gocfd_euler2d_sync.argtypes = [_vp]
    L.gocfd_euler2d_time.argtypes = [_vp, _dp, _dp, _i32p]
    L.gocfd_euler2d_set_time.argtypes = [_vp, C.c_double]
    L.gocfd_euler2d_get_edges.argtypes = [_vp, _i64p]
    L.gocfd_euler2d_get_connectivity.argtypes = [_vp, _i32p, _i32p]
    L.gocfd_euler2d_get_operators.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
    L.gocfd_euler2d_get_geometry.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.gocfd_euler2d_get_coords.argtypes = [_vp, _dp, _dp, _dp, _dp]
    L.gocfd_euler2d_kernel_time.argtypes = [_vp, _dp, _i32p]
    L.gocfd_euler2d_tiling.argtypes = [_vp, _i32p, _i32p, _i32p]
    L.gocfd_euler2d_get_shock.argtypes = [_vp, _dp, _dp, _dp]
    L.gocfd_euler2d_get_element_dt.argtypes = [_vp, _dp]
    L.gocfd_euler2d_vjp.argtypes = [_vp, C.c_int32, C.c_int32, _dp, _dp, _dp]
    L.gocfd_euler2d_interp_matrix.argtypes = [_vp, C.c_int32, _dp, _dp, _dp]
    L.gocfd_euler2d_jvp.argtypes = [_vp, C.c_int32, _dp, _dp, _dp, _dp]
    L.gocfd_ensemble_create.argtypes = [C.c_int32, C.POINTER(_vp), C.POINTER(Params), C.POINTER(_vp)]
    L.gocfd_ensemble_destroy.argtypes = [_vp]
    L.gocfd_ensemble_destroy.restype = None
    L.gocfd_ensemble_case_dims.argtypes = [_vp, C.c_int32, _i32p, _i32p]
    L.gocfd_ensemble_step.argtypes = [_vp, C.c_int32]
    L.gocfd_ensemble_get_state.argtypes = [_vp, C.c_int32, _dp]
    L.gocfd_ensemble_time.argtypes = [_vp, C.c_int32, _dp, _dp, _i32p]
    L.gocfd_ensemble_kernel_time.argtypes = [_vp, _dp, _i32p]
    L.gocfd_ensemble_set_launch.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32]
    L.gocfd_ensemble_get_launch.argtypes = [_vp, _i32p, _i32p, _i32p]
    L.gocfd_ensemble_tune.argtypes = [_vp, C.c_int32]
    for name in EXPORTS:  # ctypes default restype is int: every other entry returns a status
        getattr(L, name)
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != OK:
        raise GocfdError(rc, load().gocfd_last_error().decode(errors="replace"))
