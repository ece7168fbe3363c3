"""ctypes binding of libc3d.so (the C ABI declared in include/c3d.h).

The library is the only compute path: if it is missing or cannot be loaded
this module raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libc3d.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)

C3D_OK = 0
ERRORS = {-1: "C3D_EINVAL", -2: "C3D_EHIP", -3: "C3D_ENOMEM", -4: "C3D_ENOCONV", -5: "C3D_ENAN"}
PREC_FP64 = 0
PREC_MIXED = 1
RESTORE_P, RESTORE_STATE, RESTORE_FORCE, RESTORE_ALL = 1, 2, 4, 7


class c3d_desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("grid_type", C.c_int),
        ("Lxp", C.c_double), ("Lzp", C.c_double), ("Ly", C.c_double), ("alpha", C.c_double),
        ("nu", C.c_double), ("precision", C.c_int), ("min_buffer_width", C.c_double),
    ]


# (name, restype, argtypes) of every entry point in include/c3d.h
SIGNATURES = [
    ("c3d_create", C.c_int, [C.POINTER(c3d_desc), C.POINTER(C.c_void_p)]),
    ("c3d_destroy", C.c_int, [C.c_void_p]),
    ("c3d_last_error", C.c_char_p, [C.c_void_p]),
    ("c3d_grid_get", C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int]),
    ("c3d_set_fields", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    ("c3d_get_fields", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    ("c3d_apply_boundary_conditions", C.c_int, [C.c_void_p]),
    ("c3d_divergence", C.c_int, [C.c_void_p, _dp]),
    ("c3d_gradient", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    ("c3d_apply_boundary_conditions_arrays", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("c3d_divergence_arrays", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    ("c3d_solve_poisson_equation", C.c_int, [C.c_void_p, _dp, C.c_int, C.c_double]),
    ("c3d_compute_non_ib_projection", C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    ("c3d_check_divergence", C.c_int, [C.c_void_p, _dp]),
    ("c3d_wall_shear", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    ("c3d_plane_statistics", C.c_int, [C.c_void_p, _dp]),
    ("c3d_monitor_noalloc", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    ("c3d_monitor", C.c_int, [C.c_void_p, _dp]),
    ("c3d_set_pressure_gradient", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    ("c3d_get_pressure_gradient", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("c3d_save_state", C.c_int, [C.c_void_p]),
    ("c3d_compute_rhs", C.c_int, [C.c_void_p, C.c_int]),
    ("c3d_get_rhs", C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    ("c3d_compute_rhs_arrays", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("c3d_rk3_update", C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    ("c3d_compute_dt", C.c_int, [C.c_void_p, C.c_double, _dp]),
    ("c3d_mean_mass_flow_u", C.c_int, [C.c_void_p, _dp]),
    ("c3d_mean_mass_flow_u_array", C.c_int, [C.c_void_p, _dp, _dp]),
    ("c3d_add_to_u", C.c_int, [C.c_void_p, C.c_double]),
    ("c3d_time_step_rk3", C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_double]),
    ("c3d_set_rhs_exact", C.c_int, [C.c_void_p, C.c_int]),
    ("c3d_time_projection", C.c_int, [C.c_void_p, C.c_int, C.c_double, _dp, _dp, C.c_int]),
    ("c3d_copy_bandwidth", C.c_int, [C.c_void_p, C.c_size_t, C.c_int, _dp]),
    ("c3d_stage_count", C.c_int, []),
    ("c3d_stage_name", C.c_char_p, [C.c_int]),
    ("c3d_set_refinement", C.c_int, [C.c_void_p, C.c_int]),
    ("c3d_refinement_residuals", C.c_int, [C.c_void_p, _dp, C.c_int]),
    ("c3d_ib_setup", C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                               C.c_int]),
    ("c3d_ib_setup_ref", C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                                   C.c_int, C.c_int, C.c_int]),
    ("c3d_ib_setup_wave", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double]),
    ("c3d_ib_move", C.c_int, [C.c_void_p, C.c_double]),
    ("c3d_ib_get_markers", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
    ("c3d_set_time", C.c_int, [C.c_void_p, C.c_double]),
    ("c3d_get_time", C.c_int, [C.c_void_p, _dp]),
    ("c3d_write_snapshot", C.c_int, [C.c_void_p, C.c_char_p, C.c_double, _dp, _dp]),
    ("c3d_read_snapshot", C.c_int, [C.c_void_p, C.c_char_p, C.c_int, _dp]),
    ("c3d_ib_get_stencil", C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp]),
    ("c3d_ib_get_heaviside", C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("c3d_ib_interpolate", C.c_int, [C.c_void_p, C.c_int, _dp]),
    ("c3d_ib_interpolate_arrays", C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]),
    ("c3d_ib_compute_heaviside", C.c_int, [C.c_void_p]),
    ("c3d_ib_interpolate_c", C.c_int, [C.c_void_p, _dp, _dp]),
    ("c3d_ib_spread", C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    ("c3d_ib_remove_mean", C.c_int, [C.c_void_p, _dp]),
    ("c3d_ib_schur", C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_int]),
    ("c3d_ib_bicgstab", C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_int,
                                  _ip, _dp]),
    ("c3d_ib_set_solver", C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int]),
    ("c3d_ib_set_graphs", C.c_int, [C.c_void_p, C.c_int]),
    ("c3d_ib_set_force", C.c_int, [C.c_void_p, _dp]),
    ("c3d_ib_get_force", C.c_int, [C.c_void_p, _dp]),
    ("c3d_compute_ib_projection", C.c_int, [C.c_void_p, C.c_double, C.c_int, _ip, _dp]),
    ("c3d_check_slip", C.c_int, [C.c_void_p, _dp]),
    ("c3d_ib_stats", C.c_int, [C.c_void_p, _ip, C.POINTER(C.c_longlong), _dp]),
    ("c3d_ib_set_exterior_pressure_gradient", C.c_int, [C.c_void_p, C.c_int]),
    ("c3d_mlp_param_count", C.c_int, []),
    ("c3d_mlp_init", C.c_int, [C.c_void_p, C.c_ulonglong]),
    ("c3d_mlp_set_scales", C.c_int, [C.c_void_p, _fp, _fp]),
    ("c3d_mlp_get_params", C.c_int, [C.c_void_p, _fp]),
    ("c3d_mlp_set_params", C.c_int, [C.c_void_p, _fp]),
    ("c3d_mlp_forward", C.c_int, [C.c_void_p, _fp, C.c_int, _fp]),
    ("c3d_mlp_gradients", C.c_int, [C.c_void_p, _fp, _fp, C.c_int, _fp, _fp]),
    ("c3d_mlp_train", C.c_int, [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float,
                                C.c_ulonglong, _fp]),
    ("c3d_mlp_gemm_test", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp,
                                    _fp, _fp]),
    ("c3d_ib_last_pair", C.c_int, [C.c_void_p, _fp, _dp]),
]

_lib = None


def load() -> C.CDLL:
    """Load libc3d.so; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m channel3d_parallel_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """Pointer to a float64 numpy array (None passes NULL)."""
    if a is None:
        return None
    if a.dtype.kind != "f" or a.dtype.itemsize != 8:
        raise TypeError("float64 array required")
    if not (a.flags["F_CONTIGUOUS"] or a.ndim == 1):
        raise ValueError("arrays must be Fortran-contiguous")
    return a.ctypes.data_as(_dp)


def iptr(a):
    """Pointer to an int32 numpy array."""
    if a.dtype != "int32" or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("contiguous int32 array required")
    return a.ctypes.data_as(_ip)


def fptr(a):
    """Pointer to a contiguous float32 numpy array (None passes NULL)."""
    if a is None:
        return None
    if a.dtype != "float32" or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("contiguous float32 array required")
    return a.ctypes.data_as(_fp)
