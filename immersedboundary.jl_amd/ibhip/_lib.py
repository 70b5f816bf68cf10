"""ctypes binding of libibhip.so (the C ABI declared in include/ibh.h).

The library is loaded from the package directory.  There is no fallback: if the shared
object is missing or no GPU is visible, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libibhip.so")

OK, EINVAL, EBOUNDS, ENONFINITE, ENOCONV, EDEVICE, ESINGULAR = range(7)
NOALLOC_MAX_NG = 1048576  # IBH_NOALLOC_MAX_NG of include/ibh.h
STATUS_NAMES = {
    EINVAL: "invalid argument",
    EBOUNDS: "body outside the finest grid",
    ENONFINITE: "infinite flux",
    ENOCONV: "CG did not converge",
    EDEVICE: "HIP device error",
    ESINGULAR: "singular B matrix",
}

(F_Q, F_Q0, F_GAMMA, F_PSI, F_FB, F_NONLIN1, F_NONLIN2, F_CD, F_CL, F_FBSTRESS, F_POINTS_R,
 F_POINTS_V) = range(12)

EXPORTED = [
    "ibh_version", "ibh_create", "ibh_destroy", "ibh_last_error", "ibh_init_state", "ibh_setup",
    "ibh_set_body_points", "ibh_advance", "ibh_advance_n", "ibh_sync", "ibh_get_field",
    "ibh_set_field", "ibh_dims", "ibh_time", "ibh_cg_iterations", "ibh_solver_stats", "ibh_lowrank_stats",
    "ibh_op_lapinv", "ibh_op_dst2",
    "ibh_op_rot", "ibh_op_curl", "ibh_op_coarsify", "ibh_op_get_bc", "ibh_op_apply_bc",
    "ibh_op_nonlinear", "ibh_op_vort2flux", "ibh_op_reg_update", "ibh_op_E", "ibh_op_ET",
    "ibh_op_B", "ibh_op_lapinv_noalloc", "ibh_profile_enable", "ibh_profile_read", "ibh_profile_read_events",
    "ibh_profile_launches", "ibh_tune_report",
    "ibh_batch_create", "ibh_batch_destroy", "ibh_batch_last_error", "ibh_batch_advance",
    "ibh_batch_advance_n", "ibh_batch_sync", "ibh_batch_launches_per_step",
    "ibh_batch_profile_enable", "ibh_batch_profile_read", "ibh_batch_profile_read_events",
    "ibh_batch_profile_launches", "ibh_batch_tune_report",
]


class IbhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libibhip error {code} ({STATUS_NAMES.get(code, '?')}): {msg}")
        self.code = code


class CGNotConverged(RuntimeWarning):
    """The force solve's CG stopped at cg_maxiter.  The step completed with the last iterate,
    as IterativeSolvers' cg! does when the reference calls it (cnab/ops.jl:110)."""


class CaseDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nlev", C.c_int32),
        ("h", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
        ("Re", C.c_double), ("dt", C.c_double), ("beta", C.c_double * 2),
        ("nbodies", C.c_int32),
        ("npts", C.POINTER(C.c_int32)), ("moving", C.POINTER(C.c_int32)),
        ("xb", C.POINTER(C.c_double)), ("ds", C.POINTER(C.c_double)),
        ("precision", C.c_int32), ("force_solver", C.c_int32),
        ("cg_maxiter", C.c_int32), ("cg_reltol", C.c_double),
        ("pc_refresh", C.c_int32), ("refine_sweeps", C.c_int32),
        ("refine_levels", C.c_int32),
    ]


_lib = None


def load():
    """Load libibhip.so and declare its signatures (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C immersedboundary.jl_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback"
        )
    L = C.CDLL(LIB_PATH)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_void_p
    L.ibh_version.restype = C.c_int
    L.ibh_create.argtypes = [C.POINTER(CaseDesc), C.POINTER(vp)]
    L.ibh_destroy.argtypes = [vp]
    L.ibh_destroy.restype = None
    L.ibh_last_error.argtypes = [vp]
    L.ibh_last_error.restype = C.c_char_p
    L.ibh_init_state.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.ibh_setup.argtypes = [vp]
    L.ibh_set_body_points.argtypes = [vp, dp, dp]
    L.ibh_advance.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.ibh_advance_n.argtypes = [vp, C.c_double, C.c_int32, C.c_double, C.c_double]
    L.ibh_sync.argtypes = [vp]
    L.ibh_get_field.argtypes = [vp, C.c_int32, dp]
    L.ibh_set_field.argtypes = [vp, C.c_int32, dp]
    L.ibh_dims.argtypes = [vp, ip]
    L.ibh_time.argtypes = [vp]
    L.ibh_time.restype = C.c_double
    L.ibh_cg_iterations.argtypes = [vp]
    L.ibh_solver_stats.argtypes = [vp, dp]
    L.ibh_lowrank_stats.argtypes = [vp, dp]
    L.ibh_op_lapinv.argtypes = [vp, C.c_int32, C.c_int32, dp, dp]
    L.ibh_op_dst2.argtypes = [vp, dp, dp]
    L.ibh_op_rot.argtypes = [vp, dp, dp]
    L.ibh_op_curl.argtypes = [vp, dp, dp, dp]
    L.ibh_op_coarsify.argtypes = [vp, dp, dp]
    L.ibh_op_get_bc.argtypes = [vp, dp, dp]
    L.ibh_op_apply_bc.argtypes = [vp, dp, dp, C.c_double]
    L.ibh_op_nonlinear.argtypes = [vp, C.c_int32, dp, dp, dp, dp, dp]
    L.ibh_op_vort2flux.argtypes = [vp, dp, dp, dp]
    L.ibh_op_reg_update.argtypes = [vp, dp]
    L.ibh_op_E.argtypes = [vp, dp, dp]
    L.ibh_op_ET.argtypes = [vp, dp, dp]
    L.ibh_op_B.argtypes = [vp, dp, dp]
    L.ibh_op_lapinv_noalloc.argtypes = [vp, C.c_int32, C.c_int32, vp, vp]  # device pointers
    L.ibh_profile_enable.argtypes = [vp, C.c_int32]
    L.ibh_profile_read.argtypes = [vp, C.c_int32, dp, C.POINTER(C.c_int64), dp]
    L.ibh_profile_read_events.argtypes = [vp, C.c_int32, dp, C.POINTER(C.c_int64), dp]
    L.ibh_profile_launches.argtypes = [vp, C.c_int32, C.c_int32, dp, dp, dp]
    L.ibh_tune_report.argtypes = [vp, C.c_char_p, C.c_int32]
    L.ibh_batch_create.argtypes = [C.c_int32, C.POINTER(vp), C.POINTER(vp)]
    L.ibh_batch_destroy.argtypes = [vp]
    L.ibh_batch_destroy.restype = None
    L.ibh_batch_last_error.argtypes = [vp]
    L.ibh_batch_last_error.restype = C.c_char_p
    L.ibh_batch_advance.argtypes = [vp, C.c_double, C.c_double]
    L.ibh_batch_advance_n.argtypes = [vp, C.c_int32, C.c_double, C.c_double]
    L.ibh_batch_sync.argtypes = [vp]
    L.ibh_batch_launches_per_step.argtypes = [vp]
    L.ibh_batch_profile_enable.argtypes = [vp, C.c_int32]
    L.ibh_batch_profile_read.argtypes = [vp, C.c_int32, dp, C.POINTER(C.c_int64), dp]
    L.ibh_batch_profile_read_events.argtypes = [vp, C.c_int32, dp, C.POINTER(C.c_int64), dp]
    L.ibh_batch_profile_launches.argtypes = [vp, C.c_int32, C.c_int32, dp, dp, dp]
    L.ibh_batch_tune_report.argtypes = [vp, C.c_char_p, C.c_int32]
    _lib = L
    return L


def dptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a contiguous float64 array")
    return a.ctypes.data_as(C.POINTER(C.c_double))


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))
