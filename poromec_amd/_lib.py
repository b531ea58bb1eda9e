"""ctypes binding of libpmhip.so, the HIP implementation behind the C ABI of
include/poromec_hip.h.  The library is loaded from this package directory
only; if it is missing the import fails (there is no fallback)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libpmhip.so"

PM_OK = 0
PM_ERR = {
    -1: "PM_ERR_HIP", -2: "PM_ERR_ARG", -3: "PM_ERR_ZERO_DIAG", -4: "PM_ERR_MAXIT",
    -5: "PM_ERR_BREAKDOWN", -6: "PM_ERR_JACOBIAN", -7: "PM_ERR_NLIT", -8: "PM_ERR_STATE",
    -9: "PM_ERR_UNSUPPORTED",
}


class SolverCfg(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("precond", C.c_int), ("maxit", C.c_int), ("nkrylov", C.c_int),
        ("tol", C.c_double), ("precision", C.c_int), ("ir_inner_tol", C.c_double),
        ("ir_maxouter", C.c_int), ("maxnlit", C.c_int), ("nltol", C.c_double),
        ("stop_crit", C.c_int), ("maxnlit_accept", C.c_int),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("iters", C.c_int), ("outer", C.c_int), ("used_solver", C.c_int), ("resid", C.c_double),
        ("true_rel", C.c_double), ("ms", C.c_double), ("spmv64", C.c_int), ("spmv32", C.c_int),
    ]


class StepInfo(C.Structure):
    _fields_ = [
        ("nlit", C.c_int), ("nsolve", C.c_int), ("iters", C.c_int * 16), ("used", C.c_int * 16),
        ("true_rel", C.c_double * 16), ("ms_total", C.c_double), ("ms_solve", C.c_double),
        ("ms_element", C.c_double),
    ]


class PmError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        name = PM_ERR.get(code, str(code))
        detail = ""
        if code == -1 and _lib is not None:
            detail = ": " + _lib.pm_last_error().decode(errors="replace")
        super().__init__(f"{where}: {name}{detail}")


_vp = C.c_void_p
_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m poromec_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    L.pm_error_string.restype = C.c_char_p
    L.pm_last_error.restype = C.c_char_p
    L.pm_ctx_create.argtypes = [C.POINTER(_vp), C.c_int]
    L.pm_ctx_destroy.argtypes = [_vp]
    L.pm_ctx_destroy.restype = None
    L.pm_mesh_setup.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int,
                                C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.pm_mesh_setup_nen.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int64)]
    L.pm_mesh_numbering.argtypes = [C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.pm_mesh_numbering_nen.argtypes = [C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int,
                                        _vp, _vp, _vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int64)]
    L.pm_get_numbering.argtypes = [_vp, _vp, _vp, _vp]
    L.pm_set_materials.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double]
    L.pm_set_nodal_loads.argtypes = [_vp, _vp, _vp]
    L.pm_pform.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]
    L.pm_solve.argtypes = [_vp, C.POINTER(SolverCfg), _vp, _vp, C.POINTER(SolveInfo)]
    L.pm_solve_noalloc.argtypes = [_vp, C.POINTER(SolverCfg), _vp, _vp, C.POINTER(SolveInfo)]
    L.pm_struct_upload.argtypes = [_vp, C.c_int, C.c_int64, _vp, _vp]
    L.pm_values_upload.argtypes = [_vp, _vp, _vp]
    L.pm_spmv.argtypes = [_vp, _vp, _vp]
    L.pm_export_csrc.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.pm_set_state.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double]
    L.pm_get_state.argtypes = [_vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double)]
    L.pm_get_vectors.argtypes = [_vp, _vp, _vp, _vp]
    L.pm_set_vectors.argtypes = [_vp, _vp, _vp, _vp]
    L.pm_step.argtypes = [_vp, C.c_double, C.POINTER(SolverCfg), C.POINTER(StepInfo)]
    L.pm_tform.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.pm_set_vprop.argtypes = [_vp, _vp, _vp]
    L.pm_get_vprop.argtypes = [_vp, _vp]
    L.pm_batch_define.argtypes = [_vp, C.c_int, _vp]
    L.pm_batch_iters.argtypes = [_vp, _vp]
    L.pm_batch_graphs.argtypes = [_vp, C.c_int]
    L.pm_set_ordered.argtypes = [_vp, C.c_int]
    L.pm_spmv_timing.argtypes = [_vp, C.c_int, _vp]
    L.pm_tune_spmv.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_int), _vp]
    L.pm_set_spmv_shape.argtypes = [_vp, C.c_int, C.c_int]
    L.pm_adjoint.argtypes = [_vp, C.c_int, C.c_double, C.POINTER(SolverCfg), C.c_int, _vp, _vp,
                             C.c_int, C.POINTER(C.c_double), _vp, _vp]
    L.pm_matrix_info.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64)]
    L.pm_matrix_runs.argtypes = [_vp, C.POINTER(C.c_int64)]
    L.pm_bench_spmv.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.pm_bench_copy.argtypes = [_vp, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    L.pm_bench_stream.argtypes = [_vp, C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    _lib = L
    return L


def check(rc, where=""):
    if rc != PM_OK:
        raise PmError(rc, where)


def ptr(a):
    """data pointer of a contiguous numpy array (or None)"""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to libpmhip must be contiguous"
    return a.ctypes.data_as(_vp)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
