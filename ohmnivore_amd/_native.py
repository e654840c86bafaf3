"""ctypes binding of include/ohm_solver.h (libohm_hip.so, built in-tree).

The library is required: every entry point of this package goes through it,
and importing a solver without it raises instead of falling back to a CPU
path.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libohm_hip.so")
HEADER_DIR = os.path.join(os.path.dirname(HERE), "include")

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_i32p = ctypes.POINTER(ctypes.c_int32)
vp = ctypes.c_void_p

STATUS_NAMES = {
    0: "Ok", 1: "Parse", 2: "Compile", 3: "Solve", 4: "NewtonNotConverged",
    5: "NewtonNumericalError", 6: "Analysis", 7: "Convergence", 8: "Timestep", 9: "Io",
    10: "NoDevice", 11: "Invalid",
}

PREC_NONE, PREC_JACOBI, PREC_ISAI, PREC_AUTO = 0, 1, 2, 3
SHADOW_RHS, SHADOW_HASHED = 0, 1
FP64, FP32_IR = 0, 1


class OhmnivoreError(RuntimeError):
    """OhmnivoreError (src/error.rs:4-36): .kind names the variant, str() is
    the reference's Display text."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code
        self.kind = STATUS_NAMES.get(code, str(code))


class BicgstabParams(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("tolerance", ctypes.c_double),
                ("shadow", ctypes.c_int32), ("warm_start", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("check_interval", ctypes.c_int32),
                ("inner_tolerance", ctypes.c_double), ("max_refinements", ctypes.c_int64),
                ("breakdown", ctypes.c_int32), ("tol_mode", ctypes.c_int32),
                ("tol_floor", ctypes.c_double)]


class BicgstabResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("refinements", ctypes.c_int64),
                ("residual_norm", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("dispatches", ctypes.c_int64), ("readbacks", ctypes.c_int64),
                ("fp64_fallback", ctypes.c_int64)]


class NewtonParams(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("abs_tol", ctypes.c_double),
                ("bicgstab_max_iterations", ctypes.c_int64), ("bicgstab_tolerance", ctypes.c_double),
                ("precision", ctypes.c_int32), ("precond", ctypes.c_int32),
                ("initial_guess", c_f64p), ("bicgstab_floor", ctypes.c_double),
                ("forcing", ctypes.c_int32), ("forcing_safeguard", ctypes.c_int32),
                ("forcing_max", ctypes.c_double), ("linear_solver", ctypes.c_int32)]


class NewtonResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("bicgstab_iterations", ctypes.c_int64),
                ("max_diff", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("fp64_fallbacks", ctypes.c_int64)]


class DcStep(ctypes.Structure):
    _fields_ = [("stage", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("level", ctypes.c_double), ("iterations", ctypes.c_int64),
                ("retry", ctypes.c_int32), ("_pad", ctypes.c_int32)]


class DcStats(ctypes.Structure):
    _fields_ = [("stage", ctypes.c_int32), ("_pad", ctypes.c_int32), ("runs", ctypes.c_int64),
                ("newton_iterations", ctypes.c_int64), ("bicgstab_iterations", ctypes.c_int64),
                ("device_ms", ctypes.c_double)]


DC_STAGES = {0: "linear", 1: "direct", 2: "source", 3: "gmin", 4: "final", 5: "ptc",
             6: "ptc-final"}


class BatchSystem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("rowptr", c_u32p), ("col", c_u32p),
                ("g_val", c_f64p), ("c_val", c_f64p), ("sigma_re", ctypes.c_double),
                ("sigma_im", ctypes.c_double)]


class BatchResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("refinements", ctypes.c_int64),
                ("residual_norm", ctypes.c_double), ("status", ctypes.c_int32),
                ("reason", ctypes.c_int32)]


SWEEP_DEC, SWEEP_OCT, SWEEP_LIN = 0, 1, 2
AC_AUTO, AC_ITERATIVE, AC_DIRECT = 0, 1, 2


class AcCircuit(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("g_rowptr", c_u32p), ("g_col", c_u32p),
                ("g_val", c_f64p), ("c_rowptr", c_u32p), ("c_col", c_u32p), ("c_val", c_f64p),
                ("b_ac", c_f64p), ("x", c_f64p)]

_lib = None


def lib() -> ctypes.CDLL:
    """Load the HIP backend; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OhmnivoreError(10, f"libohm_hip.so is not built ({LIB_PATH}); "
                                     "run `python -m ohmnivore_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.ohm_last_error.restype = ctypes.c_char_p
        L.ohm_version.restype = ctypes.c_char_p
        L.ohm_buffer_len.restype = ctypes.c_size_t
        L.ohm_precond_kind.restype = ctypes.c_int
        L.ohm_batch_rows.restype = ctypes.c_int64
        L.ohm_generate_frequencies.restype = ctypes.c_int64
        L.ohm_generate_frequencies.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_double,
                                               ctypes.c_double, c_f64p, ctypes.c_int64]
        for name in ("ohm_backend_destroy", "ohm_buffer_free", "ohm_matrix_free",
                     "ohm_precond_free", "ohm_newton_free", "ohm_mna_free",
                     "ohm_bicgstab_default_params", "ohm_newton_default_params",
                     "ohm_batch_free", "ohm_transient_free", "ohm_tran_default_params",
                     "ohm_ac_plan_free"):
            getattr(L, name).restype = None
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise OhmnivoreError(rc, lib().ohm_last_error().decode())


def declared_symbols() -> list[str]:
    """Every function declared in include/*.h."""
    names = []
    for f in sorted(os.listdir(HEADER_DIR)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(ohm_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)
