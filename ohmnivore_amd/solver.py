"""Host-side mirror of the reference's solver interfaces over the HIP backend.

- ``compile``        compiler::compile            (src/compiler.rs:194)
- ``HipBackend``     SolverBackend                (src/solver/backend.rs:51-96)
- ``GpuSolver``      LinearSolver                 (src/solver/mod.rs:28-34)
- ``newton_solve``   newton_solve / NewtonParams  (src/solver/newton.rs:34-78)

Every call runs on the MI355X through libohm_hip.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import circuit as ir
from ._native import OhmnivoreError, check, lib, ptr


# ---------------------------------------------------------------- compile --
def compile(ckt: ir.Circuit) -> ir.MnaSystem:
    """compiler::compile (src/compiler.rs:194-754), native C++."""
    cc = ir.CCircuit(ckt)
    m = ir.MnaC()
    check(lib().ohm_compile(cc.ref(), ctypes.byref(m)))
    try:
        return ir.mna_from_c(m, ckt)
    finally:
        lib().ohm_mna_free(ctypes.byref(m))


# ---------------------------------------------------------------- backend --
class Buffer:
    """Device f64 vector (WgpuBuffer, src/solver/backend.rs:30-34)."""

    def __init__(self, backend: "HipBackend", n: int):
        self.backend = backend
        self.n = n
        self.h = ctypes.c_void_p()
        check(lib().ohm_buffer_new(backend.h, ctypes.c_size_t(n), ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and not sys.is_finalizing():
            lib().ohm_buffer_free(self.h)
            self.h = None


class Matrix:
    """Device CSR (GpuCsrMatrix, src/solver/backend.rs:20-27)."""

    def __init__(self, backend: "HipBackend", values, col_indices, row_pointers, n: int):
        self.backend = backend
        self.n = n
        v = np.ascontiguousarray(values, dtype=np.float64)
        c = np.ascontiguousarray(col_indices, dtype=np.uint32)
        r = np.ascontiguousarray(row_pointers, dtype=np.uint32)
        self.nnz = int(v.shape[0])
        self.h = ctypes.c_void_p()
        check(lib().ohm_upload_matrix(backend.h, ptr(v, N.c_f64p), ptr(c, N.c_u32p),
                                      ptr(r, N.c_u32p), ctypes.c_size_t(n),
                                      ctypes.c_size_t(self.nnz), ctypes.byref(self.h)))

    def set_values(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().ohm_matrix_set_values(self.backend.h, self.h, ptr(v, N.c_f64p)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and not sys.is_finalizing():
            lib().ohm_matrix_free(self.h)
            self.h = None


class Preconditioner:
    def __init__(self, backend: "HipBackend", a: Matrix, kind: int = N.PREC_AUTO, level: int = 1):
        self.backend = backend
        self.h = ctypes.c_void_p()
        check(lib().ohm_precond_create(backend.h, a.h, int(kind), int(level), ctypes.byref(self.h)))

    @property
    def kind(self) -> int:
        return lib().ohm_precond_kind(self.h)

    def set_ilu_sweeps(self, sweeps: int):
        check(lib().ohm_precond_set_ilu_sweeps(self.h, int(sweeps)))

    def update(self, a: Matrix):
        check(lib().ohm_precond_update(self.backend.h, self.h, a.h))

    def apply(self, inp: Buffer, out: Buffer):
        check(lib().ohm_precond_apply(self.backend.h, self.h, inp.h, out.h))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and not sys.is_finalizing():
            lib().ohm_precond_free(self.h)
            self.h = None


def bicgstab_params(max_iterations=10_000, tolerance=1e-5, shadow="hashed", warm_start=False,
                    precision="fp64", check_interval=0, inner_tolerance=1e-4,
                    max_refinements=30, breakdown="scaled", tol_mode="rhs",
                    tol_floor=1e-15) -> N.BicgstabParams:
    """BiCgStabParams (src/solver/bicgstab.rs:15-27) plus the backend's knobs."""
    p = N.BicgstabParams()
    lib().ohm_bicgstab_default_params(ctypes.byref(p))
    p.max_iterations = max_iterations
    p.tolerance = tolerance
    p.shadow = N.SHADOW_HASHED if shadow == "hashed" else N.SHADOW_RHS
    p.warm_start = int(bool(warm_start))
    p.precision = N.FP32_IR if precision in ("fp32_ir", "mixed") else N.FP64
    p.check_interval = check_interval
    p.inner_tolerance = inner_tolerance
    p.max_refinements = max_refinements
    p.breakdown = 1 if breakdown == "absolute" else 0
    p.tol_mode = 1 if tol_mode == "initial" else 0
    p.tol_floor = tol_floor
    return p


@dataclass
class SolveStats:
    iterations: int = 0
    refinements: int = 0
    residual_norm: float = float("nan")
    device_ms: float = 0.0
    dispatches: int = 0
    readbacks: int = 0
    fp64_fallback: bool = False

    @classmethod
    def from_c(cls, r: N.BicgstabResult) -> "SolveStats":
        return cls(r.iterations, r.refinements, r.residual_norm, r.device_ms, r.dispatches,
                   r.readbacks, bool(r.fp64_fallback))


class HipBackend:
    """SolverBackend (src/solver/backend.rs:51-96) on one MI355X."""

    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        check(lib().ohm_backend_create(int(device), ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and not sys.is_finalizing():
            lib().ohm_backend_destroy(self.h)
            self.h = None

    def new_buffer(self, n: int) -> Buffer:
        return Buffer(self, n)

    def upload_vec(self, data, buf: Buffer):
        d = np.ascontiguousarray(data, dtype=np.float64)
        check(lib().ohm_upload_vec(self.h, ptr(d, N.c_f64p), ctypes.c_size_t(d.shape[0]), buf.h))

    def download_vec(self, buf: Buffer, n: int | None = None) -> np.ndarray:
        n = buf.n if n is None else n
        out = np.empty(n, dtype=np.float64)
        check(lib().ohm_download_vec(self.h, buf.h, ptr(out, N.c_f64p), ctypes.c_size_t(n)))
        return out

    def upload_matrix(self, values, col_indices, row_pointers, n: int) -> Matrix:
        return Matrix(self, values, col_indices, row_pointers, n)

    def spmv(self, a: Matrix, x: Buffer, y: Buffer):
        check(lib().ohm_spmv(self.h, a.h, x.h, y.h))

    def spmv_transpose(self, a: Matrix, x: Buffer, y: Buffer):
        check(lib().ohm_spmv_transpose(self.h, a.h, x.h, y.h))

    def dot(self, x: Buffer, y: Buffer) -> float:
        out = ctypes.c_double()
        check(lib().ohm_dot(self.h, x.h, y.h, ctypes.byref(out)))
        return out.value

    def dot_n(self, x: Buffer, y: Buffer, n: int) -> float:
        out = ctypes.c_double()
        check(lib().ohm_dot_n(self.h, x.h, y.h, ctypes.c_size_t(n), ctypes.byref(out)))
        return out.value

    def axpy(self, alpha: float, x: Buffer, y: Buffer):
        check(lib().ohm_axpy(self.h, ctypes.c_double(alpha), x.h, y.h))

    def copy(self, src: Buffer, dst: Buffer):
        check(lib().ohm_copy(self.h, src.h, dst.h))

    def scale(self, alpha: float, x: Buffer):
        check(lib().ohm_scale(self.h, ctypes.c_double(alpha), x.h))

    def preconditioner(self, a: Matrix, kind=N.PREC_AUTO, level=1) -> Preconditioner:
        return Preconditioner(self, a, kind, level)

    def bicgstab(self, a: Matrix, b: Buffer, x: Buffer, precond: Preconditioner | None = None,
                 params: N.BicgstabParams | None = None) -> SolveStats:
        """bicgstab_with_params (src/solver/bicgstab.rs:56-189) on device buffers."""
        res = N.BicgstabResult()
        check(lib().ohm_bicgstab(self.h, a.h, b.h, x.h, precond.h if precond else None,
                                 ctypes.byref(params) if params is not None else None,
                                 ctypes.byref(res)))
        return SolveStats.from_c(res)

    def sync(self):
        check(lib().ohm_backend_sync(self.h))

    def set_graphs(self, enable: bool):
        """Replay iteration chunks from captured hipGraphs (default) or not."""
        check(lib().ohm_backend_set_graphs(self.h, int(bool(enable))))

    def counters(self) -> tuple[int, int]:
        d, r = ctypes.c_int64(), ctypes.c_int64()
        check(lib().ohm_backend_counters(self.h, ctypes.byref(d), ctypes.byref(r)))
        return d.value, r.value


class GpuSolver:
    """LinearSolver::solve_real (src/solver/mod.rs:28-34; GpuSolver src/solver/gpu.rs:107)."""

    def __init__(self, backend: HipBackend | None = None, precond: int = N.PREC_AUTO,
                 params: N.BicgstabParams | None = None):
        self.backend = backend or HipBackend()
        self.precond = precond
        self.params = params
        self.last_stats: SolveStats | None = None

    def solve_real(self, a: ir.CsrMatrix, b) -> np.ndarray:
        n = a.nrows
        bb = np.ascontiguousarray(b, dtype=np.float64)
        if a.ncols != n or bb.shape[0] != n:
            raise OhmnivoreError(3, f"Solve error: dimension mismatch: matrix is {a.nrows}x"
                                    f"{a.ncols}, rhs length is {bb.shape[0]}")
        x = np.zeros(n, dtype=np.float64)
        if n == 0:
            return x
        v = np.ascontiguousarray(a.values, dtype=np.float64)
        c = np.ascontiguousarray(a.col_indices, dtype=np.uint32)
        r = np.ascontiguousarray(a.row_pointers, dtype=np.uint32)
        res = N.BicgstabResult()
        check(lib().ohm_solve_real(self.backend.h, ctypes.c_size_t(n), ctypes.c_size_t(v.shape[0]),
                                   ptr(v, N.c_f64p), ptr(c, N.c_u32p), ptr(r, N.c_u32p),
                                   ptr(bb, N.c_f64p), ptr(x, N.c_f64p), int(self.precond),
                                   ctypes.byref(self.params) if self.params is not None else None,
                                   ctypes.byref(res)))
        self.last_stats = SolveStats.from_c(res)
        return x


# ----------------------------------------------------------------- Newton --
@dataclass
class NewtonStats:
    """The Newton counters of Stats (src/stats.rs:9-29)."""
    newton_iterations: int = 0
    bicgstab_iterations: int = 0
    bicgstab_iters_per_newton: list = field(default_factory=list)
    max_diff: float = 0.0
    device_ms: float = 0.0
    fp64_fallbacks: int = 0


def newton_params(max_iterations=50, abs_tol=1e-4, bicgstab_max_iterations=10_000,
                  bicgstab_tolerance=1e-5, precision="fp64", precond=N.PREC_AUTO,
                  initial_guess=None, bicgstab_floor=1e-15):
    """NewtonParams (src/solver/newton.rs:34-55). bicgstab_tolerance is
    relative to the Newton residual ||F(x_k)|| (inexact Newton)."""
    p = N.NewtonParams()
    lib().ohm_newton_default_params(ctypes.byref(p))
    p.max_iterations = max_iterations
    p.abs_tol = abs_tol
    p.bicgstab_max_iterations = bicgstab_max_iterations
    p.bicgstab_tolerance = bicgstab_tolerance
    p.bicgstab_floor = bicgstab_floor
    p.precision = N.FP32_IR if precision in ("fp32_ir", "mixed") else N.FP64
    p.precond = precond
    keep = None
    if initial_guess is not None:
        keep = np.ascontiguousarray(initial_guess, dtype=np.float64)
        p.initial_guess = ptr(keep, N.c_f64p)
    return p, keep


def _descs(a, dtype):
    if a is None or len(a) == 0:
        return np.zeros(1, dtype=dtype), 0
    return np.ascontiguousarray(a, dtype=dtype), len(a)


class NewtonContext:
    """Device-resident Newton problem: upload once, solve many times."""

    def __init__(self, backend: HipBackend, g: ir.CsrMatrix, base_b, diodes=None, bjts=None,
                 mosfets=None):
        self.backend = backend
        self.n = g.nrows
        self.nnz = g.nnz
        self._v = np.ascontiguousarray(g.values, dtype=np.float64)
        self._c = np.ascontiguousarray(g.col_indices, dtype=np.uint32)
        self._r = np.ascontiguousarray(g.row_pointers, dtype=np.uint32)
        self._b = np.ascontiguousarray(base_b, dtype=np.float64)
        d, nd = _descs(diodes, ir.DIODE_DESC_DTYPE)
        q, nq = _descs(bjts, ir.BJT_DESC_DTYPE)
        m, nm = _descs(mosfets, ir.MOSFET_DESC_DTYPE)
        self.h = ctypes.c_void_p()
        check(lib().ohm_newton_create(
            backend.h, ptr(self._v, N.c_f64p), ptr(self._c, N.c_u32p), ptr(self._r, N.c_u32p),
            ptr(self._b, N.c_f64p), vp(d), ctypes.c_size_t(nd), vp(q), ctypes.c_size_t(nq),
            vp(m), ctypes.c_size_t(nm), ctypes.c_size_t(self.n), ctypes.c_size_t(self.nnz),
            ctypes.byref(self.h)))

    def run(self, **kw) -> NewtonStats:
        p, keep = newton_params(**kw)
        res = N.NewtonResult()
        per = np.zeros(max(p.max_iterations, 1), dtype=np.int32)
        check(lib().ohm_newton_run(self.h, ctypes.byref(p), ctypes.byref(res), ptr(per, N.c_i32p)))
        del keep
        return NewtonStats(res.iterations, res.bicgstab_iterations,
                           per[: res.iterations].tolist(), res.max_diff, res.device_ms,
                           res.fp64_fallbacks)

    def x(self) -> np.ndarray:
        out = np.empty(self.n)
        check(lib().ohm_newton_get_x(self.h, ptr(out, N.c_f64p)))
        return out

    def system(self):
        v = np.empty(self.nnz)
        b = np.empty(self.n)
        check(lib().ohm_newton_get_system(self.h, ptr(v, N.c_f64p), ptr(b, N.c_f64p)))
        return v, b

    def assemble_at(self, x):
        xx = np.ascontiguousarray(x, dtype=np.float64)
        v = np.empty(self.nnz)
        b = np.empty(self.n)
        check(lib().ohm_newton_assemble_at(self.h, ptr(xx, N.c_f64p), ptr(v, N.c_f64p),
                                           ptr(b, N.c_f64p)))
        return v, b

    def limit(self, x_old, x_new):
        xo = np.ascontiguousarray(x_old, dtype=np.float64)
        xn = np.array(x_new, dtype=np.float64)
        check(lib().ohm_newton_limit(self.h, ptr(xo, N.c_f64p), ptr(xn, N.c_f64p)))
        return xn

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and not sys.is_finalizing():
            lib().ohm_newton_free(self.h)
            self.h = None


def vp(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def newton_solve(backend: HipBackend, g: ir.CsrMatrix, base_b, diodes=None, bjts=None,
                 mosfets=None, **params):
    """newton_solve (src/solver/newton.rs:65-78). Returns (x, NewtonStats)."""
    ctx = NewtonContext(backend, g, base_b, diodes, bjts, mosfets)
    stats = ctx.run(**params)
    return ctx.x(), stats
