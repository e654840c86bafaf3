"""Host-side mirror of the reference's solver interfaces over the HIP backend.

- ``compile``        compiler::compile            (src/compiler.rs:194)
- ``HipBackend``     SolverBackend                (src/solver/backend.rs:51-96)
- ``GpuSolver``      LinearSolver                 (src/solver/mod.rs:28-34)
- ``newton_solve``   newton_solve / NewtonParams  (src/solver/newton.rs:34-78)
- ``Batch``          batched (G_k + sigma_k C_k) x_k = b_k solves (no reference
                     counterpart: AC points, Monte Carlo, uneven circuits)
- ``ac_sweep``       ac::run / generate_frequencies (src/analysis/ac.rs:25-109)

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

    def autotune(self) -> tuple[int, int]:
        """ohm_matrix_autotune: fastest (rows per workgroup, staged products)."""
        r, z = ctypes.c_int32(), ctypes.c_int32()
        check(lib().ohm_matrix_autotune(self.backend.h, self.h, ctypes.byref(r), ctypes.byref(z)))
        return r.value, z.value

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

    def bicgstab_noalloc(self, a: Matrix, b: Buffer, x: Buffer,
                         precond: Preconditioner | None = None,
                         params: N.BicgstabParams | None = None) -> SolveStats:
        """bicgstab() without device allocation (ohm_bicgstab_noalloc): the
        solve's own vectors live in the library's static scratch array;
        FP64 only, systems up to about 700,000 unknowns."""
        res = N.BicgstabResult()
        check(lib().ohm_bicgstab_noalloc(self.h, a.h, b.h, x.h, precond.h if precond else None,
                                         ctypes.byref(params) if params is not None else None,
                                         ctypes.byref(res)))
        return SolveStats.from_c(res)

    def bicgstab_transpose(self, a: Matrix, b: Buffer, x: Buffer,
                           precond: Preconditioner | None = None,
                           params: N.BicgstabParams | None = None) -> SolveStats:
        """A^T x = b (adjoint systems) with the transposed preconditioner."""
        res = N.BicgstabResult()
        check(lib().ohm_bicgstab_transpose(self.h, a.h, b.h, x.h, precond.h if precond else None,
                                           ctypes.byref(params) if params is not None else None,
                                           ctypes.byref(res)))
        return SolveStats.from_c(res)

    def sync(self):
        check(lib().ohm_backend_sync(self.h))

    def set_autotune(self, enable: bool):
        check(lib().ohm_backend_set_autotune(self.h, int(bool(enable))))

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
        rc = lib().ohm_solve_real(self.backend.h, ctypes.c_size_t(n), ctypes.c_size_t(v.shape[0]),
                                  ptr(v, N.c_f64p), ptr(c, N.c_u32p), ptr(r, N.c_u32p),
                                  ptr(bb, N.c_f64p), ptr(x, N.c_f64p), int(self.precond),
                                  ctypes.byref(self.params) if self.params is not None else None,
                                  ctypes.byref(res))
        self.last_stats = SolveStats.from_c(res)
        check(rc)
        return x

    def solve_complex(self, a: ir.CsrMatrix, b) -> np.ndarray:
        """LinearSolver::solve_complex (src/solver/mod.rs:33; gpu.rs:210-735)."""
        n = a.nrows
        bb = np.ascontiguousarray(b, dtype=np.complex128)
        if a.ncols != n or bb.shape[0] != n:
            raise OhmnivoreError(3, f"Solve error: dimension mismatch: matrix is {a.nrows}x"
                                    f"{a.ncols}, rhs length is {bb.shape[0]}")
        x = np.zeros(n, dtype=np.complex128)
        if n == 0:
            return x
        v = np.ascontiguousarray(a.values, dtype=np.complex128)
        c = np.ascontiguousarray(a.col_indices, dtype=np.uint32)
        r = np.ascontiguousarray(a.row_pointers, dtype=np.uint32)
        res = N.BicgstabResult()
        prec = N.PREC_JACOBI if self.precond == N.PREC_AUTO else self.precond
        rc = lib().ohm_solve_complex(
            self.backend.h, ctypes.c_size_t(n), ctypes.c_size_t(v.shape[0]),
            v.ctypes.data_as(N.c_f64p), ptr(c, N.c_u32p), ptr(r, N.c_u32p),
            bb.ctypes.data_as(N.c_f64p), x.ctypes.data_as(N.c_f64p), int(prec),
            ctypes.byref(self.params) if self.params is not None else None, ctypes.byref(res))
        self.last_stats = SolveStats.from_c(res)
        check(rc)
        return x


# ---------------------------------------------------------------- batches --
@dataclass
class BatchStats:
    iterations: int
    refinements: int
    residual_norm: float
    status: int
    reason: int

    @property
    def converged(self) -> bool:
        return self.status == 0


class Batch:
    """K independent systems (G_k + sigma_k C_k) x_k = b_k solved together.

    ``systems`` is a list of ``(rowptr, col, g, c_or_None, sigma)``; arrays
    that are the same numpy object are uploaded once. Vectors are the
    systems' vectors concatenated (complex dtype for complex batches)."""

    def __init__(self, backend: HipBackend, systems, complex_: bool):
        self.backend = backend
        self.complex = bool(complex_)
        self._keep = []
        cache = {}

        def arr(a, dt):
            if a is None:
                return None
            key = id(a)
            if key not in cache:
                cache[key] = np.ascontiguousarray(a, dtype=dt)
                self._keep.append((a, cache[key]))
            return cache[key]

        k = len(systems)
        cs = (N.BatchSystem * k)()
        self.sizes = []
        for i, (rp, cl, g, c, sig) in enumerate(systems):
            rp_, cl_, g_, c_ = (arr(rp, np.uint32), arr(cl, np.uint32), arr(g, np.float64),
                                arr(c, np.float64))
            n = int(rp_.shape[0] - 1)
            sig = complex(sig)
            cs[i] = N.BatchSystem(n, rp_.ctypes.data_as(N.c_u32p), cl_.ctypes.data_as(N.c_u32p),
                                  g_.ctypes.data_as(N.c_f64p),
                                  c_.ctypes.data_as(N.c_f64p) if c_ is not None else None,
                                  sig.real, sig.imag)
            self.sizes.append(n)
        self.k = k
        self.h = ctypes.c_void_p()
        check(lib().ohm_batch_create(backend.h, int(self.complex), ctypes.c_int64(k), cs,
                                     ctypes.byref(self.h)))
        self.rows = int(lib().ohm_batch_rows(self.h))
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.last_results: list[BatchStats] = []

    @property
    def dtype(self):
        return np.complex128 if self.complex else np.float64

    def set_sigma(self, sigma):
        s = np.asarray(sigma, dtype=np.complex128)
        buf = np.ascontiguousarray(np.stack([s.real, s.imag], axis=1).ravel())
        check(lib().ohm_batch_set_sigma(self.h, buf.ctypes.data_as(N.c_f64p)))

    def spmv(self, x) -> np.ndarray:
        xx = np.ascontiguousarray(x, dtype=self.dtype)
        y = np.zeros(self.rows, dtype=self.dtype)
        check(lib().ohm_batch_spmv(self.h, xx.ctypes.data_as(N.c_f64p), y.ctypes.data_as(N.c_f64p)))
        return y

    def solve(self, rhs, x0=None, params: N.BicgstabParams | None = None,
              precond: int = N.PREC_JACOBI, raise_on_failure: bool = True) -> np.ndarray:
        b = np.ascontiguousarray(rhs, dtype=self.dtype)
        if b.shape[0] != self.rows:
            raise OhmnivoreError(3, f"Solve error: dimension mismatch: batch has {self.rows} "
                                    f"rows, rhs length is {b.shape[0]}")
        prm = params if params is not None else bicgstab_params()
        x = (np.ascontiguousarray(x0, dtype=self.dtype).copy() if x0 is not None
             else np.zeros(self.rows, dtype=self.dtype))
        if x0 is not None:
            prm = _with(prm, warm_start=1)
        res = (N.BatchResult * self.k)()
        rc = lib().ohm_batch_solve(self.h, b.ctypes.data_as(N.c_f64p), x.ctypes.data_as(N.c_f64p),
                                   ctypes.byref(prm), int(precond), res)
        self.last_results = [BatchStats(r.iterations, r.refinements, r.residual_norm, r.status,
                                        r.reason) for r in res]
        if rc != 0 and raise_on_failure:
            check(rc)
        return x

    def split(self, v) -> list:
        return [v[self.offsets[i]:self.offsets[i + 1]] for i in range(self.k)]

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and not sys.is_finalizing():
            lib().ohm_batch_free(self.h)
            self.h = None


def _with(p: N.BicgstabParams, **kw) -> N.BicgstabParams:
    q = N.BicgstabParams()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(p))
    for k, v in kw.items():
        setattr(q, k, v)
    return q


# --------------------------------------------------------------------- AC --
SWEEPS = {"dec": N.SWEEP_DEC, "oct": N.SWEEP_OCT, "lin": N.SWEEP_LIN}


def generate_frequencies(sweep: str, n_points: int, f_start: float, f_stop: float) -> np.ndarray:
    """generate_frequencies (src/analysis/ac.rs:80-109), native."""
    t = SWEEPS[sweep.lower()]
    cnt = lib().ohm_generate_frequencies(t, n_points, f_start, f_stop, None, 0)
    out = np.zeros(max(cnt, 0), dtype=np.float64)
    if cnt > 0:
        lib().ohm_generate_frequencies(t, n_points, f_start, f_stop,
                                       out.ctypes.data_as(N.c_f64p), cnt)
    return out


@dataclass
class AcResult:
    """AcResult (src/analysis/mod.rs): per-node / per-branch complex phasors."""
    frequencies: np.ndarray
    node_voltages: list
    branch_currents: list
    x: np.ndarray                    # [n_freqs, n] complex
    stats: list = field(default_factory=list)


def ac_sweep(backend: HipBackend, mna: ir.MnaSystem, freqs, batch: int = 0,
             params: N.BicgstabParams | None = None, precond: int = N.PREC_JACOBI,
             raise_on_failure: bool = True) -> AcResult:
    """ac::run (src/analysis/ac.rs:25-77): every frequency point solved on
    the GPU, ``batch`` points per launch sequence (0 = all at once)."""
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    n = mna.size
    g, c = mna.g, mna.c
    gv = np.ascontiguousarray(g.values, dtype=np.float64)
    gc = np.ascontiguousarray(g.col_indices, dtype=np.uint32)
    gr = np.ascontiguousarray(g.row_pointers, dtype=np.uint32)
    cv = np.ascontiguousarray(c.values, dtype=np.float64)
    cc = np.ascontiguousarray(c.col_indices, dtype=np.uint32)
    cr = np.ascontiguousarray(c.row_pointers, dtype=np.uint32)
    bac = np.ascontiguousarray(mna.b_ac, dtype=np.complex128)
    x = np.zeros((f.shape[0], n), dtype=np.complex128)
    res = (N.BatchResult * max(f.shape[0], 1))()
    prm = params if params is not None else bicgstab_params()
    rc = lib().ohm_ac_sweep(backend.h, ctypes.c_int64(n), ptr(gr, N.c_u32p), ptr(gc, N.c_u32p),
                            ptr(gv, N.c_f64p), ptr(cr, N.c_u32p), ptr(cc, N.c_u32p),
                            ptr(cv, N.c_f64p), bac.ctypes.data_as(N.c_f64p),
                            ctypes.c_int64(f.shape[0]), ptr(f, N.c_f64p), ctypes.c_int64(batch),
                            ctypes.byref(prm), int(precond), x.ctypes.data_as(N.c_f64p), res)
    stats = [BatchStats(r.iterations, r.refinements, r.residual_norm, r.status, r.reason)
             for r in res[:f.shape[0]]]
    if rc != 0 and raise_on_failure:
        check(rc)
    nn = mna.n_nodes
    nodes = [(name, x[:, i].copy()) for i, name in enumerate(mna.node_names)]
    branches = [(name, x[:, nn + i].copy()) for i, name in enumerate(mna.branch_names)]
    return AcResult(f, nodes, branches, x, stats)


AC_METHODS = {"auto": N.AC_AUTO, "iterative": N.AC_ITERATIVE, "direct": N.AC_DIRECT}


class AcSweepResult:
    """Solutions and per-system figures of a multi-circuit AC sweep.

    x: per circuit [n_freqs, n] complex (None when not kept). records:
    [n_circuits, n_freqs] structured array of ohm_batch_result (iterations,
    refinements, residual_norm, status, reason). stats: the same records as
    per-circuit lists of BatchStats, built on first access."""

    def __init__(self, x: list, records: np.ndarray, device_ms: float):
        self.x = x
        self.records = records
        self.device_ms = device_ms
        self._stats = None

    @property
    def stats(self) -> list:
        if self._stats is None:
            self._stats = [[BatchStats(int(r["iterations"]), int(r["refinements"]),
                                       float(r["residual_norm"]), int(r["status"]),
                                       int(r["reason"])) for r in row] for row in self.records]
        return self._stats

    @property
    def residual_norms(self) -> np.ndarray:
        return self.records["residual_norm"]

    @property
    def converged(self) -> np.ndarray:
        return self.records["status"] == 0


def _ac_circuits(mnas, nf: int, keep_x: bool):
    """ohm_ac_circuit array over compiled circuits (+ the arrays it points
    into, and per-circuit x outputs when keep_x)."""
    k = len(mnas)
    cs = (N.AcCircuit * max(k, 1))()
    keep, xs = [], []
    for i, m in enumerate(mnas):
        arrs = [np.ascontiguousarray(m.g.row_pointers, np.uint32),
                np.ascontiguousarray(m.g.col_indices, np.uint32),
                np.ascontiguousarray(m.g.values, np.float64),
                np.ascontiguousarray(m.c.row_pointers, np.uint32),
                np.ascontiguousarray(m.c.col_indices, np.uint32),
                np.ascontiguousarray(m.c.values, np.float64),
                np.ascontiguousarray(m.b_ac, np.complex128)]
        keep.append(arrs)
        x = np.zeros((nf, m.size), dtype=np.complex128) if keep_x else None
        xs.append(x)
        cs[i] = N.AcCircuit(m.size, ptr(arrs[0], N.c_u32p), ptr(arrs[1], N.c_u32p),
                            ptr(arrs[2], N.c_f64p), ptr(arrs[3], N.c_u32p), ptr(arrs[4], N.c_u32p),
                            ptr(arrs[5], N.c_f64p), arrs[6].ctypes.data_as(N.c_f64p),
                            x.ctypes.data_as(N.c_f64p) if x is not None else None)
    return cs, keep, xs


_RESULT_DTYPE = None


def _records(res, k: int, nf: int) -> np.ndarray:
    """The ohm_batch_result array as a [k, nf] structured numpy array: a view
    of the ctypes array, which it keeps alive (a 25,856-system sweep's
    conversion went from 0.25 ms to a few microseconds)."""
    global _RESULT_DTYPE
    if _RESULT_DTYPE is None:
        _RESULT_DTYPE = np.dtype(N.BatchResult)
    if k * nf == 0:
        return np.zeros((k, nf), dtype=_RESULT_DTYPE)
    return np.frombuffer(res, dtype=_RESULT_DTYPE, count=k * nf).reshape(k, nf)


def ac_sweep_circuits(backend: HipBackend, mnas, freqs, method: str = "auto",
                      params: N.BicgstabParams | None = None, precond: int = N.PREC_JACOBI,
                      keep_x: bool = True, raise_on_failure: bool = True) -> AcSweepResult:
    """ac::run over many circuits at once (ohm_ac_sweep_circuits): banded
    direct solves for chain-like circuits, batched BiCGSTAB for the rest."""
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    nf = f.shape[0]
    k = len(mnas)
    cs, keep, xs = _ac_circuits(mnas, nf, keep_x)
    res = (N.BatchResult * max(k * nf, 1))()
    ms = ctypes.c_double(0.0)
    prm = params if params is not None else bicgstab_params()
    rc = lib().ohm_ac_sweep_circuits(backend.h, ctypes.c_int64(k), cs, ctypes.c_int64(nf),
                                     ptr(f, N.c_f64p), AC_METHODS[method], ctypes.byref(prm),
                                     int(precond), res, ctypes.byref(ms))
    del keep
    if rc != 0 and raise_on_failure:
        check(rc)
    return AcSweepResult(xs, _records(res, k, nf), ms.value)


class AcPlan:
    """The direct AC sweep split into its symbolic half (ohm_ac_plan_create:
    pattern union, RCM and band rows per circuit, uploaded once) and its
    numeric half (solve: every (circuit, frequency) system by banded LU on
    the device). Solutions of a sweep that fits the device in one chunk stay
    resident for x()."""

    def __init__(self, backend: HipBackend, mnas):
        self.backend = backend
        self.sizes = [m.size for m in mnas]
        cs, keep, _ = _ac_circuits(mnas, 0, False)
        self.h = ctypes.c_void_p()
        check(lib().ohm_ac_plan_create(backend.h, ctypes.c_int64(len(mnas)), cs,
                                       ctypes.byref(self.h)))
        del keep
        self.n_freqs = 0

    def solve(self, freqs, keep_x: bool = False, raise_on_failure: bool = True) -> AcSweepResult:
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        nf = f.shape[0]
        k = len(self.sizes)
        xs = [np.zeros((nf, n), dtype=np.complex128) for n in self.sizes] if keep_x else None
        xp = None
        if keep_x:
            xp = (N.c_f64p * max(k, 1))(*[x.ctypes.data_as(N.c_f64p) for x in xs])
        res = (N.BatchResult * max(k * nf, 1))()
        ms = ctypes.c_double(0.0)
        rc = lib().ohm_ac_plan_solve(self.h, ctypes.c_int64(nf), ptr(f, N.c_f64p), xp, res,
                                     ctypes.byref(ms))
        self.n_freqs = nf
        if rc != 0 and raise_on_failure:
            check(rc)
        return AcSweepResult(xs if keep_x else [None] * k, _records(res, k, nf), ms.value)

    def stats(self) -> dict:
        """ohm_ac_plan_stats of the last solve: device ms per phase, rows per
        part (0: one lane per whole system), systems redone in one lane,
        chunks."""
        ms = (ctypes.c_double * 6)()
        rows, redo, chunks = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        check(lib().ohm_ac_plan_stats(self.h, ms, ctypes.byref(rows), ctypes.byref(redo),
                                      ctypes.byref(chunks)))
        names = ("relations", "reduced", "part_solve", "residual", "redo", "unpermute")
        return {"phase_ms": dict(zip(names, list(ms))), "rows_per_part": rows.value,
                "redone_systems": redo.value, "chunks": chunks.value}

    def x(self, circuit: int, f_first: int = 0, f_count: int | None = None) -> np.ndarray:
        """Resident solutions of one circuit: [f_count, n] complex."""
        if f_count is None:
            f_count = self.n_freqs - f_first
        # an out-of-range circuit is reported by ohm_ac_plan_get_x itself
        n = self.sizes[circuit] if 0 <= circuit < len(self.sizes) else 0
        out = np.zeros((max(f_count, 0), n), dtype=np.complex128)
        check(lib().ohm_ac_plan_get_x(self.h, ctypes.c_int64(circuit), ctypes.c_int64(f_first),
                                      ctypes.c_int64(f_count), out.ctypes.data_as(N.c_f64p)))
        return out

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and not sys.is_finalizing():
            lib().ohm_ac_plan_free(self.h)
            self.h = None


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
                  initial_guess=None, bicgstab_floor=1e-15, forcing="fixed", forcing_max=0.1,
                  forcing_safeguard="published", linear_solver="bicgstab"):
    """NewtonParams (src/solver/newton.rs:34-55). bicgstab_tolerance is
    relative to the Newton residual ||F(x_k)|| (inexact Newton); forcing
    "ew" (Eisenstat-Walker choice 2) or "ew1" (choice 1) adapts it per
    iteration down to that value; forcing_safeguard "always" applies the
    forcing safeguard at every iteration instead of only above 0.1.
    linear_solver "dense" solves every Newton step by GPU dense LU (systems
    of at most 4096 unknowns)."""
    p = N.NewtonParams()
    lib().ohm_newton_default_params(ctypes.byref(p))
    p.max_iterations = max_iterations
    p.abs_tol = abs_tol
    p.bicgstab_max_iterations = bicgstab_max_iterations
    p.bicgstab_tolerance = bicgstab_tolerance
    p.bicgstab_floor = bicgstab_floor
    p.precision = N.FP32_IR if precision in ("fp32_ir", "mixed") else N.FP64
    p.precond = precond
    p.forcing = {"fixed": 0, "ew": 1, "ew2": 1, "ew1": 2}[forcing]
    p.forcing_safeguard = {"published": 0, "always": 1}[forcing_safeguard]
    p.forcing_max = forcing_max
    p.linear_solver = {"bicgstab": 0, "dense": 1, "dense_lu": 1}[linear_solver]
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
        self.n_mosfets = nm
        self._counts = (nd, nq, nm)
        self._descs_keep = (d, q, m)
        self.h = ctypes.c_void_p()
        check(lib().ohm_newton_create(
            backend.h, ptr(self._v, N.c_f64p), ptr(self._c, N.c_u32p), ptr(self._r, N.c_u32p),
            ptr(self._b, N.c_f64p), vp(d), ctypes.c_size_t(nd), vp(q), ctypes.c_size_t(nq),
            vp(m), ctypes.c_size_t(nm), ctypes.c_size_t(self.n), ctypes.c_size_t(self.nnz),
            ctypes.byref(self.h)))

    def update(self, values=None, base_b=None, diodes=None, bjts=None, mosfets=None):
        """ohm_newton_update: new base matrix values / RHS / device
        parameters on the same topology (None keeps the current ones); the
        preconditioner plan built by the first run is reused."""
        def arr(a, dtype):
            return None if a is None else np.ascontiguousarray(a, dtype=dtype)
        v, b = arr(values, np.float64), arr(base_b, np.float64)
        if v is not None and v.shape[0] != self.nnz:
            raise ValueError(f"values: expected {self.nnz} entries, got {v.shape[0]}")
        if b is not None and b.shape[0] != self.n:
            raise ValueError(f"base_b: expected {self.n} entries, got {b.shape[0]}")
        descs = []
        for a, dtype, count in zip((diodes, bjts, mosfets), (ir.DIODE_DESC_DTYPE, ir.BJT_DESC_DTYPE,
                                                             ir.MOSFET_DESC_DTYPE), self._counts):
            d = arr(a, dtype)
            if d is not None and len(d) != count:
                raise ValueError(f"descriptor arrays must keep the counts given at create "
                                 f"({count}, got {len(d)})")
            descs.append(d)
        check(lib().ohm_newton_update(
            self.h, ptr(v, N.c_f64p) if v is not None else None,
            ptr(b, N.c_f64p) if b is not None else None,
            *[vp(d) if d is not None and len(d) else None for d in descs]))
        if v is not None:
            self._v = v
        if b is not None:
            self._b = b

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

    def adjoint(self, dJdx, res_nodes=None, params: N.BicgstabParams | None = None):
        """ohm_adjoint_dc at the operating point of the last run: returns
        (dObj/dg per resistor, dObj/dkp per MOSFET, lambda, SolveStats)."""
        g = np.ascontiguousarray(dJdx, dtype=np.float64)
        rn = (np.zeros((0, 2), np.uint32) if res_nodes is None
              else np.ascontiguousarray(res_nodes, dtype=np.uint32).reshape(-1, 2))
        gg = np.zeros(max(rn.shape[0], 1))
        nm = int(getattr(self, "n_mosfets", 0))
        gk = np.zeros(max(nm, 1))
        lam = np.zeros(self.n)
        res = N.BicgstabResult()
        check(lib().ohm_adjoint_dc(self.h, ptr(g, N.c_f64p),
                                   ctypes.byref(params) if params is not None else None,
                                   ctypes.c_int64(rn.shape[0]), ptr(rn, N.c_u32p),
                                   ptr(gg, N.c_f64p), ptr(gk, N.c_f64p), ptr(lam, N.c_f64p),
                                   ctypes.byref(res)))
        return gg[: rn.shape[0]], gk[:nm], lam, SolveStats.from_c(res)

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
    """The reference's free function newton_solve (src/solver/newton.rs:65-78)
    through its one-shot C entry point ohm_newton_solve: base matrix (values
    and CSR pattern), RHS and device descriptors in; (x, NewtonStats) out.
    Errors carry the reference's texts (NewtonNotConverged,
    NewtonNumericalError)."""
    v = np.ascontiguousarray(g.values, dtype=np.float64)
    c = np.ascontiguousarray(g.col_indices, dtype=np.uint32)
    r = np.ascontiguousarray(g.row_pointers, dtype=np.uint32)
    b = np.ascontiguousarray(base_b, dtype=np.float64)
    n = g.nrows
    if b.shape[0] != n:
        raise ValueError(f"base_b: expected {n} entries, got {b.shape[0]}")
    d, nd = _descs(diodes, ir.DIODE_DESC_DTYPE)
    q, nq = _descs(bjts, ir.BJT_DESC_DTYPE)
    m, nm = _descs(mosfets, ir.MOSFET_DESC_DTYPE)
    p, keep = newton_params(**params)
    res = N.NewtonResult()
    per = np.zeros(max(p.max_iterations, 1), dtype=np.int32)
    x = np.zeros(n)
    check(lib().ohm_newton_solve(
        backend.h, ptr(v, N.c_f64p), ptr(c, N.c_u32p), ptr(r, N.c_u32p), ptr(b, N.c_f64p),
        vp(d), ctypes.c_size_t(nd), vp(q), ctypes.c_size_t(nq), vp(m), ctypes.c_size_t(nm),
        ctypes.c_size_t(n), ctypes.c_size_t(v.shape[0]), ctypes.byref(p), ptr(x, N.c_f64p),
        ctypes.byref(res), ptr(per, N.c_i32p)))
    del keep
    return x, NewtonStats(res.iterations, res.bicgstab_iterations, per[: res.iterations].tolist(),
                          res.max_diff, res.device_ms, res.fp64_fallbacks)
