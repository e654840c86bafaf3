"""CPU oracle for the ohmnivore solve path — TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``oracle/libohm_oracle.so`` (built from ohm_oracle.c by
oracle/Makefile). Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this package, and only as the reference answer
or the timed CPU baseline. Each function restates the reference algorithm cited in
ohm_oracle.c / ohm_oracle.h.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from ohmnivore_amd import circuit as _ir

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libohm_oracle.so")
_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
f64p = ctypes.POINTER(ctypes.c_double)
i64p = ctypes.POINTER(ctypes.c_int64)


class OracleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oo_last_error.restype = ctypes.c_char_p
        _lib.oo_from_triplets.restype = ctypes.c_int64
        _lib.oo_pnjlim.restype = ctypes.c_double
        _lib.oo_eval_pulse.restype = ctypes.c_double
        _lib.oo_source_value.restype = ctypes.c_double
        _lib.oo_pnjlim.argtypes = [ctypes.c_double] * 4
        _lib.oo_eval_pulse.argtypes = [ctypes.c_double] * 8
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise OracleError(rc, lib().oo_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(t)


def _csr_args(a: _ir.CsrMatrix):
    rp = np.ascontiguousarray(a.row_pointers, dtype=np.uint32)
    cl = np.ascontiguousarray(a.col_indices, dtype=np.uint32)
    if np.iscomplexobj(a.values):
        v = np.ascontiguousarray(a.values, dtype=np.complex128).view(np.float64)
    else:
        v = np.ascontiguousarray(a.values, dtype=np.float64)
    return rp, cl, v


def from_triplets(nrows: int, ncols: int, triplets) -> _ir.CsrMatrix:
    """CsrMatrix::from_triplets (src/sparse.rs:52-100)."""
    t = list(triplets)
    cplx = any(isinstance(v, complex) for _, _, v in t)
    rows = np.array([r for r, _, _ in t], dtype=np.int64)
    cols = np.array([c for _, c, _ in t], dtype=np.int64)
    vals = np.array([v for _, _, v in t], dtype=np.complex128 if cplx else np.float64)
    return from_triplet_arrays(nrows, ncols, rows, cols, vals)


def from_triplet_arrays(nrows, ncols, rows, cols, vals) -> _ir.CsrMatrix:
    n = rows.shape[0]
    cplx = np.iscomplexobj(vals)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    v = np.ascontiguousarray(vals, dtype=np.complex128 if cplx else np.float64)
    vin = v.view(np.float64)
    rp = np.zeros(nrows + 1, dtype=np.int64)
    oc = np.zeros(max(n, 1), dtype=np.int64)
    ov = np.zeros(max(n, 1) * (2 if cplx else 1), dtype=np.float64)
    nnz = lib().oo_from_triplets(ctypes.c_int64(nrows), ctypes.c_int64(n), _p(rows, i64p),
                                 _p(cols, i64p), _p(vin, f64p), int(cplx), _p(rp, i64p),
                                 _p(oc, i64p), _p(ov, f64p))
    vals_out = ov.view(np.complex128)[:nnz].copy() if cplx else ov[:nnz].copy()
    return _ir.CsrMatrix(nrows, ncols, vals_out, oc[:nnz].astype(np.uint32),
                         rp.astype(np.uint32))


def compile(ckt: _ir.Circuit) -> _ir.MnaSystem:
    """compiler::compile (src/compiler.rs:194-754)."""
    cc = _ir.CCircuit(ckt)
    m = _ir.MnaC()
    L = lib()
    _check(L.oo_compile(cc.ref(), ctypes.byref(m)))
    try:
        return _ir.mna_from_c(m, ckt)
    finally:
        L.oo_mna_free(ctypes.byref(m))


def spmv(a: _ir.CsrMatrix, x: np.ndarray) -> np.ndarray:
    rp, cl, v = _csr_args(a)
    cplx = np.iscomplexobj(a.values)
    xx = np.ascontiguousarray(x, dtype=np.complex128 if cplx else np.float64)
    y = np.zeros_like(xx)
    lib().oo_spmv(ctypes.c_int64(a.nrows), _p(rp, u32p), _p(cl, u32p), _p(v, f64p),
                  _p(xx.view(np.float64), f64p), _p(y.view(np.float64), f64p), int(cplx))
    return y


def dense_solve(a: _ir.CsrMatrix, b) -> np.ndarray:
    """CpuSolver::solve_real / solve_complex (src/solver/cpu.rs:28-170)."""
    rp, cl, v = _csr_args(a)
    cplx = np.iscomplexobj(a.values) or np.iscomplexobj(b)
    if cplx and not np.iscomplexobj(a.values):
        v = np.ascontiguousarray(a.values.astype(np.complex128)).view(np.float64)
    bb = np.ascontiguousarray(b, dtype=np.complex128 if cplx else np.float64)
    x = np.zeros(max(a.nrows, 1), dtype=bb.dtype)
    _check(lib().oo_dense_solve(ctypes.c_int64(a.nrows), ctypes.c_int64(a.ncols),
                                ctypes.c_int64(bb.shape[0]), _p(rp, u32p), _p(cl, u32p),
                                _p(v, f64p), _p(bb.view(np.float64), f64p),
                                _p(x.view(np.float64), f64p), int(cplx)))
    return x[: a.nrows]


def sparse_lu_solve(a: _ir.CsrMatrix, b) -> np.ndarray:
    """Sparse direct LU standing in for faer sp_lu (src/solver/sparse_direct.rs:12-61)."""
    rp, cl, v = _csr_args(a)
    cplx = np.iscomplexobj(a.values)
    bb = np.ascontiguousarray(b, dtype=np.complex128 if cplx else np.float64)
    x = np.zeros_like(bb)
    _check(lib().oo_sparse_lu_solve(ctypes.c_int64(a.nrows), _p(rp, u32p), _p(cl, u32p),
                                    _p(v, f64p), _p(bb.view(np.float64), f64p),
                                    _p(x.view(np.float64), f64p), int(cplx)))
    return x


def jacobi_inv_diag(a: _ir.CsrMatrix) -> np.ndarray:
    rp, cl, v = _csr_args(a)
    inv = np.zeros(a.nrows)
    lib().oo_jacobi_inv_diag(ctypes.c_int64(a.nrows), _p(rp, u32p), _p(cl, u32p), _p(v, f64p),
                             _p(inv, f64p))
    return inv


class _Prec(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("inv_diag", f64p), ("ml_rowptr", u32p),
                ("ml_col", u32p), ("ml_val", f64p), ("mu_rowptr", u32p), ("mu_col", u32p),
                ("mu_val", f64p)]


def bicgstab(a: _ir.CsrMatrix, b, precond=None, max_iters=10_000, tol=1e-5, shadow="hashed",
             breakdown="scaled"):
    """bicgstab_with_params (src/solver/bicgstab.rs:56-189). precond is None,
    ("jacobi", inv_diag) or ("isai", M_L, M_U). Returns (x, iterations)."""
    rp, cl, v = _csr_args(a)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(bb)
    p = _Prec()
    keep = []
    if precond is None:
        p.kind = 0
    elif precond[0] == "jacobi":
        p.kind = 1
        inv = np.ascontiguousarray(precond[1], dtype=np.float64)
        keep.append(inv)
        p.inv_diag = _p(inv, f64p)
    else:
        p.kind = 2
        ml, mu = precond[1], precond[2]
        a1, a2 = _csr_args(ml), _csr_args(mu)
        keep += [a1, a2]
        p.ml_rowptr, p.ml_col, p.ml_val = _p(a1[0], u32p), _p(a1[1], u32p), _p(a1[2], f64p)
        p.mu_rowptr, p.mu_col, p.mu_val = _p(a2[0], u32p), _p(a2[1], u32p), _p(a2[2], f64p)
    it = ctypes.c_int64(0)
    _check(lib().oo_bicgstab(ctypes.c_int64(a.nrows), _p(rp, u32p), _p(cl, u32p), _p(v, f64p),
                             _p(bb, f64p), _p(x, f64p), ctypes.byref(p), ctypes.c_int64(max_iters),
                             ctypes.c_double(tol), int(shadow == "hashed"),
                             int(breakdown == "scaled"), ctypes.byref(it)))
    return x, it.value


def _take_csr(n, nnz, rp, cl, vv):
    L = lib()
    out = _ir.CsrMatrix(n, n, np.ctypeslib.as_array(vv, shape=(max(nnz.value, 1),))[: nnz.value].copy(),
                        np.ctypeslib.as_array(cl, shape=(max(nnz.value, 1),))[: nnz.value].copy(),
                        np.ctypeslib.as_array(rp, shape=(n + 1,)).copy())
    for q in (rp, cl, vv):
        L.oo_free(ctypes.cast(q, ctypes.c_void_p))
    return out


def _mk_out():
    return ctypes.c_int64(0), u32p(), u32p(), f64p()


def lu_reference(a: _ir.CsrMatrix):
    """ilu0 (src/solver/preconditioner.rs:64-153): returns (L, U, perm)."""
    rp, cl, v = _csr_args(a)
    l = _mk_out()
    u = _mk_out()
    perm = i64p()
    _check(lib().oo_lu_reference(ctypes.c_int64(a.nrows), _p(rp, u32p), _p(cl, u32p), _p(v, f64p),
                                 ctypes.byref(l[0]), ctypes.byref(l[1]), ctypes.byref(l[2]),
                                 ctypes.byref(l[3]), ctypes.byref(u[0]), ctypes.byref(u[1]),
                                 ctypes.byref(u[2]), ctypes.byref(u[3]), ctypes.byref(perm)))
    n = a.nrows
    pm = np.ctypeslib.as_array(perm, shape=(max(n, 1),))[:n].copy()
    lib().oo_free(ctypes.cast(perm, ctypes.c_void_p))
    return _take_csr(n, *l), _take_csr(n, *u), pm


def isai_reference(a: _ir.CsrMatrix, level: int = 1):
    """compute_isai (src/solver/preconditioner.rs:239-277): returns (M_L, M_U, perm)."""
    rp, cl, v = _csr_args(a)
    ml = _mk_out()
    mu = _mk_out()
    perm = i64p()
    _check(lib().oo_isai_reference(ctypes.c_int64(a.nrows), _p(rp, u32p), _p(cl, u32p),
                                   _p(v, f64p), int(level), ctypes.byref(ml[0]),
                                   ctypes.byref(ml[1]), ctypes.byref(ml[2]), ctypes.byref(ml[3]),
                                   ctypes.byref(mu[0]), ctypes.byref(mu[1]), ctypes.byref(mu[2]),
                                   ctypes.byref(mu[3]), ctypes.byref(perm)))
    n = a.nrows
    pm = np.ctypeslib.as_array(perm, shape=(max(n, 1),))[:n].copy()
    lib().oo_free(ctypes.cast(perm, ctypes.c_void_p))
    return _take_csr(n, *ml), _take_csr(n, *mu), pm


def ilu0_isai(a: _ir.CsrMatrix, level: int = 1):
    """Exact ILU(0)+ISAI(level) — the backend's preconditioner. Returns (L, U, M_L, M_U)."""
    rp, cl, v = _csr_args(a)
    outs = [_mk_out() for _ in range(4)]
    args = []
    for o in outs:
        args += [ctypes.byref(o[0]), ctypes.byref(o[1]), ctypes.byref(o[2]), ctypes.byref(o[3])]
    _check(lib().oo_ilu0_isai(ctypes.c_int64(a.nrows), _p(rp, u32p), _p(cl, u32p), _p(v, f64p),
                              int(level), *args))
    return tuple(_take_csr(a.nrows, *o) for o in outs)


def _desc_ptr(arr, dtype):
    arr = np.ascontiguousarray(arr, dtype=dtype) if len(arr) else np.zeros(1, dtype=dtype)
    return arr, ctypes.c_void_p(arr.ctypes.data)


def diode_eval(desc, x):
    d = np.ascontiguousarray(np.atleast_1d(desc), dtype=_ir.DIODE_DESC_DTYPE)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    i, g = ctypes.c_double(), ctypes.c_double()
    lib().oo_diode_eval(ctypes.c_void_p(d.ctypes.data), _p(xx, f64p), ctypes.byref(i),
                        ctypes.byref(g))
    return i.value, g.value


def bjt_eval(desc, x):
    d = np.ascontiguousarray(np.atleast_1d(desc), dtype=_ir.BJT_DESC_DTYPE)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(6)
    lib().oo_bjt_eval(ctypes.c_void_p(d.ctypes.data), _p(xx, f64p), _p(out, f64p))
    return out


def mosfet_core(vgs, vds, vto, kp, lam, polarity):
    i, gm, gds = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().oo_mosfet_core(ctypes.c_double(vgs), ctypes.c_double(vds), ctypes.c_double(vto),
                         ctypes.c_double(kp), ctypes.c_double(lam), ctypes.c_double(polarity),
                         ctypes.byref(i), ctypes.byref(gm), ctypes.byref(gds))
    return i.value, gm.value, gds.value


def pnjlim(vnew, vold, vt, vcrit):
    return lib().oo_pnjlim(vnew, vold, vt, vcrit)


def eval_pulse(v1, v2, td, tr, tf, pw, per, t):
    return lib().oo_eval_pulse(v1, v2, td, tr, tf, pw, per, t)


class _NewtonParams(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("abs_tol", ctypes.c_double),
                ("linear_mode", ctypes.c_int), ("initial_guess", f64p)]


def _newton_arrays(g, diodes, bjts, mosfets):
    rp, cl, v = _csr_args(g)
    d, dp = _desc_ptr(diodes, _ir.DIODE_DESC_DTYPE)
    q, qp = _desc_ptr(bjts, _ir.BJT_DESC_DTYPE)
    m, mp = _desc_ptr(mosfets, _ir.MOSFET_DESC_DTYPE)
    return rp, cl, v, (d, dp), (q, qp), (m, mp)


def newton_solve(g: _ir.CsrMatrix, b, diodes=(), bjts=(), mosfets=(), max_iterations=50,
                 abs_tol=1e-4, linear_mode="sparse_lu", initial_guess=None):
    """newton_solve (src/solver/newton.rs:65-513) with a direct-LU linear step.
    Returns (x, iterations, max_diff)."""
    rp, cl, v, dd, qq, mm = _newton_arrays(g, diodes, bjts, mosfets)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(g.nrows)
    prm = _NewtonParams(max_iterations, abs_tol, 1 if linear_mode == "dense" else 0, None)
    if initial_guess is not None:
        ig = np.ascontiguousarray(initial_guess, dtype=np.float64)
        prm.initial_guess = _p(ig, f64p)
    it, md = ctypes.c_int64(), ctypes.c_double()
    _check(lib().oo_newton_solve(
        ctypes.c_int64(g.nrows), ctypes.c_int64(g.nnz), _p(rp, u32p), _p(cl, u32p), _p(v, f64p),
        _p(bb, f64p), ctypes.c_int64(len(diodes)), dd[1], ctypes.c_int64(len(bjts)), qq[1],
        ctypes.c_int64(len(mosfets)), mm[1], ctypes.byref(prm), _p(x, f64p), ctypes.byref(it),
        ctypes.byref(md)))
    return x, it.value, md.value


def newton_assemble(g: _ir.CsrMatrix, b, x, diodes=(), bjts=(), mosfets=()):
    rp, cl, v, dd, qq, mm = _newton_arrays(g, diodes, bjts, mosfets)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    vo = np.zeros(g.nnz)
    bo = np.zeros(g.nrows)
    lib().oo_newton_assemble(ctypes.c_int64(g.nrows), ctypes.c_int64(g.nnz), _p(v, f64p),
                             _p(bb, f64p), ctypes.c_int64(len(diodes)), dd[1],
                             ctypes.c_int64(len(bjts)), qq[1], ctypes.c_int64(len(mosfets)), mm[1],
                             _p(xx, f64p), _p(vo, f64p), _p(bo, f64p))
    return vo, bo


def newton_limit(n, x_old, x_new, diodes=(), bjts=(), mosfets=()):
    d, dp = _desc_ptr(diodes, _ir.DIODE_DESC_DTYPE)
    q, qp = _desc_ptr(bjts, _ir.BJT_DESC_DTYPE)
    m, mp = _desc_ptr(mosfets, _ir.MOSFET_DESC_DTYPE)
    xo = np.ascontiguousarray(x_old, dtype=np.float64)
    xn = np.array(x_new, dtype=np.float64)
    lib().oo_newton_limit(ctypes.c_int64(n), ctypes.c_int64(len(diodes)), dp,
                          ctypes.c_int64(len(bjts)), qp, ctypes.c_int64(len(mosfets)), mp,
                          _p(xo, f64p), _p(xn, f64p))
    return xn


def transient(mna: _ir.MnaSystem, sources, pwl=None, method="be", tstep=1e-12, tstop=1e-10,
              tstart=0.0, x0=None, probes=None, max_records=None):
    """transient::run (src/analysis/transient.rs:25-199) with sparse LU solves.
    Returns (times, waves[records, probes], x_final, accepted, rejected)."""
    n = mna.size
    grp, gcl, gv = _csr_args(mna.g)
    crp, ccl, cv = _csr_args(mna.c)
    src = np.ascontiguousarray(sources, dtype=_ir.TRAN_SOURCE_DTYPE)
    pw = np.ascontiguousarray(pwl if pwl is not None else np.zeros((0, 2)), dtype=np.float64)
    pr = np.ascontiguousarray(np.arange(n) if probes is None else probes, dtype=np.uint32)
    if max_records is None:
        max_records = int(np.ceil(tstop / tstep * (1 if method == "be" else 4))) + 2
    times = np.zeros(max_records)
    waves = np.zeros((max_records, pr.shape[0]))
    xf = np.zeros(max(n, 1))
    x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    nrec, acc, rej = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _check(lib().oo_transient(
        ctypes.c_int64(n), _p(grp, u32p), _p(gcl, u32p), _p(gv, f64p), _p(crp, u32p),
        _p(ccl, u32p), _p(cv, f64p), ctypes.c_int64(src.shape[0]), src.ctypes.data_as(ctypes.c_void_p),
        _p(pw, f64p), 0 if method == "be" else 1, ctypes.c_double(tstep), ctypes.c_double(tstop),
        ctypes.c_double(tstart), _p(x0a, f64p) if x0a is not None else None,
        ctypes.c_int64(pr.shape[0]), _p(pr, u32p), ctypes.c_int64(max_records), _p(times, f64p),
        _p(waves, f64p), ctypes.byref(nrec), _p(xf, f64p), ctypes.byref(acc), ctypes.byref(rej)))
    r = nrec.value
    return times[:r], waves[:r], xf[:n], acc.value, rej.value


def source_value(src_entry, t: float, pwl=None) -> float:
    """evaluate (src/analysis/transient_source.rs:8-37) for one ohm_tran_source."""
    e = np.ascontiguousarray(np.array([src_entry], dtype=_ir.TRAN_SOURCE_DTYPE))
    pw = np.ascontiguousarray(pwl if pwl is not None else np.zeros((0, 2)), dtype=np.float64)
    return lib().oo_source_value(e.ctypes.data_as(ctypes.c_void_p), _p(pw, f64p), ctypes.c_double(t))


def transient_nonlinear(mna: _ir.MnaSystem, sources, pwl=None, method="adaptive", tstep=1e-9,
                        tstop=1e-7, tstart=0.0, x0=None, abs_tol=1e-12, max_iterations=100):
    """Transient with the device currents in the time loop (no reference
    counterpart: the reference steps G and C only). Each step is the
    oracle's LU Newton (newton_solve) on
        BE:   (G + C/h) x + i(x) = b(t+h) + (C/h) x_n
        TRAP: (G + 2C/h) x + i(x) = b(t+h) + (2C/h) x_n + b(t) - G x_n - i(x_n)
    with the step control of transient::run (src/analysis/transient.rs:84-178):
    BE first step, TRAP with a BE companion for the LTE estimate, h in
    [tstep/1e4, tstep], LTE tolerance 1e-3; a Newton failure rejects the step
    and halves h. Without devices this is transient() exactly.
    Returns (times, waves[records, n], x_final, accepted, rejected)."""
    n = mna.size
    g, c = mna.g, mna.c
    descs = (mna.diode_descriptors, mna.bjt_descriptors, mna.mosfet_descriptors)
    grp = g.row_pointers.astype(np.int64)
    cg = np.zeros(g.nnz)
    for r in range(n):
        cols = g.col_indices[grp[r]:grp[r + 1]]
        for q in range(int(c.row_pointers[r]), int(c.row_pointers[r + 1])):
            pos = int(np.searchsorted(cols, c.col_indices[q]))
            if pos >= cols.shape[0] or cols[pos] != c.col_indices[q]:
                raise OracleError(11, "C entry outside G's pattern")
            cg[grp[r] + pos] += c.values[q]
    src = np.atleast_1d(np.asarray(sources, dtype=_ir.TRAN_SOURCE_DTYPE))

    def b_at(t):
        b = np.zeros(n)
        for e in src:
            v = source_value(e, t, pwl)
            for j in range(int(e["n_idx"])):
                b[int(e["idx"][j])] += float(e["sign"][j]) * v
        return b

    def csr(vals):
        return _ir.CsrMatrix(n, n, vals, g.col_indices, g.row_pointers)

    def newton(sig, rhs, guess):
        x, _, _ = newton_solve(csr(g.values + sig * cg), rhs, *descs, abs_tol=abs_tol,
                               max_iterations=max_iterations, initial_guess=guess)
        return x

    def be_step(x, t_next, h):
        return newton(1.0 / h, b_at(t_next) + spmv(csr(cg), x) / h, x)

    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    times, waves = [], []
    if tstart <= 0.0:
        times.append(0.0)
        waves.append(x.copy())
    h_min, h_max, lte_tol = tstep / 10000.0, tstep, 1e-3
    t, h = 0.0, min(tstep, h_max)
    acc = rej = 0
    use_be, fails = True, 0
    while t < tstop:
        if t + h > tstop:
            h = tstop - t
        if h < h_min * 0.5:
            break
        t_next = t + h
        lte, ok = 0.0, True
        try:
            if method == "be" or use_be:
                x_new = be_step(x, t_next, h)
            else:
                vals, bb = newton_assemble(g, np.zeros(n), x, *descs)
                d = b_at(t) + bb - spmv(csr(vals), x)
                x_tr = newton(2.0 / h, b_at(t_next) + 2.0 * spmv(csr(cg), x) / h + d, x)
                x_be = be_step(x, t_next, h)
                lte = (2.0 / 3.0) * float(np.max(np.abs(x_tr - x_be), initial=0.0))
                x_new = x_tr
        except OracleError:
            if method == "be":
                raise
            ok = False
        if method != "be" and (not ok or lte > lte_tol):
            h_new = 0.5 * h if not ok else h * min(2.0, max(0.5, 0.9 * np.sqrt(lte_tol / lte)))
            h = max(h_new, h_min)
            fails += 1
            if fails >= 10 and h <= h_min * 1.01:
                raise OracleError(8, "Timestep error: consecutive failures at minimum timestep")
            rej += 1
            use_be = True
            continue
        if method != "be":
            if use_be:
                use_be = False
            else:
                grow = 2.0 if lte == 0.0 else min(2.0, max(0.5, 0.9 * np.sqrt(lte_tol / lte)))
                h = min(h_max, max(h_min, h * grow))  # sqrt(tol / 0) = inf -> 2 in C
        fails = 0
        acc += 1
        x = x_new
        t = t_next
        if t >= tstart:
            times.append(t)
            waves.append(x.copy())
    return np.array(times), np.array(waves).reshape(len(times), n), x, acc, rej


def ac_sweep(mna: _ir.MnaSystem, freqs) -> np.ndarray:
    """ac::run over one circuit (src/analysis/ac.rs:56-70) with the sparse LU
    as the linear solve: x[f] = (G + j 2 pi f C)^-1 b_ac for every f
    (oo_ac_sweep; the ordering is computed once per circuit). Releases the
    GIL while it runs, so circuits can be swept on host threads."""
    n = mna.size
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    g, c = mna.g, mna.c
    grp, gcl, gv = _csr_args(g)
    crp, ccl, cv = _csr_args(c) if c is not None else (None, None, None)
    bb = np.ascontiguousarray(mna.b_ac, dtype=np.complex128)
    x = np.zeros((f.shape[0], n), dtype=np.complex128)
    _check(lib().oo_ac_sweep(ctypes.c_int64(n), _p(grp, u32p), _p(gcl, u32p), _p(gv, f64p),
                             _p(crp, u32p) if c is not None else None,
                             _p(ccl, u32p) if c is not None else None,
                             _p(cv, f64p) if c is not None else None,
                             _p(bb.view(np.float64), f64p), ctypes.c_int64(f.shape[0]),
                             _p(f, f64p), _p(x.view(np.float64), f64p)))
    return x
