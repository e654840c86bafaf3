"""Transient analysis on the GPU (transient::run, src/analysis/transient.rs:25-199).

``source_table`` restates collect_source_info (:211-262); ``Transient`` steps
one circuit, or K Monte Carlo instances of it, through libohm_hip.so: the
companion matrix, right-hand side, sources, solve and probe gather all run
on the device, so a timestep never returns to the host.
"""
from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass

import numpy as np

from . import _native as N
from . import circuit as ir
from ._native import check, lib
from .solver import HipBackend, _with, bicgstab_params


class TranSource(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n_idx", ctypes.c_int32),
                ("idx", ctypes.c_uint32 * 2), ("sign", ctypes.c_double * 2),
                ("dc", ctypes.c_double), ("param", ctypes.c_double * 7),
                ("pwl_offset", ctypes.c_int64), ("pwl_count", ctypes.c_int64)]


class TranParams(ctypes.Structure):
    _fields_ = [("tstep", ctypes.c_double), ("tstop", ctypes.c_double),
                ("tstart", ctypes.c_double), ("method", ctypes.c_int32),
                ("precond", ctypes.c_int32), ("solver", N.BicgstabParams)]


class TranStats(ctypes.Structure):
    _fields_ = [("steps_accepted", ctypes.c_int64), ("steps_rejected", ctypes.c_int64),
                ("linear_solves", ctypes.c_int64), ("bicgstab_iterations", ctypes.c_int64),
                ("records", ctypes.c_int64), ("t_end", ctypes.c_double),
                ("device_ms", ctypes.c_double)]


assert ctypes.sizeof(TranSource) == ir.TRAN_SOURCE_DTYPE.itemsize
METHODS = {"be": 0, "fixed_be": 0, "adaptive": 1, "trap": 1}


def source_table(ckt: ir.Circuit, mna: ir.MnaSystem):
    """collect_source_info (src/analysis/transient.rs:211-262): one entry per
    V source (its branch row, sign +1) and I source (-1 at n+, +1 at n-,
    ground omitted), in netlist order; inductors advance the branch count."""
    comps = ckt.components
    index = {name: i for i, name in enumerate(mna.node_names)}
    names = ckt.node_names
    n_nodes = mna.n_nodes
    out = []
    pwl = []  # (time, value) rows; a PWL entry names its run by offset and count
    branch = 0
    for k in range(comps.shape[0]):
        kind = int(comps["kind"][k])
        if kind not in (ir.V, ir.I, ir.L):
            continue
        if kind == ir.L:
            branch += 1
            continue
        e = np.zeros(1, dtype=ir.TRAN_SOURCE_DTYPE)[0]
        tr = ckt.tran.get(k)
        flags = int(comps["flags"][k])
        if isinstance(tr, ir.Pulse):
            e["kind"] = ir.SRC_PULSE
            e["param"] = [tr.v1, tr.v2, tr.td, tr.tr, tr.tf, tr.pw, tr.per]
        elif isinstance(tr, ir.Sin):
            e["kind"] = ir.SRC_SIN
            e["param"][:5] = [tr.vo, tr.va, tr.freq, tr.td, tr.theta]
        elif isinstance(tr, ir.Exp):
            e["kind"] = ir.SRC_EXP
            e["param"][:6] = [tr.v1, tr.v2, tr.td1, tr.tau1, tr.td2, tr.tau2]
        elif isinstance(tr, ir.Pwl):
            e["kind"] = ir.SRC_PWL
            e["pwl_offset"] = len(pwl)
            e["pwl_count"] = len(tr.pairs)
            pwl.extend((float(t), float(v)) for t, v in tr.pairs)
        elif tr is not None:
            raise TypeError(f"component {k}: unknown waveform {type(tr).__name__}")
        else:
            e["kind"] = ir.SRC_DC
            e["dc"] = float(comps["value"][k]) if flags & ir.HAS_DC else 0.0
        if kind == ir.V:
            e["n_idx"] = 1
            e["idx"][0] = n_nodes + branch
            e["sign"][0] = 1.0
            branch += 1
        else:
            j = 0
            for node, sign in ((int(comps["node"][k][0]), -1.0), (int(comps["node"][k][1]), 1.0)):
                nm = names[node]
                if ir.is_ground(nm):
                    continue
                e["idx"][j] = index[nm]
                e["sign"][j] = sign
                j += 1
            e["n_idx"] = j
        out.append(e)
    table = np.array(out, dtype=ir.TRAN_SOURCE_DTYPE) if out else np.zeros(0, ir.TRAN_SOURCE_DTYPE)
    return table, np.array(pwl, dtype=np.float64).reshape(-1, 2)


def c_on_g(mna: ir.MnaSystem) -> np.ndarray:
    """C's values scattered onto G's pattern (the compiler zero-fills every C
    position into G, src/compiler.rs:577-579)."""
    g, c = mna.g, mna.c
    out = np.zeros(g.values.shape[0])
    grp = g.row_pointers.astype(np.int64)
    crp = c.row_pointers.astype(np.int64)
    for r in range(mna.size):
        gc = g.col_indices[grp[r]:grp[r + 1]]
        for q in range(crp[r], crp[r + 1]):
            pos = int(np.searchsorted(gc, c.col_indices[q]))
            if pos >= gc.shape[0] or gc[pos] != c.col_indices[q]:
                raise ValueError(f"C entry ({r}, {c.col_indices[q]}) is not in G's pattern")
            out[grp[r] + pos] += c.values[q]
    return out


@dataclass
class TranResult:
    times: np.ndarray
    waves: np.ndarray          # [records, instances, probes]
    x_final: np.ndarray        # [instances, n]
    steps_accepted: int
    steps_rejected: int
    linear_solves: int
    bicgstab_iterations: int
    device_ms: float


class Transient:
    """GPU transient of one circuit or K instances sharing its pattern and
    sources. ``instances``: list of (g_values, c_on_g_values); default is the
    circuit's own values."""

    def __init__(self, backend: HipBackend, mna: ir.MnaSystem, sources, pwl=None,
                 instances=None):
        self.backend = backend
        self.n = mna.size
        rp = np.ascontiguousarray(mna.g.row_pointers, dtype=np.uint32)
        cl = np.ascontiguousarray(mna.g.col_indices, dtype=np.uint32)
        if instances is None:
            instances = [(mna.g.values, c_on_g(mna))]
        self._keep = [rp, cl]
        k = len(instances)
        cs = (N.BatchSystem * k)()
        for i, (g, c) in enumerate(instances):
            g_ = np.ascontiguousarray(g, dtype=np.float64)
            c_ = np.ascontiguousarray(c, dtype=np.float64)
            self._keep += [g_, c_]
            cs[i] = N.BatchSystem(self.n, rp.ctypes.data_as(N.c_u32p), cl.ctypes.data_as(N.c_u32p),
                                  g_.ctypes.data_as(N.c_f64p), c_.ctypes.data_as(N.c_f64p), 0.0, 0.0)
        self.k = k
        src = np.ascontiguousarray(sources, dtype=ir.TRAN_SOURCE_DTYPE)
        pw = np.ascontiguousarray(pwl if pwl is not None else np.zeros((0, 2)), dtype=np.float64)
        self.h = ctypes.c_void_p()
        check(lib().ohm_transient_create(
            backend.h, ctypes.c_int64(k), cs, ctypes.c_int64(src.shape[0]),
            src.ctypes.data_as(ctypes.POINTER(TranSource)), ctypes.c_int64(pw.shape[0]),
            pw.ctypes.data_as(N.c_f64p), ctypes.byref(self.h)))

    def run(self, tstep: float, tstop: float, tstart: float = 0.0, method: str = "be",
            x0=None, probes=None, params: N.BicgstabParams | None = None,
            precond: int = N.PREC_JACOBI, max_records: int | None = None) -> TranResult:
        p = TranParams()
        lib().ohm_tran_default_params(ctypes.byref(p))
        p.tstep, p.tstop, p.tstart = tstep, tstop, tstart
        p.method = METHODS[method]
        p.precond = precond
        p.solver = params if params is not None else _with(bicgstab_params(tolerance=1e-10))
        pr = np.ascontiguousarray(np.arange(self.n) if probes is None else probes, dtype=np.uint32)
        if max_records is None:
            max_records = int(np.ceil(tstop / tstep * (1.0 if method in ("be", "fixed_be") else 4.0))) + 2
        times = np.zeros(max_records)
        waves = np.zeros((max_records, self.k, pr.shape[0]))
        xf = np.zeros((self.k, self.n))
        x0a = None
        if x0 is not None:
            x0a = np.ascontiguousarray(np.broadcast_to(x0, (self.k, self.n)), dtype=np.float64)
        st = TranStats()
        check(lib().ohm_transient_run(
            self.h, x0a.ctypes.data_as(N.c_f64p) if x0a is not None else None, ctypes.byref(p),
            ctypes.c_int64(pr.shape[0]), pr.ctypes.data_as(N.c_u32p), ctypes.c_int64(max_records),
            times.ctypes.data_as(N.c_f64p), waves.ctypes.data_as(N.c_f64p),
            xf.ctypes.data_as(N.c_f64p), ctypes.byref(st)))
        r = st.records
        return TranResult(times[:r], waves[:r], xf, st.steps_accepted, st.steps_rejected,
                          st.linear_solves, st.bicgstab_iterations, st.device_ms)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and not sys.is_finalizing():
            lib().ohm_transient_free(self.h)
            self.h = None


def uic_system(ckt: ir.Circuit, mna: ir.MnaSystem):
    """compute_uic_initial_conditions (src/analysis/transient.rs:374-474):
    inductor branch rows become I_L = 0, a capacitor's node row becomes
    V(i) - V(j) = 0 (first unconstrained end). Returns (G_uic, b_uic)."""
    g = mna.g
    vals = g.values.copy()
    b = mna.b_dc.copy()
    rp = g.row_pointers.astype(np.int64)
    n_nodes = mna.n_nodes
    index = {name: i for i, name in enumerate(mna.node_names)}
    names = ckt.node_names
    comps = ckt.components

    def idx(node):
        nm = names[int(node)]
        return None if ir.is_ground(nm) else index[nm]

    def slot(r, c):
        cols = g.col_indices[rp[r]:rp[r + 1]]
        hit = np.nonzero(cols == c)[0]
        return int(rp[r] + hit[0]) if hit.shape[0] else None

    def zero_row(r):
        vals[rp[r]:rp[r + 1]] = 0.0

    branch = 0
    for k in range(comps.shape[0]):
        kind = int(comps["kind"][k])
        if kind == ir.L:
            bk = n_nodes + branch
            zero_row(bk)
            d = slot(bk, bk)
            if d is not None:
                vals[d] = 1.0
            b[bk] = 0.0
            branch += 1
        elif kind == ir.V:
            branch += 1
    constrained = np.zeros(n_nodes, dtype=bool)
    for k in range(comps.shape[0]):
        if int(comps["kind"][k]) != ir.C:
            continue
        i, j = idx(comps["node"][k][0]), idx(comps["node"][k][1])
        for a, o in ((i, j), (j, i)):
            if a is None or constrained[a]:
                continue
            zero_row(a)
            d = slot(a, a)
            if d is not None:
                vals[d] = 1.0
            if o is not None:
                off = slot(a, o)
                if off is not None:
                    vals[off] = -1.0
            b[a] = 0.0
            constrained[a] = True
            break
    return ir.CsrMatrix(g.nrows, g.ncols, vals, g.col_indices, g.row_pointers), b


@dataclass
class TranAnalysis:
    """TranResult (src/analysis/mod.rs): waveforms per node and branch."""
    times: np.ndarray
    node_voltages: list
    branch_currents: list
    result: TranResult


def has_devices(mna: ir.MnaSystem) -> bool:
    return any(d is not None and len(d) for d in
               (mna.diode_descriptors, mna.bjt_descriptors, mna.mosfet_descriptors))


def run_nonlinear(backend: HipBackend, ckt: ir.Circuit, mna: ir.MnaSystem, tstep: float,
                  tstop: float, tstart: float = 0.0, x0=None, method: str = "adaptive",
                  newton_kw: dict | None = None, probes=None,
                  max_records: int | None = None) -> TranResult:
    """ohm_transient_nonlinear: a Newton solve on the GPU per timestep, with
    the diode / BJT / MOSFET currents in the time loop (BE, or the
    reference's adaptive TRAP/BE step control). ``newton_kw``: newton_params
    keywords for the per-step solves (default abs_tol 1e-9, sub-solves to
    1e-10). ``x0`` None = zero state."""
    from .solver import newton_params
    kw = dict(abs_tol=1e-9, bicgstab_tolerance=1e-10)
    kw.update(newton_kw or {})
    npar, _keep = newton_params(**kw)
    src, pwl = source_table(ckt, mna)
    src = np.ascontiguousarray(src, dtype=ir.TRAN_SOURCE_DTYPE)
    pw = np.ascontiguousarray(pwl if pwl is not None else np.zeros((0, 2)), dtype=np.float64)
    view = ir.MnaView(mna)
    n = mna.size
    pr = np.ascontiguousarray(np.arange(n) if probes is None else probes, dtype=np.uint32)
    p = TranParams()
    lib().ohm_tran_default_params(ctypes.byref(p))
    p.tstep, p.tstop, p.tstart = tstep, tstop, tstart
    p.method = METHODS[method]
    x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    cap = max_records or int(np.ceil(tstop / tstep * (1.0 if method in ("be", "fixed_be") else 4.0))) + 2
    bound = int(np.ceil(tstop / tstep * 20000.0)) + 2
    while True:
        times = np.zeros(cap)
        waves = np.zeros((cap, pr.shape[0]))
        xf = np.zeros(n)
        st = TranStats()
        check(lib().ohm_transient_nonlinear(
            backend.h, view.ref(), ctypes.c_int64(src.shape[0]),
            src.ctypes.data_as(ctypes.POINTER(TranSource)), ctypes.c_int64(pw.shape[0]),
            pw.ctypes.data_as(N.c_f64p), ctypes.byref(p), ctypes.byref(npar),
            x0a.ctypes.data_as(N.c_f64p) if x0a is not None else None,
            ctypes.c_int64(pr.shape[0]), pr.ctypes.data_as(N.c_u32p), ctypes.c_int64(cap),
            times.ctypes.data_as(N.c_f64p), waves.ctypes.data_as(N.c_f64p),
            xf.ctypes.data_as(N.c_f64p), ctypes.byref(st)))
        if max_records or st.records < cap or cap >= bound:
            break
        cap = min(cap * 4, bound)
    r = st.records
    return TranResult(times[:r], waves[:r, None, :], xf[None, :], st.steps_accepted,
                      st.steps_rejected, st.linear_solves, st.bicgstab_iterations, st.device_ms)


def run(backend: HipBackend, ckt: ir.Circuit, mna: ir.MnaSystem, tstep: float, tstop: float,
        tstart: float = 0.0, uic: bool = False, method: str = "adaptive",
        params: N.BicgstabParams | None = None, precond: int = N.PREC_JACOBI,
        dc_kw: dict | None = None, devices: bool = True) -> TranAnalysis:
    """transient::run (src/analysis/transient.rs:25-199): initial state from
    the DC operating point (dc::run, so nonlinear devices are solved by the
    Newton cascade) or the UIC system, both on the GPU, then GPU time
    stepping of every unknown. ``dc_kw``: dc.run's Newton settings for that
    point (and for the per-step Newton solves of a circuit with devices).
    The reference's time stepping uses G and C only, so diode, BJT and
    MOSFET currents enter only through the operating point; here a circuit
    with devices is stepped with them (run_nonlinear) unless ``devices`` is
    False, which reproduces the reference's linear stepping."""
    if uic:
        from .solver import GpuSolver
        s = GpuSolver(backend, precond=N.PREC_AUTO,
                      params=bicgstab_params(tolerance=1e-12, max_iterations=100000))
        a, b = uic_system(ckt, mna)
        x0 = s.solve_real(a, b)
    else:
        from . import dc
        x0 = dc.run(backend, mna, **(dc_kw or {})).x
    if devices and has_devices(mna):
        res = run_nonlinear(backend, ckt, mna, tstep, tstop, tstart, x0=x0, method=method,
                            newton_kw=dc_kw)
        w = res.waves[:, 0, :]
        nn = mna.n_nodes
        nodes = [(name, w[:, i].copy()) for i, name in enumerate(mna.node_names)]
        branches = [(name, w[:, nn + i].copy()) for i, name in enumerate(mna.branch_names)]
        return TranAnalysis(res.times, nodes, branches, res)
    src, pwl = source_table(ckt, mna)
    tr = Transient(backend, mna, src, pwl)
    # Every accepted step is recorded. Start from room for 4 records per
    # tstep and rerun with more if the stepper filled it (steps never fall
    # below h_min / 2 = tstep / 20000, which bounds the count).
    cap = int(np.ceil(tstop / tstep * (1.0 if method in ("be", "fixed_be") else 4.0))) + 2
    bound = int(np.ceil(tstop / tstep * 20000.0)) + 2
    while True:
        res = tr.run(tstep, tstop, tstart, method=method, x0=x0, params=params, precond=precond,
                     max_records=cap)
        if res.times.shape[0] < cap or cap >= bound:
            break
        cap = min(cap * 4, bound)
    w = res.waves[:, 0, :]
    nn = mna.n_nodes
    nodes = [(name, w[:, i].copy()) for i, name in enumerate(mna.node_names)]
    branches = [(name, w[:, nn + i].copy()) for i, name in enumerate(mna.branch_names)]
    return TranAnalysis(res.times, nodes, branches, res)
