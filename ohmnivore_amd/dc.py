"""DC operating point (dc::run, src/analysis/dc.rs:25-650) on the GPU.

One native call, ``ohm_dc_run``: linear circuits are one solve_real;
nonlinear circuits run the reference's continuation cascade (run_nonlinear,
:60-293) on one device-resident Newton context — a direct Newton attempt,
source stepping when MOSFETs are present (:480-640), adaptive GMIN stepping
(:152-260, gmin_newton_step :297-332) and pseudo-transient continuation
(:340-470) when the final true-GMIN solve fails. Where the reference retries
a failed step on the CPU with a sparse LU, the step is retried on the GPU in
FP64 with a larger BiCGSTAB budget: there is no CPU path in the product.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import circuit as ir
from ._native import check, lib
from .solver import HipBackend, newton_params

TRACE_CAP = 4096


@dataclass
class DcResult:
    """DcResult (src/analysis/mod.rs): node voltages and branch currents."""
    node_voltages: list
    branch_currents: list
    x: np.ndarray
    # Newton runs of the cascade: (stage, level, converged, iterations, retry)
    path: list = field(default_factory=list)
    stage: str = "linear"          # the stage whose run produced x
    newton_iterations: int = 0
    bicgstab_iterations: int = 0
    device_ms: float = 0.0


class _MnaView:
    """An ``ohm_mna`` pointing into an MnaSystem's arrays (kept alive here)."""

    def __init__(self, mna: ir.MnaSystem):
        g = mna.g
        self.keep = [np.ascontiguousarray(g.row_pointers, np.uint32),
                     np.ascontiguousarray(g.col_indices, np.uint32),
                     np.ascontiguousarray(g.values, np.float64),
                     np.ascontiguousarray(mna.b_dc, np.float64)]
        descs = []
        for arr, dt in ((mna.diode_descriptors, ir.DIODE_DESC_DTYPE),
                        (mna.bjt_descriptors, ir.BJT_DESC_DTYPE),
                        (mna.mosfet_descriptors, ir.MOSFET_DESC_DTYPE)):
            a = np.ascontiguousarray(arr if arr is not None else np.zeros(0, dt), dtype=dt)
            descs.append(a)
        self.keep += descs
        m = ir.MnaC()
        m.size, m.n_nodes = mna.size, mna.n_nodes
        m.n_branches = mna.size - mna.n_nodes
        m.g_nnz = self.keep[2].shape[0]
        m.g_rowptr = self.keep[0].ctypes.data_as(N.c_u32p)
        m.g_col = self.keep[1].ctypes.data_as(N.c_u32p)
        m.g_val = self.keep[2].ctypes.data_as(N.c_f64p)
        m.b_dc = self.keep[3].ctypes.data_as(N.c_f64p)
        m.n_diodes, m.n_bjts, m.n_mosfets = (len(d) for d in descs)
        m.diodes, m.bjts, m.mosfets = (d.ctypes.data if len(d) else None for d in descs)
        self.struct = m


def run(backend: HipBackend, mna: ir.MnaSystem, **newton_kw) -> DcResult:
    """dc::run (src/analysis/dc.rs:25-52). ``newton_kw``: solver.newton_params
    keywords applied to every Newton run (max_iterations caps the direct
    attempt; the cascade sets the other steps' caps as the reference does)."""
    view = _MnaView(mna)
    p, _keep = newton_params(**newton_kw)
    x = np.zeros(max(mna.size, 1))
    st = N.DcStats()
    trace = (N.DcStep * TRACE_CAP)()
    check(lib().ohm_dc_run(backend.h, ctypes.byref(view.struct), ctypes.byref(p),
                           x.ctypes.data_as(N.c_f64p), ctypes.byref(st), trace,
                           ctypes.c_int64(TRACE_CAP)))
    x = x[:mna.size]
    path = [(N.DC_STAGES[t.stage], t.level, bool(t.converged), int(t.iterations), bool(t.retry))
            for t in trace[:min(st.runs, TRACE_CAP)]]
    nn = mna.n_nodes
    nodes = [(name, float(x[i])) for i, name in enumerate(mna.node_names)]
    branches = [(name, float(x[nn + i])) for i, name in enumerate(mna.branch_names)]
    return DcResult(nodes, branches, x, path, N.DC_STAGES[st.stage], int(st.newton_iterations),
                    int(st.bicgstab_iterations), float(st.device_ms))


def run_nonlinear(backend: HipBackend, mna: ir.MnaSystem, **newton_kw) -> tuple[np.ndarray, list]:
    """run_nonlinear (src/analysis/dc.rs:60-293): (x, path)."""
    r = run(backend, mna, **newton_kw)
    return r.x, r.path
