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


def run(backend: HipBackend, mna: ir.MnaSystem, **newton_kw) -> DcResult:
    """dc::run (src/analysis/dc.rs:25-52). ``newton_kw``: solver.newton_params
    keywords applied to every Newton run (max_iterations caps the direct
    attempt; the cascade sets the other steps' caps as the reference does)."""
    view = ir.MnaView(mna)
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
