"""DC operating point (dc::run, src/analysis/dc.rs:25-650) on the GPU.

Linear circuits are one solve_real. Nonlinear circuits follow the
reference's continuation cascade (run_nonlinear, :60-293): a direct Newton
attempt, source stepping when MOSFETs are present (:480-640), adaptive GMIN
stepping (:152-260, gmin_newton_step :297-332) and pseudo-transient
continuation (:340-470) when the final true-GMIN solve fails. Every Newton
solve runs on the GPU. Where the reference retries a failed GPU step on the
CPU with a sparse LU, this module retries it on the GPU in FP64 with a
larger BiCGSTAB budget: there is no CPU path in the product.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import circuit as ir
from .solver import GpuSolver, HipBackend, NewtonContext, bicgstab_params


@dataclass
class DcResult:
    """DcResult (src/analysis/mod.rs): node voltages and branch currents."""
    node_voltages: list
    branch_currents: list
    x: np.ndarray
    path: list = field(default_factory=list)  # continuation steps taken


def node_diagonal_positions(mna: ir.MnaSystem) -> np.ndarray:
    """find_node_diagonal_positions (src/analysis/dc.rs:340-350)."""
    g = mna.g
    out = []
    for row in range(mna.n_nodes):
        cols = g.col_indices[g.row_pointers[row]:g.row_pointers[row + 1]]
        hit = np.nonzero(cols == row)[0]
        if hit.shape[0]:
            out.append(int(g.row_pointers[row]) + int(hit[0]))
    return np.array(out, dtype=np.int64)


class _Stepper:
    def __init__(self, backend: HipBackend, mna: ir.MnaSystem, newton_kw: dict):
        self.be = backend
        self.mna = mna
        self.diag = node_diagonal_positions(mna)
        self.kw = newton_kw
        self.path = []

    def step(self, b, extra_gmin: float, initial_guess=None, max_iterations=50,
             bicgstab_max_iterations=10_000, label=""):
        """gmin_newton_step (src/analysis/dc.rs:297-332) on the GPU; the
        reference's CPU sparse-LU retry becomes a GPU retry with a larger
        BiCGSTAB budget."""
        vals = self.mna.g.values.copy()
        vals[self.diag] += extra_gmin
        g = ir.CsrMatrix(self.mna.size, self.mna.size, vals, self.mna.g.col_indices,
                         self.mna.g.row_pointers)
        ctx = NewtonContext(self.be, g, b, self.mna.diode_descriptors, self.mna.bjt_descriptors,
                            self.mna.mosfet_descriptors)
        last = None
        for budget in (bicgstab_max_iterations, max(10 * bicgstab_max_iterations, 20_000)):
            kw = dict(self.kw)
            kw.update(max_iterations=max_iterations, bicgstab_max_iterations=budget,
                      initial_guess=initial_guess)
            try:
                st = ctx.run(**kw)
                self.path.append((label, extra_gmin, True, st.newton_iterations))
                return ctx.x()
            except N.OhmnivoreError as e:
                last = e
        self.path.append((label, extra_gmin, False, 0))
        raise last


def run_nonlinear(backend: HipBackend, mna: ir.MnaSystem, **newton_kw) -> tuple[np.ndarray, list]:
    """run_nonlinear (src/analysis/dc.rs:60-293)."""
    s = _Stepper(backend, mna, newton_kw)
    b = mna.b_dc
    try:
        return s.step(b, 0.0, max_iterations=newton_kw.get("max_iterations", 50),
                      bicgstab_max_iterations=200, label="direct"), s.path
    except N.OhmnivoreError:
        pass
    if mna.mosfet_descriptors is not None and len(mna.mosfet_descriptors):
        try:
            return _source_stepping(s), s.path
        except N.OhmnivoreError:
            pass
    # adaptive GMIN stepping: 1000x jumps, 10x subdivisions on failure
    x_prev = None
    last_good, target, subdivisions = 1.0, 1e-1, 0
    while target > 1e-13:
        try:
            x_prev = s.step(b, target, x_prev, max_iterations=15, bicgstab_max_iterations=300,
                            label="gmin")
            last_good, subdivisions = target, 0
            target *= 1e-3
        except N.OhmnivoreError:
            subdivisions += 1
            if subdivisions > 13:
                raise
            target = last_good * 10.0 ** (-subdivisions)
    try:
        return s.step(b, 0.0, x_prev, label="final"), s.path
    except N.OhmnivoreError:
        if x_prev is None:
            raise
        return _pseudo_transient(s, x_prev), s.path


def _source_stepping(s: _Stepper) -> np.ndarray:
    """run_source_stepping (src/analysis/dc.rs:480-640)."""
    x_prev = None
    last_good, target, subdivisions = 0.0, 0.1, 0
    while target <= 1.0:
        try:
            x = s.step(target * s.mna.b_dc, 0.0, x_prev, max_iterations=30,
                       bicgstab_max_iterations=200, label="source")
            x_prev = x
            prev_step = target - last_good
            last_good, subdivisions = target, 0
            if abs(target - 1.0) < 1e-12:
                return x
            target = min(target + prev_step * 2.0, 1.0)
        except N.OhmnivoreError:
            subdivisions += 1
            if subdivisions > 20:
                raise
            target = last_good + (target - last_good) / 2.0
            if target - last_good < 1e-6:
                raise
    raise N.OhmnivoreError(3, "Solve error: Source stepping completed without reaching alpha=1.0")


def _pseudo_transient(s: _Stepper, seed: np.ndarray) -> np.ndarray:
    """run_pseudo_transient_continuation (src/analysis/dc.rs:352-470):
    alpha I on node rows anchored at b + alpha x_prev, alpha -> 0."""
    n_nodes = s.mna.n_nodes
    x_prev = seed
    last_good, target, subdivisions = 1.0, 1e-1, 0
    while target > 1e-8:
        b = s.mna.b_dc.copy()
        b[:n_nodes] += target * x_prev[:n_nodes]
        try:
            x_prev = s.step(b, target, x_prev, max_iterations=20, bicgstab_max_iterations=500,
                            label="ptc")
            last_good, subdivisions = target, 0
            target *= 1e-2
        except N.OhmnivoreError:
            subdivisions += 1
            if subdivisions > 7:
                raise
            target = last_good * 10.0 ** (-subdivisions)
    return s.step(s.mna.b_dc, 0.0, x_prev, max_iterations=25, bicgstab_max_iterations=1500,
                  label="ptc-final")


def run(backend: HipBackend, mna: ir.MnaSystem, **newton_kw) -> DcResult:
    """dc::run (src/analysis/dc.rs:25-52)."""
    nonlinear = any(d is not None and len(d) for d in
                    (mna.diode_descriptors, mna.bjt_descriptors, mna.mosfet_descriptors))
    path = []
    if nonlinear:
        x, path = run_nonlinear(backend, mna, **newton_kw)
    else:
        s = GpuSolver(backend, precond=N.PREC_AUTO,
                      params=bicgstab_params(tolerance=newton_kw.get("tolerance", 1e-12),
                                             max_iterations=200_000))
        x = s.solve_real(mna.g, mna.b_dc)
    nn = mna.n_nodes
    nodes = [(name, float(x[i])) for i, name in enumerate(mna.node_names)]
    branches = [(name, float(x[nn + i])) for i, name in enumerate(mna.branch_names)]
    return DcResult(nodes, branches, x, path)
