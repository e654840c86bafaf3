"""Adjoint DC sensitivity (no counterpart in the reference).

At a DC operating point x*(p) with F(x*, p) = 0 (the MNA residual, whose
Jacobian is the assembled Newton matrix J), the gradient of an objective
Obj(x*) is  dObj/dp = -lambda^T dF/dp  with  J^T lambda = dObj/dx.
The GPU does the Newton solve, the transposed BiCGSTAB (transposed ISAI or
Jacobi preconditioner) and the lambda^T dF/dp scatter; ``DcOperatingPoint``
wraps them as a torch.autograd.Function so the operating point can sit inside
any torch computation. Parameters: resistor conductances and per-MOSFET kp
(= KP W/L).
"""
from __future__ import annotations

import numpy as np
import torch

from . import circuit as ir
from .solver import HipBackend, NewtonContext, bicgstab_params

SENT = ir.SENTINEL


class DcProblem:
    """A nonlinear circuit whose resistor conductances and MOSFET kp values
    can be changed without recompiling: G(g) = G0 + sum_r (g_r - g0_r) S_r,
    with S_r the resistor's +1/-1 stamp on its (a,a), (b,b), (a,b), (b,a)
    slots (src/compiler.rs:308-329)."""

    def __init__(self, backend: HipBackend, ckt: ir.Circuit, mna: ir.MnaSystem,
                 newton_kw: dict | None = None):
        self.backend = backend
        self.ckt, self.mna = ckt, mna
        self.n = mna.size
        comps = ckt.components
        index = {name: i for i, name in enumerate(mna.node_names)}
        names = ckt.node_names
        rsel = np.nonzero(comps["kind"] == ir.R)[0]
        self.res_components = rsel
        nodes = np.full((rsel.shape[0], 2), SENT, dtype=np.uint32)
        for j, k in enumerate(rsel):
            for t in range(2):
                nm = names[int(comps["node"][k][t])]
                if not ir.is_ground(nm):
                    nodes[j, t] = index[nm]
        self.res_nodes = nodes
        self.g0 = 1.0 / comps["value"][rsel].astype(np.float64)
        g = mna.g
        rp = g.row_pointers.astype(np.int64)

        def slot(r, c):
            if r == SENT or c == SENT:
                return -1
            cols = g.col_indices[rp[r]:rp[r + 1]]
            pos = int(np.searchsorted(cols, c))
            return int(rp[r] + pos) if pos < cols.shape[0] and cols[pos] == c else -1

        self.slots = np.array([[slot(a, a), slot(b, b), slot(a, b), slot(b, a)]
                               for a, b in nodes], dtype=np.int64).reshape(-1, 4)
        self.mosfets = mna.mosfet_descriptors.copy()
        self.kp0 = self.mosfets["kp"].astype(np.float64)
        self.newton_kw = dict(max_iterations=100, abs_tol=1e-12, bicgstab_tolerance=1e-10,
                              bicgstab_max_iterations=50000)
        if newton_kw:
            self.newton_kw.update(newton_kw)
        self.ctx: NewtonContext | None = None
        self.last_stats = None

    def g_values(self, g_res: np.ndarray) -> np.ndarray:
        vals = self.mna.g.values.copy()
        dg = np.asarray(g_res, dtype=np.float64) - self.g0
        for col, sign in ((0, 1.0), (1, 1.0), (2, -1.0), (3, -1.0)):
            s = self.slots[:, col]
            ok = s >= 0
            np.add.at(vals, s[ok], sign * dg[ok])
        return vals

    def solve(self, g_res=None, kp=None) -> np.ndarray:
        """Operating point for the given conductances and kp values. The
        first call builds the device context; later calls update its values
        and parameters in place (same topology, so the preconditioner plan is
        reused) and run Newton again from x = 0."""
        g = self.g_values(self.g0 if g_res is None else g_res)
        m = self.mosfets.copy()
        if kp is not None:
            m["kp"] = np.asarray(kp, dtype=np.float64).astype(np.float32)
        if self.ctx is None:
            gm = ir.CsrMatrix(self.n, self.n, g, self.mna.g.col_indices, self.mna.g.row_pointers)
            self.ctx = NewtonContext(self.backend, gm, self.mna.b_dc, self.mna.diode_descriptors,
                                     self.mna.bjt_descriptors, m)
        else:
            self.ctx.update(values=g, mosfets=m if len(m) else None)
        self.last_stats = self.ctx.run(**self.newton_kw)
        return self.ctx.x()

    def adjoint(self, dJdx, tolerance: float = 1e-12):
        """(dObj/dg per resistor, dObj/dkp per MOSFET) at the last operating point."""
        prm = bicgstab_params(tolerance=tolerance, max_iterations=100000)
        gg, gk, lam, st = self.ctx.adjoint(dJdx, self.res_nodes, prm)
        self.last_adjoint = st
        return gg, gk


class DcOperatingPoint(torch.autograd.Function):
    """x* = DcOperatingPoint.apply(g_res, kp, problem): float64 tensors of
    resistor conductances and MOSFET kp values in, node voltages and branch
    currents out; backward runs the GPU adjoint."""

    @staticmethod
    def forward(ctx, g_res, kp, problem: DcProblem):
        x = problem.solve(g_res.detach().cpu().numpy(), kp.detach().cpu().numpy())
        ctx.problem = problem
        return torch.from_numpy(x).to(g_res.device)

    @staticmethod
    def backward(ctx, grad_x):
        gg, gk = ctx.problem.adjoint(grad_x.detach().cpu().numpy().astype(np.float64))
        return (torch.from_numpy(gg).to(grad_x.device), torch.from_numpy(gk).to(grad_x.device),
                None)


def output_nodes(ckt: ir.Circuit, mna: ir.MnaSystem, prefix: str = "d") -> np.ndarray:
    """Matrix indices of the amplifier outputs (drain / collector nodes)."""
    return np.array([i for i, nm in enumerate(mna.node_names)
                     if nm.startswith(prefix) and nm[len(prefix):].isdigit()], dtype=np.int64)


def objective(x: torch.Tensor, outs, target: float = 0.9) -> torch.Tensor:
    """J = sum (V_out - target)^2 over the outputs."""
    return ((x[outs] - target) ** 2).sum()
