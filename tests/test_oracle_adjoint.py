"""The adjoint sensitivity pinned on the CPU: lambda from the oracle's LU on
J^T, the lambda^T dF/dp scatter for resistor conductances and MOSFET kp,
against central finite differences of the oracle's LU-Newton operating point
(amplifier-array circuit family, 8 stages). Agreement 1e-5 relative, plus
the differences' own noise floor 1e-13 / h (the operating point is resolved
to ~1e-13 V; a stage in cut-off has a true gradient ~1e-13)."""
import numpy as np
import pytest

import oracle
from ohmnivore_amd import adjoint as A
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver

SENT = ir.SENTINEL


def op_point(mna, g_vals, mos):
    g = ir.CsrMatrix(mna.size, mna.size, g_vals, mna.g.col_indices, mna.g.row_pointers)
    x, it, md = oracle.newton_solve(g, mna.b_dc, mna.diode_descriptors, mna.bjt_descriptors, mos,
                                    max_iterations=200, abs_tol=1e-14)
    return x


def transpose(a):
    n = a.nrows
    rows = np.repeat(np.arange(n), np.diff(a.row_pointers.astype(np.int64)))
    return oracle.from_triplet_arrays(n, n, a.col_indices.astype(np.int64), rows, a.values)


def xv(v, i):
    return 0.0 if i == SENT else v[i]


@pytest.fixture(scope="module")
def problem():
    ckt = ir.amplifier_array(8, seed=5)
    mna = solver.compile(ckt)
    prob = A.DcProblem.__new__(A.DcProblem)  # host bookkeeping only (no GPU)
    A.DcProblem.__init__(prob, None, ckt, mna)
    return ckt, mna, prob


def objective(x, outs):
    return float(((x[outs] - 0.9) ** 2).sum())


def test_adjoint_formula_matches_finite_differences(problem):
    ckt, mna, prob = problem
    outs = A.output_nodes(ckt, mna)
    assert outs.shape[0] == 8
    mos = prob.mosfets
    x = op_point(mna, mna.g.values, mos)
    jv, _ = oracle.newton_assemble(mna.g, mna.b_dc, x, mna.diode_descriptors,
                                   mna.bjt_descriptors, mos)
    jac = ir.CsrMatrix(mna.size, mna.size, jv, mna.g.col_indices, mna.g.row_pointers)
    dj = np.zeros(mna.size)
    dj[outs] = 2 * (x[outs] - 0.9)
    lam = oracle.sparse_lu_solve(transpose(jac), dj)
    # scatter (the kernels' formulas)
    grad_g = np.array([-(xv(lam, a) - xv(lam, b)) * (xv(x, a) - xv(x, b))
                       for a, b in prob.res_nodes])
    grad_kp = []
    for d in mos:
        vd, vg, vs = xv(x, d["drain_idx"]), xv(x, d["gate_idx"]), xv(x, d["source_idx"])
        iu = oracle.mosfet_core(vg - vs, vd - vs, float(d["vto"]), 1.0, float(d["lambda"]),
                                float(d["polarity"]))[0]
        grad_kp.append(-(xv(lam, d["b_idx"][0]) - xv(lam, d["b_idx"][1])) * iu)
    grad_kp = np.array(grad_kp)
    # central differences: conductances (relative step 1e-4) ...
    for r in range(prob.g0.shape[0]):
        h = 1e-4 * prob.g0[r]
        gp, gm = prob.g0.copy(), prob.g0.copy()
        gp[r] += h
        gm[r] -= h
        fd = (objective(op_point(mna, prob.g_values(gp), mos), outs)
              - objective(op_point(mna, prob.g_values(gm), mos), outs)) / (2 * h)
        assert abs(grad_g[r] - fd) <= 1e-5 * abs(fd) + 1e-13 / h, (r, grad_g[r], fd)
    # ... and kp, stepped by an exact number of f32 ulps (descriptors hold f32)
    for m in range(len(mos)):
        kp = np.float32(mos["kp"][m])
        ulp = np.spacing(kp)
        step = np.float32(2 ** 10) * ulp
        mp, mm = mos.copy(), mos.copy()
        mp["kp"][m] = kp + step
        mm["kp"][m] = kp - step
        h = float(np.float64(kp + step) - np.float64(kp))
        fd = (objective(op_point(mna, mna.g.values, mp), outs)
              - objective(op_point(mna, mna.g.values, mm), outs)) / (2 * h)
        assert abs(grad_kp[m] - fd) <= 1e-5 * abs(fd) + 1e-13 / h, (m, grad_kp[m], fd)
