"""Writes the full-size golden fixtures of tests/golden/ (run from the
repository root: python tests/golden/make_golden.py).

Every reference value comes from a direct FP64 solve on the CPU: the
oracle's sparse LU where its banded factorisation is cheap, SciPy's SuperLU
(splu) where the band is wide (the 512x512 grid, the amplifier array whose
VDD row couples every stage). Device models, assembly and limiting come from
the oracle, so the nonlinear references follow the same Newton algorithm
(newton_solve, src/solver/newton.rs:153-488) with a direct linear step.

Fixtures (numpy .npz, FP64):
  ladder_10k.npz        x of the 10,002-unknown resistor ladder (seed 1)
  diode_mesh_256.npz    every 4th unknown of the 256x256 diode-mesh operating point
  rc_grid_512.npz       64 probe voltages over the first 50 backward-Euler
                        steps (h = 1 ps) of the 512x512 RC grid from its DC point
                        (waves), and at every 50th step through step 1,000
                        (waves_every)
  amplifier_4096.npz    the 4,096 stage outputs of the amplifier array and the
                        adjoint gradient for 64 seeded parameters
"""
from __future__ import annotations

import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ohmnivore_amd import adjoint as A  # noqa: E402
from ohmnivore_amd import circuit as ir  # noqa: E402
from ohmnivore_amd import solver  # noqa: E402
from ohmnivore_amd import transient as T  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
SENT = ir.SENTINEL


def csr(m, vals=None):
    return sp.csr_matrix((m.values if vals is None else vals, m.col_indices, m.row_pointers),
                         shape=(m.nrows, m.ncols))


def ladder():
    mna = solver.compile(ir.resistor_ladder(10_000, seed=1))
    x = oracle.sparse_lu_solve(mna.g, mna.b_dc)
    np.savez_compressed(os.path.join(OUT, "ladder_10k.npz"), x=x)


def diode_mesh():
    mna = solver.compile(ir.diode_mesh(256, 256, seed=2))
    x, its, md = oracle.newton_solve(mna.g, mna.b_dc, mna.diode_descriptors, abs_tol=1e-12,
                                     max_iterations=100)
    np.savez_compressed(os.path.join(OUT, "diode_mesh_256.npz"), x_every4=x[::4],
                        newton_iterations=its)


def rc_grid(steps=1000, h=1e-12, dense_steps=50, every=50):
    ckt = ir.rc_grid(512, 512, 1024, seed=3)
    mna = solver.compile(ckt)
    g = csr(mna.g).tocsc()
    x = sla.splu(g).solve(mna.b_dc)
    src, _ = T.source_table(ckt, mna)
    cg = T.c_on_g(mna)
    cm = csr(mna.g, cg)
    lu = sla.splu((csr(mna.g) + cm / h).tocsc())
    probes = np.arange(0, mna.size, 4099)
    waves = [x[probes]]
    sparse_waves = [x[probes]]
    t = 0.0
    for k in range(1, steps + 1):
        t_next = t + h  # t accumulates as in transient.rs:84-178
        b = np.zeros(mna.size)
        for e in src:  # evaluate_b_at_time (src/analysis/transient.rs:265-280)
            v = oracle.source_value(e, t_next)
            for j in range(int(e["n_idx"])):
                b[int(e["idx"][j])] += float(e["sign"][j]) * v
        x = lu.solve(b + (1.0 / h) * (cm @ x))
        if k <= dense_steps:
            waves.append(x[probes])
        if k % every == 0:
            sparse_waves.append(x[probes])
        t = t_next
    np.savez_compressed(os.path.join(OUT, "rc_grid_512.npz"), probes=probes,
                        waves=np.array(waves), h=h, waves_every=np.array(sparse_waves),
                        every=every)


def amplifier():
    ckt = ir.amplifier_array(4096, seed=5)
    mna = solver.compile(ckt)
    prob = A.DcProblem.__new__(A.DcProblem)
    A.DcProblem.__init__(prob, None, ckt, mna)
    mos = prob.mosfets
    x = np.zeros(mna.size)
    for it in range(200):  # Newton with the oracle's assembly / limiting, SuperLU solves
        v, b = oracle.newton_assemble(mna.g, mna.b_dc, x, mna.diode_descriptors,
                                      mna.bjt_descriptors, mos)
        xn = sla.spsolve(csr(mna.g, v).tocsc(), b)
        xn = oracle.newton_limit(mna.size, x, xn, mna.diode_descriptors, mna.bjt_descriptors, mos)
        diff = np.abs(xn - x).max()
        x = xn
        if diff < 1e-13:
            break
    outs = A.output_nodes(ckt, mna)
    v, _ = oracle.newton_assemble(mna.g, mna.b_dc, x, mna.diode_descriptors,
                                  mna.bjt_descriptors, mos)
    dj = np.zeros(mna.size)
    dj[outs] = 2 * (x[outs] - 0.9)
    lam = sla.spsolve(csr(mna.g, v).T.tocsc(), dj)
    rng = np.random.default_rng(64)
    rsel = rng.choice(prob.res_nodes.shape[0], size=48, replace=False)
    msel = rng.choice(len(mos), size=16, replace=False)

    def xv(a, i):
        return 0.0 if i == SENT else a[i]

    g_grad = np.array([-(xv(lam, a) - xv(lam, b)) * (xv(x, a) - xv(x, b))
                       for a, b in prob.res_nodes[rsel]])
    k_grad = []
    for d in mos[msel]:
        vd, vg, vs = xv(x, d["drain_idx"]), xv(x, d["gate_idx"]), xv(x, d["source_idx"])
        iu = oracle.mosfet_core(vg - vs, vd - vs, float(d["vto"]), 1.0, float(d["lambda"]),
                                float(d["polarity"]))[0]
        k_grad.append(-(xv(lam, d["b_idx"][0]) - xv(lam, d["b_idx"][1])) * iu)
    np.savez_compressed(os.path.join(OUT, "amplifier_4096.npz"), outputs=outs, v_out=x[outs],
                        newton_iterations=it + 1, res_index=rsel, grad_g=g_grad,
                        mos_index=msel, grad_kp=np.array(k_grad))


if __name__ == "__main__":
    which = sys.argv[1:] or ["ladder", "diode_mesh", "rc_grid", "amplifier"]
    for w in which:
        print("writing", w, flush=True)
        globals()[w]()
