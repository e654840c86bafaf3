"""Full-size parity against the committed golden fixtures (tests/golden/,
written by tests/golden/make_golden.py from FP64 direct solves): every
workload at the size it is benchmarked at, 1e-6 relative / 1e-9 absolute
(gradients 1e-5 relative)."""
import os

import numpy as np
import pytest

from ohmnivore_amd import _native as N
from ohmnivore_amd import adjoint as A
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver
from ohmnivore_amd import transient as T

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def close(x, ref, rel=1e-6, abs_=1e-9):
    err = np.abs(x - ref)
    return bool((err <= np.maximum(rel * np.abs(ref), abs_)).all()), float(err.max(initial=0.0))


@pytest.fixture(scope="module")
def be():
    return solver.HipBackend()


def test_ladder_10k_golden(be):
    mna = solver.compile(ir.resistor_ladder(10_000, seed=1))
    s = solver.GpuSolver(be, precond=N.PREC_JACOBI,
                         params=solver.bicgstab_params(tolerance=1e-13, max_iterations=200_000))
    ok, err = close(s.solve_real(mna.g, mna.b_dc), load("ladder_10k.npz")["x"])
    assert ok, err


@pytest.mark.parametrize("precision,forcing,floor", [("fp64", "fixed", 1e-15),
                                                     ("fp32_ir", "fixed", 1e-15),
                                                     ("fp64", "ew1", 1e-18),
                                                     ("fp64", "ew", 1e-18)])
def test_diode_mesh_256_golden(be, precision, forcing, floor):
    """The headline workload at full size against the LU fixture, in the
    fixed-tolerance forms and in the benchmark's form (Eisenstat-Walker
    forcing with the safeguard at every iteration, sub-solve floor
    1e-18 ||b||)."""
    mna = solver.compile(ir.diode_mesh(256, 256, seed=2))
    x, st = solver.newton_solve(be, mna.g, mna.b_dc, mna.diode_descriptors, precision=precision,
                                abs_tol=1e-10, max_iterations=100, bicgstab_tolerance=1e-6,
                                bicgstab_max_iterations=20_000, forcing=forcing,
                                forcing_safeguard="always", bicgstab_floor=floor)
    ok, err = close(x[::4], load("diode_mesh_256.npz")["x_every4"])
    assert ok, (err, st)


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_rc_grid_512_golden(be, precision):
    """The transient workload at full length (1,000 backward-Euler steps of
    1 ps, warm-started ISAI(1)-BiCGSTAB at the benchmark's 1e-9 relative
    residual) against the SuperLU fixture: every step of the first 50 and
    every 50th step through step 1,000."""
    g = load("rc_grid_512.npz")
    ckt = ir.rc_grid(512, 512, 1024, seed=3)
    mna = solver.compile(ckt)
    dc = solver.GpuSolver(be, precond=N.PREC_AUTO,
                          params=solver.bicgstab_params(tolerance=1e-13, max_iterations=100_000))
    x0 = dc.solve_real(mna.g, mna.b_dc)
    src, _ = T.source_table(ckt, mna)
    every = int(g["every"])
    steps = (g["waves_every"].shape[0] - 1) * every
    assert steps == 1000
    res = T.Transient(be, mna, src).run(
        float(g["h"]), steps * float(g["h"]), method="be", x0=x0, probes=g["probes"],
        params=solver.bicgstab_params(tolerance=1e-9, max_iterations=20_000, precision=precision),
        precond=N.PREC_ISAI)
    assert res.steps_accepted == steps
    w = res.waves[:, 0, :]
    dense = g["waves"].shape[0]
    ok, err = close(w[:dense], g["waves"])
    assert ok, err
    ok, err = close(w[::every], g["waves_every"])
    assert ok, err


def test_amplifier_4096_golden(be):
    g = load("amplifier_4096.npz")
    ckt = ir.amplifier_array(4096, seed=5)
    mna = solver.compile(ckt)
    prob = A.DcProblem(be, ckt, mna, newton_kw=dict(abs_tol=1e-13, bicgstab_tolerance=1e-12))
    x = prob.solve()
    outs = g["outputs"]
    ok, err = close(x[outs], g["v_out"])
    assert ok, err
    dj = np.zeros(mna.size)
    dj[outs] = 2 * (x[outs] - 0.9)
    gg, gk = prob.adjoint(dj, tolerance=1e-13)
    ok, err = close(gg[g["res_index"]], g["grad_g"], rel=1e-5, abs_=1e-12 * np.abs(g["grad_g"]).max())
    assert ok, err
    ok, err = close(gk[g["mos_index"]], g["grad_kp"], rel=1e-5, abs_=1e-12 * np.abs(g["grad_kp"]).max())
    assert ok, err
