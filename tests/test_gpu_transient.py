"""GPU transient (transient::run, src/analysis/transient.rs:25-199) against
the oracle's LU-based stepping: fixed-step backward Euler and the
reference's adaptive TRAP/BE, FP64 and FP32 + FP64 refinement, one circuit
and Monte Carlo instances. Tolerance 1e-6 relative / 1e-9 absolute per
recorded sample."""
import math

import numpy as np
import pytest

import oracle
from ohmnivore_amd import _native as N
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver
from ohmnivore_amd import transient as T

pytestmark = pytest.mark.gpu

REL, ABS = 1e-6, 1e-9


@pytest.fixture(scope="module")
def be():
    return solver.HipBackend()


def close(x, ref, rel=REL, abs_=ABS):
    err = np.abs(x - ref)
    return bool((err <= np.maximum(rel * np.abs(ref), abs_)).all()), float(err.max(initial=0.0))


def rc_circuit():
    ckt = ir.Circuit()
    ckt.vsource("V1", "1", "0", dc=5.0)
    ckt.resistor("R1", "1", "2", 1e3)
    ckt.capacitor("C1", "2", "0", 1e-6)
    return ckt


def rl_circuit():
    ckt = ir.Circuit()
    ckt.vsource("V1", "1", "0", dc=5.0)
    ckt.resistor("R1", "1", "2", 100.0)
    ckt.inductor("L1", "2", "0", 10e-3)
    return ckt


TIGHT = dict(tolerance=1e-13, max_iterations=50000)


def test_be_recurrence_rc(be):
    ckt = rc_circuit()
    mna = solver.compile(ckt)
    a, b = T.uic_system(ckt, mna)
    x0 = oracle.dense_solve(a, b)
    src, _ = T.source_table(ckt, mna)
    res = T.Transient(be, mna, src).run(20e-6, 1e-3, method="be", x0=x0,
                                        params=solver.bicgstab_params(**TIGHT))
    i2 = mna.node_names.index("2")
    v, al = 0.0, 20e-6 / 1e-3
    assert res.steps_accepted == 50 and res.times.shape[0] == 51
    for k in range(1, res.times.shape[0]):
        v = (v + al * 5.0) / (1 + al)
        assert abs(res.waves[k, 0, i2] - v) <= REL * abs(v) + ABS


@pytest.mark.parametrize("which", ["rc", "rl"])
def test_adaptive_matches_oracle_and_closed_form(be, which):
    """test_rc_charging / test_rl_step_response (tests/transient_integration_test.rs:52-132)
    through transient.run, and sample-for-sample against the oracle."""
    ckt = rc_circuit() if which == "rc" else rl_circuit()
    mna = solver.compile(ckt)
    tstep, tstop = (10e-6, 5e-3) if which == "rc" else (1e-6, 500e-6)
    out = T.run(be, ckt, mna, tstep, tstop, uic=True, method="adaptive",
                params=solver.bicgstab_params(**TIGHT))
    src, _ = T.source_table(ckt, mna)
    a, b = T.uic_system(ckt, mna)
    times, waves, _, acc, rej = oracle.transient(mna, src, method="adaptive", tstep=tstep,
                                                 tstop=tstop, x0=oracle.dense_solve(a, b))
    assert out.times.shape == times.shape and np.allclose(out.times, times, rtol=1e-12)
    ok, err = close(out.result.waves[:, 0, :], waves)
    assert ok, err
    if which == "rc":
        v2 = dict(out.node_voltages)["2"]
        for t, v in zip(out.times, v2):
            if t > 0:
                assert abs(v - 5 * (1 - math.exp(-t / 1e-3))) < 0.02 * 5 + 1e-6
    else:
        il = dict(out.branch_currents)["L1"]
        for t, i in zip(out.times, il):
            if t > 1e-6:
                assert abs(i - 0.05 * (1 - math.exp(-t / 1e-4))) < 0.02 * 0.05 + 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_rc_grid_be_vs_oracle(be, precision):
    """The RC power-grid workload's circuit family (24x24, 12 PULSE sinks),
    300 fixed 1 ps backward-Euler steps from the DC operating point."""
    ckt = ir.rc_grid(24, 24, 12, seed=3)
    mna = solver.compile(ckt)
    src, _ = T.source_table(ckt, mna)
    x0 = oracle.sparse_lu_solve(mna.g, mna.b_dc)
    prm = solver.bicgstab_params(tolerance=1e-13, max_iterations=20000, precision=precision)
    res = T.Transient(be, mna, src).run(1e-12, 300e-12, method="be", x0=x0, params=prm)
    times, waves, xf, acc, _ = oracle.transient(mna, src, method="be", tstep=1e-12,
                                                tstop=300e-12, x0=x0)
    assert res.steps_accepted == acc == 300
    assert np.array_equal(res.times, times)
    ok, err = close(res.waves[:, 0, :], waves)
    assert ok, err
    ok, err = close(res.x_final[0], xf)
    assert ok, err


@pytest.mark.parametrize("precond", [N.PREC_ISAI, N.PREC_JACOBI], ids=["isai", "jacobi"])
def test_captured_segments_match_eager(be, precond, monkeypatch):
    """Graph-captured BE timesteps (segments of up to 16 steps replayed as
    one graph, one segment-word read per segment) give bitwise the eager
    loop's samples, final state and iteration counts; so do 3-step segments
    whose 8 captured iterations per step are too few, so that every step
    pauses and the host finishes its solve. Recording starts at tstart and
    stops at max_records inside the segments."""
    ckt = ir.rc_grid(24, 24, 12, seed=3)
    mna = solver.compile(ckt)
    src, _ = T.source_table(ckt, mna)
    x0 = oracle.sparse_lu_solve(mna.g, mna.b_dc)
    prm = solver.bicgstab_params(tolerance=1e-12, max_iterations=20000)
    probes = np.array([0, 5, 100, mna.size - 1])
    out = {}
    for name, steps, iters in (("eager", "0", None), ("segments", "16", None),
                               ("paused", "3", "8")):
        monkeypatch.setenv("OHM_TRAN_SEGMENT_STEPS", steps)
        if iters:
            monkeypatch.setenv("OHM_TRAN_SEGMENT_ITERS", iters)
        else:
            monkeypatch.delenv("OHM_TRAN_SEGMENT_ITERS", raising=False)
        tr = T.Transient(be, mna, src)
        r = tr.run(1e-12, 120e-12, tstart=20e-12, method="be", x0=x0, params=prm,
                   precond=precond, probes=probes, max_records=90)
        out[name] = (r.times.tobytes(), r.waves.tobytes(), r.x_final.tobytes(),
                     r.steps_accepted, r.linear_solves, r.bicgstab_iterations)
    assert out["segments"] == out["eager"]
    assert out["paused"] == out["eager"]
    assert out["eager"][3] == 120 and len(np.frombuffer(out["eager"][0])) == 90


def test_breakdown_retried_inside_transient(be):
    """The reference's shadow r_hat = b breaks BiCGSTAB down on a resistor
    ladder (the plain solve reports it, as the reference does); inside the
    transient analysis a broken-down step is solved again from its starting
    iterate with the other shadow vector, so the run completes and matches
    the oracle's LU stepping."""
    ckt = ir.resistor_ladder(200, seed=1)
    n_nodes = 201
    for i in range(1, n_nodes + 1):  # 1 pF to ground at every node: C/h << G at h = 1 ms
        ckt.capacitor(f"C{i}", str(i), "0", 1e-12)
    mna = solver.compile(ckt)
    src, _ = T.source_table(ckt, mna)
    h = 1e-3
    prm = solver.bicgstab_params(tolerance=1e-13, max_iterations=20000, shadow="rhs")
    cg = T.c_on_g(mna)
    gs = solver.GpuSolver(be, precond=N.PREC_JACOBI, params=prm)
    a = ir.CsrMatrix(mna.size, mna.size, mna.g.values + cg / h, mna.g.col_indices,
                     mna.g.row_pointers)
    with pytest.raises(N.OhmnivoreError, match="BiCGSTAB breakdown"):
        gs.solve_real(a, mna.b_dc)
    res = T.Transient(be, mna, src).run(h, 5 * h, method="be", params=prm,
                                         precond=N.PREC_JACOBI)
    times, waves, xf, acc, _ = oracle.transient(mna, src, method="be", tstep=h, tstop=5 * h)
    assert res.steps_accepted == acc == 5
    ok, err = close(res.waves[:, 0, :], waves)
    assert ok, err


def test_breakdown_retried_inside_batched_transient(be, monkeypatch):
    """The same ladder as three instances stepped together (scaled G and C):
    with the reference's shadow r_hat = b each batched step breaks down, and
    without a saved start (OHM_TRAN_PREDICT=0) the run reports it; by default
    every fixed BE step keeps a copy of its extrapolated start, a broken-down
    system is solved again from it with the other shadow vector, and each
    instance matches the oracle's LU stepping."""
    ckt = ir.resistor_ladder(200, seed=1)
    for i in range(1, 202):
        ckt.capacitor(f"C{i}", str(i), "0", 1e-12)
    mna = solver.compile(ckt)
    src, _ = T.source_table(ckt, mna)
    cg = T.c_on_g(mna)
    h = 1e-3
    prm = solver.bicgstab_params(tolerance=1e-13, max_iterations=20000, shadow="rhs")
    inst, mnas = [], []
    for sc in (1.0, 1.5, 0.7):
        inst.append((mna.g.values * sc, cg * sc))
        mnas.append(ir.MnaSystem(g=ir.CsrMatrix(mna.size, mna.size, mna.g.values * sc,
                                                mna.g.col_indices, mna.g.row_pointers),
                                 c=ir.CsrMatrix(mna.size, mna.size, cg * sc, mna.g.col_indices,
                                                mna.g.row_pointers),
                                 b_dc=mna.b_dc, b_ac=mna.b_ac, size=mna.size,
                                 node_names=mna.node_names, branch_names=mna.branch_names,
                                 diode_descriptors=None, bjt_descriptors=None,
                                 mosfet_descriptors=None, n_nodes=mna.n_nodes))
    probes = np.arange(0, mna.size, 5)
    monkeypatch.setenv("OHM_TRAN_PREDICT", "0")
    with pytest.raises(N.OhmnivoreError, match="BiCGSTAB breakdown"):
        T.Transient(be, mna, src, instances=inst).run(h, 5 * h, method="be", params=prm,
                                                       probes=probes, precond=N.PREC_JACOBI)
    monkeypatch.delenv("OHM_TRAN_PREDICT")
    res = T.Transient(be, mna, src, instances=inst).run(h, 5 * h, method="be", params=prm,
                                                         probes=probes, precond=N.PREC_JACOBI)
    assert res.steps_accepted == 5
    for k, m in enumerate(mnas):
        _, waves, _, acc, _ = oracle.transient(m, src, method="be", tstep=h, tstop=5 * h,
                                               probes=probes)
        assert acc == 5
        ok, err = close(res.waves[:, k, :], waves)
        assert ok, (k, err)


def test_predicted_start_matches_plain_warm_start(be, monkeypatch):
    """Fixed BE steps start their solve from the solution extrapolated from
    the last three steps (OHM_TRAN_PREDICT=2, the default; 1: from the last
    two; 0: from x_n). On a 32 x 32 RC grid every run stays within the parity
    bar of the oracle's LU stepping, and each extrapolated start needs fewer
    BiCGSTAB iterations than the plain warm start."""
    ckt = ir.rc_grid(32, 32, 40, seed=3)
    mna = solver.compile(ckt)
    src, _ = T.source_table(ckt, mna)
    x0 = oracle.sparse_lu_solve(mna.g, mna.b_dc)
    probes = np.arange(0, mna.size, 11)
    prm = solver.bicgstab_params(tolerance=1e-10, max_iterations=20000)
    _, ref, _, _, _ = oracle.transient(mna, src, method="be", tstep=1e-12, tstop=200e-12, x0=x0,
                                       probes=probes)
    its = {}
    for pred in ("0", "1", "2"):
        monkeypatch.setenv("OHM_TRAN_PREDICT", pred)
        res = T.Transient(be, mna, src).run(1e-12, 200e-12, method="be", x0=x0, probes=probes,
                                             params=prm, precond=N.PREC_ISAI)
        ok, err = close(res.waves[:, 0, :], ref)
        assert ok, (pred, err)
        its[pred] = res.bicgstab_iterations
    assert its["1"] < its["0"] and its["2"] < its["0"], its


def test_monte_carlo_instances_vs_oracle(be):
    """K instances stepped together, each with its own conductance scale and
    per-node capacitances (stable perturbations); each against its own
    oracle run."""
    ckt = ir.rc_grid(16, 16, 6, seed=3)
    mna = solver.compile(ckt)
    src, _ = T.source_table(ckt, mna)
    cg = T.c_on_g(mna)
    rng = np.random.default_rng(3)
    inst, mnas = [], []
    for k in range(5):
        gs = mna.g.values * (1 + 0.1 * (rng.random() - 0.5))
        cs = cg * (1 + 0.3 * rng.random(cg.shape[0]))
        inst.append((gs, cs))
        # the same instance as an MnaSystem for the oracle (C on G's pattern)
        m = ir.MnaSystem(g=ir.CsrMatrix(mna.size, mna.size, gs, mna.g.col_indices,
                                        mna.g.row_pointers),
                         c=ir.CsrMatrix(mna.size, mna.size, cs, mna.g.col_indices,
                                        mna.g.row_pointers),
                         b_dc=mna.b_dc, b_ac=mna.b_ac, size=mna.size, node_names=mna.node_names,
                         branch_names=mna.branch_names, diode_descriptors=None,
                         bjt_descriptors=None, mosfet_descriptors=None, n_nodes=mna.n_nodes)
        mnas.append(m)
    x0 = np.stack([oracle.sparse_lu_solve(m.g, m.b_dc) for m in mnas])
    probes = np.arange(0, mna.size, 7)
    res = T.Transient(be, mna, src, instances=inst).run(
        1e-12, 120e-12, method="be", x0=x0, probes=probes,
        params=solver.bicgstab_params(tolerance=1e-13, max_iterations=20000))
    for k, m in enumerate(mnas):
        times, waves, _, _, _ = oracle.transient(m, src, method="be", tstep=1e-12,
                                                 tstop=120e-12, x0=x0[k], probes=probes)
        ok, err = close(res.waves[:, k, :], waves)
        assert ok, (k, err)
