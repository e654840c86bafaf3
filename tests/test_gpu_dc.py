"""dc::run on the GPU (src/analysis/dc.rs:25-650): linear operating points,
the direct Newton attempt and the continuation cascade (source stepping,
GMIN stepping, pseudo-transient continuation), each against the oracle's LU
answers."""
import numpy as np
import pytest

import oracle
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import dc
from ohmnivore_amd import solver
from ohmnivore_amd import _native as N

pytestmark = pytest.mark.gpu
TIGHT = dict(abs_tol=1e-12, bicgstab_tolerance=1e-10)


@pytest.fixture(scope="module")
def be():
    return solver.HipBackend()


def close(x, ref, rel=1e-6, abs_=1e-9):
    err = np.abs(x - ref)
    return bool((err <= np.maximum(rel * np.abs(ref), abs_)).all()), float(err.max(initial=0.0))


def test_linear_divider(be):
    """test_voltage_divider (tests/integration_test.rs): V2 = 10 * 2k / 3k."""
    ckt = ir.Circuit()
    ckt.vsource("V1", "1", "0", dc=10.0)
    ckt.resistor("R1", "1", "2", 1e3)
    ckt.resistor("R2", "2", "0", 2e3)
    res = dc.run(be, solver.compile(ckt))
    assert abs(dict(res.node_voltages)["2"] - 20.0 / 3.0) < 1e-6  # the reference test's epsilon


def oracle_op(mna):
    x, _, _ = oracle.newton_solve(mna.g, mna.b_dc, mna.diode_descriptors, mna.bjt_descriptors,
                                  mna.mosfet_descriptors, abs_tol=1e-13, max_iterations=200)
    return x


@pytest.mark.parametrize("ckt", [ir.diode_mesh(32, 32, seed=2), ir.amplifier_array(64, seed=5)])
def test_direct_newton(be, ckt):
    mna = solver.compile(ckt)
    res = dc.run(be, mna, max_iterations=100, **TIGHT)
    assert res.path[0][0] == "direct" and res.path[0][2]
    ok, err = close(res.x, oracle_op(mna))
    assert ok, err


def test_gmin_stepping_path(be):
    """Direct Newton capped at one iteration fails; GMIN stepping (no
    MOSFETs) carries the solve to the same operating point."""
    mna = solver.compile(ir.diode_mesh(24, 24, seed=2))
    x, path = dc.run_nonlinear(be, mna, max_iterations=1, **TIGHT)
    labels = [p[0] for p in path]
    assert labels[0] == "direct" and not path[0][2] and "gmin" in labels and labels[-1] == "final"
    ok, err = close(x, oracle_op(mna))
    assert ok, err


def test_pseudo_transient_path(be, monkeypatch):
    """The final true-GMIN Newton capped at zero iterations (forced from the
    host) fails after GMIN stepping, so the cascade runs pseudo-transient
    continuation from the last GMIN point (src/analysis/dc.rs:274-290,
    341-500): its damped steps, then the final alpha = 0 solve, landing on
    the oracle's LU operating point."""
    monkeypatch.setenv("OHM_DC_FINAL_MAX_ITERATIONS", "0")
    mna = solver.compile(ir.diode_mesh(24, 24, seed=2))
    res = dc.run(be, mna, max_iterations=1, **TIGHT)
    labels = [p[0] for p in res.path]
    assert labels[0] == "direct" and not res.path[0][2] and "gmin" in labels
    finals = [p for p in res.path if p[0] == "final"]
    assert len(finals) == 1 and not finals[0][2] and finals[0][3] == 0
    assert "ptc" in labels and labels[-1] == "ptc-final" and res.path[-1][2]
    ptc_levels = [p[1] for p in res.path if p[0] == "ptc" and p[2]]
    assert ptc_levels and ptc_levels == sorted(ptc_levels, reverse=True) and ptc_levels[0] <= 0.1
    assert res.stage == "ptc-final"
    assert res.newton_iterations == sum(p[3] for p in res.path)
    ok, err = close(res.x, oracle_op(mna))
    assert ok, err


def test_source_stepping_path(be):
    """With MOSFETs present the cascade tries source stepping first."""
    mna = solver.compile(ir.amplifier_array(16, seed=5))
    x, path = dc.run_nonlinear(be, mna, max_iterations=1, **TIGHT)
    labels = [p[0] for p in path]
    assert labels[0] == "direct" and "source" in labels
    ok, err = close(x, oracle_op(mna))
    assert ok, err


def cmos_chain(stages):
    """The reference's cmos_inverter_chain_dc netlist family: one PMOS/NMOS
    pair per stage, VIN = VDD = 5."""
    ckt = ir.Circuit()
    ckt.add_mosfet_model("NMOD", 0.7, 1.1e-4, 0.04, True)
    ckt.add_mosfet_model("PMOD", -0.7, 5.5e-5, 0.04, False)
    ckt.vsource("VDD", "vdd", "0", dc=5.0)
    ckt.vsource("VIN", "in", "0", dc=5.0)
    prev = "in"
    for k in range(1, stages + 1):
        ckt.mosfet(f"MP{k}", f"out_{k}", prev, "vdd", "PMOD")
        ckt.mosfet(f"MN{k}", f"out_{k}", prev, "0", "NMOD")
        prev = f"out_{k}"
    return ckt


def oracle_source_stepping(mna):
    """The oracle's LU Newton along the cascade's source-stepping schedule
    (steps converged to 1e-10, the last one polished to 1e-12)."""
    def newton(scale, x0, tol, iters):
        x, _, _ = oracle.newton_solve(mna.g, scale * mna.b_dc, mna.diode_descriptors,
                                      mna.bjt_descriptors, mna.mosfet_descriptors,
                                      abs_tol=tol, max_iterations=iters, initial_guess=x0)
        return x
    x, last, target = None, 0.0, 0.1
    while True:
        try:
            x = newton(target, x, 1e-10, 30)
        except oracle.OracleError:
            target = last + (target - last) / 2.0
            assert target - last > 1e-6
            continue
        if target == 1.0:
            try:
                return newton(1.0, x, 1e-12, 50)
            except oracle.OracleError:
                return x
        target, last = min(target + 2.0 * (target - last), 1.0), target


@pytest.mark.parametrize("stages", [10, 200])
def test_cmos_chain_cascade(be, stages):
    """A CMOS inverter chain needs the cascade (direct Newton from zero does
    not converge); the native cascade's stages, counters and answer."""
    mna = solver.compile(cmos_chain(stages))
    res = dc.run(be, mna, **TIGHT)
    stages_run = [p[0] for p in res.path]
    assert stages_run[0] == "direct" and not res.path[0][2]
    assert res.stage in ("source", "final", "ptc-final") and res.path[-1][2]
    assert res.newton_iterations == sum(p[3] for p in res.path)
    ok, err = close(res.x, oracle_source_stepping(mna))
    assert ok, err
    v = dict(res.node_voltages)
    for k in range(1, stages + 1):
        assert (v[f"out_{k}"] < 0.5) if k % 2 else (v[f"out_{k}"] > 4.5)


def test_dense_lu_newton_matches_oracle(be):
    """OHM_LINEAR_DENSE_LU: every Newton step by the GPU dense LU (the
    reference's CpuSparseLu mode) lands on the oracle's operating point."""
    for ckt in (cmos_chain(10), ir.amplifier_array(16, seed=5)):
        mna = solver.compile(ckt)
        ref = oracle_source_stepping(mna)
        ctx = solver.NewtonContext(be, mna.g, mna.b_dc, mna.diode_descriptors,
                                   mna.bjt_descriptors, mna.mosfet_descriptors)
        # from the oracle's answer perturbed: the direct steps must pull it back
        st = ctx.run(abs_tol=1e-12, linear_solver="dense", max_iterations=50,
                     initial_guess=ref * (1 + 1e-3))
        ok, err = close(ctx.x(), ref)
        assert ok, err
        assert st.bicgstab_iterations == 0


def test_dense_lu_reports_singular_system(be):
    """Two ideal sources in parallel make the MNA matrix singular."""
    ckt = ir.Circuit()
    ckt.add_diode_model("D", 1e-14, 1.0)
    ckt.vsource("V1", "1", "0", dc=1.0)
    ckt.vsource("V2", "1", "0", dc=2.0)
    ckt.diode("D1", "1", "0", "D")
    mna = solver.compile(ckt)
    ctx = solver.NewtonContext(be, mna.g, mna.b_dc, mna.diode_descriptors)
    with pytest.raises(N.OhmnivoreError, match="singular"):
        ctx.run(linear_solver="dense")


def test_long_cmos_chain_cascade(be):
    """1,000 inverter stages: BiCGSTAB sub-solves break down on the chain's
    nearly floating nodes, so the cascade's retries solve each Newton step
    by dense LU (the reference's CPU sparse-LU retry)."""
    mna = solver.compile(cmos_chain(1000))
    res = dc.run(be, mna, **TIGHT)
    assert any(p[4] for p in res.path)  # a retry ran
    ok, err = close(res.x, oracle_source_stepping(mna))
    assert ok, err
