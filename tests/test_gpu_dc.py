"""dc::run on the GPU (src/analysis/dc.rs:25-650): linear operating points,
the direct Newton attempt and the continuation cascade (source stepping,
GMIN stepping), each against the oracle's LU answers."""
import numpy as np
import pytest

import oracle
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import dc
from ohmnivore_amd import solver

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


def test_source_stepping_path(be):
    """With MOSFETs present the cascade tries source stepping first."""
    mna = solver.compile(ir.amplifier_array(16, seed=5))
    x, path = dc.run_nonlinear(be, mna, max_iterations=1, **TIGHT)
    labels = [p[0] for p in path]
    assert labels[0] == "direct" and "source" in labels
    ok, err = close(x, oracle_op(mna))
    assert ok, err
