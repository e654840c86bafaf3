"""Transient oracle pinned by the reference's own known answers:
transient_source.rs:120-361 (source functions), transient_integration_test.rs
(RC charging / RL step closed forms, 2 %) and the exact discrete recurrence
of backward Euler on an RC stage. Also checks the product's host-side
source table (collect_source_info) and UIC system."""
import math

import numpy as np
import pytest

import oracle
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver
from ohmnivore_amd import transient as T


def src(kind, param, pwl=None):
    e = np.zeros(1, dtype=ir.TRAN_SOURCE_DTYPE)[0]
    e["kind"] = kind
    e["param"][: len(param)] = param
    if pwl is not None:
        e["pwl_offset"], e["pwl_count"] = 0, len(pwl)
    return e


PULSE_CASES = [  # (v1 v2 td tr tf pw per), t, expected
    ((0, 5, 1, .1, .1, 1, 3), .5, 0.0), ((0, 5, 0, 1, 1, 2, 6), .5, 2.5),
    ((0, 5, 0, 1, 1, 2, 6), 2.0, 5.0), ((0, 5, 0, 1, 1, 2, 6), 3.5, 2.5),
    ((0, 5, 0, 1, 1, 2, 6), 5.0, 0.0), ((0, 5, 0, 0, 0, 1, 2), 2.5, 5.0),
    ((0, 5, 0, 0, 0, 1, 2), 3.5, 0.0)]
SIN_CASES = [((1, 2, 1, 1, 0), .5, 1.0), ((0, 1, 1, 0, 0), .25, 1.0), ((0, 1, 1, 0, 0), .5, 0.0),
             ((0, 1, 1, 0, 1), .25, math.exp(-.25)), ((2.5, 1, 1, 0, 0), .25, 3.5)]
EXP_CASES = [((0, 5, 1, 1, 3, 1), .5, 0.0, 1e-12),
             ((0, 5, 0, 1, 10, 1), 1.0, 5 * (1 - math.exp(-1)), 1e-10),
             ((0, 5, 0, 1, 5, 1), 6.0, 5 * (1 - math.exp(-6)) - 5 * (1 - math.exp(-1)), 1e-10)]
PWL_CASES = [([(1, 0), (2, 5)], 0.0, 0.0), ([(1, 0), (2, 5)], 3.0, 5.0),
             ([(0, 0), (1, 10), (2, 10), (3, 0)], .5, 5.0),
             ([(0, 0), (1, 10), (2, 10), (3, 0)], 1.5, 10.0),
             ([(0, 0), (1, 10), (2, 10), (3, 0)], 2.5, 5.0),
             ([(0, 0), (1, 5), (2, 3)], 1.0, 5.0)]


@pytest.mark.parametrize("p,t,want", PULSE_CASES)
def test_pulse_known_answers(p, t, want):
    assert abs(oracle.source_value(src(ir.SRC_PULSE, p), t) - want) < 1e-12


@pytest.mark.parametrize("p,t,want", SIN_CASES)
def test_sin_known_answers(p, t, want):
    assert abs(oracle.source_value(src(ir.SRC_SIN, p), t) - want) < 1e-12


@pytest.mark.parametrize("p,t,want,tol", EXP_CASES)
def test_exp_known_answers(p, t, want, tol):
    assert abs(oracle.source_value(src(ir.SRC_EXP, p), t) - want) < tol


@pytest.mark.parametrize("pairs,t,want", PWL_CASES)
def test_pwl_known_answers(pairs, t, want):
    e = src(ir.SRC_PWL, [], pwl=pairs)
    assert abs(oracle.source_value(e, t, np.array(pairs, float)) - want) < 1e-12


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


def uic_x0(ckt, mna):
    a, b = T.uic_system(ckt, mna)
    return oracle.dense_solve(a, b)


def test_source_table_matches_collect_source_info():
    ckt = ir.rc_grid(8, 8, 5, seed=3)
    mna = solver.compile(ckt)
    table, _ = T.source_table(ckt, mna)
    assert table.shape[0] == 4 + 5
    assert np.all(table["kind"][:4] == ir.SRC_DC) and np.all(table["dc"][:4] == 1.0)
    assert np.all(table["idx"][:4, 0] == mna.n_nodes + np.arange(4))
    sinks = table[4:]
    assert np.all(sinks["n_idx"] == 1) and np.all(sinks["sign"][:, 0] == -1.0)
    for k, e in enumerate(sinks):
        p = ckt.tran[ckt.components.shape[0] - 5 + k]
        for t in (0.0, 1e-10, 3.3e-10, 7e-10):
            assert oracle.source_value(e, t) == p(t)


def test_uic_system_rc_and_rl():
    ckt = rc_circuit()
    mna = solver.compile(ckt)
    x0 = uic_x0(ckt, mna)
    i2 = mna.node_names.index("2")
    assert x0[i2] == 0.0 and abs(x0[mna.node_names.index("1")] - 5.0) < 1e-12
    ckt = rl_circuit()
    mna = solver.compile(ckt)
    x0 = uic_x0(ckt, mna)
    assert x0[mna.n_nodes + 1] == 0.0  # I_L = 0 (L1 is the second branch)


def test_oracle_rc_charging_closed_form():
    """test_rc_charging (tests/transient_integration_test.rs:52-87)."""
    ckt = rc_circuit()
    mna = solver.compile(ckt)
    table, _ = T.source_table(ckt, mna)
    i2 = mna.node_names.index("2")
    times, waves, _, acc, rej = oracle.transient(mna, table, method="adaptive", tstep=10e-6,
                                                 tstop=5e-3, x0=uic_x0(ckt, mna))
    assert acc >= 500 and times[-1] == pytest.approx(5e-3)
    for t, v in zip(times, waves[:, i2]):
        if t > 0:
            assert abs(v - 5 * (1 - math.exp(-t / 1e-3))) < 0.02 * 5 + 1e-6


def test_oracle_rl_step_closed_form():
    """test_rl_step_response (tests/transient_integration_test.rs:89-132)."""
    ckt = rl_circuit()
    mna = solver.compile(ckt)
    table, _ = T.source_table(ckt, mna)
    times, waves, _, acc, rej = oracle.transient(mna, table, method="adaptive", tstep=1e-6,
                                                 tstop=500e-6, x0=uic_x0(ckt, mna))
    il = waves[:, mna.n_nodes + 1]
    for t, i in zip(times, il):
        if t > 1e-6:
            assert abs(i - 0.05 * (1 - math.exp(-t / 1e-4))) < 0.02 * 0.05 + 1e-6


def test_oracle_backward_euler_recurrence():
    """Fixed-step BE on one RC stage is v_{n+1} = (v_n + a V) / (1 + a),
    a = h / RC, exactly (up to rounding)."""
    ckt = rc_circuit()
    mna = solver.compile(ckt)
    table, _ = T.source_table(ckt, mna)
    h = 20e-6
    times, waves, _, acc, _ = oracle.transient(mna, table, method="be", tstep=h, tstop=1e-3,
                                               x0=uic_x0(ckt, mna))
    i2 = mna.node_names.index("2")
    v, a = 0.0, h / 1e-3
    assert acc == 50
    for k in range(1, times.shape[0]):
        v = (v + a * 5.0) / (1 + a)
        assert abs(waves[k, i2] - v) < 1e-9 * 5


# ------------------------------------------- transient with device currents --
def _rlc_pulse():
    ckt = ir.Circuit()
    ckt.vsource("V1", "1", "0", tran=ir.Pulse(0.0, 5.0, 0.0, 1e-6, 1e-6, 5e-6, 20e-6))
    ckt.resistor("R1", "1", "2", 1e3)
    ckt.capacitor("C1", "2", "0", 1e-9)
    ckt.inductor("L1", "2", "3", 1e-3)
    ckt.resistor("R2", "3", "0", 50.0)
    return ckt


def test_nonlinear_oracle_without_devices_is_the_linear_oracle():
    """With no devices the per-step Newton is one LU solve of the reference's
    companion system: fixed-step BE agrees sample for sample; the adaptive
    run takes the same accepted / rejected steps, its step sizes drifting
    only by the LTE estimate's rounding."""
    ckt = _rlc_pulse()
    mna = solver.compile(ckt)
    src, pwl = T.source_table(ckt, mna)
    a = oracle.transient(mna, src, pwl, method="be", tstep=1e-6, tstop=40e-6)
    b = oracle.transient_nonlinear(mna, src, pwl, method="be", tstep=1e-6, tstop=40e-6)
    assert np.array_equal(a[0], b[0]) and np.allclose(a[1], b[1], rtol=1e-12, atol=1e-14)
    a = oracle.transient(mna, src, pwl, method="adaptive", tstep=1e-6, tstop=40e-6,
                         max_records=5000)
    b = oracle.transient_nonlinear(mna, src, pwl, method="adaptive", tstep=1e-6, tstop=40e-6)
    assert (a[3], a[4]) == (b[3], b[4]) and a[0].shape == b[0].shape
    assert np.allclose(a[0], b[0], rtol=1e-6, atol=0) and np.abs(a[1] - b[1]).max() < 1e-4


def test_nonlinear_oracle_diode_discharge_closed_form():
    """C dv/dt = -Is e^(v/nVt) for a capacitor discharging through a diode
    from 0.65 V: v(t) = -nVt ln(e^(-v0/nVt) + Is t/(C nVt)) while v >> nVt,
    against backward Euler at the reference's 2 % transient tolerance."""
    ckt = ir.Circuit()
    ckt.add_diode_model("D", 1e-14, 1.0)
    ckt.capacitor("C1", "1", "0", 1e-6)
    ckt.diode("D1", "1", "0", "D")
    mna = solver.compile(ckt)
    d = mna.diode_descriptors[0]
    i_s, nvt = float(d["is_val"]), float(d["n_vt"])
    src, pwl = T.source_table(ckt, mna)
    times, waves, _, acc, _ = oracle.transient_nonlinear(mna, src, pwl, method="be", tstep=2e-6,
                                                         tstop=4e-3, x0=[0.65])
    expect = -nvt * np.log(np.exp(-0.65 / nvt) + i_s * times / (1e-6 * nvt))
    assert acc == 2000 and expect[-1] > 0.45
    assert np.all(np.abs(waves[:, 0] - expect) < 0.02 * 0.65)
    # TRAP (adaptive) is second order: tighter than BE on the same span
    times2, waves2, _, _, _ = oracle.transient_nonlinear(mna, src, pwl, method="adaptive",
                                                         tstep=2e-6, tstop=4e-3, x0=[0.65])
    e2 = -nvt * np.log(np.exp(-0.65 / nvt) + i_s * times2 / (1e-6 * nvt))
    assert np.abs(waves2[:, 0] - e2).max() < np.abs(waves[:, 0] - expect).max()
