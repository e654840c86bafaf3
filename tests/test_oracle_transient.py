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
