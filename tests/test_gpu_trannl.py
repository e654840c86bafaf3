"""Transient with the device currents in the time loop
(ohm_transient_nonlinear; the reference's transient steps G and C only).
Checked against the oracle's per-step LU Newton (oracle.transient_nonlinear,
itself pinned to the linear transient oracle and a closed form in
tests/test_oracle_transient.py) at 1e-6 relative / 1e-9 absolute, against a
closed form at the reference's 2 % transient tolerance, and on the
reference's CMOS inverter-chain transient netlist, which has no capacitors
and so must follow the DC operating point of every input level."""
import io
import os

import numpy as np
import pytest

import oracle
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import cli, dc
from ohmnivore_amd import netlist as nl
from ohmnivore_amd import solver
from ohmnivore_amd import transient as T

pytestmark = pytest.mark.gpu
REL, ABS = 1e-6, 1e-9
TIGHT = dict(abs_tol=1e-12, bicgstab_tolerance=1e-12)
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def be():
    return solver.HipBackend()


def close(x, ref, rel=REL, abs_=ABS):
    err = np.abs(np.asarray(x) - np.asarray(ref))
    return bool((err <= np.maximum(rel * np.abs(ref), abs_)).all()), float(err.max(initial=0.0))


def rlc_pulse():
    ckt = ir.Circuit()
    ckt.vsource("V1", "1", "0", tran=ir.Pulse(0.0, 5.0, 0.0, 1e-6, 1e-6, 5e-6, 20e-6))
    ckt.resistor("R1", "1", "2", 1e3)
    ckt.capacitor("C1", "2", "0", 1e-9)
    ckt.inductor("L1", "2", "3", 1e-3)
    ckt.resistor("R2", "3", "0", 50.0)
    return ckt


def diode_rc():
    """A diode clamp behind an RC, driven by a 0 -> 5 V pulse train."""
    ckt = ir.Circuit()
    ckt.add_diode_model("D", 1e-14, 1.0)
    ckt.vsource("V1", "in", "0", tran=ir.Pulse(0.0, 5.0, 1e-6, 0.5e-6, 0.5e-6, 4e-6, 10e-6))
    ckt.resistor("R1", "in", "out", 1e3)
    ckt.capacitor("C1", "out", "0", 1e-9)
    ckt.diode("D1", "out", "0", "D")
    ckt.resistor("R2", "out", "0", 10e3)
    return ckt


def test_without_devices_equals_the_linear_stepper(be):
    """No devices: every step is one linear companion solve, so fixed-step
    BE reproduces the linear oracle sample for sample."""
    ckt = rlc_pulse()
    mna = solver.compile(ckt)
    src, pwl = T.source_table(ckt, mna)
    res = T.run_nonlinear(be, ckt, mna, 1e-6, 40e-6, method="be", newton_kw=TIGHT)
    times, waves, _, acc, _ = oracle.transient(mna, src, pwl, method="be", tstep=1e-6, tstop=40e-6)
    assert res.steps_accepted == acc == 40 and np.array_equal(res.times, times)
    ok, err = close(res.waves[:, 0, :], waves)
    assert ok, err


@pytest.mark.parametrize("method", ["be", "adaptive"])
def test_diode_rc_against_oracle(be, method):
    ckt = diode_rc()
    mna = solver.compile(ckt)
    src, pwl = T.source_table(ckt, mna)
    res = T.run_nonlinear(be, ckt, mna, 0.1e-6, 20e-6, method=method, newton_kw=TIGHT)
    times, waves, _, acc, rej = oracle.transient_nonlinear(mna, src, pwl, method=method,
                                                           tstep=0.1e-6, tstop=20e-6)
    w = res.waves[:, 0, :]
    if method == "be":
        assert np.array_equal(res.times, times)
    else:
        # LTE-driven step sizes carry the solves' rounding into the times;
        # compare up to the first place the two step sequences part
        n = min(times.shape[0], res.times.shape[0])
        same = np.isclose(res.times[:n], times[:n], rtol=1e-8, atol=0)
        k = n if same.all() else int(np.argmin(same))
        assert k >= n // 2, (k, n)
        times, waves, w = times[:k], waves[:k], w[:k]
    ok, err = close(w, waves)
    assert ok, err
    # the clamp holds the output near a diode drop while the input is high
    out = mna.node_names.index("out")
    k = np.searchsorted(times, 4.5e-6)
    assert 0.5 < w[k, out] < 0.8


def test_diode_discharge_closed_form(be):
    """C dv/dt = -Is e^(v/nVt) from 0.65 V (tests/test_oracle_transient.py):
    GPU BE against the closed form at 2 % and the oracle at 1e-6."""
    ckt = ir.Circuit()
    ckt.add_diode_model("D", 1e-14, 1.0)
    ckt.capacitor("C1", "1", "0", 1e-6)
    ckt.diode("D1", "1", "0", "D")
    mna = solver.compile(ckt)
    d = mna.diode_descriptors[0]
    i_s, nvt = float(d["is_val"]), float(d["n_vt"])
    src, pwl = T.source_table(ckt, mna)
    res = T.run_nonlinear(be, ckt, mna, 2e-6, 4e-3, x0=[0.65], method="be", newton_kw=TIGHT)
    v = res.waves[:, 0, 0]
    expect = -nvt * np.log(np.exp(-0.65 / nvt) + i_s * res.times / (1e-6 * nvt))
    assert np.all(np.abs(v - expect) < 0.02 * 0.65)
    times, waves, _, _, _ = oracle.transient_nonlinear(mna, src, pwl, method="be", tstep=2e-6,
                                                       tstop=4e-3, x0=[0.65])
    assert np.array_equal(res.times, times)
    ok, err = close(v, waves[:, 0])
    assert ok, err


def test_cmos_chain_transient_follows_operating_points(be):
    """tests/regression/circuits/cmos_inverter_chain_tran.spice: no
    capacitors, so every recorded point is the DC operating point of the
    chain at that input voltage; the stage outputs alternate with VIN."""
    with open(os.path.join(ROOT, "tests", "golden", "netlists",
                           "cmos_inverter_chain_tran.spice")) as f:
        text = f.read()
    out = io.StringIO()
    res = cli.simulate(be, text, out, newton=dict(abs_tol=1e-12, bicgstab_tolerance=1e-10))
    a, r = res[0]
    assert isinstance(a, nl.Tran) and r.times[-1] == pytest.approx(a.tstop)
    v = dict(r.node_voltages)
    vin = v["in"]
    # the input follows its PULSE(0 5 10n 1n 1n 50n 100n)
    # t = 0 holds the UIC state (device-free solve, floating stage outputs at 0)
    lo, hi = (vin < 0.01) & (r.times > 0), vin > 4.99
    assert lo.sum() > 10 and hi.sum() > 10
    for k in range(1, 11):
        o = v[f"out_{k}"]
        high_when_in_low = (k % 2 == 1)
        assert np.all(o[lo] > 4.5) if high_when_in_low else np.all(o[lo] < 0.5)
        assert np.all(o[hi] < 0.5) if high_when_in_low else np.all(o[hi] > 4.5)
    # a quasi-static point: the DC operating point with VIN held at 5 V
    c = nl.parse(text.replace("VIN in 0 PULSE(0 5 10n 1n 1n 50n 100n)", "VIN in 0 DC 5"))
    mna = solver.compile(c.to_circuit())
    op = dc.run(be, mna, abs_tol=1e-12, bicgstab_tolerance=1e-10)
    k = int(np.argmax(hi))
    x = r.result.waves[k, 0, :]
    ok, err = close(x[:mna.n_nodes], op.x[:mna.n_nodes], rel=1e-6, abs_=1e-6)
    assert ok, err
    assert out.getvalue().startswith("time,V(vdd),V(in),V(out_1)")
