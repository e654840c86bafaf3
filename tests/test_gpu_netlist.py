"""Netlist text -> GPU analyses -> CSV, end to end (src/main.rs:73-205).

The netlists are the reference's regression circuits
(tests/regression/circuits/*.spice), example netlists (examples/*.spice)
and integration-test netlists (tests/integration_test.rs,
tests/transient_integration_test.rs), kept here as fixture text. Each
result is checked against closed forms where the circuit has one and
against the oracle (LU-based DC / Newton / AC / transient) at 1e-6 relative
/ 1e-9 absolute; the closed-form checks of the transient keep the
reference's 2 % integration tolerance (the stepper's own error, not the
solver's).
"""
import io
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from ohmnivore_amd import cli
from ohmnivore_amd import netlist as nl
from ohmnivore_amd import solver as S
from ohmnivore_amd import transient as T

pytestmark = pytest.mark.gpu
REL, ABS = 1e-6, 1e-9
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

VOLTAGE_DIVIDER = "* Voltage Divider\nV1 1 0 DC 10\nR1 1 2 1k\nR2 2 0 1k\n.DC\n.END\n"
CURRENT_SOURCE = "* Current Source with Resistor\nI1 0 1 DC 0.001\nR1 1 0 1k\n.DC\n.END\n"
RC_LOWPASS = "* RC Low-Pass Filter\nV1 1 0 AC 1 0\nR1 1 2 1k\nC1 2 0 1u\n.AC DEC 10 1 1000000\n.END\n"
DIODE_FORWARD = ("* Diode Forward Bias\nV1 1 0 DC 5\nR1 1 2 1k\n.MODEL DMOD D(IS=1e-14 N=1.0)\n"
                 "D1 2 0 DMOD\n.DC\n.END\n")
NPN_CE = ("* NPN Common Emitter\nV1 vcc 0 DC 5\nV2 vb 0 DC 1.0\nR1 vcc vc 1k\nR2 vb base 100k\n"
          ".MODEL Q2N2222 NPN(IS=1e-14 BF=200 BR=2 NF=1.0 NR=1.0)\nQ1 vc base 0 Q2N2222\n.DC\n.END\n")
NMOS_CS = ("* NMOS Common Source\nV1 vdd 0 DC 5\nV2 gate 0 DC 2.0\nR1 vdd drain 1k\n"
           ".MODEL NMOD NMOS(VTO=0.7 KP=1.1e-4 LAMBDA=0.04)\nM1 drain gate 0 0 NMOD\n.DC\n.END\n")
DIODE_CLAMP = ("* Diode Voltage Clamp\nV1 in 0 DC 5\nR1 in out 1k\n"
               ".MODEL D1N4148 D(IS=2.52e-9 N=1.752)\nD1 out 0 D1N4148\n\n.OP\n.PRINT OP V(out)\n.END\n")
TRAN_RC = "* RC Charging\nV1 1 0 DC 5\nR1 1 2 1k\nC1 2 0 1u\n.TRAN 10u 5m 0 UIC\n.END\n"
TRAN_PULSE = ("* Pulse through RC filter\nV1 1 0 PULSE(0 5 0 1n 1n 0.5m 1m)\nR1 1 2 1k\nC1 2 0 1u\n"
              ".TRAN 10u 1m 0 UIC\n.END\n")
TRAN_DIVIDER = ("* RC Voltage Divider Step\nV1 1 0 DC 5\nR1 1 2 1k\nR2 2 0 1k\nC1 2 0 1u\n"
                ".TRAN 10u 5m 0 UIC\n.END\n")


def cmos_chain(stages=10, analysis=".DC"):
    """cmos_inverter_chain_dc.spice: a PMOS/NMOS pair per stage."""
    lines = ["* CMOS inverter chain", ".MODEL NMOD NMOS(VTO=0.7 KP=1.1e-4 LAMBDA=0.04)",
             ".MODEL PMOD PMOS(VTO=-0.7 KP=5.5e-5 LAMBDA=0.04)", "VDD vdd 0 DC 5", "VIN in 0 DC 5"]
    prev = "in"
    for k in range(1, stages + 1):
        lines += [f"* Stage {k}", f"MP{k} out_{k} {prev} vdd vdd PMOD", f"MN{k} out_{k} {prev} 0 0 NMOD"]
        prev = f"out_{k}"
    return "\n".join(lines + [analysis, ".END", ""])


@pytest.fixture(scope="module")
def be():
    return S.HipBackend()


def close(x, ref, rel=REL, abs_=ABS):
    x, ref = np.asarray(x), np.asarray(ref)
    err = np.abs(x - ref)
    return bool((err <= np.maximum(rel * np.abs(ref), abs_)).all()), float(err.max(initial=0.0))


TIGHT = dict(abs_tol=1e-12, bicgstab_tolerance=1e-10)  # Newton to the oracle's accuracy


def run(be, text, newton=TIGHT):
    out = io.StringIO()
    res = cli.simulate(be, text, out, newton=newton)
    return res, out.getvalue()


def csv_table(text):
    lines = text.strip().splitlines()
    head = lines[0].split(",")
    rows = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])
    return head, rows


def system_of(text):
    c = nl.parse(text).to_circuit()
    return c, S.compile(c)


# ------------------------------------------------------------------ DC --
def test_voltage_divider_csv(be):
    """test_voltage_divider_dc + test_dc_csv_output (tests/integration_test.rs:50-245)."""
    res, text = run(be, VOLTAGE_DIVIDER)
    lines = text.splitlines()
    assert lines[0] == "Variable,Value"
    assert [ln.split(",")[0] for ln in lines[1:]] == ["V(1)", "V(2)", "I(V1)"]
    v = [float(ln.split(",")[1]) for ln in lines[1:]]
    assert abs(v[0] - 10.0) < 1e-6 and abs(v[1] - 5.0) < 1e-6 and abs(v[2] + 0.005) < 1e-9
    _, m = system_of(VOLTAGE_DIVIDER)
    assert close(res[0][1].x, oracle.dense_solve(m.g, m.b_dc))[0]


@pytest.mark.parametrize("text,node,value", [
    (CURRENT_SOURCE, "1", 1.0),
    ("V1 1 0 DC 10\nR1 1 2 1k\nR2 2 0 2k\n.DC\n.END\n", "2", 20.0 / 3.0),
    ("* Voltage Divider\nV1 in 0 DC 10\nR1 in mid 1k\nR2 mid 0 1k\n\n.OP\n.PRINT OP V(mid)\n.END\n",
     "mid", 5.0),
])
def test_linear_operating_points(be, text, node, value):
    """test_current_source_dc / test_voltage_divider_unequal (tests/integration_test.rs:89-127),
    examples/voltage_divider.spice."""
    res, _ = run(be, text)
    # closed form without the compiler's 1e-12 S node shunts: the reference's epsilon
    assert abs(dict(res[0][1].node_voltages)[node] - value) < 1e-6
    _, m = system_of(text)
    assert close(res[0][1].x, oracle.dense_solve(m.g, m.b_dc))[0]


def oracle_op(m):
    """The oracle's LU Newton from zero; where that fails (the CMOS chain),
    the same continuation the GPU cascade runs first for MOSFET circuits:
    adaptive source stepping (src/analysis/dc.rs:480-640), each point's
    Newton seeded by the last converged one."""
    def newton(b, x0, iters=200):
        x, _, _ = oracle.newton_solve(m.g, b, m.diode_descriptors, m.bjt_descriptors,
                                      m.mosfet_descriptors, abs_tol=1e-13, max_iterations=iters,
                                      initial_guess=x0)
        return x
    try:
        return newton(m.b_dc, None)
    except oracle.OracleError:
        pass
    x, last, target = None, 0.0, 0.1
    while True:
        try:
            x = newton(target * m.b_dc, x, iters=30)
        except oracle.OracleError:
            target = last + (target - last) / 2.0
            assert target - last > 1e-6, "oracle source stepping stalled"
            continue
        if target == 1.0:
            return newton(m.b_dc, x)
        target, last = min(target + 2.0 * (target - last), 1.0), target


@pytest.mark.parametrize("name,text", [
    ("diode_forward", DIODE_FORWARD), ("diode_clamp", DIODE_CLAMP), ("npn_common_emitter", NPN_CE),
    ("nmos_common_source", NMOS_CS), ("cmos_inverter_chain_dc", cmos_chain()),
])
def test_nonlinear_operating_points_match_oracle(be, name, text):
    """Regression and example netlists through the Newton cascade on the GPU
    against the oracle's LU Newton."""
    res, csv = run(be, text)
    _, m = system_of(text)
    ok, err = close(res[0][1].x, oracle_op(m))
    assert ok, (name, err)
    assert csv.splitlines()[0] == "Variable,Value"


def test_reference_newton_defaults(be):
    """The command line's defaults (abs_tol 1e-4 on the Newton step) still
    land within the step tolerance's square of the root."""
    res, _ = run(be, DIODE_FORWARD, newton=None)
    _, m = system_of(DIODE_FORWARD)
    assert close(res[0][1].x, oracle_op(m), rel=1e-6, abs_=1e-6)[0]


def test_diode_forward_closed_form(be):
    """5 V = R I + n Vt ln(1 + I / Is) at node 2 (bisection on the scalar
    equation, with the descriptor's single-precision model values)."""
    res, _ = run(be, DIODE_FORWARD)
    _, m = system_of(DIODE_FORWARD)
    d = m.diode_descriptors[0]
    i_s, n_vt = float(d["is_val"]), float(d["n_vt"])
    gmin = 1e-12  # the compiler's shunt at each node
    lo, hi = 0.0, 5.0
    for _ in range(200):
        v = 0.5 * (lo + hi)
        f = (5.0 - v) / 1e3 - i_s * math.expm1(v / n_vt) - gmin * v
        lo, hi = (v, hi) if f > 0 else (lo, v)
    v2 = dict(res[0][1].node_voltages)["2"]
    assert abs(v2 - 0.5 * (lo + hi)) < 1e-9 + 1e-6 * v2


def test_nmos_common_source_closed_form(be):
    """Saturation: Id = kp/2 (Vgs - Vt)^2 (1 + lambda Vds), Vds = 5 - 1k Id,
    which is linear in Id."""
    res, _ = run(be, NMOS_CS)
    _, m = system_of(NMOS_CS)
    d = m.mosfet_descriptors[0]
    kp, vto, lam = float(d["kp"]), float(d["vto"]), float(d["lambda"])
    a = 0.5 * kp * (2.0 - vto) ** 2
    i_d = a * (1 + lam * 5.0) / (1 + a * lam * 1e3)
    vd = dict(res[0][1].node_voltages)["drain"]
    assert vd > 2.0 - vto  # saturated
    assert abs(vd - (5.0 - 1e3 * i_d)) < 1e-6 * vd


def test_cmos_chain_logic_levels(be):
    """test_cmos_inverter_chain_10_dc (tests/transistor_integration_test.rs:553-607):
    stage outputs alternate low / high from VIN = 5."""
    res, _ = run(be, cmos_chain())
    v = dict(res[0][1].node_voltages)
    for k in range(1, 11):
        if k % 2:
            assert v[f"out_{k}"] < 0.5
        else:
            assert v[f"out_{k}"] > 4.5


# ------------------------------------------------------------------ AC --
def test_rc_lowpass_ac(be):
    """rc_lowpass.spice / test_rc_lowpass_ac_sweep: H = 1 / (1 + j w RC) at
    every frequency, the oracle's complex solve, and the CSV columns."""
    res, text = run(be, RC_LOWPASS)
    r = res[0][1]
    f = r.frequencies
    assert f.shape[0] == 60 and f[0] == 1.0
    v2 = dict(r.node_voltages)["2"]
    h = 1.0 / (1.0 + 2j * np.pi * f * 1e3 * 1e-6)
    assert close(np.abs(v2), np.abs(h), rel=1e-6)[0]
    assert close(np.angle(v2), np.angle(h), abs_=1e-6)[0]
    _, m = system_of(RC_LOWPASS)
    for k in (0, 22, 59):
        a = oracle.from_triplet_arrays(*_complex_system(m, f[k]))
        ref = oracle.dense_solve(a, np.concatenate([m.b_ac.real, m.b_ac.imag]))
        xk = np.array([w[k] for _, w in r.node_voltages] + [w[k] for _, w in r.branch_currents])
        assert close(np.concatenate([xk.real, xk.imag]), ref)[0]
    head, rows = csv_table(text)
    assert head == ["Frequency", "V(1)_mag", "V(1)_phase_deg", "V(2)_mag", "V(2)_phase_deg",
                    "I(V1)_mag", "I(V1)_phase_deg"]
    assert rows.shape == (60, 7)
    assert close(rows[:, 3], np.abs(h))[0]
    assert close(rows[:, 4], np.degrees(np.angle(h)), abs_=1e-6)[0]


def _complex_system(m, freq):
    """(G + j w C) as the real 2n x 2n block [[G, -wC], [wC, G]] in triplets."""
    n = m.size
    w = 2 * np.pi * freq
    rows, cols, vals = [], [], []
    for mat, scale, blocks in ((m.g, 1.0, ((0, 0), (n, n))), (m.c, w, ((n, 0),)),
                               (m.c, -w, ((0, n),))):
        rp = mat.row_pointers.astype(np.int64)
        for r in range(n):
            for q in range(rp[r], rp[r + 1]):
                for br, bc in blocks:
                    rows.append(br + r)
                    cols.append(bc + int(mat.col_indices[q]))
                    vals.append(scale * mat.values[q])
    return 2 * n, 2 * n, np.array(rows), np.array(cols), np.array(vals)


@pytest.mark.parametrize("spec,count", [("LIN 100 1 1000", 100), ("OCT 5 100 800", 15),
                                        ("LIN 3 100 1000", 3)])
def test_ac_sweep_grids(be, spec, count):
    """test_ac_linear_sweep / test_ac_octave_sweep / test_ac_csv_output."""
    res, text = run(be, f"V1 1 0 AC 1 0\nR1 1 2 1k\nC1 2 0 1u\n.AC {spec}\n.END\n")
    r = res[0][1]
    assert r.frequencies.shape[0] == count
    h = 1.0 / (1.0 + 2j * np.pi * r.frequencies * 1e-3)
    assert close(np.abs(dict(r.node_voltages)["2"]), np.abs(h))[0]
    lines = text.strip().splitlines()
    assert len(lines) == count + 1
    assert all(ln.count(",") == lines[0].count(",") for ln in lines)


# ----------------------------------------------------------- transient --
def oracle_tran(text, tran):
    c, m = system_of(text)
    src, pwl = T.source_table(c, m)
    if tran.uic:
        a, b = T.uic_system(c, m)
        x0 = oracle.dense_solve(a, b)
    else:
        x0 = oracle_op(m)
    times, waves, _, _, _ = oracle.transient(m, src, pwl, method="adaptive", tstep=tran.tstep,
                                             tstop=tran.tstop, tstart=tran.tstart, x0=x0,
                                             max_records=int(tran.tstop / tran.tstep * 40) + 2)
    return m, times, waves


def check_against_oracle(text, r, a, common_prefix=False):
    """Sample-for-sample against the oracle's run. ``common_prefix``: a
    waveform source keeps the step controller busy, so step sizes come from
    LTE estimates (which carry the solves' 1e-13 differences into the times
    at ~1e-10) and an accept/reject decision taken at the threshold itself
    may go the other way; compare the samples up to the first time the two
    step sequences part (at least half of the run)."""
    m, times, waves = oracle_tran(text, a)
    w = r.result.waves[:, 0, :]
    if common_prefix:
        n = min(times.shape[0], r.times.shape[0])
        same = np.isclose(r.times[:n], times[:n], rtol=1e-8, atol=0)
        k = n if same.all() else int(np.argmin(same))
        assert k >= n // 2, (k, n)
        times, waves, w = times[:k], waves[:k], w[:k]
    else:
        assert r.times.shape == times.shape, (r.times.shape, times.shape)
        assert np.allclose(r.times, times, rtol=1e-12, atol=0)
    ok, err = close(w, waves)
    assert ok, err
    return m


@pytest.mark.parametrize("name,text,tau,final,node", [
    ("tran_rc_charging", TRAN_RC, 1e-3, 5.0, "2"),
    ("tran_rl_step", TRAN_DIVIDER, 0.5e-3, 2.5, "2"),
])
def test_step_responses(be, name, text, tau, final, node):
    """Regression transients: oracle sample-for-sample, closed form at the
    reference's 2 % (tests/transient_integration_test.rs:52-87)."""
    res, csv = run(be, text)
    a, r = res[0]
    check_against_oracle(text, r, a)
    v = dict(r.node_voltages)[node]
    expect = final * (1 - np.exp(-r.times / tau))
    assert np.all(np.abs(v - expect) < 0.02 * final + 1e-6)
    head, rows = csv_table(csv)
    assert head[0] == "time" and rows.shape[0] == r.times.shape[0]
    assert close(rows[:, 0], r.times, rel=1e-15)[0]


def test_pulse_rc_filter(be):
    res, _ = run(be, TRAN_PULSE)
    a, r = res[0]
    check_against_oracle(TRAN_PULSE, r, a)
    v = dict(r.node_voltages)["2"]
    k = np.searchsorted(r.times, 0.25e-3)
    assert 0.5 < v[k] < 5.0


def test_rl_branch_current(be):
    """test_rl_step_response: I(L1) = V/R (1 - exp(-t R / L))."""
    text = "V1 1 0 DC 5\nR1 1 2 100\nL1 2 0 10m\n.TRAN 1u 500u UIC\n.END\n"
    res, _ = run(be, text)
    a, r = res[0]
    check_against_oracle(text, r, a)
    i = dict(r.branch_currents)["L1"]
    t = r.times
    expect = 0.05 * (1 - np.exp(-t / 1e-4))
    assert np.all(np.abs(i - expect)[t > 1e-6] < 0.02 * 0.05 + 1e-6)


def test_lc_oscillation_from_operating_point(be):
    """test_lc_oscillation: no UIC, so the start is the DC operating point."""
    text = "V1 1 0 DC 5\nL1 1 2 1m\nC1 2 0 1u\n.TRAN 1u 200u\n.END\n"
    res, _ = run(be, text)
    a, r = res[0]
    check_against_oracle(text, r, a)
    v = dict(r.node_voltages)["2"]
    assert v.max() > 4.9 and r.times.shape[0] > 10


@pytest.mark.parametrize("src", ["SIN(0 1 10k)", "SIN(0.5 1 10k 20u 5k)", "PWL(0 0 50u 1 100u -1 150u 0)",
                                 "EXP(0 2 10u 20u 100u 30u)"])
def test_waveform_sources(be, src):
    """test_sin_source: a voltage source fixes its node, so V(1) is the
    waveform at every recorded time; an RC behind it checks the stepping."""
    text = f"V1 1 0 {src}\nR1 1 0 1k\nR2 1 2 1k\nC1 2 0 10n\n.TRAN 1u 200u\n.END\n"
    res, _ = run(be, text)
    a, r = res[0]
    m = check_against_oracle(text, r, a, common_prefix=True)
    c, _ = system_of(text)
    table, pwl = T.source_table(c, m)
    expect = np.array([oracle.source_value(table[0], t, pwl) for t in r.times])
    # t = 0 holds the operating point, where the source has no DC value (0 V)
    v1 = dict(r.node_voltages)["1"]
    assert v1[0] == pytest.approx(0.0, abs=1e-12)
    assert close(v1[1:], expect[1:], abs_=1e-9)[0]


def test_nonlinear_transient_starts_at_operating_point(be):
    """Without UIC the transient starts from dc::run's Newton answer
    (src/analysis/transient.rs:43-56); the diode stays in the time loop
    (ohm_transient_nonlinear), checked against the oracle's per-step LU
    Newton up to the first point where the LTE-driven step sequences part."""
    text = ("V1 1 0 DC 5\nR1 1 2 1k\n.MODEL DMOD D(IS=1e-14 N=1.0)\nD1 2 0 DMOD\nC1 2 0 1u\n"
            ".TRAN 10u 100u\n.END\n")
    res, _ = run(be, text)
    a, r = res[0]
    c, m = system_of(text)
    x0 = oracle_op(m)
    assert close(r.result.waves[0, 0, :], x0)[0]
    src, pwl = T.source_table(c, m)
    times, waves, _, _, _ = oracle.transient_nonlinear(m, src, pwl, method="adaptive",
                                                       tstep=a.tstep, tstop=a.tstop, x0=x0)
    n = min(times.shape[0], r.times.shape[0])
    same = np.isclose(r.times[:n], times[:n], rtol=1e-8, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= n // 2
    ok, err = close(r.result.waves[:k, 0, :], waves[:k])
    assert ok, err
    # at the operating point nothing moves
    assert np.abs(r.result.waves[:, 0, :] - x0).max() < 1e-6


def test_cli_end_to_end(tmp_path):
    """python -m ohmnivore_amd netlist --stats: CSV on stdout, stats on stderr."""
    p = tmp_path / "divider.spice"
    p.write_text(VOLTAGE_DIVIDER + "\n")
    env = dict(os.environ, PYTHONPATH=ROOT)
    out = subprocess.run([sys.executable, "-m", "ohmnivore_amd", str(p), "--stats"], cwd=ROOT,
                         env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.splitlines()[0] == "Variable,Value"
    assert "=== Ohmnivore Performance Stats ===" in out.stderr and "DC analysis" in out.stderr
    bad = tmp_path / "bad.spice"
    bad.write_text("R1 1 0 1k\nX1 1 0 what\n")
    out = subprocess.run([sys.executable, "-m", "ohmnivore_amd", str(bad)], cwd=ROOT, env=env,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 1
    assert out.stderr.strip() == "Parse error: line 2: unknown element 'X': X1 1 0 what"


NETLISTS = os.path.join(ROOT, "tests", "golden", "netlists")
# The CMOS chain's transient keeps its MOSFETs in the time loop, which the
# linear transient oracle used here cannot check: tests/test_gpu_trannl.py
# covers it.
FIXTURES = sorted(f for f in os.listdir(NETLISTS)
                  if f.endswith(".spice") and f != "cmos_inverter_chain_tran.spice")


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_netlists_end_to_end(be, name):
    """Every regression / example / fixture netlist of the reference through
    parse -> compile -> GPU analysis -> CSV, each result against the oracle."""
    with open(os.path.join(NETLISTS, name)) as f:
        text = f.read()
    res, csv = run(be, text)
    _, m = system_of(text)
    a, r = res[0]
    if isinstance(a, nl.Dc):
        ok, err = close(r.x, oracle_op(m))
        assert ok, err
        assert csv.startswith("Variable,Value\n")
    elif isinstance(a, nl.Ac):
        for k in (0, r.frequencies.shape[0] // 2, r.frequencies.shape[0] - 1):
            am = oracle.from_triplet_arrays(*_complex_system(m, r.frequencies[k]))
            ref = oracle.dense_solve(am, np.concatenate([m.b_ac.real, m.b_ac.imag]))
            xk = np.array([w[k] for _, w in r.node_voltages] + [w[k] for _, w in r.branch_currents])
            assert close(np.concatenate([xk.real, xk.imag]), ref)[0]
        assert csv.startswith("Frequency,")
    else:
        check_against_oracle(text, r, a)
        assert csv.startswith("time,")
