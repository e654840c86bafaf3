"""The reference's device integration circuits through the GPU DC cascade.

The netlists are the end-to-end circuits of tests/transistor_integration_test.rs
(NPN/PNP, NMOS/PMOS, cutoff, threshold and subthreshold bias, high-BF
limiting, 4-terminal MOSFET syntax, the unloaded CMOS inverter, the mixed
diode + BJT + MOSFET circuit) and tests/diode_integration_test.rs (two
diodes, a diode between two non-ground nodes), kept here as test text. Each
operating point is checked twice: against the oracle's LU Newton (with its
source-stepping fallback) at 1e-6 relative / 1e-9 absolute, and against the
bounds the reference's own test asserts (file:line in each case).
"""
import numpy as np
import pytest

import oracle
from tests.test_gpu_dc import oracle_source_stepping
from tests.test_gpu_netlist import close, run, system_of

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    from ohmnivore_amd import solver as S
    return S.HipBackend()


def oracle_point(m):
    """The oracle's LU Newton from zero to 1e-13; where that does not
    converge (the mixed circuit: a MOSFET step limit cycles), its Newton
    along the cascade's source-stepping schedule."""
    try:
        x, _, _ = oracle.newton_solve(m.g, m.b_dc, m.diode_descriptors, m.bjt_descriptors,
                                      m.mosfet_descriptors, abs_tol=1e-13, max_iterations=200)
        return x
    except oracle.OracleError:
        return oracle_source_stepping(m)


def nl(*lines):
    return "\n".join(lines + (".DC", ".END", ""))


CASES = {
    # tests/transistor_integration_test.rs:279-307
    "pnp_basic": (nl("* PNP Basic", "V1 vcc 0 DC 5", "V2 base 0 DC 4.3", "R1 0 vc 1k",
                     ".MODEL QPNP PNP(IS=1e-14 BF=100)", "Q1 vc base vcc QPNP"),
                  lambda v: 0.0 < v["vc"] < 5.0),
    # :309-337
    "bjt_cutoff": (nl("* BJT Cutoff - both junctions reverse biased", "V1 vcc 0 DC 5",
                      "V2 base 0 DC 0", "R1 vcc vc 1k", ".MODEL QNPN NPN(IS=1e-14 BF=200)",
                      "Q1 vc base 0 QNPN"),
                   lambda v: abs(v["base"]) < 0.1 and abs(v["vc"] - 5.0) < 0.5),
    # :366-389
    "pmos_basic": (nl("* PMOS Basic", "V1 vdd 0 DC 5", "V2 gate 0 DC 3.0", "R1 0 drain 1k",
                      ".MODEL PMOD PMOS(VTO=0.7 KP=4.5e-5 LAMBDA=0.05)", "M1 drain gate vdd PMOD"),
                   lambda v: 0.0 < v["drain"] < 5.0),
    # :391-414
    "mosfet_cutoff": (nl("* MOSFET Cutoff", "V1 vdd 0 DC 5", "V2 gate 0 DC 0.3", "R1 vdd drain 1k",
                         ".MODEL NMOD NMOS(VTO=0.7 KP=1.1e-4)", "M1 drain gate 0 NMOD"),
                      lambda v: abs(v["drain"] - 5.0) < 0.5),
    # :416-458
    "mixed_diode_bjt_mosfet": (
        nl("* Mixed circuit with all three nonlinear devices", "V1 vcc 0 DC 5", "R1 vcc n1 1k",
           ".MODEL DMOD D(IS=1e-14 N=1.0)", "D1 n1 n2 DMOD", "R2 n2 n3 10k",
           ".MODEL QMOD NPN(IS=1e-14 BF=100)", "Q1 vcc n2 n3 QMOD", "R3 n3 n4 1k",
           ".MODEL MMOD NMOS(VTO=0.7 KP=1e-4)", "M1 n4 n2 0 MMOD"),
        lambda v: all(-0.5 <= x <= 5.5 for x in v.values())),
    # :460-481
    "bjt_high_bf": (nl("* BJT with high BF tests voltage limiting", "V1 vcc 0 DC 5",
                       "V2 base 0 DC 0.7", "R1 vcc vc 1k", ".MODEL QHBF NPN(IS=1e-14 BF=1000)",
                       "Q1 vc base 0 QHBF"),
                    lambda v: 0.0 <= v["vc"] <= 5.0),
    # :483-509
    "mosfet_4terminal": (nl("* 4-terminal MOSFET syntax (bulk ignored)", "V1 vdd 0 DC 5",
                            "V2 gate 0 DC 2.0", "R1 vdd drain 1k",
                            ".MODEL NMOD NMOS(VTO=0.7 KP=1.1e-4)", "M1 drain gate 0 0 NMOD"),
                         lambda v: 4.0 < v["drain"] < 5.1),
    # :511-530
    "cmos_inverter_dc": (nl("* CMOS Inverter - no resistive load", "VDD vdd 0 DC 5",
                            "VIN in 0 DC 5", ".MODEL NMOD NMOS(VTO=0.7 KP=1.1e-4 LAMBDA=0.04)",
                            ".MODEL PMOD PMOS(VTO=-0.7 KP=5.5e-5 LAMBDA=0.04)",
                            "MP1 out in vdd vdd PMOD", "MN1 out in 0 0 NMOD"),
                         lambda v: v["out"] < 0.5),
    # :609-632
    "mosfet_at_threshold": (nl("* NMOS at threshold", "V1 vdd 0 DC 5", "V2 gate 0 DC 0.7",
                               "R1 vdd drain 10k",
                               ".MODEL NMOD NMOS(VTO=0.7 KP=1.1e-4 LAMBDA=0.04)",
                               "M1 drain gate 0 NMOD"),
                            lambda v: 3.0 < v["drain"] <= 5.01),
    # :634-657
    "mosfet_subthreshold": (nl("* NMOS subthreshold", "V1 vdd 0 DC 5", "V2 gate 0 DC 0.5",
                               "R1 vdd drain 10k",
                               ".MODEL NMOD NMOS(VTO=0.7 KP=1.1e-4 LAMBDA=0.04)",
                               "M1 drain gate 0 NMOD"),
                            lambda v: abs(v["drain"] - 5.0) < 0.5),
    # tests/diode_integration_test.rs:34-55 (compiled there; solved here)
    "two_diodes": (nl("V1 1 0 DC 5", "R1 1 2 1k", "R2 2 3 1k", ".MODEL DMOD D(IS=1e-14 N=1.0)",
                      "D1 2 0 DMOD", "D2 3 0 DMOD"),
                   lambda v: 0.3 < v["3"] < v["2"] < 1.0),
    # tests/diode_integration_test.rs:57-86
    "diode_between_nodes": (nl("V1 1 0 DC 5", "R1 1 2 1k", "R2 3 0 1k",
                               ".MODEL DMOD D(IS=1e-14 N=1.0)", "D1 2 3 DMOD"),
                            lambda v: 0.3 < v["2"] - v["3"] < 1.0 and 2.0 < v["3"] < 2.5),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_circuit_operating_point(be, name):
    text, bounds = CASES[name]
    res, csv = run(be, text)
    r = res[0][1]
    _, m = system_of(text)
    ok, err = close(r.x, oracle_point(m))
    assert ok, (name, err, r.stage)
    v = dict(r.node_voltages)
    assert bounds(v), (name, v)
    assert csv.splitlines()[0] == "Variable,Value"
    assert np.isfinite(r.x).all()


def test_pnp_mirrors_npn(be):
    """A PNP stage is the NPN stage with every voltage and current negated
    (the compiler's polarity flag, src/compiler.rs:482-522)."""
    npn = nl("V1 vcc 0 DC 5", "V2 base 0 DC 0.7", "R1 vcc vc 1k",
             ".MODEL Q NPN(IS=1e-14 BF=100)", "Q1 vc base 0 Q")
    pnp = nl("V1 vcc 0 DC -5", "V2 base 0 DC -0.7", "R1 vcc vc 1k",
             ".MODEL Q PNP(IS=1e-14 BF=100)", "Q1 vc base 0 Q")
    a, _ = run(be, npn)
    b, _ = run(be, pnp)
    xa, xb = a[0][1].x, b[0][1].x
    ok, err = close(xb, -xa)
    assert ok, err


# -------------------------------------------- tests/gpu_integration_test.rs --
def test_current_source_resistor_network(be):
    """gpu_vs_cpu_current_source_resistor_network_dc (:109-140): 1 mA into
    R1 + (R2 || R3): V(2) = 2/3 V, V(1) = 5/3 V."""
    text = nl("I1 0 1 DC 0.001", "R1 1 2 1k", "R2 2 0 2k", "R3 2 0 1k")
    res, _ = run(be, text)
    v = dict(res[0][1].node_voltages)
    assert abs(v["2"] - 2.0 / 3.0) < 1e-6 and abs(v["1"] - 5.0 / 3.0) < 1e-6
    _, m = system_of(text)
    ok, err = close(res[0][1].x, oracle.dense_solve(m.g, m.b_dc))
    assert ok, err


def test_gmin_voltage_source_system(be):
    """ds_bicgstab_ill_conditioned_matrix (:201-237): GMIN diagonals and one
    voltage-source stamp (condition number ~1e12) through solve_real."""
    from ohmnivore_amd import circuit as ir
    from ohmnivore_amd import solver as S
    gmin = 1e-12
    a = ir.CsrMatrix(3, 3, np.array([gmin, 1.0, gmin, 1.0]), np.array([0, 2, 1, 0], np.uint32),
                     np.array([0, 2, 3, 4], np.uint32))
    gs = S.GpuSolver(be, params=S.bicgstab_params(tolerance=1e-12))
    x = gs.solve_real(a, np.array([0.0, 0.0, 5.0]))
    ref = oracle.dense_solve(a, np.array([0.0, 0.0, 5.0]))
    ok, err = close(x, ref)
    assert ok, (x, ref, err)
    assert abs(x[0] - 5.0) < 1e-9 and abs(x[1]) < 1e-9


def test_rlc_series_ac(be):
    """gpu_vs_cpu_rlc_series_ac (:299-350): V(3) = Z_C / (R + Z_L + Z_C)
    over DEC 20 100 100k (the reference checks every 5th point at 5 %;
    here every point at 1e-6, and against the oracle's complex solve)."""
    from tests.test_gpu_netlist import _complex_system
    text = "V1 1 0 AC 1 0\nR1 1 2 100\nL1 2 3 10m\nC1 3 0 1u\n.AC DEC 20 100 100000\n.END\n"
    res, _ = run(be, text)
    r = res[0][1]
    f = r.frequencies
    assert f.shape[0] == 60 and f[0] == 100.0  # the reference's DEC grid stops short of f_stop
    w = 2 * np.pi * f
    zc, zl = 1.0 / (1j * w * 1e-6), 1j * w * 10e-3
    h = zc / (100.0 + zl + zc)
    v3 = dict(r.node_voltages)["3"]
    assert close(np.abs(v3), np.abs(h))[0]
    # |V_C| peaks at sqrt(1/(LC) - R^2/(2 L^2)) / (2 pi) = 1125 Hz (Q = 1): the
    # nearest point of the 20-per-decade grid
    k = int(np.argmax(np.abs(v3)))
    f_peak = np.sqrt(1.0 / (10e-3 * 1e-6) - 100.0 ** 2 / (2 * 10e-3 ** 2)) / (2 * np.pi)
    assert abs(np.log10(f[k] / f_peak)) < 1.0 / 40.0
    _, m = system_of(text)
    for q in (0, 20, 40, 59):
        a = oracle.from_triplet_arrays(*_complex_system(m, f[q]))
        ref = oracle.dense_solve(a, np.concatenate([m.b_ac.real, m.b_ac.imag]))
        xq = np.array([wv[q] for _, wv in r.node_voltages] + [wv[q] for _, wv in r.branch_currents])
        ok, err = close(np.concatenate([xq.real, xq.imag]), ref)
        assert ok, (q, err)
