"""Pins for the CPU oracle: the reference's own known-answer tests, ported.

Each test names the reference test it restates. These run on the CPU only
and check the oracle (oracle/ohm_oracle.c), which the GPU parity tests then
compare against.
"""
import cmath
import math

import numpy as np
import pytest

import oracle
from ohmnivore_amd import circuit as ir

VT = 0.02585


def ckt_with(build):
    c = ir.Circuit()
    build(c)
    return c


# --------------------------------------------------------------------------
# src/sparse.rs:184-227
# --------------------------------------------------------------------------
def test_sparse_empty_matrix():
    m = oracle.from_triplets(3, 3, [])
    assert m.nnz == 0
    assert list(m.row_pointers) == [0, 0, 0, 0]


def test_sparse_from_triplets_simple():
    m = oracle.from_triplets(2, 2, [(0, 0, 1.0), (1, 1, 1.0)])
    assert m.nnz == 2
    assert np.array_equal(m.to_dense(), [[1.0, 0.0], [0.0, 1.0]])


def test_sparse_duplicates_summed():
    m = oracle.from_triplets(2, 2, [(0, 0, 1.0), (0, 0, 2.0), (1, 1, 3.0)])
    assert m.nnz == 2
    assert np.array_equal(m.to_dense(), [[3.0, 0.0], [0.0, 3.0]])


def test_sparse_spmv():
    m = oracle.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 1, 3.0)])
    assert list(oracle.spmv(m, np.array([1.0, 2.0]))) == [4.0, 6.0]


def test_sparse_duplicate_sum_order_is_insertion_order():
    # Bit-level: (a + b) + c with a=1e16, b=1, c=-1e16 differs from other orders.
    m = oracle.from_triplets(1, 1, [(0, 0, 1e16), (0, 0, 1.0), (0, 0, -1e16)])
    assert m.values[0] == (1e16 + 1.0) + -1e16
    m2 = oracle.from_triplets(1, 1, [(0, 0, 1.0), (0, 0, 1e16), (0, 0, -1e16)])
    assert m2.values[0] == (1.0 + 1e16) + -1e16


def test_form_complex_matrix():
    # G = [[1,0],[0,2]], C = [[0,0],[0,1]], omega=1 -> diag(1, 2+1j)
    g = oracle.from_triplets(2, 2, [(0, 0, 1.0), (1, 1, 2.0)])
    c = oracle.from_triplets(2, 2, [(1, 1, 1.0)])
    rows = np.concatenate([np.repeat(np.arange(2), np.diff(g.row_pointers.astype(int))),
                           np.repeat(np.arange(2), np.diff(c.row_pointers.astype(int)))])
    cols = np.concatenate([g.col_indices, c.col_indices]).astype(np.int64)
    vals = np.concatenate([g.values.astype(complex), 1j * 1.0 * c.values])
    a = oracle.from_triplet_arrays(2, 2, rows, cols, vals)
    d = a.to_dense()
    assert d[0, 0] == 1 + 0j and d[1, 1] == 2 + 1j


# --------------------------------------------------------------------------
# src/solver/cpu.rs:187-397
# --------------------------------------------------------------------------
def test_cpu_identity_2x2():
    a = oracle.from_triplets(2, 2, [(0, 0, 1.0), (1, 1, 1.0)])
    x = oracle.dense_solve(a, [3.0, 7.0])
    assert np.allclose(x, [3, 7], atol=1e-12)


def test_cpu_known_2x2():
    a = oracle.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 0, 5.0), (1, 1, 7.0)])
    x = oracle.dense_solve(a, [11.0, 13.0])
    assert abs(x[0] - 64 / 9) < 1e-12 and abs(x[1] + 29 / 9) < 1e-12


def test_cpu_3x3_needs_pivoting():
    a = oracle.from_triplets(3, 3, [(0, 1, 2.0), (0, 2, 1.0), (1, 0, 1.0), (1, 1, 1.0),
                                    (1, 2, 1.0), (2, 0, 2.0), (2, 1, 1.0)])
    x = oracle.dense_solve(a, [5.0, 4.0, 4.0])
    assert np.allclose(x, [1, 2, 1], atol=1e-12)


def test_cpu_singular():
    a = oracle.from_triplets(2, 2, [(0, 0, 1.0), (0, 1, 2.0), (1, 0, 2.0), (1, 1, 4.0)])
    with pytest.raises(oracle.OracleError, match="singular"):
        oracle.dense_solve(a, [3.0, 6.0])


def test_cpu_dimension_mismatch():
    a = oracle.from_triplets(2, 2, [(0, 0, 1.0), (1, 1, 1.0)])
    with pytest.raises(oracle.OracleError, match="dimension"):
        oracle.dense_solve(a, [1.0, 2.0, 3.0])


def test_cpu_nonsquare():
    a = oracle.from_triplets(2, 2, [(0, 0, 1.0), (1, 1, 1.0)])
    a.ncols = 3
    with pytest.raises(oracle.OracleError):
        oracle.dense_solve(a, [1.0, 2.0])


def test_cpu_1x1_and_empty():
    a = oracle.from_triplets(1, 1, [(0, 0, 5.0)])
    assert abs(oracle.dense_solve(a, [15.0])[0] - 3.0) < 1e-12
    e = oracle.from_triplets(0, 0, [])
    assert oracle.dense_solve(e, np.zeros(0)).shape == (0,)


def test_cpu_complex_identity_and_known():
    one = 1 + 0j
    a = oracle.from_triplets(2, 2, [(0, 0, one), (1, 1, one)])
    x = oracle.dense_solve(a, np.array([3 + 1j, 7 - 2j]))
    assert np.allclose(x, [3 + 1j, 7 - 2j], atol=1e-12)
    a = oracle.from_triplets(2, 2, [(0, 0, 1 + 1j), (0, 1, 2 + 0j), (1, 1, 1 - 1j)])
    x = oracle.dense_solve(a, np.array([5 + 1j, 2 - 2j]))
    assert np.allclose(x, [1, 2], atol=1e-12)


def test_cpu_complex_singular_and_pivoting():
    a = oracle.from_triplets(2, 2, [(0, 0, 1 + 1j), (0, 1, 2 + 2j), (1, 0, 1 + 0j), (1, 1, 2 + 0j)])
    with pytest.raises(oracle.OracleError):
        oracle.dense_solve(a, np.array([1 + 0j, 1 + 0j]))
    a = oracle.from_triplets(2, 2, [(0, 1, 1 + 1j), (1, 0, 1 + 0j), (1, 1, 2 + 0j)])
    x = oracle.dense_solve(a, np.array([1 + 1j, 3 + 0j]))
    assert np.allclose(x, [1, 1], atol=1e-12)


def test_cpu_verifies_with_spmv():
    a = oracle.from_triplets(3, 3, [(0, 0, 4.0), (0, 1, -1.0), (1, 0, -1.0), (1, 1, 4.0),
                                    (1, 2, -1.0), (2, 1, -1.0), (2, 2, 4.0)])
    b = np.array([1.0, 5.0, 10.0])
    x = oracle.dense_solve(a, b)
    assert np.allclose(oracle.spmv(a, x), b, atol=1e-10)


# --------------------------------------------------------------------------
# src/solver/sparse_direct.rs:67-86 (the faer path's only pins)
# --------------------------------------------------------------------------
def test_sparse_lu_known_2x2():
    a = oracle.from_triplets(2, 2, [(0, 0, 2.0), (0, 1, 1.0), (1, 0, 5.0), (1, 1, 7.0)])
    x = oracle.sparse_lu_solve(a, np.array([11.0, 13.0]))
    assert abs(x[0] - 64 / 9) < 1e-10 and abs(x[1] + 29 / 9) < 1e-10


def test_sparse_lu_singular():
    a = oracle.from_triplets(2, 2, [(0, 0, 1.0), (1, 0, 1.0)])
    with pytest.raises(oracle.OracleError, match="Solve error"):
        oracle.sparse_lu_solve(a, np.array([1.0, 1.0]))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sparse_lu_matches_scipy_and_dense(seed):
    """Independent pin: SuperLU (scipy) and the dense GE agree with the banded LU."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    rng = np.random.default_rng(seed)
    ckt = ir.diode_mesh(12, 11, seed=seed + 10)
    m = oracle.compile(ckt)
    g = m.g
    g.values = g.values + rng.uniform(0, 1e-3, g.nnz) * (g.values != 0)
    b = rng.standard_normal(g.nrows)
    x = oracle.sparse_lu_solve(g, b)
    A = sp.csr_matrix((g.values, g.col_indices, g.row_pointers), shape=(g.nrows, g.nrows))
    xs = spl.spsolve(A.tocsc(), b)
    xd = oracle.dense_solve(g, b)
    assert np.allclose(x, xs, rtol=1e-9, atol=1e-12)
    assert np.allclose(x, xd, rtol=1e-9, atol=1e-12)


def test_sparse_lu_complex_matches_scipy():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    rng = np.random.default_rng(7)
    n = 40
    dense = np.zeros((n, n), complex)
    for i in range(n):
        dense[i, i] = 2 + 1j * rng.uniform()
        if i + 1 < n:
            dense[i, i + 1] = -1 + 0.1j
            dense[i + 1, i] = -0.5j
    dense[0, n - 1] = 0.3
    rows, cols = np.nonzero(dense)
    a = oracle.from_triplet_arrays(n, n, rows, cols, dense[rows, cols])
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = oracle.sparse_lu_solve(a, b)
    assert np.allclose(x, spl.spsolve(sp.csc_matrix(dense), b), rtol=1e-10)


# --------------------------------------------------------------------------
# src/compiler.rs:773-1981
# --------------------------------------------------------------------------
def test_compile_single_resistor_to_ground():
    m = oracle.compile(ckt_with(lambda c: c.resistor("R1", "1", "0", 1000.0)))
    assert m.size == 1 and m.node_names == ["1"] and m.branch_names == []
    assert abs(m.g.to_dense()[0, 0] - 0.001) < 1e-9


def _divider(c):
    c.vsource("V1", "1", "0", dc=5.0)
    c.resistor("R1", "1", "2", 1000.0)
    c.resistor("R2", "2", "0", 1000.0)


def test_compile_voltage_divider():
    m = oracle.compile(ckt_with(_divider))
    assert m.size == 3 and m.node_names == ["1", "2"] and m.branch_names == ["V1"]
    g = m.g.to_dense()
    assert np.allclose(g, [[0.001, -0.001, 1.0], [-0.001, 0.002, 0.0], [1.0, 0.0, 0.0]], atol=1e-9)
    assert np.allclose(m.b_dc, [0, 0, 5.0])


def test_compile_rc_circuit():
    def b(c):
        c.vsource("V1", "1", "0", ac=(1.0, 0.0))
        c.resistor("R1", "1", "2", 100.0)
        c.capacitor("C1", "2", "0", 1e-6)
    m = oracle.compile(ckt_with(b))
    g, cc = m.g.to_dense(), m.c.to_dense()
    assert abs(g[0, 0] - 0.01) < 1e-9 and abs(g[0, 1] + 0.01) < 1e-9 and abs(g[0, 2] - 1) < 1e-9
    assert abs(g[1, 0] + 0.01) < 1e-9 and abs(g[1, 1] - 0.01) < 1e-9 and abs(g[2, 0] - 1) < 1e-9
    assert abs(cc[1, 1] - 1e-6) < 1e-18 and abs(cc[0, 0]) < 1e-18 and abs(cc[0, 1]) < 1e-18
    assert abs(m.b_ac[2] - 1) < 1e-9


def test_compile_inductor_stamps():
    m = oracle.compile(ckt_with(lambda c: c.inductor("L1", "1", "2", 0.001)))
    assert m.size == 3 and m.branch_names == ["L1"]
    g = m.g.to_dense()
    assert g[0, 2] == 1 and g[2, 0] == 1 and g[1, 2] == -1 and g[2, 1] == -1
    assert g[0, 0] == 1e-12 and g[1, 1] == 1e-12 and g[2, 2] == 0  # GMIN on nodes only
    assert abs(m.c.to_dense()[2, 2] + 0.001) < 1e-9


def test_compile_current_sources():
    m = oracle.compile(ckt_with(lambda c: c.isource("I1", "1", "0", dc=2.0)))
    assert m.size == 1 and abs(m.b_dc[0] + 2.0) < 1e-9
    m = oracle.compile(ckt_with(lambda c: c.isource("I1", "1", "2", dc=3.0, ac=(1.0, 90.0))))
    assert np.allclose(m.b_dc, [-3.0, 3.0])
    ac = cmath.rect(1.0, math.radians(90.0))
    assert abs(m.b_ac[0] + ac) < 1e-9 and abs(m.b_ac[1] - ac) < 1e-9


def test_compile_ground_variants():
    def b(c):
        for k, gname in enumerate(["0", "GND", "gnd", "Gnd"]):
            c.resistor(f"R{k + 1}", str(k + 1), gname, 100.0)
    m = oracle.compile(ckt_with(b))
    assert m.size == 4 and len(m.node_names) == 4
    assert not any(ir.is_ground(n) for n in m.node_names)


def test_compile_series_and_parallel():
    def b(c):
        c.resistor("R1", "1", "2", 100.0)
        c.resistor("R2", "2", "0", 200.0)
    g = oracle.compile(ckt_with(b)).g.to_dense()
    assert np.allclose(g, [[0.01, -0.01], [-0.01, 0.015]], atol=1e-9)

    def p(c):
        c.resistor("R1", "1", "0", 100.0)
        c.resistor("R2", "1", "0", 200.0)
    assert abs(oracle.compile(ckt_with(p)).g.to_dense()[0, 0] - 0.015) < 1e-9


def test_compile_vsource_ac_and_negative_terminal():
    m = oracle.compile(ckt_with(lambda c: c.vsource("V1", "1", "0", dc=10.0, ac=(5.0, 45.0))))
    assert m.size == 2 and abs(m.b_dc[1] - 10.0) < 1e-9
    assert abs(m.b_ac[1] - cmath.rect(5.0, math.radians(45))) < 1e-9
    m = oracle.compile(ckt_with(lambda c: c.vsource("V1", "1", "2", dc=3.0)))
    g = m.g.to_dense()
    assert g[0, 2] == 1 and g[2, 0] == 1 and g[1, 2] == -1 and g[2, 1] == -1 and m.b_dc[2] == 3


def test_compile_empty_and_errors():
    m = oracle.compile(ir.Circuit())
    assert m.size == 0 and m.node_names == [] and m.b_dc.shape == (0,)
    with pytest.raises(oracle.OracleError):
        oracle.compile(ckt_with(lambda c: c.resistor("R1", "1", "0", 0.0)))
    with pytest.raises(oracle.OracleError, match="undefined model"):
        oracle.compile(ckt_with(lambda c: c.diode("D1", "1", "0", "NOMODEL")))
    with pytest.raises(oracle.OracleError, match="undefined model"):
        oracle.compile(ckt_with(lambda c: c.bjt("Q1", "1", "2", "0", "NOMODEL")))
    with pytest.raises(oracle.OracleError, match="undefined model"):
        oracle.compile(ckt_with(lambda c: c.mosfet("M1", "1", "2", "0", "NOMODEL")))


def test_compile_full_rlc():
    def b(c):
        c.vsource("V1", "1", "0", dc=1.0, ac=(1.0, 0.0))
        c.resistor("R1", "1", "2", 100.0)
        c.inductor("L1", "2", "3", 0.001)
        c.capacitor("C1", "3", "0", 1e-6)
    m = oracle.compile(ckt_with(b))
    assert m.size == 5 and m.node_names == ["1", "2", "3"] and m.branch_names == ["V1", "L1"]
    g, cc = m.g.to_dense(), m.c.to_dense()
    assert abs(g[0, 0] - 0.01) < 1e-9 and abs(g[0, 1] + 0.01) < 1e-9 and g[0, 3] == 1
    assert abs(g[1, 0] + 0.01) < 1e-9 and abs(g[1, 1] - 0.01) < 1e-9 and g[1, 4] == 1
    assert g[2, 4] == -1 and g[3, 0] == 1 and g[4, 1] == 1 and g[4, 2] == -1
    assert abs(cc[2, 2] - 1e-6) < 1e-18 and abs(cc[4, 4] + 0.001) < 1e-9
    assert m.b_dc[3] == 1 and abs(m.b_ac[3] - 1) < 1e-9
    # the C pattern is zero-filled into G (src/compiler.rs:577-579)
    assert m.g.value_index(4, 4) is not None and g[4, 4] == 0.0


def test_compile_multiple_vsources():
    def b(c):
        c.vsource("V1", "1", "0", dc=5.0)
        c.vsource("V2", "2", "0", dc=3.0)
    m = oracle.compile(ckt_with(b))
    assert m.size == 4 and m.branch_names == ["V1", "V2"] and list(m.b_dc[2:]) == [5.0, 3.0]


def test_compile_diode_descriptors():
    def b(c):
        c.add_diode_model("DMOD", 1e-14, 1.0)
        c.vsource("V1", "1", "0", dc=1.0)
        c.diode("D1", "1", "0", "DMOD")
    m = oracle.compile(ckt_with(b))
    d = m.diode_descriptors[0]
    assert d["anode_idx"] == 0 and d["cathode_idx"] == ir.SENTINEL
    assert d["is_val"] == np.float32(1e-14) and d["n_vt"] == np.float32(VT)
    assert m.g.value_index(0, 0) is not None
    assert d["g_row_col"][0] != ir.SENTINEL and all(d["g_row_col"][1:] == ir.SENTINEL)
    assert list(d["b_idx"]) == [0, ir.SENTINEL]


def test_compile_diode_between_nodes_and_slot_values():
    def b(c):
        c.add_diode_model("DMOD", 1e-14, 1.5)
        c.resistor("R1", "1", "0", 1000.0)
        c.resistor("R2", "2", "0", 1000.0)
        c.diode("D1", "1", "2", "DMOD")
    m = oracle.compile(ckt_with(b))
    d = m.diode_descriptors[0]
    assert d["anode_idx"] == 0 and d["cathode_idx"] == 1
    assert d["n_vt"] == np.float32(1.5 * VT)
    assert all(d["g_row_col"] != ir.SENTINEL)
    # [aa, cc, ac, ca] slots point at those CSR positions
    for slot, (r, cc) in zip(d["g_row_col"], [(0, 0), (1, 1), (0, 1), (1, 0)]):
        assert m.g.value_index(r, cc) == slot
    g = m.g.to_dense()
    assert abs(g[0, 0] - 0.001) < 1e-9 and abs(g[1, 1] - 0.001) < 1e-9
    assert g[0, 1] == 0 and g[1, 0] == 0


def test_compile_shared_model_and_ground_anode():
    def b(c):
        c.add_diode_model("DMOD", 1e-12, 2.0)
        c.resistor("R1", "1", "0", 100.0)
        c.resistor("R2", "2", "0", 100.0)
        c.diode("D1", "1", "0", "DMOD")
        c.diode("D2", "2", "0", "DMOD")
    m = oracle.compile(ckt_with(b))
    assert [int(d["anode_idx"]) for d in m.diode_descriptors] == [0, 1]
    assert all(d["is_val"] == np.float32(1e-12) for d in m.diode_descriptors)

    def b2(c):
        c.add_diode_model("DMOD")
        c.resistor("R1", "1", "0", 1000.0)
        c.diode("D1", "0", "1", "DMOD")
    d = oracle.compile(ckt_with(b2)).diode_descriptors[0]
    assert d["anode_idx"] == ir.SENTINEL and d["cathode_idx"] == 0
    assert list(d["g_row_col"] != ir.SENTINEL) == [False, True, False, False]


def test_compile_diode_ground_cathode_slot_value():
    def b(c):
        c.add_diode_model("DMOD")
        c.resistor("R1", "1", "0", 1000.0)
        c.diode("D1", "1", "0", "DMOD")
    m = oracle.compile(ckt_with(b))
    slot = m.diode_descriptors[0]["g_row_col"][0]
    assert abs(m.g.values[slot] - 0.001) < 1e-9


def test_compile_bjt_descriptor():
    def b(c):
        c.add_bjt_model("2N2222", is_=1e-14, bf=200.0, br=2.0)
        c.vsource("V1", "1", "0", dc=5.0)
        c.resistor("R1", "1", "2", 10000.0)
        c.resistor("R2", "1", "3", 1000.0)
        c.resistor("R3", "4", "0", 100.0)
        c.bjt("Q1", "3", "2", "4", "2N2222")
    m = oracle.compile(ckt_with(b))
    d = m.bjt_descriptors[0]
    assert (d["collector_idx"], d["base_idx"], d["emitter_idx"]) == (2, 1, 3)
    assert d["polarity"] == 1.0 and d["is_val"] == np.float32(1e-14)
    assert d["bf"] == 200.0 and d["br"] == 2.0 and d["nf_vt"] == np.float32(VT)
    assert all(d["g_row_col"] != ir.SENTINEL) and list(d["b_idx"]) == [2, 1, 3]
    assert list(d["node_fixed"]) == [0, 0, 0]


def test_compile_mosfet_descriptor_and_fixed_flags():
    def b(c):
        c.add_mosfet_model("NMOS1", vto=0.7, kp=1.1e-4, lam=0.04)
        c.vsource("V1", "1", "0", dc=5.0)
        c.resistor("R1", "1", "2", 1000.0)
        c.resistor("R2", "3", "0", 100.0)
        c.mosfet("M1", "2", "1", "3", "NMOS1")
    m = oracle.compile(ckt_with(b))
    d = m.mosfet_descriptors[0]
    assert (d["drain_idx"], d["gate_idx"], d["source_idx"]) == (1, 0, 2)
    assert d["vto"] == np.float32(0.7) and d["kp"] == np.float32(1.1e-4)
    assert d["lambda"] == np.float32(0.04) and all(d["g_row_col"] != ir.SENTINEL)
    assert list(d["b_idx"]) == [1, 2]
    assert d["node_fixed_packed"] == 0b010  # gate on the V1 node


def test_compile_bjt_mosfet_to_ground():
    def b(c):
        c.add_bjt_model("QMOD")
        c.resistor("R1", "1", "0", 1000.0)
        c.resistor("R2", "2", "0", 1000.0)
        c.bjt("Q1", "1", "2", "0", "QMOD")
    d = oracle.compile(ckt_with(b)).bjt_descriptors[0]
    assert d["emitter_idx"] == ir.SENTINEL
    assert list(d["g_row_col"] != ir.SENTINEL) == [1, 1, 0, 1, 1, 0, 0, 0, 0]
    assert list(d["b_idx"]) == [0, 1, ir.SENTINEL]

    def b2(c):
        c.add_mosfet_model("MMOD", vto=1.0, kp=2e-5, lam=0.0)
        c.resistor("R1", "1", "0", 1000.0)
        c.resistor("R2", "2", "0", 1000.0)
        c.mosfet("M1", "1", "2", "0", "MMOD")
    d = oracle.compile(ckt_with(b2)).mosfet_descriptors[0]
    assert d["source_idx"] == ir.SENTINEL
    assert list(d["g_row_col"] != ir.SENTINEL) == [1, 1, 0, 0, 0, 0]


def test_compile_pmos_vto_abs():
    def b(c):
        c.add_mosfet_model("P", vto=-0.8, kp=5e-5, lam=0.0, nmos=False)
        c.resistor("R1", "1", "0", 1000.0)
        c.mosfet("M1", "1", "2", "3", "P")
    d = oracle.compile(ckt_with(b)).mosfet_descriptors[0]
    assert d["vto"] == np.float32(0.8) and d["polarity"] == -1.0


def test_compile_mixed_devices():
    def b(c):
        c.add_diode_model("DMOD")
        c.add_bjt_model("QMOD")
        c.add_mosfet_model("MMOD", vto=0.7, kp=1.1e-4, lam=0.04)
        c.vsource("V1", "1", "0", dc=5.0)
        c.resistor("R1", "1", "2", 1e3)
        c.resistor("R2", "3", "0", 1e3)
        c.resistor("R3", "4", "0", 1e3)
        c.resistor("R4", "5", "0", 1e3)
        c.diode("D1", "2", "0", "DMOD")
        c.bjt("Q1", "3", "2", "4", "QMOD")
        c.mosfet("M1", "3", "2", "5", "MMOD")
    m = oracle.compile(ckt_with(b))
    assert (len(m.diode_descriptors), len(m.bjt_descriptors), len(m.mosfet_descriptors)) == (1, 1, 1)
    assert len(m.node_names) == 5 and m.size == 6


# --------------------------------------------------------------------------
# tests/integration_test.rs:50-180 (CPU solver end to end)
# --------------------------------------------------------------------------
def test_integration_voltage_divider_dc():
    def b(c):
        c.vsource("V1", "1", "0", dc=10.0)
        c.resistor("R1", "1", "2", 1e3)
        c.resistor("R2", "2", "0", 1e3)
    m = oracle.compile(ckt_with(b))
    x = oracle.dense_solve(m.g, m.b_dc)
    assert abs(x[0] - 10) < 1e-6 and abs(x[1] - 5) < 1e-6 and abs(x[2] + 0.005) < 1e-9


def test_integration_current_source_and_unequal_divider():
    def b(c):
        c.isource("I1", "0", "1", dc=0.001)
        c.resistor("R1", "1", "0", 1e3)
    m = oracle.compile(ckt_with(b))
    assert abs(oracle.dense_solve(m.g, m.b_dc)[0] - 1.0) < 1e-9

    def b2(c):
        c.vsource("V1", "1", "0", dc=10.0)
        c.resistor("R1", "1", "2", 1e3)
        c.resistor("R2", "2", "0", 2e3)
    m = oracle.compile(ckt_with(b2))
    assert abs(oracle.dense_solve(m.g, m.b_dc)[1] - 20 / 3) < 1e-6


def test_integration_rc_lowpass_cutoff():
    def b(c):
        c.vsource("V1", "1", "0", ac=(1.0, 0.0))
        c.resistor("R1", "1", "2", 1e3)
        c.capacitor("C1", "2", "0", 1e-6)
    m = oracle.compile(ckt_with(b))
    fc = 1 / (2 * math.pi * 1e3 * 1e-6)
    for f, expect, tol in [(1.0, 1.0, 0.01), (fc, 2 ** -0.5, 1e-9)]:
        a = m.g.values.astype(complex)
        a = a.copy()
        for r in range(m.size):
            for k in range(m.c.row_pointers[r], m.c.row_pointers[r + 1]):
                a[m.g.value_index(r, int(m.c.col_indices[k]))] += 2j * math.pi * f * m.c.values[k]
        A = ir.CsrMatrix(m.size, m.size, a, m.g.col_indices, m.g.row_pointers)
        x = oracle.dense_solve(A, m.b_ac)
        assert abs(abs(x[1]) - expect) < tol


# --------------------------------------------------------------------------
# src/analysis/transient_source.rs:125-212 (PULSE)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("args,t,expect", [
    ((0, 5, 1, .1, .1, 1, 3), 0.5, 0.0),
    ((0, 5, 0, 1, 1, 2, 6), 0.5, 2.5),
    ((0, 5, 0, 1, 1, 2, 6), 2.0, 5.0),
    ((0, 5, 0, 1, 1, 2, 6), 3.5, 2.5),
    ((0, 5, 0, 1, 1, 2, 6), 5.0, 0.0),
    ((0, 5, 0, 0, 0, 1, 2), 2.5, 5.0),
    ((0, 5, 0, 0, 0, 1, 2), 3.5, 0.0),
])
def test_pulse_waveform(args, t, expect):
    assert abs(oracle.eval_pulse(*args, t) - expect) < 1e-12
    assert abs(ir.Pulse(*args)(t) - expect) < 1e-12


# --------------------------------------------------------------------------
# tests/transistor_integration_test.rs:238-657, tests/diode_integration_test.rs:
# 34-86: the reference's device circuits, solved by the oracle's LU Newton,
# land inside the bounds the reference's tests assert
# --------------------------------------------------------------------------
def test_device_circuits_meet_reference_bounds():
    from ohmnivore_amd import netlist as nl
    from tests.test_gpu_device_netlists import CASES, oracle_point
    for name, (text, bounds) in sorted(CASES.items()):
        m = oracle.compile(nl.parse(text).to_circuit())
        x = oracle_point(m)
        v = {n: float(x[i]) for i, n in enumerate(m.node_names)}
        assert bounds(v), (name, v)


def test_pnp_is_the_negated_npn():
    """src/compiler.rs:482-522: PNP = NPN with the polarity flag, so the
    mirrored circuit's solution is the exact negation."""
    from ohmnivore_amd import netlist as nl
    from tests.test_gpu_device_netlists import nl as dc_netlist
    xs = []
    for pol, s in (("NPN", 1), ("PNP", -1)):
        text = dc_netlist(f"V1 vcc 0 DC {5 * s}", f"V2 base 0 DC {0.7 * s}", "R1 vcc vc 1k",
                          f".MODEL Q {pol}(IS=1e-14 BF=100)", "Q1 vc base 0 Q")
        m = oracle.compile(nl.parse(text).to_circuit())
        xs.append(oracle.newton_solve(m.g, m.b_dc, m.diode_descriptors, m.bjt_descriptors,
                                      m.mosfet_descriptors, abs_tol=1e-13, max_iterations=200)[0])
    assert np.array_equal(xs[0], -xs[1])
    assert 0.0 < xs[0][2] < 0.5  # V(vc): saturated NPN
