"""GPU adjoint DC sensitivity: transposed BiCGSTAB against the oracle's LU on
A^T, the GPU gradient against the oracle's LU adjoint (small amplifier
array) and against central differences of the GPU operating point at the
full 4,096-stage size (64 random parameters, 1e-5 relative), plus
torch.autograd.gradcheck through DcOperatingPoint."""
import numpy as np
import pytest
import torch

import oracle
from ohmnivore_amd import _native as N
from ohmnivore_amd import adjoint as A
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver

pytestmark = pytest.mark.gpu
SENT = ir.SENTINEL


@pytest.fixture(scope="module")
def be():
    return solver.HipBackend()


def transpose(a):
    n = a.nrows
    rows = np.repeat(np.arange(n), np.diff(a.row_pointers.astype(np.int64)))
    return oracle.from_triplet_arrays(n, n, a.col_indices.astype(np.int64), rows, a.values)


@pytest.mark.parametrize("kind", [N.PREC_JACOBI, N.PREC_ISAI])
def test_bicgstab_transpose_vs_lu(be, kind):
    mna = solver.compile(ir.diode_mesh(24, 24, seed=2))  # nonsymmetric once the V rows are in
    a = mna.g
    rng = np.random.default_rng(2)
    b = rng.standard_normal(a.nrows)
    ref = oracle.sparse_lu_solve(transpose(a), b)
    m = be.upload_matrix(a.values, a.col_indices, a.row_pointers, a.nrows)
    p = be.preconditioner(m, kind)
    bb, xx = be.new_buffer(a.nrows), be.new_buffer(a.nrows)
    be.upload_vec(b, bb)
    st = be.bicgstab_transpose(m, bb, xx, p, solver.bicgstab_params(tolerance=1e-13,
                                                                     max_iterations=20000))
    x = be.download_vec(xx)
    assert np.all(np.abs(x - ref) <= 1e-6 * np.abs(ref) + 1e-9), (np.abs(x - ref).max(), st)
    assert st.residual_norm < 1e-10


def xv(v, i):
    return 0.0 if i == SENT else v[i]


def test_gpu_adjoint_vs_oracle_lu_adjoint(be):
    ckt = ir.amplifier_array(64, seed=5)
    mna = solver.compile(ckt)
    prob = A.DcProblem(be, ckt, mna)
    outs = A.output_nodes(ckt, mna)
    x = prob.solve()
    dj = np.zeros(mna.size)
    dj[outs] = 2 * (x[outs] - 0.9)
    gg, gk = prob.adjoint(dj, tolerance=1e-13)
    # oracle: J at the GPU operating point, lambda by LU on J^T
    jv, _ = oracle.newton_assemble(mna.g, mna.b_dc, x, mna.diode_descriptors,
                                   mna.bjt_descriptors, prob.mosfets)
    jac = ir.CsrMatrix(mna.size, mna.size, jv, mna.g.col_indices, mna.g.row_pointers)
    lam = oracle.sparse_lu_solve(transpose(jac), dj)
    ref_g = np.array([-(xv(lam, a) - xv(lam, b)) * (xv(x, a) - xv(x, b)) for a, b in prob.res_nodes])
    scale = np.abs(ref_g).max()
    assert np.all(np.abs(gg - ref_g) <= 1e-8 * np.abs(ref_g) + 1e-12 * scale)
    ref_k = []
    for d in prob.mosfets:
        vd, vg, vs = xv(x, d["drain_idx"]), xv(x, d["gate_idx"]), xv(x, d["source_idx"])
        iu = oracle.mosfet_core(vg - vs, vd - vs, float(d["vto"]), 1.0, float(d["lambda"]),
                                float(d["polarity"]))[0]
        ref_k.append(-(xv(lam, d["b_idx"][0]) - xv(lam, d["b_idx"][1])) * iu)
    ref_k = np.array(ref_k)
    assert np.all(np.abs(gk - ref_k) <= 1e-8 * np.abs(ref_k) + 1e-12 * np.abs(ref_k).max())


def test_gpu_adjoint_full_size_vs_finite_differences(be):
    """The adjoint workload at full size: 4,096 stages, J = sum (V_out - 0.9)^2,
    64 random parameters (resistors and W/L), central differences of the
    GPU operating point. J sums 4,096 outputs, so its noise between two
    operating-point solves is ~1e-12; the check is 1e-5 relative above the
    differences' noise floor 1e-12 / h."""
    ckt = ir.amplifier_array(4096, seed=5)
    mna = solver.compile(ckt)
    assert mna.size == 8194
    prob = A.DcProblem(be, ckt, mna, newton_kw=dict(abs_tol=1e-13, bicgstab_tolerance=1e-14))
    outs = A.output_nodes(ckt, mna)
    x = prob.solve()

    def obj(xx):
        return float(((xx[outs] - 0.9) ** 2).sum())

    dj = np.zeros(mna.size)
    dj[outs] = 2 * (x[outs] - 0.9)
    gg, gk = prob.adjoint(dj, tolerance=1e-13)
    assert gg.shape == (12288,) and gk.shape == (2048,)
    rng = np.random.default_rng(64)

    def richardson(f, h):
        """Central differences at h and h/2 combined to cancel the h^2 term."""
        d1 = (f(h) - f(-h)) / (2 * h)
        d2 = (f(h / 2) - f(-h / 2)) / h
        return (4 * d2 - d1) / 3

    for r in rng.choice(12288, size=48, replace=False):
        def f_g(dh, r=r):
            g = prob.g0.copy()
            g[r] += dh
            return obj(prob.solve(g))

        h = 5e-4 * prob.g0[r]
        fd = richardson(f_g, h)
        assert abs(gg[r] - fd) <= 1e-5 * abs(fd) + 3e-12 / h, (r, gg[r], fd)
    kp = prob.mosfets["kp"]
    for m in rng.choice(2048, size=16, replace=False):
        k32 = np.float32(kp[m])
        step = np.float32(2 ** 12) * np.spacing(k32)  # kp +- step, +- step/2 are exact f32

        def f_k(dh, m=m):
            kk = kp.astype(np.float64)
            kk[m] = np.float64(k32) + dh
            return obj(prob.solve(None, kk))

        h = float(step)
        fd = richardson(f_k, h)
        assert abs(gk[m] - fd) <= 1e-5 * abs(fd) + 3e-12 / h, (m, gk[m], fd)


def test_autograd_function_gradcheck(be):
    ckt = ir.amplifier_array(6, seed=5)
    mna = solver.compile(ckt)
    prob = A.DcProblem(be, ckt, mna, newton_kw=dict(abs_tol=1e-13, bicgstab_tolerance=1e-12))
    outs = torch.from_numpy(A.output_nodes(ckt, mna))
    g0 = torch.from_numpy(prob.g0)
    k0 = torch.from_numpy(prob.kp0)

    def f(ug, uk):  # relative parametrisation keeps gradcheck's steps well scaled
        x = A.DcOperatingPoint.apply(g0 * (1 + ug), k0 * (1 + uk), prob)
        return A.objective(x, outs)

    ug = torch.zeros(g0.shape[0], dtype=torch.float64, requires_grad=True)
    uk = torch.zeros(k0.shape[0], dtype=torch.float64, requires_grad=True)
    # kp is stored in f32 descriptors: eps = 1e-3 keeps its quantisation
    # (6e-8 relative) well below the check's tolerance
    assert torch.autograd.gradcheck(f, (ug, uk), eps=1e-3, atol=1e-6, rtol=2e-3)


def test_newton_update_matches_fresh_context(be):
    """ohm_newton_update (new conductances and kp on one topology) gives the
    operating point a freshly built context gives, and rejects descriptors
    whose nodes or stamp slots differ."""
    ckt = ir.amplifier_array(256, seed=11)
    mna = solver.compile(ckt)
    prob = A.DcProblem(be, ckt, mna, newton_kw=dict(abs_tol=1e-13, bicgstab_tolerance=1e-14))
    x0 = prob.solve()
    rng = np.random.default_rng(5)
    g = prob.g0 * (1 + 0.1 * rng.random(prob.g0.shape[0]))
    kp = prob.kp0 * (1 + 0.2 * rng.random(prob.kp0.shape[0]))
    x_upd = prob.solve(g, kp)
    fresh = A.DcProblem(be, ckt, mna, newton_kw=prob.newton_kw)
    x_new = fresh.solve(g, kp)
    assert np.allclose(x_upd, x_new, rtol=1e-9, atol=1e-12), np.abs(x_upd - x_new).max()
    assert not np.allclose(x_upd, x0, rtol=1e-6, atol=1e-9)  # the update changed the answer
    bad = prob.mosfets.copy()
    bad["drain_idx"][0] = bad["gate_idx"][0]
    with pytest.raises(N.OhmnivoreError, match="create a new context"):
        prob.ctx.update(mosfets=bad)
    with pytest.raises(ValueError):
        prob.ctx.update(values=np.zeros(3))


def test_newton_update_of_the_rhs_reaches_source_rows(be):
    """A new base right-hand side replaces every entry, the voltage-source
    rows that no device stamps included: a context re-run with b scaled by
    0.3 lands where a fresh context built with that b lands."""
    mna = solver.compile(ir.amplifier_array(32, seed=11))
    kw = dict(abs_tol=1e-13, bicgstab_tolerance=1e-14)
    descs = (mna.diode_descriptors, mna.bjt_descriptors, mna.mosfet_descriptors)
    ctx = solver.NewtonContext(be, mna.g, mna.b_dc, *descs)
    ctx.run(**kw)
    ctx.update(base_b=0.3 * mna.b_dc)
    ctx.run(**kw)
    fresh = solver.NewtonContext(be, mna.g, 0.3 * mna.b_dc, *descs)
    fresh.run(**kw)
    x_upd, x_new = ctx.x(), fresh.x()
    assert np.allclose(x_upd, x_new, rtol=1e-9, atol=1e-12), np.abs(x_upd - x_new).max()
    vdd = mna.node_names.index("vdd")
    assert abs(x_upd[vdd] - 0.3 * 1.8) < 1e-12
