"""GPU parity of the one-wave BiCGSTAB (systems of at most 64 unknowns, one
launch per solve; the reference's bicgstab_fused, src/solver/ds_shaders.rs:
340-728) against the multi-kernel pipeline (OHM_FUSED_SMALL=0) and the FP64
LU oracle. The one-wave solve rounds every product, sum and update in the
pipeline's order (one row per lane, products summed per row from 0, dot
products by the same wave tree), so the two agree to the bit: same stop
status, same iteration count, identical x. Against LU: 1e-6 relative /
1e-9 absolute on circuits (a 1e-13 relative residual leaves errors of that
size in the small node voltages of a ladder), 1e-9 relative on the
well-conditioned random systems."""
import numpy as np
import pytest

import oracle
from ohmnivore_amd import _native as N
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    return solver.HipBackend()


def rand_system(n, seed):
    """Nonsymmetric, diagonally dominant, with a few empty off-diagonal rows."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for r in range(n):
        c = np.unique(np.concatenate([[r], rng.integers(0, n, size=4 if r % 7 else 0)]))
        for j in c:
            rows.append(r)
            cols.append(j)
            vals.append(6.0 + rng.random() if j == r else rng.standard_normal())
    return oracle.from_triplet_arrays(n, n, np.array(rows), np.array(cols), np.array(vals))


def diagonal(a):
    d = np.zeros(a.nrows)
    for r in range(a.nrows):
        k = slice(int(a.row_pointers[r]), int(a.row_pointers[r + 1]))
        hit = np.asarray(a.col_indices[k]) == r
        d[r] = np.asarray(a.values[k])[hit].sum()
    return d


def solve(be, a, b, prec, prm, monkeypatch, fused, x0=None):
    monkeypatch.setenv("OHM_FUSED_SMALL", "1" if fused else "0")
    m = be.upload_matrix(a.values, a.col_indices, a.row_pointers, a.nrows)
    p = be.preconditioner(m, prec, 1)
    bb, xx = be.new_buffer(a.nrows), be.new_buffer(a.nrows)
    be.upload_vec(b, bb)
    be.upload_vec(np.zeros(a.nrows) if x0 is None else x0, xx)
    try:
        s = be.bicgstab(m, bb, xx, p, prm)
        return "ok", s.iterations, s.dispatches, be.download_vec(xx)
    except N.OhmnivoreError as e:
        return str(e).split(" after")[0], None, None, None


CASES = [("ladder", lambda: solver.compile(ir.resistor_ladder(30, seed=1))),
         ("mesh7x7", lambda: solver.compile(ir.diode_mesh(7, 7, seed=2))),
         ("mesh4x4", lambda: solver.compile(ir.diode_mesh(4, 4, seed=3)))]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("prec", [N.PREC_JACOBI, N.PREC_ISAI], ids=["jacobi", "isai"])
def test_fused_matches_pipeline_on_circuits(be, name, make, prec, monkeypatch):
    m = make()
    a = m.g
    assert a.nrows <= 64
    if prec == N.PREC_JACOBI and np.any(diagonal(a) == 0):
        pytest.skip("zero diagonal: Jacobi is not defined")
    ref = oracle.sparse_lu_solve(a, m.b_dc)
    prm = solver.bicgstab_params(tolerance=1e-13, max_iterations=2000)
    st1, it1, d1, x1 = solve(be, a, m.b_dc, prec, prm, monkeypatch, True)
    st0, it0, d0, x0 = solve(be, a, m.b_dc, prec, prm, monkeypatch, False)
    assert st1 == st0 == "ok", (st1, st0)
    assert it1 == it0 and x1.tobytes() == x0.tobytes(), (it1, it0, np.abs(x1 - x0).max())
    assert d1 == 1 and d0 > 1  # one launch against the pipeline's several
    assert np.allclose(x1, ref, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("n", [1, 2, 17, 63, 64])
@pytest.mark.parametrize("prec", [N.PREC_NONE, N.PREC_JACOBI, N.PREC_ISAI],
                         ids=["none", "jacobi", "isai"])
def test_fused_sizes_and_preconditioners(be, n, prec, monkeypatch):
    a = rand_system(n, seed=n)
    b = np.random.default_rng(n + 1).standard_normal(n)
    ref = oracle.sparse_lu_solve(a, b)
    prm = solver.bicgstab_params(tolerance=1e-13, max_iterations=500)
    st1, it1, _, x1 = solve(be, a, b, prec, prm, monkeypatch, True)
    st0, it0, _, x0 = solve(be, a, b, prec, prm, monkeypatch, False)
    assert st1 == st0 == "ok"
    assert it1 == it0 and x1.tobytes() == x0.tobytes(), (it1, it0, np.abs(x1 - x0).max())
    assert np.allclose(x1, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_fused_warm_start_tol_modes_and_stops(be, monkeypatch):
    """Warm start with the 'initial' tolerance, an already-converged warm
    start (0 iterations), zero right-hand side, the iteration cap, the
    reference's own shadow r_hat = b breaking down on a ladder, and n = 65
    (the pipeline) all give the pipeline's status and iterations."""
    m = solver.compile(ir.resistor_ladder(30, seed=1))
    a, b = m.g, m.b_dc
    ref = oracle.sparse_lu_solve(a, b)
    x0 = np.linspace(0.0, 0.3, a.nrows)
    checks = [
        (b, solver.bicgstab_params(tolerance=1e-6, tol_mode="initial", warm_start=True), x0),
        (b, solver.bicgstab_params(tolerance=1e-8, warm_start=True), ref),
        (np.zeros(a.nrows), solver.bicgstab_params(), None),
        (b, solver.bicgstab_params(tolerance=1e-14, max_iterations=3), None),
        (b, solver.bicgstab_params(shadow="rhs"), None),
    ]
    for rhs, prm, start in checks:
        r1 = solve(be, a, rhs, N.PREC_ISAI, prm, monkeypatch, True, start)
        r0 = solve(be, a, rhs, N.PREC_ISAI, prm, monkeypatch, False, start)
        assert r1[0] == r0[0], (r1[0], r0[0])
        if r1[0] == "ok":
            assert r1[1] == r0[1] and r1[3].tobytes() == r0[3].tobytes(), (r1[1], r0[1])
    assert solve(be, a, b, N.PREC_ISAI, checks[1][1], monkeypatch, True, ref)[1] == 0
    assert not solve(be, a, np.zeros(a.nrows), N.PREC_ISAI, checks[2][1], monkeypatch, True)[3].any()
    big = rand_system(65, seed=65)
    rb = np.ones(65)
    st, _, d, x = solve(be, big, rb, N.PREC_ISAI, solver.bicgstab_params(tolerance=1e-12),
                        monkeypatch, True)
    assert st == "ok" and d > 1
    assert np.allclose(x, oracle.sparse_lu_solve(big, rb), rtol=1e-8, atol=1e-12)


def test_fused_newton_with_forcing(be, monkeypatch):
    """A Newton run on a 53-unknown diode mesh with Eisenstat-Walker forcing
    (warm-started sub-solves, 'initial' tolerance) takes the pipeline's
    iterates to the bit with the one-wave solve."""
    m = solver.compile(ir.diode_mesh(7, 7, seed=5))
    kw = dict(abs_tol=1e-10, max_iterations=100, bicgstab_tolerance=1e-6, forcing="ew1",
              forcing_max=0.9, forcing_safeguard="always", bicgstab_floor=1e-18)
    out = {}
    for fused in (True, False):
        monkeypatch.setenv("OHM_FUSED_SMALL", "1" if fused else "0")
        ctx = solver.NewtonContext(be, m.g, m.b_dc, m.diode_descriptors)
        st = ctx.run(**kw)
        out[fused] = (ctx.x().tobytes(), st.bicgstab_iters_per_newton)
    assert out[True] == out[False]
    assert len(out[True][1]) > 2


def test_fused_dense_matrices_stay_in_global_memory(be, monkeypatch):
    """A dense 64 x 64 system with ISAI: the matrix and its two factors need
    more than the 60 KiB of LDS the one-wave solve stages, so their rows are
    read from global memory; the Jacobi run of the same system (48 KiB) is
    staged. Both reach the LU solution. (The pipeline splits this matrix into
    two row blocks, so its dot products sum in another order: no bitwise
    comparison here.)"""
    rng = np.random.default_rng(11)
    n = 64
    d = rng.standard_normal((n, n)) + np.diag(np.full(n, 3.0 * n))
    rows, cols = np.nonzero(np.ones((n, n)))
    a = oracle.from_triplet_arrays(n, n, rows, cols, d[rows, cols])
    b = rng.standard_normal(n)
    ref = np.linalg.solve(d, b)
    prm = solver.bicgstab_params(tolerance=1e-13, max_iterations=500)
    for prec in (N.PREC_ISAI, N.PREC_JACOBI):
        st, it, disp, x = solve(be, a, b, prec, prm, monkeypatch, True)
        assert st == "ok" and disp == 1 and it > 0
        assert np.allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
