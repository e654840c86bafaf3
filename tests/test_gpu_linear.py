"""GPU parity: SolverBackend ops, SpMV, preconditioners and BiCGSTAB against
the CPU oracle. Tolerances: 1e-6 relative / 1e-9 absolute on solutions
(versus FP64 LU); SpMV and BLAS-1 within a few ulps of the FP64 oracle."""
import numpy as np
import pytest

import oracle
from ohmnivore_amd import _native as N
from ohmnivore_amd import circuit as ir
from ohmnivore_amd import solver

pytestmark = pytest.mark.gpu

REL, ABS = 1e-6, 1e-9


@pytest.fixture(scope="module")
def be():
    return solver.HipBackend()


def close(x, ref, rel=REL, abs_=ABS):
    err = np.abs(x - ref)
    ok = err <= np.maximum(rel * np.abs(ref), abs_)
    return bool(ok.all()), float(err.max(initial=0.0))


def rand_csr(n, per_row, rng, long_rows=(), empty_rows=()):
    rows, cols = [], []
    for r in range(n):
        if r in empty_rows:
            continue
        k = 3000 if r in long_rows else per_row
        c = np.unique(rng.integers(0, n, size=k))
        rows += [r] * len(c)
        cols += list(c)
    vals = rng.standard_normal(len(rows))
    return oracle.from_triplet_arrays(n, n, np.array(rows), np.array(cols), vals)


@pytest.mark.parametrize("n,per_row,long_rows", [(1, 1, ()), (257, 5, ()), (5000, 7, (3, 4000)),
                                                 (20000, 3, (0,))])
def test_spmv_matches_oracle(be, n, per_row, long_rows):
    rng = np.random.default_rng(n)
    a = rand_csr(n, per_row, rng, long_rows=long_rows, empty_rows=(1,) if n > 2 else ())
    x = rng.standard_normal(n)
    ref = oracle.spmv(a, x)
    m = be.upload_matrix(a.values, a.col_indices, a.row_pointers, n)
    bx, by = be.new_buffer(n), be.new_buffer(n)
    be.upload_vec(x, bx)
    be.spmv(m, bx, by)
    y = be.download_vec(by)
    scale = np.abs(a.to_dense()).dot(np.abs(x)) if n <= 5000 else np.abs(ref) + 1
    assert np.all(np.abs(y - ref) <= 1e-13 * scale + 1e-300)


def test_spmv_transpose(be):
    rng = np.random.default_rng(3)
    a = rand_csr(3000, 6, rng, long_rows=(5,))
    x = rng.standard_normal(3000)
    m = be.upload_matrix(a.values, a.col_indices, a.row_pointers, 3000)
    bx, by = be.new_buffer(3000), be.new_buffer(3000)
    be.upload_vec(x, bx)
    be.spmv_transpose(m, bx, by)
    ref = a.to_dense().T @ x
    assert np.allclose(be.download_vec(by), ref, rtol=1e-12, atol=1e-12)


def test_blas1_ops(be):
    rng = np.random.default_rng(1)
    n = 100_003
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    bx, by = be.new_buffer(n), be.new_buffer(n)
    be.upload_vec(x, bx)
    be.upload_vec(y, by)
    assert abs(be.dot(bx, by) - x @ y) < 1e-10 * np.abs(x * y).sum()
    assert abs(be.dot_n(bx, by, 77) - x[:77] @ y[:77]) < 1e-12 * np.abs(x[:77] * y[:77]).sum()
    be.axpy(0.5, bx, by)
    assert np.allclose(be.download_vec(by), y + 0.5 * x, rtol=1e-15, atol=1e-15)
    be.scale(-2.0, by)
    be.copy(by, bx)
    assert np.allclose(be.download_vec(bx), -2 * (y + 0.5 * x), rtol=1e-15, atol=1e-15)


def test_dot_is_deterministic(be):
    rng = np.random.default_rng(2)
    n = 1 << 20
    x = rng.standard_normal(n)
    bx = be.new_buffer(n)
    be.upload_vec(x, bx)
    vals = {be.dot(bx, bx) for _ in range(5)}
    assert len(vals) == 1


def lu_solution(m: ir.MnaSystem):
    return oracle.sparse_lu_solve(m.g, m.b_dc)


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_ladder_jacobi_matches_lu(be, precision):
    """10,000-section ladder, Jacobi-BiCGSTAB. A 1e-10 relative residual
    leaves ~2e-7 V errors (KCL rows are in amperes at mS scale), so parity
    needs ~1e-13 — reachable with the scaled breakdown tests."""
    m = solver.compile(ir.resistor_ladder(10_000, seed=1))
    ref = lu_solution(m)
    gs = solver.GpuSolver(be, precond=N.PREC_JACOBI,
                          params=solver.bicgstab_params(tolerance=1e-13, precision=precision))
    x = gs.solve_real(m.g, m.b_dc)
    ok, err = close(x, ref)
    assert ok, (err, gs.last_stats)
    assert gs.last_stats.residual_norm < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_mesh_isai_matches_lu(be, precision):
    """Linear part of the diode mesh (diodes at zero bias) with ISAI(1)."""
    m = solver.compile(ir.diode_mesh(48, 40, seed=7))
    ref = lu_solution(m)
    gs = solver.GpuSolver(be, precond=N.PREC_ISAI,
                          params=solver.bicgstab_params(tolerance=1e-12, precision=precision,
                                                        max_iterations=20_000))
    x = gs.solve_real(m.g, m.b_dc)
    ok, err = close(x, ref)
    assert ok, (err, gs.last_stats)


def test_auto_preconditioner_rule(be):
    m = solver.compile(ir.resistor_ladder(100, seed=1))  # V-source row -> zero diagonal
    a = be.upload_matrix(m.g.values, m.g.col_indices, m.g.row_pointers, m.size)
    assert be.preconditioner(a, N.PREC_AUTO).kind == N.PREC_ISAI
    c = ir.Circuit()
    c.isource("I1", "0", "1", dc=1e-3)
    c.resistor("R1", "1", "0", 1e3)
    c.resistor("R2", "1", "2", 1e3)
    c.resistor("R3", "2", "0", 1e3)
    m2 = solver.compile(c)
    a2 = be.upload_matrix(m2.g.values, m2.g.col_indices, m2.g.row_pointers, m2.size)
    assert be.preconditioner(a2, N.PREC_AUTO).kind == N.PREC_JACOBI


def test_isai_matches_exact_ilu0_isai(be):
    """With the ILU(0) fixed point run to convergence, the device ISAI(1)
    equals the oracle's exact ILU(0)+ISAI(1) (apply compared on random r)."""
    m = solver.compile(ir.diode_mesh(9, 8, seed=5))
    g = m.g
    rng = np.random.default_rng(0)
    g.values = g.values + (g.values != 0) * rng.uniform(0, 1e-3, g.nnz)
    _, _, ml, mu = oracle.ilu0_isai(g, 1)
    a = be.upload_matrix(g.values, g.col_indices, g.row_pointers, g.nrows)
    p = be.preconditioner(a, N.PREC_ISAI, 1)
    p.set_ilu_sweeps(200)
    p.update(a)
    r = rng.standard_normal(g.nrows)
    br, bz = be.new_buffer(g.nrows), be.new_buffer(g.nrows)
    be.upload_vec(r, br)
    p.apply(br, bz)
    ref = oracle.spmv(mu, oracle.spmv(ml, r))
    assert np.allclose(be.download_vec(bz), ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_reference_shadow_breaks_down(be):
    m = solver.compile(ir.resistor_ladder(200, seed=1))
    gs = solver.GpuSolver(be, precond=N.PREC_JACOBI,
                          params=solver.bicgstab_params(shadow="rhs"))
    with pytest.raises(N.OhmnivoreError, match="BiCGSTAB breakdown") as e:
        gs.solve_real(m.g, m.b_dc)
    assert e.value.kind == "Solve"


def test_not_converged_and_param_errors(be):
    m = solver.compile(ir.resistor_ladder(2000, seed=1))
    gs = solver.GpuSolver(be, precond=N.PREC_JACOBI,
                          params=solver.bicgstab_params(max_iterations=3, tolerance=1e-14))
    with pytest.raises(N.OhmnivoreError, match="did not converge within 3 iterations"):
        gs.solve_real(m.g, m.b_dc)
    for kw, msg in (({"max_iterations": 0}, "max_iterations must be > 0"),
                    ({"tolerance": float("nan")}, "tolerance must be finite")):
        gs = solver.GpuSolver(be, precond=N.PREC_JACOBI, params=solver.bicgstab_params(**kw))
        with pytest.raises(N.OhmnivoreError, match=msg):
            gs.solve_real(m.g, m.b_dc)


def test_zero_rhs_empty_and_dimension_errors(be):
    m = solver.compile(ir.resistor_ladder(50, seed=1))
    gs = solver.GpuSolver(be, precond=N.PREC_JACOBI)
    x = gs.solve_real(m.g, np.zeros(m.size))
    assert not x.any() and gs.last_stats.iterations == 0
    e = ir.CsrMatrix(0, 0, np.zeros(0), np.zeros(0, np.uint32), np.zeros(1, np.uint32))
    assert gs.solve_real(e, np.zeros(0)).shape == (0,)
    with pytest.raises(N.OhmnivoreError, match="dimension"):
        gs.solve_real(m.g, np.zeros(m.size + 1))


@pytest.mark.parametrize("precision,tol_mode", [("fp64", "rhs"), ("fp32_ir", "rhs"),
                                                ("fp64", "initial")])
def test_graph_replay_matches_eager(be, precision, tol_mode):
    """Graph-captured iteration chunks, replayed many times with eager solves
    and host syncs in between, give bitwise the eager result every time.
    tol_mode 'initial' with a warm start takes the 4-iteration opening
    chunk (the inexact-Newton sub-solve form) before the 8-iteration ones."""
    m = solver.compile(ir.diode_mesh(40, 40, seed=9))
    a = be.upload_matrix(m.g.values, m.g.col_indices, m.g.row_pointers, m.size)
    p = be.preconditioner(a, N.PREC_ISAI, 1)
    b, x = be.new_buffer(m.size), be.new_buffer(m.size)
    be.upload_vec(m.b_dc, b)
    warm = tol_mode == "initial"
    prm = solver.bicgstab_params(tolerance=1e-11, precision=precision, tol_mode=tol_mode,
                                 warm_start=warm)
    x0 = np.linspace(0.0, 0.5, m.size) if warm else np.zeros(m.size)
    results = []
    for mode in (False, True, False, True, True):
        be.set_graphs(mode)
        be.upload_vec(x0, x)
        s = be.bicgstab(a, b, x, p, prm)
        be.sync()
        results.append((be.download_vec(x).tobytes(), s.iterations))
    be.set_graphs(True)
    assert all(r == results[0] for r in results)
    assert results[0][1] > 8  # several replays per solve


def test_graph_replay_matches_eager_in_newton(be):
    """A Newton run with Eisenstat-Walker forcing (warm-started sub-solves
    with the opening 4-iteration chunk, no poll before it) gives bitwise the
    same iterate and per-Newton BiCGSTAB counts with and without graphs."""
    m = solver.compile(ir.diode_mesh(48, 40, seed=5))
    ctx = solver.NewtonContext(be, m.g, m.b_dc, m.diode_descriptors)
    kw = dict(abs_tol=1e-10, max_iterations=100, bicgstab_tolerance=1e-6, forcing="ew1",
              forcing_max=0.9, forcing_safeguard="always", bicgstab_floor=1e-18)
    out = []
    for mode in (False, True, True, False):
        be.set_graphs(mode)
        st = ctx.run(**kw)
        out.append((ctx.x().tobytes(), st.bicgstab_iters_per_newton))
    be.set_graphs(True)
    assert all(o == out[0] for o in out)
    assert len(out[0][1]) > 3 and sum(out[0][1]) > 8


def test_warm_start_and_determinism(be):
    m = solver.compile(ir.diode_mesh(32, 32, seed=3))
    a = be.upload_matrix(m.g.values, m.g.col_indices, m.g.row_pointers, m.size)
    p = be.preconditioner(a, N.PREC_ISAI, 1)
    b, x = be.new_buffer(m.size), be.new_buffer(m.size)
    be.upload_vec(m.b_dc, b)
    prm = solver.bicgstab_params(tolerance=1e-10)
    s1 = be.bicgstab(a, b, x, p, prm)
    x1 = be.download_vec(x)
    be.upload_vec(np.zeros(m.size), x)
    s2 = be.bicgstab(a, b, x, p, prm)
    assert be.download_vec(x).tobytes() == x1.tobytes() and s1.iterations == s2.iterations
    warm = solver.bicgstab_params(tolerance=1e-10, warm_start=True)
    s3 = be.bicgstab(a, b, x, p, warm)
    assert s3.iterations == 0  # already converged from x1


def test_autotune_keeps_results(be):
    """The autotuner only changes the row-block shape: SpMV and a BiCGSTAB
    solve agree with the untuned matrix."""
    mna = solver.compile(ir.diode_mesh(96, 96, seed=2))
    a = mna.g
    rng = np.random.default_rng(9)
    x = rng.standard_normal(a.nrows)
    m0 = be.upload_matrix(a.values, a.col_indices, a.row_pointers, a.nrows)
    m1 = be.upload_matrix(a.values, a.col_indices, a.row_pointers, a.nrows)
    rows, nnz = m1.autotune()
    assert rows in (32, 64, 128, 256) and nnz in (256, 512, 768, 1024, 1280, 1536, 2048)
    bx, y0, y1 = be.new_buffer(a.nrows), be.new_buffer(a.nrows), be.new_buffer(a.nrows)
    be.upload_vec(x, bx)
    be.spmv(m0, bx, y0)
    be.spmv(m1, bx, y1)
    assert np.allclose(be.download_vec(y1), be.download_vec(y0), rtol=1e-14, atol=1e-14)
    b = rng.standard_normal(a.nrows)
    ref = oracle.sparse_lu_solve(a, b)
    for m in (m0, m1):
        p = be.preconditioner(m, N.PREC_ISAI)
        bb, xx = be.new_buffer(a.nrows), be.new_buffer(a.nrows)
        be.upload_vec(b, bb)
        be.bicgstab(m, bb, xx, p, solver.bicgstab_params(tolerance=1e-13, max_iterations=20000))
        ok, err = close(be.download_vec(xx), ref)
        assert ok, err


@pytest.mark.parametrize("precond", [N.PREC_JACOBI, N.PREC_ISAI])
def test_wide_grid_ticket_reductions(be, precond):
    """600,000 rows: the SpMV kernels run more than 2,048 workgroups, so their
    dot products fold in the producer (ticket) instead of being re-summed by
    every consumer workgroup. The solve still matches the banded LU."""
    n = 600_000
    rng = np.random.default_rng(21)
    lo, up = -rng.uniform(0.5, 1.0, n - 1), -rng.uniform(0.5, 1.0, n - 1)
    diag = 2.5 + rng.uniform(0.0, 1.0, n)
    rows = np.concatenate([np.arange(1, n), np.arange(n), np.arange(n - 1)])
    cols = np.concatenate([np.arange(n - 1), np.arange(n), np.arange(1, n)])
    vals = np.concatenate([lo, diag, up])
    a = oracle.from_triplet_arrays(n, n, rows, cols, vals)
    b = rng.standard_normal(n)
    ref = oracle.sparse_lu_solve(a, b)
    m = be.upload_matrix(a.values, a.col_indices, a.row_pointers, n)
    p = be.preconditioner(m, precond)
    bb, xx = be.new_buffer(n), be.new_buffer(n)
    be.upload_vec(b, bb)
    s = be.bicgstab(m, bb, xx, p, solver.bicgstab_params(tolerance=1e-13, max_iterations=5000))
    ok, err = close(be.download_vec(xx), ref)
    assert ok, (err, s.iterations)
    assert s.iterations > 3


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_isai_product_apply_matches_two_factor(be, precision, monkeypatch):
    """OHM_ISAI_PRODUCT=1 applies M = M_U M_L formed on the device (one SpMV,
    long pair lists summed by whole workgroups); the default applies M_L then
    M_U. Both give the LU solution, and the product apply equals the
    two-factor apply to rounding."""
    mna = solver.compile(ir.amplifier_array(64, seed=3))  # VDD row: long product entries
    a = mna.g
    rng = np.random.default_rng(4)
    b = rng.standard_normal(a.nrows)
    ref = oracle.sparse_lu_solve(a, b)
    outs = {}
    for prod in ("0", "1"):
        monkeypatch.setenv("OHM_ISAI_PRODUCT", prod)
        m = be.upload_matrix(a.values, a.col_indices, a.row_pointers, a.nrows)
        p = be.preconditioner(m, N.PREC_ISAI, 1)
        br, bz = be.new_buffer(a.nrows), be.new_buffer(a.nrows)
        be.upload_vec(b, br)
        p.apply(br, bz)
        bb, xx = be.new_buffer(a.nrows), be.new_buffer(a.nrows)
        be.upload_vec(b, bb)
        be.bicgstab(m, bb, xx, p, solver.bicgstab_params(tolerance=1e-13, max_iterations=20000,
                                                         precision=precision))
        x = be.download_vec(xx)
        ok, err = close(x, ref)
        assert ok, (prod, err)
        outs[prod] = be.download_vec(bz)
    scale = np.abs(outs["0"]).max()
    assert np.allclose(outs["1"], outs["0"], rtol=1e-10, atol=1e-13 * scale)
