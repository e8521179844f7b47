"""GPU solvers against the oracle: the reference GS solver (iterations, RMS,
solution), symmetric-GS PCG (relative residual, dense solution), mixed
precision refinement and the block-diagonal batch."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def _solve(name, m, b, nch=1, max_iters=200, target=1e-6, interval=1, x0=None, dtype=np.float64, **opts):
    from nsfem_amd import cu
    s = cu.createSolver(name, nch, m, max_iters, target, interval, dtype=dtype)
    for k, v in opts.items():
        s.setOption(k, v)
    s.getRhs().upload(b)
    if x0 is None:
        s.getSol().memsetZero()
    else:
        s.getSol().upload(x0)
    s.solve()
    return s, s.getSol().download()


@pytest.mark.parametrize("interval", [1, 3])
def test_gs_matches_oracle_one_channel(oracle, octx_h025, interval):
    m = octx_h025["Kp_int"]
    rng = np.random.default_rng(21)
    b = rng.uniform(-1, 1, m.shape[0])
    s, x = _solve("gs", m, b, max_iters=60, target=1e-4, interval=interval)
    xo, mse, it = oracle.gs_solve(m, b, np.zeros_like(b), 60, 1e-4, interval)
    assert s.getLastIterations() == it
    assert abs(s.getLastMse() - mse) <= 1e-10 * mse
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))


def test_gs_converges_and_reports_iteration_of_break(oracle, octx_c3):
    m = octx_c3["M"]
    rng = np.random.default_rng(22)
    b = rng.uniform(-1, 1, m.shape[0]) * 1e-3
    s, x = _solve("gs", m, b, max_iters=200, target=1e-9)
    xo, mse, it = oracle.gs_solve(m, b, np.zeros_like(b), 200, 1e-9, 1)
    assert it < 200 and s.getLastIterations() == it
    assert s.getLastMse() < 1e-9
    assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))


def test_gs_two_channels_stop_together(oracle, octx_c3):
    m = octx_c3["M"]
    n = m.shape[0]
    rng = np.random.default_rng(23)
    b = np.concatenate([rng.uniform(-1, 1, n) * 1e-3, rng.uniform(-1, 1, n) * 1e-6])
    s, x = _solve("gs", m, b, nch=2, max_iters=100, target=1e-10)
    L = oracle.Mcgs(m)
    # channel-wise oracle with a shared stop: both below target
    it_ref = None
    xs = [np.zeros(n), np.zeros(n)]
    for it in range(100):
        xs = [oracle.gs_solve(m, b[c * n:(c + 1) * n], xs[c], 1, 0.0, 1, L)[0] for c in range(2)]
        r = [np.sqrt(np.sum((m @ xs[c] - b[c * n:(c + 1) * n]) ** 2) / n) for c in range(2)]
        if r[0] < 1e-10 and r[1] < 1e-10:
            it_ref = it
            break
    assert s.getLastIterations() == it_ref
    assert abs(s.getLastMse(-1) - np.hypot(s.getLastMse(0), s.getLastMse(1))) < 1e-20
    assert np.max(np.abs(x - np.concatenate(xs))) <= 1e-12 * np.max(np.abs(x))


def test_gs_float32_matches_float64_to_single_precision(octx_c3):
    m = octx_c3["Kp_int"]
    rng = np.random.default_rng(24)
    b = rng.uniform(-1, 1, m.shape[0])
    _, x64 = _solve("gs", m, b, max_iters=40, target=0.0)
    _, x32 = _solve("gs", m, b.astype(np.float32), max_iters=40, target=0.0, dtype=np.float32)
    assert x32.dtype == np.float32
    assert np.max(np.abs(x32 - x64)) <= 1e-4 * np.max(np.abs(x64))


@pytest.mark.parametrize("which", ["Kp_int", "M"])
def test_pcg_matches_oracle_and_direct(oracle, octx_h025, which):
    m = octx_h025[which]
    rng = np.random.default_rng(25)
    b = rng.uniform(-1, 1, m.shape[0])
    s, x = _solve("pcg", m, b, max_iters=3000, target=1e-12)
    xo, rel, it = oracle.pcg_solve(m, b, np.zeros_like(b), 1e-12, 3000)
    assert abs(s.getLastIterations() - it) <= 1
    assert s.getLastRelRes() <= 1e-12
    true_rel = np.linalg.norm(b - m @ x) / np.linalg.norm(b)
    assert true_rel <= 1e-11
    xd = spla.spsolve(m.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    assert np.max(np.abs(x - xo)) <= 1e-9 * np.max(np.abs(xo))


def test_pcg_two_channels_and_warm_start(octx_h025):
    m = octx_h025["M"]
    n = m.shape[0]
    rng = np.random.default_rng(26)
    b = rng.uniform(-1, 1, 2 * n)
    s, x = _solve("pcg", m, b, nch=2, max_iters=500, target=1e-10)
    for c in range(2):
        bc = b[c * n:(c + 1) * n]
        assert np.linalg.norm(bc - m @ x[c * n:(c + 1) * n]) <= 1.0001e-10 * np.linalg.norm(bc)
    it_cold = s.getLastIterations()
    # warm start from the solution: converged at set-up, zero iterations
    s.getSol().upload(x)
    s.solve()
    assert s.getLastIterations() <= 1
    assert it_cold > 3


def test_pcg_zero_rhs_gives_zero(octx_c3):
    m = octx_c3["Kp_int"]
    s, x = _solve("pcg", m, np.zeros(m.shape[0]), max_iters=100, target=1e-10)
    assert not x.any() and s.getLastIterations() == 0


def test_pcg_float32(octx_h025):
    m = octx_h025["Kp_int"]
    rng = np.random.default_rng(27)
    b = rng.uniform(-1, 1, m.shape[0])
    s, x = _solve("pcg", m, b.astype(np.float32), max_iters=2000, target=1e-5, dtype=np.float32)
    # the recursive FP32 residual meets the target; the true FP64 residual of
    # an FP32 solution is bounded by the rounding of the matrix (~kappa * 6e-8)
    assert s.getLastRelRes() <= 1e-5
    assert np.linalg.norm(b - m @ x.astype(np.float64)) <= 1e-3 * np.linalg.norm(b)
    xd = spla.spsolve(m.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-4 * np.linalg.norm(xd)


@pytest.mark.parametrize("which,nch", [("Kp_int", 1), ("M", 2)])
def test_pcg_mixed_reaches_fp64_tolerance(octx_h025, which, nch):
    m = octx_h025[which]
    n = m.shape[0]
    rng = np.random.default_rng(28)
    b = rng.uniform(-1, 1, nch * n)
    s, x = _solve("pcg_mixed", m, b, nch=nch, max_iters=3000, target=1e-10)
    for c in range(nch):
        bc = b[c * n:(c + 1) * n]
        rel = np.linalg.norm(bc - m @ x[c * n:(c + 1) * n]) / np.linalg.norm(bc)
        assert rel <= 1e-10
        assert s.getLastRelRes(c) <= 1e-10
    xd = spla.spsolve(m.tocsc(), b[:n])
    assert np.linalg.norm(x[:n] - xd) <= 1e-7 * np.linalg.norm(xd)


def test_batch_of_uneven_systems_matches_individual_solves(oracle):
    from nsfem_amd import cu, fem, mesher
    mats = []
    for k, n in enumerate([600, 2500, 1300]):
        m = mesher.channel_mesh(n, seed=100 + k)
        ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), fem.DfgConditions(0.01, 1.0, 1.0, 4.0))
        mats.append(ctx.scipy("pressureStiffnessInternal"))
    s = cu.createBatchSolver("pcg", 1, mats, 3000, 1e-11, 1)
    rng = np.random.default_rng(29)
    bs = [rng.uniform(-1, 1, m.shape[0]) for m in mats]
    s.getRhs().upload(np.concatenate(bs))
    s.getSol().memsetZero()
    s.solve()
    x = s.getSol().download()
    for k, (m, b) in enumerate(zip(mats, bs)):
        o = s.system_offsets[k]
        xk = x[o:o + m.shape[0]]
        assert np.linalg.norm(b - m @ xk) <= 1.0001e-11 * np.linalg.norm(b)
        xo, rel, it = oracle.pcg_solve(m, b, np.zeros_like(b), 1e-11, 3000)
        assert abs(s.getLastIterations(k) - it) <= 1
        assert s.getLastRelRes(0, k) <= 1e-11
    # colourings of each system equal the single-system colouring
    for k, m in enumerate(mats):
        col, ps = s.coloring(k, m.shape[0])
        oc, ops, _ = oracle.mcgs_layout(m)
        assert np.array_equal(col, oc) and np.array_equal(ps, ops)


def test_batch_gs_per_system_iterations(oracle):
    from nsfem_amd import cu
    rng = np.random.default_rng(30)
    mats = []
    for n in (50, 400):
        d = sp.diags([-1, 4.0, -1], [-1, 0, 1], shape=(n, n)).tocsr()
        mats.append(d)
    s = cu.createBatchSolver("gs", 1, mats, 100, 1e-8, 1)
    bs = [rng.uniform(-1, 1, m.shape[0]) for m in mats]
    s.getRhs().upload(np.concatenate(bs))
    s.getSol().memsetZero()
    s.solve()
    x = s.getSol().download()
    for k, (m, b) in enumerate(zip(mats, bs)):
        xo, mse, it = oracle.gs_solve(m, b, np.zeros_like(b), 100, 1e-8, 1)
        assert s.getLastIterations(k) == it
        o = s.system_offsets[k]
        assert np.max(np.abs(x[o:o + m.shape[0]] - xo)) <= 1e-12 * np.max(np.abs(xo))


def test_zero_diagonal_rejected():
    from nsfem_amd import _lib, cu
    m = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 2.0]]))
    with pytest.raises(_lib.NsfemInvalidArgument, match="diagonal"):
        cu.createSolver("gs", 1, m, 10, 1e-6, 1)


@pytest.mark.parametrize("name", ["pcg", "pcg_mixed"])
def test_pure_neumann_laplacian_with_mean_projection(oracle, octx_h025, name):
    """Singular P1 Laplacian (all-Neumann), zero-mean RHS: the solution is the
    zero-mean one (pseudo-inverse), to 1e-10 relative residual."""
    from nsfem_amd import cu
    m = octx_h025["Kp"]
    n = m.shape[0]
    rng = np.random.default_rng(31)
    b = rng.uniform(-1, 1, n)
    b -= b.mean()
    s = cu.createSolver(name, 1, m, 5000, 1e-10, 1)
    s.setOption("projectMean", 1)
    s.getRhs().upload(b)
    s.getSol().memsetZero()
    s.solve()
    x = s.getSol().download()
    assert abs(x.mean()) <= 1e-12 * np.max(np.abs(x))
    assert np.linalg.norm(b - m @ x) <= 1.01e-10 * np.linalg.norm(b)
    xp = np.linalg.lstsq(m.toarray(), b, rcond=None)[0]
    xp -= xp.mean()
    assert np.linalg.norm(x - xp) <= 1e-7 * np.linalg.norm(xp)
    xo, rel, it = oracle.pcg_solve(m, b, np.zeros(n), 1e-10, 5000)
    assert np.linalg.norm((xo - xo.mean()) - xp) <= 1e-7 * np.linalg.norm(xp)


def test_kernel_timing_hooks_report_per_launch_figures(octx_h025):
    """The measurement hook replays one iteration's launches: positive times
    and byte counts, 2C-3 sweep launches per iteration (2C-4 when the update
    kernel also sweeps colour 1: fused reductions on one system), and the
    solver still solves afterwards."""
    m = octx_h025["Kp_int"]
    rng = np.random.default_rng(32)
    b = rng.uniform(-1, 1, m.shape[0])
    s, _ = _solve("pcg", m, b, max_iters=2000, target=1e-10, rowsUpdate=64)
    C = s.numColors()
    nnz = m.nnz
    # one system: the reductions are fused by default (no finalize launches)
    for fused in (-1, 0, 1):
        s.setOption("fusedReductions", fused)
        expect = {"direction": 1, "sweep": 2 * C - 3 - (fused != 0), "rho": 1, "update": 1,
                  "finalize": 2 if fused == 0 else 0}
        for k, launches in expect.items():
            ms, by, per_it = s.timeKernel(k, reps=3)
            assert ms > 0 and by > 0 and per_it == launches, (k, fused)
    s.setOption("fusedReductions", -1)
    # a direction launch streams the whole off-diagonal matrix once
    assert s.timeKernel("direction", reps=1)[1] >= (nnz - m.shape[0]) * 12
    s.getRhs().upload(b)
    s.getSol().memsetZero()
    s.solve()
    x = s.getSol().download()
    assert np.linalg.norm(b - m @ x) <= 1.0001e-10 * np.linalg.norm(b)


@pytest.fixture(scope="module")
def channel20k_kp_int():
    from nsfem_amd import fem, mesher
    m = mesher.channel_mesh(20000, seed=5)
    ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), fem.DfgConditions(0.01, 1.0, 1.0, 4.0))
    return ctx.scipy("pressureStiffnessInternal")


@pytest.mark.parametrize("restart", [0, 1])
def test_pcg_mixed_iteration_counts(channel20k_kp_int, restart):
    """Default pcg_mixed (FP64 Krylov vectors and products, FP32 symmetric GS
    preconditioner) converges in the FP64 solver's iteration count (within
    2 %), to the same FP64 relative residual, on a 20k-row pressure system
    (below ~5k rows the FP32 preconditioner's rounding adds up to ~30 % to
    the count for some right-hand sides).  restart = 1 is iterative refinement: its FP32
    inner solves from zero cost extra iterations."""
    m = channel20k_kp_int
    rng = np.random.default_rng(33)
    b = rng.uniform(-1, 1, m.shape[0])
    s64, x64 = _solve("pcg", m, b, max_iters=5000, target=1e-10)
    smx, x = _solve("pcg_mixed", m, b, max_iters=5000, target=1e-10, mixedRestart=restart)
    assert np.linalg.norm(b - m @ x) <= 1.0001e-10 * np.linalg.norm(b)
    assert smx.getLastRelRes() <= 1e-10
    it64, itmx = s64.getLastIterations(), smx.getLastIterations()
    if restart == 0:
        assert abs(itmx - it64) <= max(2, 0.02 * it64), (itmx, it64)
        assert smx.getLastOuterIterations() == 0
        assert np.linalg.norm(x - x64) <= 1e-8 * np.linalg.norm(x64)
    else:
        assert itmx > it64
        assert smx.getLastOuterIterations() >= 2


def test_reshape_and_autotune_keep_results(octx_h025):
    """Launch shapes change only the reduction order: every shape converges
    in the same iterations (+-1) to the same solution; autotune picks a valid
    shape and leaves a working solver."""
    m = octx_h025["Kp_int"]
    rng = np.random.default_rng(34)
    b = rng.uniform(-1, 1, m.shape[0])
    s, x0 = _solve("pcg", m, b, max_iters=3000, target=1e-11)
    it0 = s.getLastIterations()
    for lanes, rows in ((1, 256), (2, 128), (4, 64), (8, 32), (16, 16), (8, 256), (0, 0)):
        s.reshape(lanes, rows)
        L, R = s.shape()
        assert L in (1, 2, 4, 8, 16) and R % (256 // L) == 0 and R >= 256 // L
        assert s.sweepShape() == (L, R)
        s.getRhs().upload(b)
        s.getSol().memsetZero()
        s.solve()
        x = s.getSol().download()
        assert abs(s.getLastIterations() - it0) <= 1
        assert np.linalg.norm(x - x0) <= 1e-9 * np.linalg.norm(x0)
    # the colour sweeps on shapes of their own
    for lanes, rows in ((2, 64), (16, 512), (4, 0)):
        s.reshapeSweeps(lanes, rows)
        SL, SR = s.sweepShape()
        slab = 256 // lanes  # rows per slab; a workgroup holds whole slabs
        assert SL == lanes and SR % slab == 0 and (rows == 0 or SR == max(slab, rows // slab * slab))
        s.getRhs().upload(b)
        s.getSol().memsetZero()
        s.solve()
        x = s.getSol().download()
        assert abs(s.getLastIterations() - it0) <= 1
        assert np.linalg.norm(x - x0) <= 1e-9 * np.linalg.norm(x0)
    L, R = s.autotune(reps=2)
    assert (L, R) == s.shape()
    SL, SR = s.sweepShape()
    assert SL in (1, 2, 4, 8, 16) and SR % (256 // SL) == 0
    s.getRhs().upload(b)
    s.getSol().memsetZero()
    s.solve()
    assert np.linalg.norm(b - m @ s.getSol().download()) <= 1.0001e-11 * np.linalg.norm(b)
    from nsfem_amd import _lib
    with pytest.raises(_lib.NsfemInvalidArgument):
        s.reshape(3, 0)


@pytest.mark.parametrize("name,nch", [("pcg", 1), ("pcg", 2), ("pcg_mixed", 1), ("pcg_mixed", 2), ("pcg", 8)])
def test_fused_reductions_match_finalize_kernels(octx_h025, name, nch):
    """Fused reductions (the direction kernel finishes rho and the
    convergence test, the update kernel the step length, every block from
    the partial sums) against the one-block finalize kernels: the sums
    differ only in order, so both converge in the same iterations (+-1) to
    the same solution, and the fused mode repeats bit for bit."""
    m = octx_h025["Kp_int"]
    n = m.shape[0]
    rng = np.random.default_rng(35)
    b = rng.uniform(-1, 1, nch * n)
    runs = {}
    for fused in (0, 1, 1):
        s, x = _solve(name, m, b, nch=nch, max_iters=3000, target=1e-10, fusedReductions=fused)
        its = s.getLastIterations()
        if fused in runs:
            assert its == runs[fused][1] and np.array_equal(x, runs[fused][0])
        runs[fused] = (x, its)
        for c in range(nch):
            bc = b[c * n:(c + 1) * n]
            assert np.linalg.norm(bc - m @ x[c * n:(c + 1) * n]) <= 1.0001e-10 * np.linalg.norm(bc)
            assert s.getLastRelRes(c) <= 1e-10
    (x0, it0), (x1, it1) = runs[0], runs[1]
    assert abs(it0 - it1) <= 1
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)


@pytest.mark.parametrize("fused", [0, 1])
def test_iteration_cap_records_the_last_update(octx_h025, fused):
    """A solve that stops at maxIters: the iteration count is the cap and the
    reported relative residual is that of the returned x (the last update's
    convergence test is recorded also when no direction kernel follows)."""
    m = octx_h025["Kp_int"]
    rng = np.random.default_rng(40)
    b = rng.uniform(-1, 1, m.shape[0])
    for cap in (1, 2, 17):
        s, x = _solve("pcg", m, b, max_iters=cap, target=1e-12, fusedReductions=fused)
        assert s.getLastIterations() == cap
        rel = np.linalg.norm(b - m @ x) / np.linalg.norm(b)
        assert abs(s.getLastRelRes() - rel) <= 1e-9 * rel


def test_fused_reduction_row_counts(octx_h025):
    """Own rows per workgroup for the direction, update and rho kernels:
    results repeat within the iteration tolerance, the shape is reported."""
    m = octx_h025["Kp_int"]
    rng = np.random.default_rng(41)
    b = rng.uniform(-1, 1, m.shape[0])
    s, x0 = _solve("pcg_mixed", m, b, max_iters=3000, target=1e-10)
    it0 = s.getLastIterations()
    L, R = s.shape()
    for rows in ((4 * R, 2 * R, 8 * R), (R, 16 * R, 2 * R), (0, 0, 0)):
        for key, v in zip(("rowsDirection", "rowsUpdate", "rowsRho"), rows):
            s.setOption(key, v)
        assert s.kindRows() == tuple(v if v else R for v in rows)
        s.getRhs().upload(b)
        s.getSol().memsetZero()
        s.solve()
        x = s.getSol().download()
        assert abs(s.getLastIterations() - it0) <= 1
        assert np.linalg.norm(x - x0) <= 1e-9 * np.linalg.norm(x0)


def test_fused_reductions_option_on_a_batch(oracle):
    """Fused reductions on a batch: every workgroup sums its own system's
    partials and each system's first workgroup records its decision; every
    system converges as with the finalize kernels (the oracle's iterations
    +-1) to the same solution."""
    from nsfem_amd import cu, fem, mesher
    mats = []
    for k, n in enumerate([600, 2500, 1300]):
        m = mesher.channel_mesh(n, seed=100 + k)
        ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), fem.DfgConditions(0.01, 1.0, 1.0, 4.0))
        mats.append(ctx.scipy("pressureStiffnessInternal"))
    rng = np.random.default_rng(36)
    bs = [rng.uniform(-1, 1, m.shape[0]) for m in mats]
    xs = {}
    for fused in (0, 1):
        s = cu.createBatchSolver("pcg", 1, mats, 3000, 1e-11, 1)
        s.setOption("fusedReductions", fused)
        s.getRhs().upload(np.concatenate(bs))
        s.getSol().memsetZero()
        s.solve()
        xs[fused] = s.getSol().download()
        for k, (m, b) in enumerate(zip(mats, bs)):
            o = s.system_offsets[k]
            assert np.linalg.norm(b - m @ xs[fused][o:o + m.shape[0]]) <= 1.0001e-11 * np.linalg.norm(b)
            xo, rel, it = oracle.pcg_solve(m, b, np.zeros_like(b), 1e-11, 3000)
            assert abs(s.getLastIterations(k) - it) <= 1, (fused, k)
            assert s.getLastRelRes(0, k) <= 1e-11
    assert np.linalg.norm(xs[1] - xs[0]) <= 1e-8 * np.linalg.norm(xs[0])


@pytest.mark.parametrize("name", ["pcg", "pcg_mixed"])
@pytest.mark.parametrize("nch", [4, 8, 32])
def test_wide_channel_tiles_match_single_channel_solves(octx_h025, name, nch):
    """4 to 32 right-hand sides sharing the matrix, one lane per channel:
    every column reaches the tolerance and equals its own one-channel solve
    (columns scaled over six decades, so each stops on its own test).  FP64
    pcg needs the iterations of the slowest column; pcg_mixed counts jitter
    by up to ~30 % on a system this small (the FP32 preconditioner's
    rounding), so only its results are compared."""
    m = octx_h025["Kp_int"]
    n = m.shape[0]
    rng = np.random.default_rng(37)
    scale = np.repeat(10.0 ** rng.uniform(-3, 3, nch), n)
    b = rng.uniform(-1, 1, nch * n) * scale
    s, x = _solve(name, m, b, nch=nch, max_iters=3000, target=1e-10)
    for c in range(nch):
        bc = b[c * n:(c + 1) * n]
        assert np.linalg.norm(bc - m @ x[c * n:(c + 1) * n]) <= 1.0001e-10 * np.linalg.norm(bc), c
        assert s.getLastRelRes(c) <= 1e-10
    its = []
    for c in (0, nch // 2, nch - 1):
        s1, x1 = _solve(name, m, b[c * n:(c + 1) * n], max_iters=3000, target=1e-10)
        its.append(s1.getLastIterations())
        assert np.linalg.norm(x[c * n:(c + 1) * n] - x1) <= 1e-8 * np.linalg.norm(x1), c
    if name == "pcg":
        assert s.getLastIterations() >= max(its) - 1


@pytest.mark.parametrize("nch", [4, 16])
def test_wide_channel_gs_matches_oracle_sweeps(oracle, octx_c3, nch):
    """Wide-tile GS sweeps: 25 sweeps of every column equal the oracle's
    sequential GS in colouring order.  The RMS residual (~1e-10 of |b|) is
    a difference of O(1) terms: it agrees to ~1e-6 relative."""
    m = octx_c3["M"]
    n = m.shape[0]
    rng = np.random.default_rng(38)
    b = rng.uniform(-1, 1, nch * n)
    s, x = _solve("gs", m, b, nch=nch, max_iters=25, target=1e-30, interval=5)
    L = oracle.Mcgs(m)
    for c in range(nch):
        xo, mse, it = oracle.gs_solve(m, b[c * n:(c + 1) * n], np.zeros(n), 25, 0.0, 5, L)
        assert s.getLastIterations() == it
        assert abs(s.getLastMse(c) - mse) <= 1e-6 * mse
        assert np.max(np.abs(x[c * n:(c + 1) * n] - xo)) <= 1e-12 * np.max(np.abs(xo)), c


def test_wide_channel_pure_neumann_batch(octx_h025):
    """8 zero-mean right-hand sides on the singular Laplacian with mean
    projection (the multi-right-hand-side stress configuration in small):
    pcg_mixed to 1e-8 per column, zero-mean solutions; autotuning keeps the
    one-lane-per-channel shape."""
    from nsfem_amd import cu
    m = octx_h025["Kp"]
    n = m.shape[0]
    nch = 8
    rng = np.random.default_rng(39)
    b = rng.uniform(-1, 1, (nch, n))
    b -= b.mean(axis=1, keepdims=True)
    s = cu.createSolver("pcg_mixed", nch, m, 5000, 1e-8, 1)
    s.setOption("projectMean", 1)
    L, R = s.autotune(reps=2)
    assert L == nch and R % (256 // nch) == 0
    s.getRhs().upload(b.ravel())
    s.getSol().memsetZero()
    s.solve()
    x = s.getSol().download().reshape(nch, n)
    for c in range(nch):
        assert abs(x[c].mean()) <= 1e-12 * np.max(np.abs(x[c]))
        assert np.linalg.norm(b[c] - m @ x[c]) <= 1.0001e-8 * np.linalg.norm(b[c])


def test_batch_of_uneven_systems_with_wide_tiles():
    """A block-diagonal batch of uneven systems, each with 4 right-hand sides
    (one lane per channel, per-system workgroup table): every column of every
    system equals its own one-system solve."""
    from nsfem_amd import cu, fem, mesher
    mats = []
    for k, n in enumerate([700, 1800]):
        m = mesher.channel_mesh(n, seed=200 + k)
        ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), fem.DfgConditions(0.01, 1.0, 1.0, 4.0))
        mats.append(ctx.scipy("pressureStiffnessInternal"))
    nch = 4
    rng = np.random.default_rng(42)
    bs = [rng.uniform(-1, 1, nch * m.shape[0]) for m in mats]
    for name in ("pcg", "pcg_mixed"):
        s = cu.createBatchSolver(name, nch, mats, 3000, 1e-10, 1)
        # io layout: channel-planar over the stacked systems
        total = sum(m.shape[0] for m in mats)
        rhs = np.zeros(nch * total)
        for k, (m, b) in enumerate(zip(mats, bs)):
            o = s.system_offsets[k]
            for c in range(nch):
                rhs[c * total + o:c * total + o + m.shape[0]] = b[c * m.shape[0]:(c + 1) * m.shape[0]]
        s.getRhs().upload(rhs)
        s.getSol().memsetZero()
        s.solve()
        x = s.getSol().download()
        for k, (m, b) in enumerate(zip(mats, bs)):
            o, n = s.system_offsets[k], m.shape[0]
            s1, x1 = _solve(name, m, b, nch=nch, max_iters=3000, target=1e-10)
            for c in range(nch):
                xc = x[c * total + o:c * total + o + n]
                bc = b[c * n:(c + 1) * n]
                assert np.linalg.norm(bc - m @ xc) <= 1.0001e-10 * np.linalg.norm(bc), (name, k, c)
                assert np.linalg.norm(xc - x1[c * n:(c + 1) * n]) <= 1e-8 * np.linalg.norm(xc), (name, k, c)


@pytest.mark.parametrize("nch", [1, 2])
def test_direction_kernel_lanes(octx_h025, nch):
    """The direction kernel on its own lanes per row: same iterations (+-1)
    and solutions as the main shape's."""
    m = octx_h025["Kp_int"]
    n = m.shape[0]
    rng = np.random.default_rng(43)
    b = rng.uniform(-1, 1, nch * n)
    s, x0 = _solve("pcg", m, b, nch=nch, max_iters=3000, target=1e-10)
    it0 = s.getLastIterations()
    for lanes in (1, 2, 4, 8, 16, 0):
        s.setOption("lanesDirection", lanes)
        s.getRhs().upload(b)
        s.getSol().memsetZero()
        s.solve()
        x = s.getSol().download()
        assert abs(s.getLastIterations() - it0) <= 1, lanes
        assert np.linalg.norm(x - x0) <= 1e-9 * np.linalg.norm(x0), lanes


@pytest.mark.parametrize("name,nch", [("pcg", 1), ("pcg", 2), ("pcg_mixed", 2), ("pcg", 4)])
def test_region_major_direction_map(octx_h025, name, nch):
    """The direction kernel on the region-major map (each XCD's workgroups
    take one stretch of every colour; the default): same iterations (+-1) and
    solutions as the row-order map, also with larger direction workgroups, and the
    repeated solve is bit-identical (its partial sums keep a fixed order)."""
    m = octx_h025["Kp_int"]
    n = m.shape[0]
    rng = np.random.default_rng(47)
    b = rng.uniform(-1, 1, nch * n)
    s, x0 = _solve(name, m, b, nch=nch, max_iters=3000, target=1e-10, regionDirection=0)
    it0 = s.getLastIterations()
    for rows in (0, 512):
        s.setOption("rowsDirection", rows)
        s.setOption("regionDirection", 1)
        xs = []
        for _ in range(2):
            s.getRhs().upload(b)
            s.getSol().memsetZero()
            s.solve()
            xs.append(s.getSol().download())
            assert abs(s.getLastIterations() - it0) <= 1, rows
        assert np.array_equal(xs[0], xs[1]), rows
        assert np.linalg.norm(xs[0] - x0) <= 1e-9 * np.linalg.norm(x0), rows
        res = np.linalg.norm(b.reshape(nch, n).T - m @ xs[0].reshape(nch, n).T, axis=0)
        assert np.all(res <= 1.01e-10 * np.linalg.norm(b.reshape(nch, n), axis=1)), rows
        s.setOption("regionDirection", 0)


@pytest.mark.parametrize("name,nch", [("pcg", 1), ("pcg", 2), ("pcg_mixed", 1), ("pcg_mixed", 2)])
def test_update_fused_with_first_sweep(octx_h025, name, nch):
    """The update kernel fused with the colour-1 forward sweep (the default
    on one system): same iterations (+-1) and solutions as the separate
    launches, with several update workgroup sizes, and the repeated solve is
    bit-identical; with the iteration count fixed (an unreachable target) the
    two give bit-identical solutions: the sweep computes its colour-0 values
    with the arithmetic of the stored ones."""
    m = octx_h025["Kp_int"]
    n = m.shape[0]
    rng = np.random.default_rng(53)
    b = rng.uniform(-1, 1, nch * n)
    fixed = []
    for fz in (0, 1):
        sf, xf = _solve(name, m, b, nch=nch, max_iters=40, target=1e-30, rowsUpdate=64, fuseUpdateSweep=fz)
        assert sf.getLastIterations() == 40
        assert sf.timeKernel("sweep", reps=1)[2] == 2 * sf.numColors() - 3 - fz
        fixed.append(xf)
    assert np.array_equal(fixed[0], fixed[1])
    s, x0 = _solve(name, m, b, nch=nch, max_iters=3000, target=1e-10, fuseUpdateSweep=0)
    it0 = s.getLastIterations()
    for rows in (0, 64, 1024):
        s.setOption("rowsUpdate", rows)
        s.setOption("fuseUpdateSweep", 1)
        xs = []
        for _ in range(2):
            s.getRhs().upload(b)
            s.getSol().memsetZero()
            s.solve()
            xs.append(s.getSol().download())
            assert abs(s.getLastIterations() - it0) <= 1, rows
        assert np.array_equal(xs[0], xs[1]), rows
        assert np.linalg.norm(xs[0] - x0) <= 1e-9 * np.linalg.norm(x0), rows
        res = np.linalg.norm(b.reshape(nch, n).T - m @ xs[0].reshape(nch, n).T, axis=0)
        assert np.all(res <= 1.01e-10 * np.linalg.norm(b.reshape(nch, n), axis=1)), rows
        s.setOption("fuseUpdateSweep", 0)
