"""The batched (uneven instances in one block-diagonal step) and multi-RHS
configurations against the FP64 oracle, and at their full sizes through
properties that hold at any size.

Tolerances: per-instance and per-column oracle parity is held to 1e-10
relative with every solve at a 1e-13 relative residual on both sides (the
plain FP64 bar); the full-size runs check their stated residual targets
(1e-10 per step, 1e-8 per column), exact Dirichlet values, zero means and
bit-repeatability."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _oracle_context(oracle, m, nu, peak):
    tri = {"xy": m["xy"].astype(np.float64), "triangles": m["triangles"],
           "border": oracle.border_from_edges(m["triangles"], m["border"]), "groups": m["groups"]}
    return oracle.build_context(tri, nu, peak, channel_height=m["height"], profile_scale=4.0)


def _instance_slices(batch, u, p, k, n, v):
    vo, po = batch.system_offsets()
    Nt = batch.numVelocityNodes
    return np.concatenate([u[vo[k]:vo[k] + n], u[Nt + vo[k]:Nt + vo[k] + n]]), p[po[k]:po[k] + v]


def test_batched_step_matches_oracle_per_instance(oracle):
    """Three uneven instances (different meshes, hole radii, inflow peaks and
    viscosities) stepped as one block-diagonal batch: every instance's
    velocity and pressure equal the oracle's run of that instance alone."""
    from nsfem_amd import fem, mesher
    specs = [(1500, 0.10, 1.0, 0.02), (4000, 0.14, 0.6, 0.05), (2500, 0.08, 1.8, 0.015)]
    raw, meshes, conds = [], [], []
    for k, (n, r, peak, nu) in enumerate(specs):
        m = mesher.channel_mesh(n, radius=r, seed=300 + k)
        raw.append((m, nu, peak))
        meshes.append(mesher.to_triangle_mesh(m))
        conds.append(fem.DfgConditions(nu, peak, m["height"], 4.0))
    batch = fem.ChorinContext.batch(meshes, conds)
    cfg = fem.ChorinConfig(fem.SolverConfig("pcg", 2000, 1e-13, 1), fem.SolverConfig("pcg", 8000, 1e-13, 1))
    ch = fem.Chorin(batch, cfg)
    tau = min(2e-4, ch.time_step_info()["tauMax"])
    ch.set_time_step(tau)
    orcs = []
    for m, nu, peak in raw:
        octx = _oracle_context(oracle, m, nu, peak)
        orcs.append(oracle.ChorinOracle(octx, tau, vel=("pcg", 2000, 1e-13, 1), pres=("pcg", 8000, 1e-13, 1)))
    for step in range(3):
        st = ch.step()
        assert st["relTent"] <= 1e-13 and st["relPressure"] <= 1e-13 and st["relFinal"] <= 1e-13
        u, p = ch.download()
        for k, orc in enumerate(orcs):
            uo, po = orc.step()
            uk, pk = _instance_slices(batch, u, p, k, orc.ctx["N"], orc.ctx["V"])
            assert _rel(uk, uo) <= 1e-10, (step, k, _rel(uk, uo))
            assert _rel(pk, po) <= 1e-10, (step, k, _rel(pk, po))


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_batch_of_48_uneven_instances(oracle):
    """The batched configuration at full size: 48 seeded instances, P1 DOFs
    log-uniform in [10k, 250k], radii in [0.08, 0.16], peaks in [0.5, 2],
    Re in [20, 100], one block-diagonal step (one launch per kernel for all
    instances, per-instance colouring and convergence).  Two FP64 steps at
    the batch's clamped time step: every solve of every instance at its
    1e-13 relative residual, Dirichlet values exact, a rerun from the same
    state bit-identical; the smallest instance equals the oracle's run of it
    to 1e-10; the mixed path agrees with FP64 within what 1e-10 solves allow."""
    import time
    from nsfem_amd import fem, mesher
    t0 = time.perf_counter()
    inst = mesher.uneven_instances(48, 20241103)
    meshes = [mesher.to_triangle_mesh(d["mesh"]) for d in inst]
    conds = [fem.DfgConditions(d["viscosity"], d["peak"], d["mesh"]["height"], 4.0) for d in inst]
    batch = fem.ChorinContext.batch(meshes, conds, gpu_assembly=True)
    sizes = [len(d["mesh"]["xy"]) for d in inst]
    print(f"48 instances: {sum(sizes)} pressure DOFs ({min(sizes)}-{max(sizes)}), set-up {time.perf_counter() - t0:.1f} s",
          flush=True)
    assert batch.num_systems() == 48 and batch.numPressureNodes == sum(sizes)
    assert min(sizes) >= 9_000 and max(sizes) <= 260_000
    N = batch.numVelocityNodes
    rng = np.random.default_rng(1)
    u0 = rng.uniform(-1, 1, 2 * N) * 1e-3
    ids, vals = batch.dirichlet("vx")
    idy, valy = batch.dirichlet("vy")
    u0[ids] = vals
    u0[idy + N] = valy

    cfg = fem.ChorinConfig(fem.SolverConfig("pcg", 5000, 1e-13, 1), fem.SolverConfig("pcg", 50000, 1e-13, 1),
                           timeStepPolicy=fem.TAU_CLAMP)
    ch = fem.Chorin(batch, cfg)
    ch.set_time_step(5e-4)
    tau = ch.time_step_info()["tauUsed"]
    runs = []
    for rep in range(2):
        ch.reset()
        ch.upload_velocity(u0)
        for _ in range(2):
            st = ch.step()
            assert st["relTent"] <= 1e-13 and st["relPressure"] <= 1e-13 and st["relFinal"] <= 1e-13, st
        u, p = ch.download()
        assert np.isfinite(u).all() and np.isfinite(p).all()
        assert np.array_equal(u[ids], vals) and np.array_equal(u[idy + N], valy)
        runs.append((u, p))
        print(f"run {rep}: pressure iterations {st['itersPressure']}, tau {tau:.4g}", flush=True)
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    del ch

    # the smallest instance against the oracle, from the same initial state
    k = int(np.argmin(sizes))
    d = inst[k]
    octx = _oracle_context(oracle, d["mesh"], d["viscosity"], d["peak"])
    orc = oracle.ChorinOracle(octx, tau, vel=("pcg", 2000, 1e-13, 1), pres=("pcg", 20000, 1e-13, 1))
    vo, _ = batch.system_offsets()
    n = octx["N"]
    orc.u[:] = np.concatenate([u0[vo[k]:vo[k] + n], u0[N + vo[k]:N + vo[k] + n]])
    for _ in range(2):
        uo, po = orc.step()
    uk, pk = _instance_slices(batch, *runs[0], k, n, octx["V"])
    assert _rel(uk, uo) <= 1e-10 and _rel(pk, po) <= 1e-10, (_rel(uk, uo), _rel(pk, po))

    cfgm = fem.ChorinConfig(fem.SolverConfig("pcg_mixed", 5000, 1e-10, 1), fem.SolverConfig("pcg_mixed", 50000, 1e-10, 1),
                            timeStepPolicy=fem.TAU_UNCHECKED)
    mx = fem.Chorin(batch, cfgm)
    mx.set_time_step(tau)
    mx.upload_velocity(u0)
    for _ in range(2):
        st = mx.step()
        assert st["relTent"] <= 1e-10 and st["relPressure"] <= 1e-10 and st["relFinal"] <= 1e-10
    um, pm = mx.download()
    du = np.linalg.norm(um - runs[0][0]) / np.linalg.norm(runs[0][0])
    dp = np.linalg.norm(pm - runs[0][1]) / np.linalg.norm(runs[0][1])
    print(f"48 instances: mixed vs FP64 after two steps: velocity {du:.2e}, pressure {dp:.2e}")
    assert du <= 1e-9 and dp <= 1e-6


def _p1_laplacian(oracle, m):
    tri = {"xy": m["xy"].astype(np.float64), "triangles": m["triangles"], "border": np.zeros((0, 3), np.int64),
           "groups": m["groups"]}
    B, det = oracle.geometry(tri["xy"], tri["triangles"])
    T = tri["triangles"]
    rp = np.repeat(T[:, :, None], 3, 2)
    cp = np.repeat(T[:, None, :], 3, 1)
    return oracle._coo(rp, cp, oracle.local_stiffness(B, det, 1), (len(tri["xy"]), len(tri["xy"])))


@pytest.mark.parametrize("width", [4, 8])
def test_multi_rhs_columns_match_oracle(oracle, width):
    """A tile of right-hand sides sharing the matrix (the multi-RHS
    configuration's solver): every column equals the oracle's SGS-PCG of that
    column alone (FP64, 1e-13) to 1e-10, with the same iteration count +-1;
    the pure-Neumann case (mean projection) against the oracle's solution
    made zero-mean."""
    from nsfem_amd import cu, mesher
    m = mesher.channel_mesh(3000, seed=77)
    K = _p1_laplacian(oracle, m)
    n = K.shape[0]
    rng = np.random.default_rng(5)
    B = rng.uniform(-1, 1, (width, n))
    B -= B.mean(axis=1, keepdims=True)
    s = cu.createSolver("pcg", width, K, 5000, 1e-13, 1)
    s.setOption("projectMean", 1)
    s.getRhs().upload(B.ravel())
    s.getSol().memsetZero()
    s.solve()
    X = s.getSol().download().reshape(width, n)
    layout = oracle.Mcgs(K)
    its = []
    for c in range(width):
        xo, rel, it = oracle.pcg_solve(None, B[c], np.zeros(n), 1e-13, 5000, layout)
        xo -= xo.mean()
        its.append(it)
        assert rel <= 1e-13
        assert _rel(X[c], xo) <= 1e-10, (c, _rel(X[c], xo))
        assert abs(X[c].mean()) <= 1e-12 * np.max(np.abs(X[c]))
    # the tile runs until its slowest column has converged
    assert abs(s.getLastIterations() - max(its)) <= 1, (s.getLastIterations(), its)


@pytest.fixture(scope="module")
def laplacian_4m(oracle):
    """The pure-Neumann P1 Laplacian of a seeded 4M-node channel mesh and 32
    zero-mean uniform(-1, 1) right-hand sides (rows x 32)."""
    import time
    from nsfem_amd import mesher
    t0 = time.perf_counter()
    m = mesher.channel_mesh(4_000_000, seed=20241103)
    K = _p1_laplacian(oracle, m)
    del m
    n = K.shape[0]
    rng = np.random.default_rng(1)
    B = rng.uniform(-1, 1, (32, n))
    B -= B.mean(axis=1, keepdims=True)
    print(f"4M multi-RHS: {n} rows, {K.nnz} nnz, set-up {time.perf_counter() - t0:.1f} s", flush=True)
    return K, B


def _check_columns(K, B, X, tol):
    for c in range(B.shape[0]):
        rel = np.linalg.norm(B[c] - K @ X[c]) / np.linalg.norm(B[c])
        assert rel <= tol, (c, rel)
        assert abs(X[c].mean()) <= 1e-12 * np.max(np.abs(X[c])), c


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_multi_rhs_4m_rows_32_columns(laplacian_4m):
    """The multi-RHS configuration at full size with independent columns:
    pcg_mixed (FP32 preconditioner, FP64 Krylov vectors and reductions) with
    mean projection, in tiles of 8 columns: every column's true relative
    residual |b - A x| / |b| (FP64, on the host) is at most 1e-8 and its
    mean is zero."""
    import time
    from nsfem_amd import cu
    K, B = laplacian_4m
    n = K.shape[0]
    assert 3_900_000 <= n <= 4_100_000
    width = 8
    s = cu.createSolver("pcg_mixed", width, K, 20000, 1e-8, 1)
    s.setOption("projectMean", 1)
    X = np.empty_like(B)
    for first in range(0, 32, width):
        t = time.perf_counter()
        s.getRhs().upload(B[first:first + width].ravel())
        s.getSol().memsetZero()
        s.solve()
        X[first:first + width] = s.getSol().download().reshape(width, n)
        print(f"columns {first}-{first + width - 1}: {s.getLastIterations()} iterations, {time.perf_counter() - t:.1f} s", flush=True)
    _check_columns(K, B, X, 1e-8 * (1 + 1e-6))


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_block_pcg_4m_rows_32_columns(laplacian_4m):
    """The same 32 right-hand sides with block PCG (one Krylov space for all
    32 columns, 32 x 32 step matrices; FP32 preconditioner, FP64 Krylov
    vectors and reductions): every column's true relative residual at most
    1e-8 (the solver stops on its recursive residual, which drifts from the
    true one by rounding: allowed 1 % on top) and zero mean, in far fewer
    iterations than a single column needs."""
    import time
    from nsfem_amd import cu
    K, B = laplacian_4m
    n = K.shape[0]
    s = cu.createSolver("bpcg_mixed", 32, K, 20000, 1e-8, 1)
    s.setOption("projectMean", 1)
    t = time.perf_counter()
    s.getRhs().upload(B.ravel())
    s.getSol().memsetZero()
    s.solve()
    X = s.getSol().download().reshape(32, n)
    its = s.getLastIterations()
    print(f"block PCG, 32 columns: {its} iterations, {time.perf_counter() - t:.1f} s", flush=True)
    _check_columns(K, B, X, 1.01e-8)
    assert its < 2000


@pytest.mark.timeout(120)
@pytest.mark.parametrize("method", ["pcg", "pcg_mixed"])
def test_batch_stability_limit_is_the_tightest_instance(method):
    """The time-step guard on a batch whose instances' lambda_max(M^-1 nu K)
    differ by a factor of about 400 (viscosities 100x apart on meshes 4x
    apart): the batch's lambda is the largest instance's.  The power
    iteration normalises every instance on its own, so the smaller ones do
    not decay into the FP32 underflow of a mixed preconditioner."""
    from nsfem_amd import fem, mesher
    specs = [(6000, 0.05), (1500, 0.0005)]
    meshes, conds, singles = [], [], []
    cfg = fem.ChorinConfig(fem.SolverConfig(method, 2000, 1e-10, 1), fem.SolverConfig(method, 8000, 1e-10, 1))
    for k, (n, nu) in enumerate(specs):
        m = mesher.channel_mesh(n, seed=500 + k)
        meshes.append(mesher.to_triangle_mesh(m))
        conds.append(fem.DfgConditions(nu, 1.0, m["height"], 4.0))
        singles.append(fem.Chorin(fem.ChorinContext(meshes[-1], conds[-1]), cfg).time_step_info()["lambdaNu"])
    assert max(singles) / min(singles) > 100
    lam = fem.Chorin(fem.ChorinContext.batch(meshes, conds), cfg).time_step_info()["lambdaNu"]
    assert abs(lam / max(singles) - 1) < 0.05, (lam, singles)



def _laplacian_3000(oracle):
    from nsfem_amd import mesher
    return _p1_laplacian(oracle, mesher.channel_mesh(3000, seed=77))


@pytest.mark.parametrize("method,width,tol", [("bpcg", 4, 1e-13), ("bpcg", 8, 1e-13), ("bpcg", 32, 1e-13),
                                              ("bpcg_mixed", 8, 1e-8), ("bpcg_mixed", 16, 1e-8), ("bpcg_mixed", 32, 1e-8)])
def test_block_pcg_matches_oracle(oracle, method, width, tol):
    """Block PCG (one Krylov space for the tile, width x width step matrices,
    the symmetric multicolour GS preconditioner; bpcg_mixed: the
    preconditioner in FP32) on the pure-Neumann P1 Laplacian with mean
    projection.  The solve ends on verified true residuals, so every column's
    true relative residual (FP64, host) is at most tol (1 % for the
    difference in summation order); the column equals the oracle's
    single-column SGS-PCG solution (1e-13, made zero-mean) within what the
    residuals allow (cond(A) ~ 1e4 times tol: 1e-9 at 1e-13, 1e-3 at 1e-8);
    fewer iterations than the oracle's slowest column at the same
    tolerance.  The mixed variant is tested at the multi-RHS
    configuration's 1e-8."""
    from nsfem_amd import cu
    K = _laplacian_3000(oracle)
    n = K.shape[0]
    rng = np.random.default_rng(9)
    B = rng.uniform(-1, 1, (width, n))
    B -= B.mean(axis=1, keepdims=True)
    s = cu.createSolver(method, width, K, 5000, tol, 1)
    s.setOption("projectMean", 1)
    s.getRhs().upload(B.ravel())
    s.getSol().memsetZero()
    s.solve()
    X = s.getSol().download().reshape(width, n)
    layout = oracle.Mcgs(K)
    its = []
    bound = 1e-9 if tol <= 1e-13 else 1e-3
    for c in range(width):
        rel = np.linalg.norm(B[c] - K @ X[c]) / np.linalg.norm(B[c])
        assert rel <= 1.01 * tol, (c, rel)
        assert s.getLastRelRes(c) <= tol
        its.append(oracle.pcg_solve(None, B[c], np.zeros(n), tol, 5000, layout)[2])
        xo = oracle.pcg_solve(None, B[c], np.zeros(n), 1e-13, 5000, layout)[0]
        xo -= xo.mean()
        assert _rel(X[c], xo) <= bound, (c, _rel(X[c], xo))
        assert abs(X[c].mean()) <= 1e-12 * np.max(np.abs(X[c]))
    print(f"{method} x{width} at {tol:g}: {s.getLastIterations()} block iterations, {s.getLastOuterIterations()} more cycles; "
          f"single columns {min(its)}-{max(its)}")
    assert s.getLastIterations() < max(its)


def test_block_pcg_mixed_at_a_tight_tolerance(oracle):
    """bpcg_mixed well below its usual tolerance: the FP32 preconditioner's
    rounding (about 1e-7 relative, not exactly symmetric) degrades a wide
    block near 1e-10 - directions are dropped and cycles restarted - but the
    solve still ends on verified true residuals at the tolerance."""
    from nsfem_amd import cu
    K = _laplacian_3000(oracle)
    n = K.shape[0]
    rng = np.random.default_rng(9)
    B = rng.uniform(-1, 1, (16, n))
    B -= B.mean(axis=1, keepdims=True)
    s = cu.createSolver("bpcg_mixed", 16, K, 5000, 1e-10, 1)
    s.setOption("projectMean", 1)
    s.getRhs().upload(B.ravel())
    s.getSol().memsetZero()
    s.solve()
    X = s.getSol().download().reshape(16, n)
    print(f"bpcg_mixed x16 at 1e-10: {s.getLastIterations()} iterations, {s.getLastOuterIterations()} more cycles")
    for c in range(16):
        rel = np.linalg.norm(B[c] - K @ X[c]) / np.linalg.norm(B[c])
        assert rel <= 1.01e-10, (c, rel)
    assert s.getLastIterations() < 5000


def test_block_pcg_rank_deficient_block(oracle):
    """A tile whose right-hand sides are linearly dependent (column 1 equals
    column 0, column 3 is column 0 + column 2): the Cholesky factorisations
    drop the dependent directions and every column still converges, the
    dependent columns to the matching combinations."""
    from nsfem_amd import cu
    K = _laplacian_3000(oracle)
    n = K.shape[0]
    rng = np.random.default_rng(10)
    B = rng.uniform(-1, 1, (4, n))
    B[1] = B[0]
    B[3] = B[0] + B[2]
    B -= B.mean(axis=1, keepdims=True)
    s = cu.createSolver("bpcg", 4, K, 5000, 1e-12, 1)
    s.setOption("projectMean", 1)
    s.getRhs().upload(B.ravel())
    s.getSol().memsetZero()
    s.solve()
    X = s.getSol().download().reshape(4, n)
    for c in range(4):
        rel = np.linalg.norm(B[c] - K @ X[c]) / np.linalg.norm(B[c])
        assert rel <= 2e-12, (c, rel)
    assert _rel(X[1], X[0]) <= 1e-9
    assert _rel(X[3], X[0] + X[2]) <= 1e-9


def test_block_pcg_rejects_unsupported_shapes(oracle):
    from nsfem_amd import _lib, cu
    K = _laplacian_3000(oracle)
    with pytest.raises(_lib.NsfemInvalidArgument):
        cu.createSolver("bpcg", 2, K, 100, 1e-8, 1)
    with pytest.raises(_lib.NsfemInvalidArgument):
        cu.createSolver("bpcg", 8, K, 100, 1e-8, 1, dtype=np.float32)
