"""CG with the smoothed-aggregation V-cycle preconditioner (amg_pcg,
amg_pcg_mixed) on the pressure systems.

The preconditioner changes only the path to the solution: the stopping rule
is the pcg solver's relative residual in FP64, so the solutions are held to
the direct solve (scipy) with the same bounds as the one-level pcg, and the
Chorin steps with an AMG pressure solve to the FP64 oracle as the pcg steps
are (1e-10 at a 1e-13 residual; the mixed bounds at 1e-10)."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import mesh_path

pytestmark = pytest.mark.gpu


def _solve(name, m, b, max_iters=500, target=1e-12, x0=None, nch=1):
    from nsfem_amd import cu
    s = cu.createSolver(name, nch, m, max_iters, target, 1)
    s.getRhs().upload(b)
    if x0 is None:
        s.getSol().memsetZero()
    else:
        s.getSol().upload(x0)
    s.solve()
    return s, s.getSol().download()


def _channel_pressure(nodes, seed=7):
    """Kp_int of a seeded channel mesh (Dirichlet outflow rows removed)."""
    from nsfem_amd import fem, mesher
    m = mesher.channel_mesh(nodes, seed=seed)
    cond = fem.DfgConditions(viscosity=1e-3, peakVelocity=1.0, channelHeight=m["height"], profileScale=4.0)
    ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), cond)
    return ctx.scipy("pressureStiffnessInternal")


@pytest.mark.parametrize("name,target,bound", [("amg_pcg", 1e-12, 1e-9), ("amg_pcg_mixed", 1e-10, 1e-7)])
def test_amg_pcg_matches_direct_solve(octx_c3, name, target, bound):
    m = octx_c3["Kp_int"]
    rng = np.random.default_rng(31)
    b = rng.uniform(-1, 1, m.shape[0])
    s, x = _solve(name, m, b, target=target)
    xd = spla.spsolve(sp.csc_matrix(m), b)
    true_rel = np.linalg.norm(b - m @ x) / np.linalg.norm(b)
    assert s.getLastRelRes() <= target and true_rel <= 1.01 * target
    assert abs(true_rel - s.getLastRelRes()) <= 1e-3 * target + 1e-16
    assert np.max(np.abs(x - xd)) <= bound * np.max(np.abs(xd))
    assert 1 <= s.getLastIterations() < 60


@pytest.mark.parametrize("name,nch", [("amg_pcg", 4), ("amg_pcg_mixed", 8), ("amg_pcg", 16)])
def test_amg_pcg_channels_match_single_solves(octx_c3, name, nch):
    """Several right-hand sides share the V-cycle's matrix reads; each
    column has its own CG scalars and stopping test: columns scaled 1e-6
    apart and a zero column all meet the relative target, and each column
    equals the direct solve."""
    m = octx_c3["Kp_int"]
    n = m.shape[0]
    rng = np.random.default_rng(34)
    cols = [rng.uniform(-1, 1, n) * (1e-6 if c % 3 == 1 else 1.0) for c in range(nch)]
    cols[-1] = np.zeros(n)
    b = np.concatenate(cols)
    target = 1e-10
    # a nonzero guess for the zero column too (its solution is zero)
    s, x = _solve(name, m, b, target=target, nch=nch, x0=rng.uniform(-1, 1, n * nch) * 1e-3)
    lu = spla.splu(sp.csc_matrix(m))
    for c in range(nch):
        bc, xc = b[c * n:(c + 1) * n], x[c * n:(c + 1) * n]
        if not bc.any():
            assert not xc.any()
            continue
        rel = np.linalg.norm(bc - m @ xc) / np.linalg.norm(bc)
        assert rel <= 1.01 * target, c
        xd = lu.solve(bc)
        assert np.max(np.abs(xc - xd)) <= 1e-7 * np.max(np.abs(xd)), c
    assert s.getLastIterations() < 60


def test_amg_pcg_is_bit_repeatable_and_warm_starts(octx_c3):
    m = octx_c3["Kp_int"]
    rng = np.random.default_rng(32)
    b = rng.uniform(-1, 1, m.shape[0])
    s1, x1 = _solve("amg_pcg", m, b, target=1e-10)
    s2, x2 = _solve("amg_pcg", m, b, target=1e-10)
    assert np.array_equal(x1, x2)
    # from its own solution: converged at once
    s3, x3 = _solve("amg_pcg", m, b, target=1e-10, x0=x1)
    assert s3.getLastIterations() == 0 and np.array_equal(x3, x1)
    # a zero right-hand side gives the zero solution
    _, x0 = _solve("amg_pcg", m, np.zeros_like(b), target=1e-10)
    assert not x0.any()


@pytest.mark.parametrize("name", ["amg_pcg", "amg_pcg_mixed"])
def test_amg_pcg_on_a_200k_channel(name):
    """Iterations stay few at 200k rows where the one-level pcg needs
    hundreds; the true residual meets the target."""
    m = _channel_pressure(200_000)
    rng = np.random.default_rng(33)
    b = rng.uniform(-1, 1, m.shape[0])
    s, x = _solve(name, m, b, target=1e-10, max_iters=300)
    true_rel = np.linalg.norm(b - m @ x) / np.linalg.norm(b)
    sp_, xp = _solve("pcg", m, b, target=1e-10, max_iters=20000)
    print(f"{name}: {s.getLastIterations()} iterations (pcg {sp_.getLastIterations()}), true relres {true_rel:.2e}")
    assert true_rel <= 1.01e-10
    assert s.getLastIterations() < 60 < sp_.getLastIterations()
    assert np.max(np.abs(x - xp)) <= 1e-6 * np.max(np.abs(xp))


def test_chorin_steps_with_amg_pressure_match_oracle(oracle, octx_c3):
    from nsfem_amd import fem
    tau = 1e-3
    orc = oracle.ChorinOracle(octx_c3, tau, vel=("pcg", 2000, 1e-13, 1), pres=("pcg", 8000, 1e-13, 1))
    ctx = fem.ChorinContext(fem.TriangleMesh.load_gmsh(mesh_path("c3")), fem.DfgConditions(1e-3, 1.0))
    cfg = fem.ChorinConfig(fem.SolverConfig("pcg", 2000, 1e-13, 1), fem.SolverConfig("amg_pcg", 500, 1e-13, 1), dtype=1)
    mixed = fem.ChorinConfig(fem.SolverConfig("pcg_mixed", 2000, 1e-10, 1), fem.SolverConfig("amg_pcg_mixed", 500, 1e-10, 1),
                             dtype=1)
    ch, chm = fem.Chorin(ctx, cfg), fem.Chorin(ctx, mixed)
    ch.set_time_step(tau)
    chm.set_time_step(tau)
    for k in range(3):
        st = ch.step()
        stm = chm.step()
        uo, po = orc.step()
        u, p = ch.download()
        assert st["relPressure"] <= 1e-13 and st["itersPressure"] < 60
        assert stm["relPressure"] <= 1e-10
        assert np.max(np.abs(u - uo)) <= 1e-10 * np.max(np.abs(uo)), k
        assert np.max(np.abs(p - po)) <= 1e-10 * np.max(np.abs(po)), k
    um, pm = chm.download()
    assert np.max(np.abs(um - uo)) <= 1e-8 * np.max(np.abs(uo))
    assert np.max(np.abs(pm - po)) <= 1e-7 * np.max(np.abs(po))


@pytest.mark.parametrize("B", [1, 4])
def test_captured_steps_with_amg_pressure(oracle, B):
    """The whole step as one captured graph with the AMG-preconditioned
    pressure solve (its iterations captured with a budget, idle once
    converged): equal to regular steps and to the oracle, and a forced small
    budget rolls the step back and reruns it."""
    from nsfem_amd import fem
    tau = 1e-3
    scales = [1.0, 0.5, 1.5, 0.8][:B]
    def make(sg):
        cfg = fem.ChorinConfig(fem.SolverConfig("pcg", 2000, 1e-13, 1), fem.SolverConfig("amg_pcg", 500, 1e-13, 1), dtype=1,
                               stepGraph=sg)
        ch = fem.Chorin(fem.ChorinContext(fem.TriangleMesh.load_gmsh(mesh_path("c3")), fem.DfgConditions(1e-3, 1.0)), cfg,
                        instances=B, inflow_scale=scales)
        ch.set_time_step(tau)
        return ch
    regular, captured = make(0), make(1)
    kinds = []
    for _ in range(4):
        regular.step()
        st = captured.step()
        kinds.append(st["capturedStep"])
        assert st["relPressure"] <= 1e-13
    assert kinds[0] == 0 and sum(kinds[1:]) >= 2
    ur, pr = regular.download()
    uc, pc = captured.download()
    assert np.max(np.abs(uc - ur)) <= 1e-11 * np.max(np.abs(ur))
    assert np.max(np.abs(pc - pr)) <= 1e-10 * np.max(np.abs(pr))
    mesh = oracle.parse_gmsh(mesh_path("c3"))
    for b in range(B):
        orc = oracle.ChorinOracle(oracle.build_context(mesh, 1e-3, scales[b]), tau, vel=("pcg", 2000, 1e-13, 1),
                                  pres=("pcg", 8000, 1e-13, 1))
        for _ in range(4):
            uo, po = orc.step()
        n2, v = 2 * captured.N, captured.V
        assert np.max(np.abs(uc[b * n2:(b + 1) * n2] - uo)) <= 1e-10 * np.max(np.abs(uo))
        assert np.max(np.abs(pc[b * v:(b + 1) * v] - po)) <= 1e-10 * np.max(np.abs(po))
    # a pressure budget of 2 iterations cannot converge: rolled back and rerun
    captured.set_step_graph_budgets(40, 2, 40)
    before = captured.step_graph_info()["rollbacks"]
    st = captured.step()
    assert st["capturedStep"] == 0 and captured.step_graph_info()["rollbacks"] == before + 1
    assert st["relPressure"] <= 1e-13


def test_amg_pcg_rejects_what_it_cannot_solve(octx_c3):
    from nsfem_amd import cu
    from nsfem_amd._lib import NsfemError
    m = octx_c3["Kp_int"]
    with pytest.raises(NsfemError, match="1, 2, 4, 8 or 16 right-hand sides"):
        cu.createSolver("amg_pcg", 32, m, 100, 1e-10, 1)
    with pytest.raises(NsfemError, match="FP64 Krylov"):
        cu.createSolver("amg_pcg", 1, sp.csr_matrix(m, dtype=np.float32), 100, 1e-10, 1)



@pytest.mark.parametrize("name", ["amg_pcg", "amg_pcg_mixed"])
def test_amg_pcg_on_the_pure_neumann_laplacian(oracle, name):
    """The singular pure-Neumann Laplacian (C2's operator) with zero-mean
    right-hand sides: the coarsest level completed by a rank-one term, the
    zero-mean solution against least squares, few iterations at 60k rows."""
    from nsfem_amd import cu, fem, mesher
    m = mesher.channel_mesh(60_000, seed=9)
    ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), fem.DfgConditions(0.01, 1.0, m["height"], 4.0))
    K = ctx.scipy("pressureStiffness")
    n = K.shape[0]
    rng = np.random.default_rng(36)
    b = rng.uniform(-1, 1, n)
    b -= b.mean()
    s = cu.createSolver(name, 1, K, 300, 1e-10, 1)
    s.setOption("projectMean", 1)
    s.getRhs().upload(b)
    s.getSol().memsetZero()
    s.solve()
    x = s.getSol().download()
    assert abs(x.mean()) <= 1e-12 * np.max(np.abs(x))
    assert np.linalg.norm(b - K @ x) / np.linalg.norm(b) <= 1.01e-10
    assert s.getLastIterations() < 60
    # the zero-mean solution: node 0 grounded (K[1:, 1:] is nonsingular; b has
    # zero mean, so the grounded solution solves K x = b), then the mean removed
    K = K.tocsc()
    xd = np.concatenate([[0.0], spla.spsolve(K[1:, 1:], b[1:])])
    xd -= xd.mean()
    assert np.max(np.abs(x - xd)) <= 1e-6 * np.max(np.abs(xd))


@pytest.fixture(scope="module")
def kp_60k():
    return _channel_pressure(60_000, seed=11)


@pytest.mark.parametrize("name", ["amg_pcg", "amg_pcg_mixed"])
@pytest.mark.parametrize("smoother", [(1 << 20, 2, 0.2), (1, 2, 0.2), (0, 2, 0.3), (1, 3, 0.3), (0, 4, 0.1)])
def test_amg_smoothers_match_direct_solve(kp_60k, name, smoother):
    """Gauss-Seidel on every level, the default (GS on the finest level, two
    Chebyshev steps on the coarser ones) and Chebyshev smoothing of other
    degrees, intervals and first levels all keep the V-cycle a symmetric
    positive definite preconditioner: CG meets the target in few iterations
    and the solution matches the direct solve (60k rows: the finest level
    and one smoothing level above the dense coarsest)."""
    from nsfem_amd import cu
    m = kp_60k
    rng = np.random.default_rng(37)
    b = rng.uniform(-1, 1, m.shape[0])
    s = cu.createSolver(name, 1, m, 200, 1e-10, 1)
    for key, v in zip(("amgChebyshevFrom", "amgChebyshevDegree", "amgChebyshevFraction"), smoother):
        s.setOption(key, v)
    s.getRhs().upload(b)
    s.getSol().memsetZero()
    s.solve()
    x = s.getSol().download()
    true_rel = np.linalg.norm(b - m @ x) / np.linalg.norm(b)
    xd = spla.spsolve(sp.csc_matrix(m), b)
    print(f"{name} {smoother}: {s.getLastIterations()} iterations, true relres {true_rel:.2e}")
    assert true_rel <= 1.01e-10
    assert s.getLastIterations() < 40
    assert np.max(np.abs(x - xd)) <= 1e-6 * np.max(np.abs(xd))
    # a second solve on the same solver (the captured iteration kept) agrees bit for bit
    s.getSol().memsetZero()
    s.solve()
    assert np.array_equal(s.getSol().download(), x)


def test_amg_smoother_options_are_checked(kp_60k):
    from nsfem_amd import cu
    from nsfem_amd._lib import NsfemError
    s = cu.createSolver("amg_pcg", 1, kp_60k, 200, 1e-10, 1)
    with pytest.raises(NsfemError, match="degree must be 2 to 8"):
        s.setOption("amgChebyshevDegree", 1)
    with pytest.raises(NsfemError, match="fraction must be in"):
        s.setOption("amgChebyshevFraction", 1.5)
    with pytest.raises(NsfemError, match="must be in"):
        s.setOption("amgChebyshevFrom", -1)
    p = cu.createSolver("pcg", 1, kp_60k, 200, 1e-10, 1)
    with pytest.raises(NsfemError, match="amg_pcg solvers only"):
        p.setOption("amgChebyshevFrom", 0)
