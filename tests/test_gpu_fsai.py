"""PCG with the factorized sparse approximate inverse preconditioner
(option "preconditioner" = 1: z = G^T G r, G lower triangular on the
lower-colour pattern plus the diagonal, built on the host).

The preconditioner changes only the path to the solution: the stopping rule
is the pcg solver's relative residual in FP64, so solutions are held to the
direct solve and Chorin steps to the FP64 oracle with the bounds of the
multicolour-GS pcg."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import mesh_path

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def channel_20k():
    from nsfem_amd import fem, mesher
    m = mesher.channel_mesh(20_000, seed=5)
    ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), fem.DfgConditions(1e-3, 1.0, m["height"], 4.0))
    return ctx.scipy("velocityMass"), ctx.scipy("pressureStiffnessInternal")


def _solve(name, m, b, nch, precond, target=1e-10, max_iters=2000):
    from nsfem_amd import cu
    s = cu.createSolver(name, nch, m, max_iters, target, 1)
    s.setOption("preconditioner", precond)
    s.getRhs().upload(b)
    s.getSol().memsetZero()
    s.solve()
    return s, s.getSol().download()


@pytest.mark.parametrize("name,target,bound", [("pcg", 1e-12, 1e-9), ("pcg_mixed", 1e-10, 1e-7)])
@pytest.mark.parametrize("which", ["mass", "pressure"])
def test_fsai_pcg_matches_direct_solve(channel_20k, name, target, bound, which):
    """The P2 mass matrix with two right-hand sides (the velocity solves) and
    the P1 pressure Laplacian with one: every column meets the target in its
    true residual and equals the direct solve; on the mass matrix the
    iterations are no more than the multicolour GS preconditioner's."""
    M, K = channel_20k
    m, nch = (M, 2) if which == "mass" else (K, 1)
    n = m.shape[0]
    rng = np.random.default_rng(41)
    b = np.concatenate([m @ rng.uniform(-1, 1, n) for _ in range(nch)])
    s, x = _solve(name, m, b, nch, 1, target)
    sg, _ = _solve(name, m, b, nch, 0, target)
    lu = spla.splu(sp.csc_matrix(m))
    for c in range(nch):
        bc, xc = b[c * n:(c + 1) * n], x[c * n:(c + 1) * n]
        assert np.linalg.norm(bc - m @ xc) / np.linalg.norm(bc) <= 1.01 * target, c
        xd = lu.solve(bc)
        assert np.max(np.abs(xc - xd)) <= bound * np.max(np.abs(xd)), c
    print(f"{name} {which}: fsai {s.getLastIterations()} iterations, sgs {sg.getLastIterations()}")
    if which == "mass":
        assert s.getLastIterations() <= sg.getLastIterations()
    # bit-repeatable, and warm-started from its own solution: converged at once
    s2, x2 = _solve(name, m, b, nch, 1, target)
    assert np.array_equal(x2, x)
    s.getSol().upload(x)
    s.solve()
    assert s.getLastIterations() == 0


@pytest.mark.parametrize("name,nch,target,bound", [("pcg", 4, 1e-12, 1e-9), ("pcg_mixed", 16, 1e-10, 1e-7)])
def test_fsai_wide_tiles_match_direct_solve(channel_20k, name, nch, target, bound):
    """Wide right-hand-side tiles (one lane per column, the ensembles'
    velocity solves): every column, scaled 1e-6 apart, meets its target and
    equals the direct solve; no more iterations than the GS sweep."""
    M, _ = channel_20k
    n = M.shape[0]
    rng = np.random.default_rng(43)
    b = np.concatenate([(M @ rng.uniform(-1, 1, n)) * (1e-6 if c % 3 == 1 else 1.0) for c in range(nch)])
    s, x = _solve(name, M, b, nch, 1, target)
    sg, _ = _solve(name, M, b, nch, 0, target)
    lu = spla.splu(sp.csc_matrix(M))
    for c in range(nch):
        bc, xc = b[c * n:(c + 1) * n], x[c * n:(c + 1) * n]
        assert np.linalg.norm(bc - M @ xc) / np.linalg.norm(bc) <= 1.01 * target, c
        xd = lu.solve(bc)
        assert np.max(np.abs(xc - xd)) <= bound * np.max(np.abs(xd)), c
    print(f"{name} x{nch}: fsai {s.getLastIterations()} iterations, sgs {sg.getLastIterations()}")
    assert s.getLastIterations() <= sg.getLastIterations()


def test_fsai_matches_the_factor_of_its_definition(channel_20k):
    """The preconditioner applied once from zero (one iteration of an
    unconverged solve leaves x = alpha G^T G b): its direction is that of
    the factor computed here densely per row, M[P_i, P_i] g = e_i scaled to
    (G M G^T)_ii = 1, in the solver's own row order (recovered from the
    exported colouring)."""
    from nsfem_amd import cu
    M, _ = channel_20k
    M = sp.csr_matrix(M[:4000, :4000])  # a leading block: still SPD with the same structure kind
    n = M.shape[0]
    rng = np.random.default_rng(42)
    b = rng.uniform(-1, 1, n)
    s = cu.createSolver("pcg", 1, M, 1, 1e-30, 1)
    s.setOption("preconditioner", 1)
    s.getRhs().upload(b)
    s.getSol().memsetZero()
    s.solve()
    x = s.getSol().download()
    # the solver's colour-major order: colours in order, rows inside each
    # colour in some order of its own; G depends only on the set of lower-
    # colour neighbours of each row, so the colour classes determine it
    order, starts = s.coloring()  # rows in colour order, colour boundaries
    colors = np.empty(n, dtype=int)
    for c in range(len(starts) - 1):
        colors[order[starts[c]:starts[c + 1]]] = c
    rows, cols, vals = [], [], []
    A = M.tocsr()
    for i in range(n):
        nb = A.indices[A.indptr[i]:A.indptr[i + 1]]
        P = [j for j in nb if colors[j] < colors[i]] + [i]
        Ap = A[P][:, P].toarray()
        e = np.zeros(len(P))
        e[-1] = 1.0
        g = np.linalg.solve(Ap, e)
        g /= np.sqrt(g[-1])
        rows += [i] * len(P)
        cols += P
        vals += list(g)
    G = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    z = G.T @ (G @ b)
    # x = alpha z with alpha = (r.z) / (z.Az)
    alpha = (b @ z) / (z @ (A @ z))
    assert np.max(np.abs(x - alpha * z)) <= 1e-12 * np.max(np.abs(alpha * z))


def test_chorin_steps_with_fsai_velocity_match_oracle(oracle, octx_c3):
    from nsfem_amd import fem
    tau = 1e-3
    orc = oracle.ChorinOracle(octx_c3, tau, vel=("pcg", 2000, 1e-13, 1), pres=("pcg", 8000, 1e-13, 1))
    ctx = fem.ChorinContext(fem.TriangleMesh.load_gmsh(mesh_path("c3")), fem.DfgConditions(1e-3, 1.0))
    cfg = fem.ChorinConfig(fem.SolverConfig("pcg", 2000, 1e-13, 1), fem.SolverConfig("amg_pcg", 500, 1e-13, 1), dtype=1)
    mixed = fem.ChorinConfig(fem.SolverConfig("pcg_mixed", 2000, 1e-10, 1), fem.SolverConfig("amg_pcg_mixed", 500, 1e-10, 1),
                             dtype=1, velocityStart=fem.VELOCITY_START_EXTRAPOLATE)
    ch, chm = fem.Chorin(ctx, cfg), fem.Chorin(ctx, mixed)
    for c in (ch, chm):
        c.set_solver_option("preconditioner", 1, "velocity")
        c.set_time_step(tau)
    for k in range(3):
        st = ch.step()
        stm = chm.step()
        uo, po = orc.step()
        u, p = ch.download()
        assert st["relTent"] <= 1e-13 and st["relFinal"] <= 1e-13
        assert stm["relTent"] <= 1e-10 and stm["relFinal"] <= 1e-10
        assert np.max(np.abs(u - uo)) <= 1e-10 * np.max(np.abs(uo)), k
        assert np.max(np.abs(p - po)) <= 1e-10 * np.max(np.abs(po)), k
    um, pm = chm.download()
    assert np.max(np.abs(um - uo)) <= 1e-8 * np.max(np.abs(uo))
    assert np.max(np.abs(pm - po)) <= 1e-7 * np.max(np.abs(po))


def test_captured_steps_with_fsai_velocity():
    """The whole step captured as one graph with the approximate inverse in
    the velocity solves: equal to regular steps; a forced small velocity
    budget rolls the step back."""
    from nsfem_amd import fem
    tau = 1e-3
    def make(sg):
        cfg = fem.ChorinConfig(fem.SolverConfig("pcg", 2000, 1e-13, 1), fem.SolverConfig("amg_pcg", 500, 1e-13, 1), dtype=1,
                               stepGraph=sg)
        ch = fem.Chorin(fem.ChorinContext(fem.TriangleMesh.load_gmsh(mesh_path("c3")), fem.DfgConditions(1e-3, 1.0)), cfg)
        ch.set_solver_option("preconditioner", 1, "velocity")
        ch.set_time_step(tau)
        return ch
    regular, captured = make(0), make(1)
    kinds = []
    for _ in range(4):
        regular.step()
        st = captured.step()
        kinds.append(st["capturedStep"])
        assert st["relTent"] <= 1e-13 and st["relFinal"] <= 1e-13
    assert kinds[0] == 0 and sum(kinds[1:]) >= 2
    ur, pr = regular.download()
    uc, pc = captured.download()
    assert np.max(np.abs(uc - ur)) <= 1e-11 * np.max(np.abs(ur))
    assert np.max(np.abs(pc - pr)) <= 1e-10 * np.max(np.abs(pr))
    captured.set_step_graph_budgets(2, 40, 40)
    before = captured.step_graph_info()["rollbacks"]
    st = captured.step()
    assert st["capturedStep"] == 0 and captured.step_graph_info()["rollbacks"] == before + 1


def test_fsai_on_a_map_of_its_own(channel_20k):
    """rowsPreconditioner gives the approximate inverse's two launches a
    region-major map of their own (here 256-row workgroups against the
    direction kernel's 1,024): the same operator, so the same solution to the
    target and the same iteration count (the r.z partial sums only regroup);
    counts that are not multiples of 256 are rejected."""
    from nsfem_amd import cu
    from nsfem_amd._lib import NsfemError
    M, _ = channel_20k
    n = M.shape[0]
    rng = np.random.default_rng(47)
    b = np.concatenate([M @ rng.uniform(-1, 1, n) for _ in range(2)])
    out = []
    for rows_pre in (0, 256):
        s = cu.createSolver("pcg_mixed", 2, M, 2000, 1e-10, 1)
        s.setOption("fusedReductions", 1)
        s.setOption("rowsDirection", 1024)
        s.setOption("preconditioner", 1)
        s.setOption("rowsPreconditioner", rows_pre)
        s.getRhs().upload(b)
        s.getSol().memsetZero()
        s.solve()
        out.append((s.getLastIterations(), s.getSol().download()))
        with pytest.raises(NsfemError, match="multiple of 256"):
            s.setOption("rowsPreconditioner", 100)
    (it0, x0), (it1, x1) = out
    assert it0 == it1
    lu = spla.splu(sp.csc_matrix(M))
    for c in range(2):
        xd = lu.solve(b[c * n:(c + 1) * n])
        for x in (x0, x1):
            xc = x[c * n:(c + 1) * n]
            assert np.linalg.norm(b[c * n:(c + 1) * n] - M @ xc) / np.linalg.norm(b[c * n:(c + 1) * n]) <= 1.01e-10
            assert np.max(np.abs(xc - xd)) <= 1e-7 * np.max(np.abs(xd))


def test_fsai_rejects_what_it_does_not_cover(channel_20k):
    from nsfem_amd import cu
    from nsfem_amd._lib import NsfemError
    M, K = channel_20k
    for name, nch, opt in (("bpcg", 4, None), ("gs", 1, None), ("pcg_mixed", 1, "mixedRestart")):
        s = cu.createSolver(name, nch, M, 100, 1e-10, 1)
        if opt:
            s.setOption(opt, 1)
        with pytest.raises(NsfemError, match="approximate inverse preconditions"):
            s.setOption("preconditioner", 1)
    s = cu.createSolver("amg_pcg", 1, K, 100, 1e-10, 1)
    with pytest.raises(NsfemError, match="approximate inverse preconditions"):
        s.setOption("preconditioner", 1)
    s = cu.createSolver("pcg", 1, K, 100, 1e-10, 1)
    with pytest.raises(NsfemError, match="must be 0 or 1"):
        s.setOption("preconditioner", 2)
