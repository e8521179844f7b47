"""Pins the CPU oracle itself against closed forms, invariants and independent
computations (no GPU)."""

import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import mesh_path


def test_quadrature_exact_to_degree_4(oracle):
    # int_T x^a y^b = a! b! / (a + b + 2)! on the reference triangle
    for a in range(5):
        for b in range(5 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            q = sum(w * x ** a * y ** b for x, y, w in oracle.QUAD4)
            assert abs(q - exact) < 1e-15, (a, b)
    # and not exact at degree 5 (so the self-convection integrand, degree 5,
    # depends on this particular rule)
    q = sum(w * x ** 5 for x, y, w in oracle.QUAD4)
    assert abs(q - 120 / math.factorial(7)) > 1e-6


def test_shape_functions_are_nodal(oracle):
    nodes2 = [(0, 0), (0.5, 0), (1, 0), (0.5, 0.5), (0, 1), (0, 0.5)]
    for i, (x, y) in enumerate(nodes2):
        e = np.zeros(6)
        e[i] = 1
        assert np.allclose(oracle.p2_shape(x, y), e, atol=1e-15)
    for i, (x, y) in enumerate([(0, 0), (1, 0), (0, 1)]):
        e = np.zeros(3)
        e[i] = 1
        assert np.allclose(oracle.p1_shape(x, y), e)
    # gradients: finite differences
    h = 1e-6
    for x, y in [(0.2, 0.3), (0.6, 0.1)]:
        gx, gy = oracle.p2_grad(x, y)
        fx = (oracle.p2_shape(x + h, y) - oracle.p2_shape(x - h, y)) / (2 * h)
        fy = (oracle.p2_shape(x, y + h) - oracle.p2_shape(x, y - h)) / (2 * h)
        assert np.allclose(gx, fx, atol=1e-8) and np.allclose(gy, fy, atol=1e-8)


def _random_triangles(rng, n):
    xy = rng.uniform(-1, 1, (3 * n, 2))
    tri = np.arange(3 * n).reshape(n, 3)
    return xy, tri


def test_p1_stiffness_cotangent_formula(oracle):
    rng = np.random.default_rng(0)
    xy, tri = _random_triangles(rng, 20)
    B, det = oracle.geometry(xy, tri)
    K = oracle.local_stiffness(B, det, 1)
    for e in range(len(tri)):
        p = xy[tri[e]]
        area = abs(det[e]) / 2
        for i in range(3):
            for j in range(3):
                if i == j:
                    continue
                k = 3 - i - j  # opposite vertex
                u, v = p[i] - p[k], p[j] - p[k]
                cot = (u @ v) / abs(u[0] * v[1] - u[1] * v[0])
                assert abs(K[e, i, j] - (-0.5 * cot)) < 1e-12
        assert np.allclose(K[e].sum(1), 0, atol=1e-12)
        assert area > 0


def test_p2_mass_closed_form(oracle):
    # |T|/180 * standard P2 mass in the order corner, mid, corner, mid, corner, mid
    ref = np.zeros((6, 6))
    corners, mids = [0, 2, 4], [1, 3, 5]
    opposite = {0: 3, 2: 5, 4: 1}
    for i in corners:
        for j in corners:
            ref[i, j] = 6 if i == j else -1
        ref[i, opposite[i]] = ref[opposite[i], i] = -4
    for i in mids:
        for j in mids:
            ref[i, j] = 32 if i == j else 16
    rng = np.random.default_rng(1)
    xy, tri = _random_triangles(rng, 10)
    B, det = oracle.geometry(xy, tri)
    M = oracle.local_mass(det, 2)
    for e in range(len(tri)):
        assert np.allclose(M[e], ref * abs(det[e]) / 2 / 180, rtol=1e-12, atol=1e-15)


def test_divergence_of_linear_fields(oracle):
    # VPD [x; 0] = int psi_p (div = 1) = |T|/3 per vertex; PVD x = int phi_v dp/dx
    rng = np.random.default_rng(2)
    xy, tri = _random_triangles(rng, 8)
    B, det = oracle.geometry(xy, tri)
    X, Y = oracle.local_divergence(B, det, False)
    # P2 nodal values of u_x = x, u_y = y
    mid = np.array([[0.0, 0.0], [0.5, 0], [1, 0], [0.5, 0.5], [0, 1], [0, 0.5]])
    for e in range(len(tri)):
        p0, p1, p2 = xy[tri[e]]
        nodes = p0 + mid[:, :1] * (p1 - p0) + mid[:, 1:] * (p2 - p0)
        area = abs(det[e]) / 2
        assert np.allclose(X[e] @ nodes[:, 0], area / 3, atol=1e-13)
        assert np.allclose(Y[e] @ nodes[:, 1], area / 3, atol=1e-13)
        assert np.allclose(X[e] @ nodes[:, 1], 0, atol=1e-13)
    Xs, Ys = oracle.local_divergence(B, det, True)
    w2 = np.array([0, 1, 0, 1, 0, 1]) / 3.0  # int phi_v over T / |T|
    for e in range(len(tri)):
        p = xy[tri[e]]
        area = abs(det[e]) / 2
        assert np.allclose(Xs[e] @ p[:, 0], w2 * area, atol=1e-13)
        assert np.allclose(Ys[e] @ p[:, 1], w2 * area, atol=1e-13)


def test_self_convection_properties(oracle):
    rng = np.random.default_rng(3)
    xy, tri = _random_triangles(rng, 6)
    B, det = oracle.geometry(xy, tri)
    ux, uy = rng.standard_normal((6, 6)), rng.standard_normal((6, 6))
    C = oracle.local_self_convection(B, det, ux, uy)
    # row sums vanish (sum_c phi_c = 1)
    assert np.allclose(C.sum(2), 0, atol=1e-12)
    # linear in the flow
    C2 = oracle.local_self_convection(B, det, 2 * ux - uy, 2 * uy + ux)
    C3 = 2 * C + oracle.local_self_convection(B, det, -uy, ux)
    assert np.allclose(C2, C3, atol=1e-12)
    # constant flow (1, 0): C = int phi_r dphi_c/dx, a degree-3 integrand the
    # rule integrates exactly; compare with a 7-point degree-5 rule.
    d5 = [(1 / 3, 1 / 3, 0.225 / 2)]
    a1, b1 = 0.0597158717897698, 0.4701420641051151
    a2, b2 = 0.7974269853530873, 0.1012865073234563
    for a, b, w in ((a1, b1, 0.1323941527885062 / 2), (a2, b2, 0.1259391805448271 / 2)):
        d5 += [(a, b, w), (b, a, w), (b, b, w)]
    one = np.ones((6, 6))
    C = oracle.local_self_convection(B, det, one, 0 * one)
    for e in range(len(tri)):
        sgn = 1 if det[e] > 0 else -1
        ref = 0
        for x, y, w in d5:
            gx, gy = oracle.p2_grad(x, y)
            ref = ref + w * np.outer(oracle.p2_shape(x, y), B[e, 0, 0] * gx + B[e, 0, 1] * gy) * sgn
        assert np.allclose(C[e], ref, atol=1e-12)


@pytest.fixture(scope="module")
def h025(oracle):
    return oracle.build_context(oracle.parse_gmsh(mesh_path("h_0.025")), 1e-3, 1.0)


def test_parse_and_numbering(oracle):
    for name, nv, ne, nb in [("c3", 594, 1066, 122), ("h_0.025", 1890, 3554, 226)]:
        tri = oracle.parse_gmsh(mesh_path(name))
        assert len(tri["xy"]) == nv and len(tri["triangles"]) == ne and len(tri["border"]) == nb
        p2 = oracle.create_mesh(tri, 2)
        edges = len(p2["xy"]) - nv
        assert nv - edges + ne == 0  # Euler characteristic of a domain with one hole
        el = p2["elements"]
        xy = p2["xy"]
        for a, m, b in [(0, 1, 2), (2, 3, 4), (4, 5, 0)]:
            assert np.allclose(xy[el[:, m]], 0.5 * (xy[el[:, a]] + xy[el[:, b]]), atol=1e-15)
        assert tri["groups"] == oracle.GROUP_ORDER


def test_assembled_matrices_invariants(h025):
    ctx = h025
    N = ctx["N"]
    ones = np.ones(N)
    # 1^T M 1 = sum of triangle areas (shoelace), close to the analytic area
    xy, tri = ctx["pm"]["xy"], ctx["pm"]["elements"]
    p0, p1, p2 = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    shoelace = 0.5 * np.abs((p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])).sum()
    assert abs(ctx["M"] @ ones @ ones - shoelace) < 1e-13
    assert abs(shoelace - (2.2 * 0.41 - math.pi * 0.05 ** 2)) < 1e-3
    assert np.allclose(ctx["K"] @ ones, 0, atol=1e-10)
    assert np.allclose(ctx["Kp"] @ np.ones(ctx["V"]), 0, atol=1e-10)
    assert abs(ctx["M"] - ctx["M"].T).max() < 1e-15
    assert abs(ctx["VPD"] - ctx["PVD"].T).max() > 1e-6  # different element pairings
    # patterns: sorted unique element pairs
    for m in (ctx["M"], ctx["Kp"], ctx["VPD"], ctx["PVD"]):
        for r in range(0, m.shape[0], 97):
            seg = m.indices[m.indptr[r]:m.indptr[r + 1]]
            assert np.all(np.diff(seg) > 0)


def test_fast_convection_map_matches_element_form(oracle, octx_c3):
    ctx = octx_c3
    Int = oracle.fast_convection_map(ctx)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(2 * ctx["N"])
    assert np.allclose(Int @ u, oracle.convection_values(ctx, u), atol=1e-12)
    # nnz per row ~ 15: about 695 V integration entries over 45.5 V rows of C
    assert 10 < Int.nnz / Int.shape[0] < 20


def _graph(m):
    m = m.tocsr()
    return [set(m.indices[m.indptr[i]:m.indptr[i + 1]]) - {i} for i in range(m.shape[0])]


def test_coloring_invariants(oracle, h025):
    for m in (h025["Kp_int"], h025["M"]):
        order, color, k = oracle.sl_greedy_coloring(m)
        g = _graph(m)
        n = len(g)
        assert sorted(order) == list(range(n))
        for v in range(n):
            assert all(color[w] != color[v] for w in g[v])
        # greedy: colour = lowest not used by earlier neighbours in the order
        pos = np.empty(n, int)
        pos[order] = np.arange(n)
        for v in range(0, n, 7):
            used = {color[w] for w in g[v] if pos[w] < pos[v]}
            assert color[v] == min(set(range(k + 1)) - used)
        # smallest-last: each vertex had minimum degree among the remaining
        # ones when it was removed (removal order = reversed order)
        removal = order[::-1]
        deg = np.array([len(x) for x in g])
        alive = np.ones(n, bool)
        for v in removal[: min(n, 400)]:
            assert deg[v] == deg[alive].min()
            alive[v] = False
            for w in g[v]:
                if alive[w]:
                    deg[w] -= 1
        assert k == (4 if m.shape[0] < 3000 else 7)


def test_mcgs_equals_sequential_gs_in_colour_order(oracle, h025):
    # The known-answer structure of the reference's SolverTest
    # (gaussSeidelCustom vs gaussSeidelCuda): a multicolour sweep equals
    # sequential GS visiting rows in colouring order.
    m = h025["Kp_int"]
    L = oracle.Mcgs(m)
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, m.shape[0])
    x_seq = oracle.gs_custom_order(m, b, L.coloring, 7)
    x_mc, _, _ = oracle.gs_solve(m, b, np.zeros_like(b), 7, 0.0, 1, L)
    assert np.max(np.abs(x_seq - x_mc)) <= 1e-13 * np.max(np.abs(x_seq))


def test_pcg_matches_direct_solve(oracle, h025):
    m = h025["Kp_int"]
    rng = np.random.default_rng(6)
    b = rng.uniform(-1, 1, m.shape[0])
    x, rel, it = oracle.pcg_solve(m, b, np.zeros_like(b), 1e-12, 2000)
    xd = spla.spsolve(m.tocsc(), b)
    assert rel <= 1e-12
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    assert np.linalg.norm(b - m @ x) <= 1.01e-12 * np.linalg.norm(b)
    # the symmetric GS preconditioner must beat plain CG
    x2, rel2, it2 = oracle.pcg_solve(m, b, np.zeros_like(b), 1e-12, it)
    assert it2 == it


def test_sgs_preconditioner_is_symmetric(oracle, octx_c3):
    m = octx_c3["Kp_int"]
    L = oracle.Mcgs(m)
    rng = np.random.default_rng(7)
    x, y = rng.standard_normal((2, L.n))
    assert abs(x @ L.sgs(y) - y @ L.sgs(x)) < 1e-12 * abs(x @ L.sgs(y))


def test_chorin_step_matches_direct_solves(oracle, octx_c3):
    """The oracle step (PCG to 1e-13) against the same step with exact sparse
    direct solves, the CPU path's LDLT semantics (chorinEigen.cpp:378, :454)."""
    ctx = octx_c3
    tau = 1e-3
    orc = oracle.ChorinOracle(ctx, tau, vel=("pcg", 2000, 1e-13, 1), pres=("pcg", 5000, 1e-13, 1))
    N = ctx["N"]
    u = orc.u.copy()
    for _ in range(2):
        uo, po = orc.step()
        A = ctx["K"].copy()
        A.data = ctx["viscosity"] * ctx["K"].data + oracle.convection_values(ctx, u)
        Mf = spla.factorized(ctx["M"].tocsc())
        a = np.concatenate([Mf(A @ u[:N]), Mf(A @ u[N:])])
        u = u - tau * a
        u[orc.dir_ids] = orc.dir_vals
        d = -(1 / tau) * (ctx["VPD"] @ u)
        p = np.zeros(ctx["V"])
        p[ctx["int_p"]] = spla.spsolve(ctx["Kp_int"].tocsc(), d[ctx["int_p"]])
        g = ctx["PVD"] @ p
        u = u - tau * np.concatenate([Mf(g[:N]), Mf(g[N:])])
        u[orc.dir_ids] = orc.dir_vals
        assert np.max(np.abs(uo - u)) <= 1e-10 * np.max(np.abs(u))
        assert np.max(np.abs(po - p)) <= 1e-10 * np.max(np.abs(p))


def test_chorin_zero_inflow_stays_at_rest(oracle, octx_c3):
    ctx = dict(octx_c3)
    ctx["dvx"] = (ctx["dvx"][0], np.zeros_like(ctx["dvx"][1]))
    orc = oracle.ChorinOracle(ctx, 1e-3)
    u, p = orc.step()
    assert np.all(u == 0) and np.all(p == 0)


def test_pcg_on_consistent_singular_system(oracle, h025):
    """All-Neumann P1 Laplacian with a zero-mean RHS: SGS-PCG converges; the
    zero-mean part of the iterate is the pseudo-inverse solution."""
    m = h025["Kp"]
    rng = np.random.default_rng(8)
    b = rng.uniform(-1, 1, m.shape[0])
    b -= b.mean()
    x, rel, it = oracle.pcg_solve(m, b, np.zeros_like(b), 1e-10, 5000)
    assert rel <= 1e-10
    xp = np.linalg.lstsq(m.toarray(), b, rcond=None)[0]
    assert np.linalg.norm((x - x.mean()) - (xp - xp.mean())) <= 1e-7 * np.linalg.norm(xp)


@pytest.mark.parametrize("name", ["c3", "h_0.025", "h_0.01"])
def test_create_mesh_matches_reference_build(oracle, name):
    """P1/P2 numbering, node coordinates and border elements of the oracle
    equal the reference's own createMesh (compiled from its sources into
    oracle/_ref/) bit for bit."""
    tri = oracle.parse_gmsh(mesh_path(name))
    for order in (1, 2):
        ref = oracle.ref_create_mesh(tri, order)
        if ref is None:
            pytest.skip("reference build oracle/_ref/libref_mesh.so absent (needs the reference sources)")
        mine = oracle.create_mesh(tri, order)
        assert np.array_equal(mine["elements"], ref["elements"])
        assert np.array_equal(mine["border"], ref["border"])
        # the reference computes mid-edge nodes in float: the oracle's exact
        # double midpoints round to them
        assert np.array_equal(mine["xy"].astype(np.float32), ref["xy"])


def _divergence_ratios(oracle, ctx, measure_vpd, steps, tau=1e-3):
    orc = oracle.ChorinOracle(ctx, tau)
    ip = ctx["int_p"]
    out = []
    for _ in range(steps):
        u, _ = orc.step()
        out.append(np.linalg.norm((measure_vpd @ u)[ip]) / np.linalg.norm((measure_vpd @ orc.u_tentative)[ip]))
    return out


@pytest.mark.parametrize("name", ["c3", "h_0.025"])
def test_projection_reduces_divergence(oracle, name):
    """The pressure correction must remove most of the tentative velocity's
    discrete divergence on the internal pressure nodes (the P1 pressure
    Laplacian is an approximate projection: about 0.25-0.4 of it remains).
    The same measure on copies of the oracle with a mistake planted in its
    operators shows that the check discriminates: the correction's sign
    flipped makes the divergence grow (~1.9x), a pressure mis-scaled by tau
    leaves it unchanged (~1.0)."""
    ctx = oracle.build_context(oracle.parse_gmsh(mesh_path(name)), 1e-3, 1.0)
    vpd = ctx["VPD"]
    ok = _divergence_ratios(oracle, ctx, vpd, 4)
    assert max(ok) <= 0.5, ok
    flipped = dict(ctx)
    flipped["PVD"] = -ctx["PVD"]
    assert min(_divergence_ratios(oracle, flipped, vpd, 2)) > 1.5
    misscaled = dict(ctx)
    misscaled["VPD"] = ctx["VPD"] * 1e-3  # right-hand side -(1/tau) VPD u scaled by tau
    assert min(_divergence_ratios(oracle, misscaled, vpd, 2)) > 0.9


def test_tentative_step_closed_form_on_linear_field(oracle, octx_h025):
    """For u = (x, -y) and nu = 0 the convective acceleration (u . grad) u =
    (x, y) lies in P2 and its Galerkin integrand (degree 3) is exact under
    the degree-4 rule, so M^-1 C(u) u is its nodal interpolant and the
    tentative velocity is u - tau (x, y) at every node off the Dirichlet
    boundary.  Pins the convection's sign, the tau scaling and the planar
    [ux; uy] layout of the step."""
    ctx = dict(octx_h025)
    ctx["viscosity"] = 0.0
    N = ctx["N"]
    xy = ctx["vm"]["xy"]
    ux, uy = xy[:, 0].copy(), -xy[:, 1]
    ctx["dvx"] = (ctx["dvx"][0], ux[ctx["dvx"][0]])
    ctx["dvy"] = (ctx["dvy"][0], uy[ctx["dvy"][0]])
    tau = 1e-3
    orc = oracle.ChorinOracle(ctx, tau, vel=("pcg", 2000, 1e-14, 1))
    orc.u[:] = np.concatenate([ux, uy])
    orc.step()
    ut = orc.u_tentative
    free = np.ones(N, bool)
    free[ctx["dvx"][0]] = False
    assert free.sum() > 0.5 * N
    assert np.max(np.abs(ut[:N] - (ux - tau * xy[:, 0]))[free]) <= 1e-12
    assert np.max(np.abs(ut[N:] - (uy - tau * xy[:, 1]))[free]) <= 1e-12


@pytest.mark.parametrize("name", ["c3", "h_0.025"])
def test_cpu_baseline_step_matches_oracle(oracle, name):
    """The multithreaded C++ baseline (the bench's CPU figure) steps like the
    numpy oracle: 3 steps, every solve to 1e-13, velocity and pressure to
    1e-10 relative."""
    ctx = oracle.build_context(oracle.parse_gmsh(mesh_path(name)), 1e-3, 1.0)
    orc = oracle.ChorinOracle(ctx, 1e-3, vel=("pcg", 2000, 1e-13, 1), pres=("pcg", 8000, 1e-13, 1))
    cpu = oracle.CpuStep(ctx, 1e-3, rtol=1e-13, threads=4)
    for k in range(3):
        uo, po = orc.step()
        it = cpu.run(1)
        u, p = cpu.state()
        assert np.max(np.abs(u - uo)) <= 1e-10 * np.max(np.abs(uo)), k
        assert np.max(np.abs(p - po)) <= 1e-10 * np.max(np.abs(po)), k
        assert it[0, 1] > 0
