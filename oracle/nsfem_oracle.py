"""ORACLE - test infrastructure only.

A CPU (numpy / scipy, FP64) restatement of the reference's Chorin-projection
path, used as the checker of the HIP library.  Only tests/, the smoke check
and bench.py's CPU-baseline leg import it; the product never does.

Each function cites the reference code it restates (paths relative to
kosyonikolov/NsFemOptim).  How each one is pinned (known answers, closed
forms, invariants) is listed in its tests; parity status:
  parse_gmsh                        pinned: shipped meshes' counts, Euler characteristic
  create_mesh                       pinned: bit-exact against the reference's own createMesh
                                            (concreteMesh.cpp + Element sources built into oracle/_ref/
                                            by oracle/Makefile) on the three shipped meshes
  dirichlet_nodes                   pinned: boundary geometry, profile values
  local matrices                    pinned: closed forms (P1 cotangent stiffness, P2 mass 1/180 pattern,
                                            row sums, Kronecker property, exactness of the rule)
  self-convection                   pinned: linearity, constant-flow closed form, probing = element form
  assemble                          pinned: patterns = sorted unique element pairs, M 1 = area, K 1 = 0
  fast_convection_map               pinned: Int u = assembled C(u) for random u
  sl_greedy_coloring (C)            pinned: valid colouring, smallest-last and greedy properties
  gs_sweeps / gs_solve              pinned: sequential GS in colour order (gaussSeidelCustomOrder)
  pcg_solve                         pinned: dense solve (numpy.linalg) to 1e-10
  ChorinOracle.step                 pinned: the projection removes most of the tentative velocity's
                                            discrete divergence on the internal pressure nodes (a wrong sign
                                            or 1/tau scaling of the pressure correction does not); the same
                                            step with sparse direct solves to 1e-10
"""

import ctypes
import math
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

GROUP_ORDER = ["Right", "Top", "Left", "Bottom", "Circle", "Interior"]

# ---------------------------------------------------------------------------
# gmsh 4.1 ASCII (cpp/libs/Mesh/src/gmsh.cpp:41-499, Mesh/src/io.cpp:65-208)
# ---------------------------------------------------------------------------

TRIANGLE_SIDES = [(0, 1), (1, 0), (1, 2), (2, 1), (2, 0), (0, 2)]  # triMesh.h:16-24


def _sections(path):
    out, cur, buf = {}, None, []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\r\n")
            if not line:
                continue
            if line.startswith("$"):
                if line.startswith("$End"):
                    if cur is not None:
                        out[cur] = out.get(cur, []) + buf
                    cur, buf = None, []
                else:
                    cur, buf = line[1:], []
                continue
            if cur is not None:
                buf.append(line)
    return out


def parse_gmsh(path):
    """Returns dict(xy, triangles, border=(element, side, group), groups)."""
    s = _sections(path)
    # Nodes
    lines = iter(s["Nodes"])
    nblocks = int(next(lines).split()[0])
    nodes = []
    for _ in range(nblocks):
        _, _, _, cnt = map(int, next(lines).split())
        tags = [int(next(lines).split()[0]) for _ in range(cnt)]
        for t in tags:
            x, y, _z = map(float, next(lines).split()[:3])
            nodes.append((t, x, y))
    nodes.sort()
    ids = [n[0] for n in nodes]
    assert ids == list(range(ids[0], ids[0] + len(ids))), "Node IDs are not contiguous"
    off = ids[0]
    xy = np.array([[n[1], n[2]] for n in nodes], dtype=np.float32).astype(np.float64)
    # Elements
    lines = iter(s["Elements"])
    nblocks = int(next(lines).split()[0])
    elems = []
    for _ in range(nblocks):
        _, tag, _, cnt = map(int, next(lines).split())
        for _ in range(cnt):
            v = list(map(int, next(lines).split()))
            elems.append((v[0], tag, v[1:]))
    elems.sort(key=lambda e: e[0])
    # Physical names and curve entities
    lines = iter(s["PhysicalNames"])
    nn = int(next(lines))
    names = []
    for _ in range(nn):
        d, tag, name = next(lines).split(maxsplit=2)
        names.append((int(tag), name.strip()[1:-1]))
    names.sort()
    goff = names[0][0]
    groups = [n for _, n in names]
    lines = iter(s["Entities"])
    npnt, ncur, _, _ = map(int, next(lines).split())
    for _ in range(npnt):
        next(lines)
    curve_group = {}
    for _ in range(ncur):
        v = next(lines).split()
        nphys = int(v[7])
        if nphys > 0:
            curve_group[int(v[0])] = int(v[8]) - goff
    edges = {}
    tris = []
    for _, tag, pts in elems:
        if len(pts) == 2:
            edges.setdefault(pts[0] - off, []).append([pts[1] - off, curve_group.get(tag, 0), False])
        elif len(pts) == 3:
            tris.append([p - off for p in pts])
        else:
            raise RuntimeError("Not all elements have been parsed")
    tris = np.array(tris, dtype=np.int64)
    border = []
    for i, t in enumerate(tris):
        for side, (a, b) in enumerate(TRIANGLE_SIDES):
            for e in edges.get(int(t[a]), []):
                if e[0] == t[b]:
                    assert not e[2]
                    e[2] = True
                    border.append((i, side, e[1]))
                    break
    return {"xy": xy, "triangles": tris, "border": np.array(border, dtype=np.int64).reshape(-1, 3), "groups": groups}


def border_from_edges(tris, edges):
    """Attach oriented border edges (from, to, group) to triangles like the gmsh path."""
    emap = {}
    for a, b, g in edges:
        emap.setdefault(int(a), []).append([int(b), int(g), False])
    border = []
    for i, t in enumerate(tris):
        for side, (a, b) in enumerate(TRIANGLE_SIDES):
            for e in emap.get(int(t[a]), []):
                if e[0] == t[b]:
                    e[2] = True
                    border.append((i, side, e[1]))
                    break
    return np.array(border, dtype=np.int64).reshape(-1, 3)


# ---------------------------------------------------------------------------
# P1 / P2 numbering (cpp/libs/Mesh/src/concreteMesh.cpp:141-356)
# ---------------------------------------------------------------------------

def create_mesh(tri, order):
    """order 1 or 2 -> dict(xy, elements, border=(element, side, group, pts...))."""
    xy = tri["xy"]
    T = tri["triangles"]
    nv = len(xy)
    if order == 1:
        elements = T.copy()
        border = []
        for e, side, g in tri["border"]:
            a, b = TRIANGLE_SIDES[side]
            border.append((e, side, g, T[e, a], T[e, b]))
        return {"xy": xy.copy(), "elements": elements, "border": np.array(border, dtype=np.int64), "groups": tri["groups"]}
    # sides keyed by the smaller vertex, in first-seen order
    graph = [[] for _ in range(nv)]
    for t in T:
        for a in range(3):
            pa, pb = int(t[a]), int(t[(a + 1) % 3])
            lo, hi = min(pa, pb), max(pa, pb)
            if hi not in graph[lo]:
                graph[lo].append(hi)
    mid = {}
    new_xy = [xy]
    nxt = nv
    extra = []
    for lo in range(nv):
        for hi in graph[lo]:
            mid[(lo, hi)] = nxt
            extra.append(0.5 * (xy[lo] + xy[hi]))
            nxt += 1
    if extra:
        new_xy.append(np.array(extra))
    allxy = np.concatenate(new_xy)
    E = len(T)
    elements = np.empty((E, 6), dtype=np.int64)
    for i, t in enumerate(T):
        for c in range(3):
            elements[i, 2 * c] = t[c]
        for a in range(3):
            pa, pb = int(t[a]), int(t[(a + 1) % 3])
            elements[i, 2 * a + 1] = mid[(min(pa, pb), max(pa, pb))]
    border = []
    for e, side, g in tri["border"]:
        a, b = TRIANGLE_SIDES[side]
        ia, ib = int(T[e, a]), int(T[e, b])
        border.append((e, side, g, ia, mid[(min(ia, ib), max(ia, ib))], ib))
    return {"xy": allxy, "elements": elements, "border": np.array(border, dtype=np.int64), "groups": tri["groups"]}


REF_LIB = os.path.join(HERE, "_ref", "libref_mesh.so")


def ref_create_mesh(tri, order):
    """The reference's own createMesh (concreteMesh.cpp:78-293 with el::P1 /
    el::P2), built from its sources into oracle/_ref/ by oracle/Makefile, on
    the parsed triangle mesh.  Returns dict(xy float32, elements, border) or
    None when the reference build is absent (it needs the reference tree)."""
    if not os.path.exists(REF_LIB):
        return None
    lib = ctypes.CDLL(REF_LIB)
    xy = np.ascontiguousarray(tri["xy"], dtype=np.float32)
    T = np.ascontiguousarray(tri["triangles"], dtype=np.int32)
    B = np.ascontiguousarray(tri["border"], dtype=np.int32).reshape(-1, 3)
    E, nb = len(T), len(B)
    per = 3 if order == 1 else 6
    cap = len(xy) + 3 * E
    oxy = np.zeros((cap, 2), dtype=np.float32)
    oel = np.zeros((E, per), dtype=np.int32)
    obd = np.zeros((nb, 3 + order + 1), dtype=np.int32)
    n = ctypes.c_int(0)
    rc = lib.ref_create_mesh(ctypes.c_int(order), ctypes.c_int(len(xy)), _p(xy), ctypes.c_int(E), _p(T), ctypes.c_int(nb), _p(B),
                             ctypes.c_int(len(tri["groups"])), ctypes.c_int(cap), _p(oxy), _p(oel), _p(obd), ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"reference createMesh failed ({rc})")
    return {"xy": oxy[:n.value].copy(), "elements": oel.astype(np.int64), "border": obd.astype(np.int64)}


def dirichlet_nodes(mesh, group_names, value_fn):
    """First group that claims a node decides its value; sorted by id
    (execs/NavierStokes/include/NavierStokes/borders.h:12-47)."""
    gids = [mesh["groups"].index(g) for g in group_names]
    seen = {}
    for row in mesh["border"]:
        g = int(row[2])
        if g not in gids:
            continue
        for nid in row[3:]:
            nid = int(nid)
            if nid not in seen:
                seen[nid] = value_fn(mesh, nid, g)
    ids = np.array(sorted(seen), dtype=np.int64)
    return ids, np.array([seen[i] for i in ids], dtype=np.float64)


def internal_nodes(n, dirichlet_ids):
    mask = np.ones(n, bool)
    mask[dirichlet_ids] = False
    return np.nonzero(mask)[0]


# ---------------------------------------------------------------------------
# Elements and local matrices (cpp/libs/Element/src/p1.cpp, p2.cpp,
# triangleIntegrator.cpp:10-293), vectorised over elements, FP64.
# ---------------------------------------------------------------------------

# The published 6-point degree-4 rule (weights include the area 1/2).
QUAD4 = np.array([
    [9.15762135097710067155318824916321e-02, 9.15762135097710067155318824916321e-02, 5.49758718276610602870846378209535e-02],
    [9.15762135097710067155318824916321e-02, 8.16847572980457958813360619387822e-01, 5.49758718276610602870846378209535e-02],
    [8.16847572980457958813360619387822e-01, 9.15762135097710067155318824916321e-02, 5.49758718276610602870846378209535e-02],
    [4.45948490915965001235576892213430e-01, 4.45948490915965001235576892213430e-01, 1.11690794839005624883299105931656e-01],
    [4.45948490915965001235576892213430e-01, 1.08103018168069997528846215573139e-01, 1.11690794839005624883299105931656e-01],
    [1.08103018168069997528846215573139e-01, 4.45948490915965001235576892213430e-01, 1.11690794839005624883299105931656e-01],
])


def p1_shape(x, y):
    return np.array([1 - x - y, x, y])


def p1_grad(x, y):
    return np.array([-1.0, 1.0, 0.0]), np.array([-1.0, 0.0, 1.0])


def p2_shape(x, y):
    return np.array([(1 - x - y) * (1 - 2 * x - 2 * y), 4 * x * (1 - x - y), x * (2 * x - 1), 4 * x * y, y * (2 * y - 1),
                     4 * y * (1 - x - y)])


def p2_grad(x, y):
    gx = np.array([4 * x + 4 * y - 3, 4 - 8 * x - 4 * y, 4 * x - 1, 4 * y, 0.0, -4 * y])
    gy = np.array([4 * x + 4 * y - 3, -4 * x, 0.0, 4 * x, 4 * y - 1, 4 - 4 * x - 8 * y])
    return gx, gy


def geometry(xy, tri3):
    """Per element: m (affine map, x = m xi + p0), B = [[m11,-m10],[-m01,m00]], det."""
    p0, p1, p2 = xy[tri3[:, 0]], xy[tri3[:, 1]], xy[tri3[:, 2]]
    m00, m01 = p1[:, 0] - p0[:, 0], p2[:, 0] - p0[:, 0]
    m10, m11 = p1[:, 1] - p0[:, 1], p2[:, 1] - p0[:, 1]
    det = m00 * m11 - m01 * m10
    B = np.stack([np.stack([m11, -m10], -1), np.stack([-m01, m00], -1)], -2)  # (E,2,2)
    return B, det


def local_mass(det, order):
    shape = p2_shape if order == 2 else p1_shape
    ref = sum(w * np.outer(shape(x, y), shape(x, y)) for x, y, w in QUAD4)
    return np.abs(det)[:, None, None] * ref[None]


def local_stiffness(B, det, order):
    grad = p2_grad if order == 2 else p1_grad
    out = 0
    for x, y, w in QUAD4:
        G = np.stack(grad(x, y))  # (2, n)
        phys = np.einsum("eab,bn->ean", B, G)  # B grad_ref, (E,2,n)
        out = out + w * np.einsum("ean,eam->enm", phys, phys)
    return out / np.abs(det)[:, None, None]


def local_divergence(B, det, swap):
    """swap False: (E,3,6) X/Y with X[r][c] = int dphi2_c/dx psi1_r;
    swap True: (E,6,3) with X[r][c] = int dpsi1_c/dx phi2_r."""
    sgn = np.where(det < 0, -1.0, 1.0)
    dm, sm = (p1_grad, p2_shape) if swap else (p2_grad, p1_shape)
    X = Y = 0
    for x, y, w in QUAD4:
        G = np.stack(dm(x, y))
        phys = np.einsum("eab,bn->ean", B, G)
        s = sm(x, y)
        X = X + w * np.einsum("r,ec->erc", s, phys[:, 0])
        Y = Y + w * np.einsum("r,ec->erc", s, phys[:, 1])
    return sgn[:, None, None] * X, sgn[:, None, None] * Y


def local_self_convection(B, det, ux, uy):
    """(E,6,6): C[r][c] = sum_q w sgn phi_r (B^T u_h) . grad_ref phi_c (triangleIntegrator.cpp:192-231)."""
    sgn = np.where(det < 0, -1.0, 1.0)
    out = 0
    for x, y, w in QUAD4:
        phi = p2_shape(x, y)
        gx, gy = p2_grad(x, y)
        fx = ux @ phi
        fy = uy @ phi
        lx = B[:, 0, 0] * fx + B[:, 1, 0] * fy
        ly = B[:, 0, 1] * fx + B[:, 1, 1] * fy
        gfd = lx[:, None] * gx[None] + ly[:, None] * gy[None]  # (E, c)
        out = out + w * phi[None, :, None] * gfd[:, None, :]
    return sgn[:, None, None] * out


# ---------------------------------------------------------------------------
# Assembly (cpp/libs/Fem/src/chorinCsr.cpp:61-502) and the context
# (execs/NavierStokes/src/buildContext.cpp:17-142)
# ---------------------------------------------------------------------------

def _coo(rows, cols, vals, shape):
    m = sp.coo_matrix((vals.ravel(), (rows.ravel(), cols.ravel())), shape=shape).tocsr()
    m.sum_duplicates()
    m.sort_indices()
    return m


def build_context(tri, viscosity, peak, channel_height=0.41, profile_scale=1.0):
    vm = create_mesh(tri, 2)
    pm = create_mesh(tri, 1)
    N, V = len(vm["xy"]), len(pm["xy"])
    E = len(tri["triangles"])
    B, det = geometry(tri["xy"], tri["triangles"])
    ev, ep = vm["elements"], pm["elements"]

    def inflow(mesh, nid, g):
        if mesh["groups"][g] == "Left":
            y = mesh["xy"][nid, 1]
            return profile_scale * peak * y * (channel_height - y) / (channel_height * channel_height)
        return 0.0

    vb = ["Left", "Top", "Bottom", "Circle"]
    dvx = dirichlet_nodes(vm, vb, inflow)
    dvy = dirichlet_nodes(vm, vb, lambda m, n, g: 0.0)
    dp = dirichlet_nodes(pm, ["Right"], lambda m, n, g: 0.0)
    int_p = internal_nodes(V, dp[0])

    rv = np.repeat(ev[:, :, None], 6, 2)
    cv = np.repeat(ev[:, None, :], 6, 1)
    M = _coo(rv, cv, local_mass(det, 2), (N, N))
    K = _coo(rv, cv, local_stiffness(B, det, 2), (N, N))
    rp = np.repeat(ep[:, :, None], 3, 2)
    cp = np.repeat(ep[:, None, :], 3, 1)
    Kp = _coo(rp, cp, local_stiffness(B, det, 1), (V, V))
    X, Y = local_divergence(B, det, False)  # (E,3,6)
    r = np.repeat(ep[:, :, None], 6, 2)
    c = np.repeat(ev[:, None, :], 3, 1)
    VPD = _coo(np.concatenate([r, r]), np.concatenate([c, c + N]), np.concatenate([X, Y]), (V, 2 * N))
    X, Y = local_divergence(B, det, True)  # (E,6,3)
    r = np.repeat(ev[:, :, None], 3, 2)
    c = np.repeat(ep[:, None, :], 6, 1)
    PVD = _coo(np.concatenate([r, r + N]), np.concatenate([c, c]), np.concatenate([X, Y]), (2 * N, V))
    Kp_int = Kp[int_p][:, int_p].tocsr()
    Kp_int.sort_indices()
    return {"vm": vm, "pm": pm, "N": N, "V": V, "E": E, "B": B, "det": det, "M": M, "K": K, "Kp": Kp, "Kp_int": Kp_int,
            "VPD": VPD, "PVD": PVD, "dvx": dvx, "dvy": dvy, "dp": dp, "int_p": int_p, "viscosity": viscosity}


def convection_values(ctx, u):
    """C(u) values in the velocity pattern (ctx['K'] ordering)."""
    N = ctx["N"]
    ev = ctx["vm"]["elements"]
    loc = local_self_convection(ctx["B"], ctx["det"], u[:N][ev], u[N:][ev])
    rv = np.repeat(ev[:, :, None], 6, 2)
    cv = np.repeat(ev[:, None, :], 6, 1)
    C = _coo(rv, cv, loc, (N, N))
    K = ctx["K"]
    if not (np.array_equal(C.indptr, K.indptr) and np.array_equal(C.indices, K.indices)):
        raise AssertionError("convection pattern differs from the stiffness pattern")
    return C.data.copy()


def fast_convection_map(ctx):
    """Int (nnz_C x 2N) with C.values = Int [ux; uy], by unit probes of the
    12 element velocities (cpp/libs/Fem/src/fastConvection.cpp:9-100)."""
    N = ctx["N"]
    ev = ctx["vm"]["elements"]
    K = ctx["K"]
    E = len(ev)
    # CSR position of (ids[r], ids[c]) for every element
    pos = np.empty((E, 6, 6), dtype=np.int64)
    for r in range(6):
        for cc in range(6):
            a, b = ev[:, r], ev[:, cc]
            starts = K.indptr[a]
            ends = K.indptr[a + 1]
            for e in range(E):
                seg = K.indices[starts[e]:ends[e]]
                pos[e, r, cc] = starts[e] + np.searchsorted(seg, b[e])
    rows, cols, vals = [], [], []
    z = np.zeros((E, 6))
    for k in range(12):
        ux, uy = z.copy(), z.copy()
        (uy if k >= 6 else ux)[:, k % 6] = 1.0
        loc = local_self_convection(ctx["B"], ctx["det"], ux, uy)
        col = ev[:, k % 6] + (N if k >= 6 else 0)
        rows.append(pos.reshape(E, 36))
        cols.append(np.repeat(col[:, None], 36, 1))
        vals.append(loc.reshape(E, 36))
    return _coo(np.concatenate(rows, 1), np.concatenate(cols, 1), np.concatenate(vals, 1), (K.nnz, 2 * N))


# ---------------------------------------------------------------------------
# Colouring (C helper) and multicolour layout
# (cpp/libs/CudaUtils/src/gaussSeidel.cu:140-182, LinAlg/src/gaussSeidel.cpp:9-65)
# ---------------------------------------------------------------------------

_clib = None


def build_clib():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def clib():
    global _clib
    if _clib is None:
        if not os.path.exists(LIB):
            build_clib()
        lib = ctypes.CDLL(LIB)
        lib.oracle_sl_greedy.restype = ctypes.c_int
        lib.oracle_sl_greedy.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 4
        lib.oracle_gs_custom_order.restype = None
        lib.oracle_gs_custom_order.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_int]
        _clib = lib
    return _clib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def sl_greedy_coloring(m):
    """(order, color, numColors) for the pattern of a square scipy CSR matrix."""
    m = m.tocsr()
    n = m.shape[0]
    rs = np.ascontiguousarray(m.indptr, np.int32)
    col = np.ascontiguousarray(m.indices, np.int32)
    order = np.empty(n, np.int32)
    color = np.empty(n, np.int32)
    k = clib().oracle_sl_greedy(n, _p(rs), _p(col), _p(order), _p(color))
    if k < 0:
        raise MemoryError("oracle colouring")
    return order, color, k


def mcgs_layout(m):
    """coloring (new -> old), partitionStart, permuted matrix."""
    _, color, k = sl_greedy_coloring(m)
    coloring = np.concatenate([np.sort(np.nonzero(color == c)[0]) for c in range(k)]).astype(np.int64)
    counts = np.bincount(color, minlength=k)
    ps = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    P = m.tocsr()[coloring][:, coloring].tocsr()
    P.sort_indices()
    return coloring, ps, P


def gs_custom_order(m, b, order, sweeps, x0=None):
    """Sequential GS in the given row order (C helper)."""
    m = m.tocsr()
    n = m.shape[0]
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    # keep every converted array alive for the duration of the call
    rs = np.ascontiguousarray(m.indptr, np.int32)
    col = np.ascontiguousarray(m.indices, np.int32)
    val = np.ascontiguousarray(m.data, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    od = np.ascontiguousarray(order, np.int32)
    clib().oracle_gs_custom_order(n, _p(rs), _p(col), _p(val), _p(x), _p(bb), _p(od), sweeps)
    return x


class Mcgs:
    """Multicolour GS on a permuted matrix: colour classes updated in order,
    each class at once from the current values of the others."""

    def __init__(self, m):
        self.coloring, self.ps, P = mcgs_layout(m)
        self.n = P.shape[0]
        self.diag = P.diagonal().copy()
        if np.any(self.diag == 0):
            raise ValueError("Matrix contains zeros on the diagonal")
        self.off = (P - sp.diags(self.diag)).tocsr()
        self.off.eliminate_zeros()
        self.P = P
        self.blocks = [self.off[self.ps[c]:self.ps[c + 1]] for c in range(len(self.ps) - 1)]

    def sweep(self, x, b, colors):
        for c in colors:
            a, e = self.ps[c], self.ps[c + 1]
            x[a:e] = (b[a:e] - self.blocks[c] @ x) / self.diag[a:e]
        return x

    def sgs(self, r):
        """Symmetric GS from zero: forward then backward over the colours."""
        z = np.zeros_like(r)
        k = len(self.ps) - 1
        self.sweep(z, r, range(k))
        self.sweep(z, r, range(k - 1, -1, -1))
        return z


def gs_solve(m, b, x0, max_iters, target, check_interval=1, layout=None):
    """The reference's GS solver (gaussSeidel.cu:224-310): returns x, mse, iterations."""
    L = layout or Mcgs(m)
    k = len(L.ps) - 1
    x = np.array(x0, dtype=np.float64)[L.coloring]
    bp = np.asarray(b, dtype=np.float64)[L.coloring]
    mse = -1.0
    it = 0
    done = False
    while it < max_iters:
        L.sweep(x, bp, range(k))
        if it % check_interval == 0:
            r = L.P @ x - bp
            mse = math.sqrt(float(r @ r) / L.n)
            done = mse < target
        if done:
            break
        it += 1
    out = np.empty_like(x)
    out[L.coloring] = x
    return out, mse, it


def pcg_solve(m, b, x0, tol, max_iters, layout=None):
    """CG preconditioned by one symmetric multicolour GS sweep; stop at
    |r| <= tol |b| (|r0| if b = 0).  Returns x, relres, iterations."""
    L = layout or Mcgs(m)
    A = L.P
    x = np.array(x0, dtype=np.float64)[L.coloring]
    bp = np.asarray(b, dtype=np.float64)[L.coloring]
    r = bp - A @ x
    bb = float(bp @ bp)
    rr = float(r @ r)
    ref = bb if bb > 0 else rr
    it = 0
    if rr > tol * tol * ref:
        z = L.sgs(r)
        p = z.copy()
        rho = float(r @ z)
        while it < max_iters:
            q = A @ p
            alpha = rho / float(p @ q)
            x += alpha * p
            r -= alpha * q
            it += 1
            rr = float(r @ r)
            if rr <= tol * tol * ref:
                break
            z = L.sgs(r)
            rho_new = float(r @ z)
            p = z + (rho_new / rho) * p
            rho = rho_new
    out = np.empty_like(x)
    out[L.coloring] = x
    return out, math.sqrt(rr / ref) if ref > 0 else 0.0, it


# ---------------------------------------------------------------------------
# Chorin projection step (execs/NavierStokes/src/chorinCuda.cu:236-376)
# ---------------------------------------------------------------------------

class ChorinOracle:
    """FP64 CPU restatement of the reference's time step.
    method: "gs" (reference solver semantics) or "pcg" (relative tolerance)."""

    def __init__(self, ctx, tau, vel=("pcg", 1000, 1e-12, 1), pres=("pcg", 5000, 1e-12, 1)):
        self.ctx = ctx
        self.tau = tau
        self.vel = vel
        self.pres = pres
        N, V = ctx["N"], ctx["V"]
        self.Lv = Mcgs(ctx["M"])
        self.Lp = Mcgs(ctx["Kp_int"])
        self.nuK = ctx["viscosity"] * ctx["K"].data
        self.u = np.zeros(2 * N)
        self.p = np.zeros(V)
        self.p_int = np.zeros(len(ctx["int_p"]))
        self.dir_ids = np.concatenate([ctx["dvx"][0], ctx["dvy"][0] + N])
        self.dir_vals = np.concatenate([ctx["dvx"][1], ctx["dvy"][1]])
        self.impose()

    def impose(self):
        self.u[self.dir_ids] = self.dir_vals

    def _solve(self, L, cfg, b, x0, nch):
        method, max_it, target, interval = cfg
        n = L.n
        out = np.empty_like(b)
        stats = []
        for c in range(nch):
            bc = b[c * n:(c + 1) * n]
            xc = x0[c * n:(c + 1) * n]
            if method == "gs":
                x, mse, it = gs_solve(None, bc, xc, max_it, target, interval, L)
            else:
                x, mse, it = pcg_solve(None, bc, xc, target, max_it, L)
            out[c * n:(c + 1) * n] = x
            stats.append((mse, it))
        return out, stats

    def _solve_velocity(self, rhs):
        method, max_it, target, interval = self.vel
        N = self.ctx["N"]
        if method == "gs":
            # both channels advance together and stop when both are below target
            L = self.Lv
            x = np.zeros((2, N))
            bp = rhs.reshape(2, N)[:, L.coloring]
            xp = x[:, L.coloring]
            k = len(L.ps) - 1
            it = 0
            while it < max_it:
                for ch in range(2):
                    L.sweep(xp[ch], bp[ch], range(k))
                if it % interval == 0:
                    r = (L.P @ xp.T).T - bp
                    mses = [math.sqrt(float(r[ch] @ r[ch]) / N) for ch in range(2)]
                    if mses[0] < target and mses[1] < target:
                        break
                it += 1
            out = np.empty((2, N))
            out[:, L.coloring] = xp
            return out.ravel()
        a, _ = self._solve(self.Lv, self.vel, rhs, np.zeros(2 * N), 2)
        return a

    def step(self):
        ctx = self.ctx
        N = ctx["N"]
        tau = self.tau
        A = ctx["K"].copy()
        A.data = self.nuK + convection_values(ctx, self.u)
        rhs = np.concatenate([A @ self.u[:N], A @ self.u[N:]])
        a = self._solve_velocity(rhs)
        self.u -= tau * a
        self.impose()
        self.u_tentative = self.u.copy()
        d = -(1.0 / tau) * (ctx["VPD"] @ self.u)
        if self.pres[0] == "gs":
            self.p_int, _, _ = gs_solve(None, d[ctx["int_p"]], self.p_int, self.pres[1], self.pres[2], self.pres[3], self.Lp)
        else:
            self.p_int, _, _ = pcg_solve(None, d[ctx["int_p"]], self.p_int, self.pres[2], self.pres[1], self.Lp)
        self.p = np.zeros(ctx["V"])
        self.p[ctx["int_p"]] = self.p_int
        g = ctx["PVD"] @ self.p
        a = self._solve_velocity(g)
        self.u -= tau * a
        self.impose()
        return self.u, self.p


# ---------------------------------------------------------------------------
# The CPU baseline: the same step in multithreaded C++ (oracle/cpu_step.cpp)
# ---------------------------------------------------------------------------

CPU_LIB = os.path.join(HERE, "_build", "libcpu_step.so")


class CpuStep:
    """FP64 C++/OpenMP Chorin step over a context dict with the keys of
    build_context (M, K, Kp_int, VPD, PVD, vm (elements, xy), dvx, dvy,
    int_p, viscosity, N, V): every solve SGS-PCG to rtol."""

    def __init__(self, ctx, tau, rtol=1e-10, threads=0):
        if not os.path.exists(CPU_LIB):
            raise RuntimeError(f"{CPU_LIB} is not built (make -C oracle)")
        self.lib = ctypes.CDLL(CPU_LIB)
        self.lib.cpu_step_error.restype = ctypes.c_char_p
        self.N, self.V = int(ctx["N"]), int(ctx["V"])
        keep = []

        def csr(m):
            m = m.tocsr()
            arrs = (np.ascontiguousarray(m.indptr, np.int32), np.ascontiguousarray(m.indices, np.int32),
                    np.ascontiguousarray(m.data, np.float64))
            keep.extend(arrs)
            return [_p(a) for a in arrs]

        def arr(a, dt):
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return _p(a)

        el = ctx["vm"]["elements"]
        h = ctypes.c_void_p()
        rc = self.lib.cpu_step_create(
            ctypes.c_int(self.N), ctypes.c_int(self.V), ctypes.c_int(len(el)), *csr(ctx["M"]), *csr(ctx["K"]),
            ctypes.c_double(ctx["viscosity"]), ctypes.c_int(len(ctx["int_p"])), *csr(ctx["Kp_int"]), *csr(ctx["VPD"]),
            *csr(ctx["PVD"]), arr(el, np.int32), arr(ctx["vm"]["xy"], np.float64), ctypes.c_int(len(ctx["dvx"][0])),
            arr(ctx["dvx"][0], np.int32), arr(ctx["dvx"][1], np.float64), ctypes.c_int(len(ctx["dvy"][0])),
            arr(ctx["dvy"][0], np.int32), arr(ctx["dvy"][1], np.float64), arr(ctx["int_p"], np.int32), ctypes.c_double(tau),
            ctypes.c_double(rtol), ctypes.c_int(threads), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("cpu_step_create failed")
        self.h = h

    @property
    def threads(self):
        return int(self.lib.cpu_step_threads())

    def set_velocity(self, u):
        u = np.ascontiguousarray(u, np.float64)
        self.lib.cpu_step_set_velocity(self.h, _p(u))

    def run(self, steps=1):
        """Returns the iteration counts (steps x [tentative, pressure, final])."""
        it = np.zeros((steps, 3), np.int32)
        if self.lib.cpu_step_run(self.h, ctypes.c_int(steps), _p(it)) != 0:
            raise RuntimeError("cpu step failed: " + self.lib.cpu_step_error(self.h).decode())
        return it

    def state(self):
        u = np.empty(2 * self.N)
        p = np.empty(self.V)
        self.lib.cpu_step_state(self.h, _p(u), _p(p))
        return u, p

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.cpu_step_destroy(h)
            self.h = None
