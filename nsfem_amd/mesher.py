"""Seeded unstructured triangulation of a channel with a circular hole.

The reference reads gmsh meshes of the DFG channel (cpp/assets/*.geo: a
2.2 x 0.41 box with an r = 0.05 cylinder at (0.2, 0.2)); the large
workloads need meshes gmsh is not here to make.  This mesher fills the
domain with a jittered hexagonal lattice (spacing from the requested node
count), places evenly spaced nodes on the four walls and on the circle,
Delaunay-triangulates the lot and drops the triangles inside the hole.
Border edges get the reference's group names and group order
(Right, Top, Left, Bottom, Circle, Interior - cpp/assets/h_0.01.msh).
Node and triangle order follow a Morton curve for memory locality."""

import math

import numpy as np

GROUPS = ["Right", "Top", "Left", "Bottom", "Circle", "Interior"]
RIGHT, TOP, LEFT, BOTTOM, CIRCLE = range(5)


def _morton(ix, iy):
    def spread(v):
        v = v.astype(np.uint64) & np.uint64(0xFFFFFFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x0000FFFF0000FFFF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x3333333333333333)
        v = (v | (v << np.uint64(1))) & np.uint64(0x5555555555555555)
        return v
    return spread(ix) | (spread(iy) << np.uint64(1))


def channel_mesh(num_nodes, length=5.4, height=1.0, center=(0.5, 0.49), radius=0.12, seed=20241103, jitter=0.15):
    """Returns dict(xy, triangles, border, groups, h) with about num_nodes vertices.
    xy is rounded to float32 (the reference stores coordinates as float)."""
    from scipy.spatial import Delaunay

    if num_nodes < 50:
        raise ValueError("num_nodes too small")
    cx, cy = center
    if not (radius > 0 and cx - radius > 0 and cx + radius < length and cy - radius > 0 and cy + radius < height):
        raise ValueError("the hole must lie strictly inside the channel")
    area = length * height - math.pi * radius * radius
    h = math.sqrt(2.0 * area / (math.sqrt(3.0) * num_nodes))
    rng = np.random.default_rng(seed)

    # Boundary nodes: rectangle corners and walls, circle.
    nx = max(2, int(round(length / h)))
    ny = max(2, int(round(height / h)))
    xs = np.linspace(0.0, length, nx + 1)
    ys = np.linspace(0.0, height, ny + 1)
    bottom = np.stack([xs, np.zeros_like(xs)], 1)
    top = np.stack([xs, np.full_like(xs, height)], 1)
    left = np.stack([np.zeros(ny - 1), ys[1:-1]], 1)
    right = np.stack([np.full(ny - 1, length), ys[1:-1]], 1)
    nc = max(8, int(round(2 * math.pi * radius / h)))
    th = np.arange(nc) * (2 * math.pi / nc)
    circle = np.stack([cx + radius * np.cos(th), cy + radius * np.sin(th)], 1)

    # Interior: jittered hexagonal lattice, kept half a spacing off the walls.
    dy = h * math.sqrt(3.0) / 2.0
    rows = np.arange(1, int(height / dy) + 1) * dy
    pts = []
    for j, y in enumerate(rows):
        x0 = h * (0.5 if j % 2 else 1.0)
        x = np.arange(x0, length, h)
        pts.append(np.stack([x, np.full_like(x, y)], 1))
    inner = np.concatenate(pts)
    inner = inner + rng.uniform(-jitter * h, jitter * h, inner.shape)
    m = ((inner[:, 0] > 0.5 * h) & (inner[:, 0] < length - 0.5 * h) & (inner[:, 1] > 0.5 * h) & (inner[:, 1] < height - 0.5 * h))
    d = np.hypot(inner[:, 0] - cx, inner[:, 1] - cy)
    m &= d > radius + 0.5 * h
    inner = inner[m]

    xy = np.concatenate([bottom, top, left, right, circle, inner])
    xy = xy.astype(np.float32).astype(np.float64)

    tri = Delaunay(xy, qhull_options="Qbb Qc Qz").simplices.astype(np.int64)
    if len(np.unique(tri)) != len(xy):
        raise RuntimeError("triangulation dropped input nodes")
    cen = xy[tri].mean(axis=1)
    keep = np.hypot(cen[:, 0] - cx, cen[:, 1] - cy) > radius
    tri = tri[keep]
    # Positive orientation and no degenerate triangles.
    p0, p1, p2 = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    det = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
    neg = det < 0
    tri[neg] = tri[neg][:, [0, 2, 1]]
    if np.min(np.abs(det)) < 1e-6 * h * h:
        raise RuntimeError("degenerate triangle in the channel mesh")

    # Renumber nodes along a Morton curve, then triangles by their first node.
    q = (1 << 20) / max(length, height)
    order = np.argsort(_morton((xy[:, 0] * q).astype(np.int64), (xy[:, 1] * q).astype(np.int64)), kind="stable")
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    xy = xy[order]
    tri = inv[tri]
    key = _morton((xy[tri, 0].mean(1) * q).astype(np.int64), (xy[tri, 1].mean(1) * q).astype(np.int64))
    tri = tri[np.argsort(key, kind="stable")]

    # Border edges: sides used by exactly one triangle, oriented as in it.
    sides = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    key = np.minimum(sides[:, 0], sides[:, 1]) * len(xy) + np.maximum(sides[:, 0], sides[:, 1])
    uniq, idx, cnt = np.unique(key, return_index=True, return_counts=True)
    bsides = sides[idx[cnt == 1]]
    a, b = xy[bsides[:, 0]], xy[bsides[:, 1]]
    eps = 1e-6 * h
    # walls at their float32-rounded positions (the coordinates are rounded)
    lr = float(np.float32(length))
    hr = float(np.float32(height))
    group = np.full(len(bsides), CIRCLE, np.int64)
    group[(np.abs(a[:, 0] - lr) < eps) & (np.abs(b[:, 0] - lr) < eps)] = RIGHT
    group[(np.abs(a[:, 1] - hr) < eps) & (np.abs(b[:, 1] - hr) < eps)] = TOP
    group[(np.abs(a[:, 0]) < eps) & (np.abs(b[:, 0]) < eps)] = LEFT
    group[(np.abs(a[:, 1]) < eps) & (np.abs(b[:, 1]) < eps)] = BOTTOM
    mid = 0.5 * (a + b)
    onc = group == CIRCLE
    if np.any(np.abs(np.hypot(mid[onc, 0] - cx, mid[onc, 1] - cy) - radius) > 0.1 * h):
        raise RuntimeError("unclassified border edge in the channel mesh")
    border = np.concatenate([bsides, group[:, None]], 1)
    order = np.lexsort((border[:, 0], border[:, 2]))
    border = border[order]
    return {"xy": xy, "triangles": tri.astype(np.int32), "border": border.astype(np.int32), "groups": list(GROUPS), "h": h,
            "length": length, "height": height, "center": center, "radius": radius}


def dfg_channel(num_nodes, seed=20241103):
    """The reference's unscaled DFG geometry (2.2 x 0.41, r = 0.05 at (0.2, 0.2))."""
    return channel_mesh(num_nodes, length=2.2, height=0.41, center=(0.2, 0.2), radius=0.05, seed=seed)


def to_triangle_mesh(m):
    from .fem import TriangleMesh
    return TriangleMesh.from_arrays(m["xy"], m["triangles"], m["border"], m["groups"])


def uneven_instances(count, seed):
    """The batched workload: independent channel-with-hole instances of
    uneven sizes, P1 pressure DOFs log-uniform in [10k, 250k], hole radius
    uniform in [0.08, 0.16], inflow peak uniform in [0.5, 2.0], Re uniform in
    [20, 100] (viscosity = 2/3 peak * diameter / Re), all drawn from one seed;
    instance i is meshed with seed + 1 + i.  Returns a list of dicts (mesh,
    radius, peak, re, viscosity)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nodes = int(round(math.exp(rng.uniform(math.log(10_000), math.log(250_000)))))
        radius = float(rng.uniform(0.08, 0.16))
        peak = float(rng.uniform(0.5, 2.0))
        re = float(rng.uniform(20.0, 100.0))
        m = channel_mesh(nodes, radius=radius, seed=seed + 1 + i)
        out.append({"mesh": m, "radius": radius, "peak": peak, "re": re, "viscosity": (2.0 / 3.0) * peak * 2 * radius / re})
    return out
