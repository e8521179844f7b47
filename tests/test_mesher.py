"""The seeded channel mesher: size, topology, border groups, reproducibility."""

import math

import numpy as np
import pytest


@pytest.mark.parametrize("n", [2000, 20000])
def test_channel_mesh_properties(n):
    from nsfem_amd import mesher
    m = mesher.channel_mesh(n, seed=11)
    xy, tri, border = m["xy"], m["triangles"], m["border"]
    V = len(xy)
    assert 0.85 * n < V < 1.15 * n
    # every node used, positive areas
    assert len(np.unique(tri)) == V
    p0, p1, p2 = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    det = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
    assert det.min() > 0
    area = 0.5 * det.sum()
    exact = 5.4 - math.pi * 0.12 ** 2
    assert abs(area - exact) < 2e-3 * exact
    # Euler characteristic of a domain with one hole: V - E + T = 0
    sides = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), 1)
    E = len(np.unique(sides[:, 0].astype(np.int64) * V + sides[:, 1]))
    assert V - E + len(tri) == 0
    # border groups sit on their walls
    g = m["groups"]
    for name, test in [("Left", lambda p: np.isclose(p[:, 0], 0)), ("Right", lambda p: np.isclose(p[:, 0], 5.4)),
                       ("Bottom", lambda p: np.isclose(p[:, 1], 0)), ("Top", lambda p: np.isclose(p[:, 1], 1.0)),
                       ("Circle", lambda p: np.isclose(np.hypot(p[:, 0] - 0.5, p[:, 1] - 0.49), 0.12, atol=1e-6))]:
        sel = border[border[:, 2] == g.index(name)]
        assert len(sel) > 0
        assert np.all(test(xy[sel[:, 0]])) and np.all(test(xy[sel[:, 1]]))


def test_channel_mesh_is_seeded():
    from nsfem_amd import mesher
    a = mesher.channel_mesh(3000, seed=5)
    b = mesher.channel_mesh(3000, seed=5)
    c = mesher.channel_mesh(3000, seed=6)
    assert np.array_equal(a["xy"], b["xy"]) and np.array_equal(a["triangles"], b["triangles"])
    assert not np.array_equal(a["xy"], c["xy"])


def test_generated_mesh_builds_a_context():
    from nsfem_amd import fem, mesher
    m = mesher.channel_mesh(5000, seed=3)
    ctx = fem.ChorinContext(mesher.to_triangle_mesh(m), fem.DfgConditions(0.0122, 1.5, 1.0, 4.0))
    assert 3.8 < ctx.numVelocityNodes / ctx.numPressureNodes < 4.1
    ids, vals = ctx.dirichlet("vx")
    assert abs(vals.max() - 1.5) < 0.01  # profileScale 4: peak reached mid-channel
