"""The library's host set-up (mesh numbering, Dirichlet lists, assembled
matrices, colouring) against the independent CPU oracle: patterns and
colourings bit-exact, FP64 values to rounding.  No GPU needed."""

import numpy as np
import pytest

from conftest import mesh_path

MATS = [("velocityMass", "M"), ("velocityStiffness", "K"), ("pressureStiffness", "Kp"),
        ("pressureStiffnessInternal", "Kp_int"), ("velocityPressureDiv", "VPD"), ("pressureVelocityDiv", "PVD")]


def _product(name, visc=1e-3, peak=1.0):
    from nsfem_amd import fem
    m = fem.TriangleMesh.load_gmsh(mesh_path(name))
    return fem.ChorinContext(m, fem.DfgConditions(visc, peak))


@pytest.fixture(scope="module", params=["c3", "h_0.025", "h_0.01"])
def pair(request, oracle):
    name = request.param
    o = oracle.build_context(oracle.parse_gmsh(mesh_path(name)), 1e-3, 1.0)
    return name, _product(name), o


def test_mesh_parse_matches_oracle(oracle):
    from nsfem_amd import fem
    for name in ("c3", "h_0.025", "h_0.01"):
        xy, tri, border, groups = fem.TriangleMesh.load_gmsh(mesh_path(name)).arrays()
        o = oracle.parse_gmsh(mesh_path(name))
        assert np.array_equal(xy, o["xy"])
        assert np.array_equal(tri, o["triangles"])
        assert np.array_equal(border, o["border"])
        assert groups == o["groups"]


def test_numbering_bit_exact(pair):
    name, p, o = pair
    xy, el = p.velocity_mesh()
    assert np.array_equal(el, o["vm"]["elements"])
    assert np.array_equal(xy, o["vm"]["xy"])
    xy1, el1 = p.pressure_mesh()
    assert np.array_equal(el1, o["pm"]["elements"])


def test_dirichlet_lists_exact(pair):
    name, p, o = pair
    for key, ok in (("vx", "dvx"), ("vy", "dvy"), ("p", "dp")):
        ids, vals = p.dirichlet(key)
        assert np.array_equal(ids, o[ok][0])
        assert np.array_equal(vals, o[ok][1])
    assert np.array_equal(p.internal_pressure(), o["int_p"])
    # inflow profile: peak * y (0.41 - y) / 0.41^2 on the Left border
    ids, vals = p.dirichlet("vx")
    xy, _ = p.velocity_mesh()
    # (the two inlet corners are claimed first by Top / Bottom: value 0)
    y_all = xy[ids, 1]
    left = np.isclose(xy[ids, 0], 0.0) & (y_all > 1e-6) & (y_all < 0.41 - 1e-6)
    y = xy[ids[left], 1]
    assert np.allclose(vals[left], y * (0.41 - y) / 0.41 ** 2, rtol=0, atol=1e-15)
    assert np.all(np.abs(vals[~left]) < 1e-8)


@pytest.mark.parametrize("pname,oname", MATS)
def test_matrices_pattern_exact_values_close(pair, pname, oname):
    name, p, o = pair
    rows, cols, rs, col, val = p.matrix(pname)
    A = o[oname]
    assert (rows, cols) == A.shape
    assert np.array_equal(rs, A.indptr)
    assert np.array_equal(col, A.indices)
    assert np.max(np.abs(val - A.data)) <= 1e-13 * np.max(np.abs(A.data))


def test_coloring_bit_exact(pair, oracle):
    from nsfem_amd import cu
    name, p, o = pair
    for pname, oname in (("velocityMass", "M"), ("pressureStiffnessInternal", "Kp_int")):
        rows, _, rs, col, _ = p.matrix(pname)
        coloring, ps = cu.mcgs_coloring(rows, rs, col)
        oc, ops, _ = oracle.mcgs_layout(o[oname])
        assert np.array_equal(coloring, oc)
        assert np.array_equal(ps, ops)


def test_fast_convection_map_matches_oracle(oracle, octx_c3):
    p = _product("c3")
    Int = p.fast_convection_map()
    Io = oracle.fast_convection_map(octx_c3)
    assert np.array_equal(Int.indptr, Io.indptr)
    assert np.array_equal(Int.indices, Io.indices)
    assert np.max(np.abs(Int.data - Io.data)) <= 1e-13 * np.max(np.abs(Io.data))


def test_generated_mesh_setup_matches_oracle(oracle):
    from nsfem_amd import fem, mesher
    m = mesher.dfg_channel(3000, seed=7)
    tm = mesher.to_triangle_mesh(m)
    p = fem.ChorinContext(tm, fem.DfgConditions(1e-3, 1.5))
    tri = {"xy": m["xy"], "triangles": m["triangles"].astype(np.int64),
           "border": oracle.border_from_edges(m["triangles"], m["border"]), "groups": m["groups"]}
    o = oracle.build_context(tri, 1e-3, 1.5)
    for pname, oname in MATS:
        _, _, rs, col, val = p.matrix(pname)
        assert np.array_equal(rs, o[oname].indptr) and np.array_equal(col, o[oname].indices)
        assert np.max(np.abs(val - o[oname].data)) <= 1e-13 * np.max(np.abs(o[oname].data))
    ids, vals = p.dirichlet("vx")
    assert np.array_equal(ids, o["dvx"][0]) and np.array_equal(vals, o["dvx"][1])


@pytest.mark.parametrize("name", ["c3", "h_0.025", "h_0.01"])
def test_numbering_matches_reference_build(oracle, name):
    """The library's P1/P2 element arrays and node coordinates equal the
    reference's own createMesh (built from its sources into oracle/_ref/)."""
    tri = oracle.parse_gmsh(mesh_path(name))
    ref2 = oracle.ref_create_mesh(tri, 2)
    if ref2 is None:
        pytest.skip("reference build oracle/_ref/libref_mesh.so absent (needs the reference sources)")
    ref1 = oracle.ref_create_mesh(tri, 1)
    p = _product(name)
    xy, el = p.velocity_mesh()
    assert np.array_equal(el, ref2["elements"])
    assert np.array_equal(xy.astype(np.float32), ref2["xy"])
    xy1, el1 = p.pressure_mesh()
    assert np.array_equal(el1, ref1["elements"])
    assert np.array_equal(xy1.astype(np.float32), ref1["xy"])
