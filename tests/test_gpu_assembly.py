"""GPU assembly of the Chorin matrices (values over host-built patterns)
against the host assembly: identical patterns and bit-identical values (the
host assembly is itself checked against the oracle and closed forms), and an
identical Chorin step from either context."""

import numpy as np
import pytest

from conftest import mesh_path

pytestmark = pytest.mark.gpu

MATS = ("velocityMass", "velocityStiffness", "pressureStiffness", "pressureStiffnessInternal", "velocityPressureDiv",
        "pressureVelocityDiv")


def _meshes():
    from nsfem_amd import fem, mesher
    yield "c3", fem.TriangleMesh.load_gmsh(mesh_path("c3")), fem.DfgConditions(1e-3, 1.0)
    yield "h_0.025", fem.TriangleMesh.load_gmsh(mesh_path("h_0.025")), fem.DfgConditions(1e-3, 1.0)
    m = mesher.channel_mesh(20000, seed=41, radius=0.1)
    yield "channel20k", mesher.to_triangle_mesh(m), fem.DfgConditions(0.004, 1.3, m["height"], 4.0)


@pytest.mark.parametrize("case", [0, 1, 2])
def test_gpu_assembly_is_bit_identical(case):
    from nsfem_amd import fem
    name, mesh, cond = list(_meshes())[case]
    host = fem.ChorinContext(mesh, cond)
    gpu = fem.ChorinContext(mesh, cond, gpu_assembly=True)
    for mat in MATS:
        a, b = host.matrix(mat), gpu.matrix(mat)
        assert a[:2] == b[:2], (name, mat)
        assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]), (name, mat, "pattern")
        # bitwise, so a signed zero or a last-bit difference shows
        assert np.array_equal(a[4].view(np.int64), b[4].view(np.int64)), (name, mat, np.max(np.abs(a[4] - b[4])))


def test_chorin_step_from_gpu_assembled_context_is_identical():
    from nsfem_amd import F64, fem
    mesh = fem.TriangleMesh.load_gmsh(mesh_path("h_0.025"))
    cond = fem.DfgConditions(1e-3, 1.0)
    cfg = fem.ChorinConfig(fem.SolverConfig("pcg", 500, 1e-12, 1), fem.SolverConfig("pcg", 3000, 1e-12, 1), dtype=F64)
    out = []
    for gpu in (False, True):
        ch = fem.Chorin(fem.ChorinContext(mesh, cond, gpu_assembly=gpu), cfg)
        ch.set_time_step(1e-3)
        for _ in range(3):
            ch.step()
        out.append(ch.download())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_batch_context_with_gpu_assembly():
    from nsfem_amd import fem, mesher
    meshes, conds = [], []
    for k, n in enumerate((900, 2500)):
        m = mesher.channel_mesh(n, seed=60 + k)
        meshes.append(mesher.to_triangle_mesh(m))
        conds.append(fem.DfgConditions(0.01 * (k + 1), 1.0, m["height"], 4.0))
    a = fem.ChorinContext.batch(meshes, conds)
    b = fem.ChorinContext.batch(meshes, conds, gpu_assembly=True)
    for mat in MATS:
        x, y = a.matrix(mat), b.matrix(mat)
        assert np.array_equal(x[3], y[3]) and np.array_equal(x[4].view(np.int64), y[4].view(np.int64)), mat


OMATS = dict(zip(MATS, ("M", "K", "Kp", "Kp_int", "VPD", "PVD")))


@pytest.mark.parametrize("name", ["c3", "h_0.025", "h_0.01"])
def test_gpu_assembly_matches_oracle(oracle, name):
    """GPU-assembled values against the independent FP64 oracle's assembly
    (closed-form-pinned element matrices, scipy COO summation): identical CSR
    patterns, values to 1e-13 relative to the matrix's largest entry."""
    from nsfem_amd import fem
    gpu = fem.ChorinContext(fem.TriangleMesh.load_gmsh(mesh_path(name)), fem.DfgConditions(1e-3, 1.0), gpu_assembly=True)
    o = oracle.build_context(oracle.parse_gmsh(mesh_path(name)), 1e-3, 1.0)
    for mat in MATS:
        rows, cols, rs, col, val = gpu.matrix(mat)
        A = o[OMATS[mat]]
        assert (rows, cols) == A.shape, mat
        assert np.array_equal(rs, A.indptr) and np.array_equal(col, A.indices), (name, mat)
        err = np.max(np.abs(val - A.data)) / np.max(np.abs(A.data))
        assert err <= 1e-13, (name, mat, err)
