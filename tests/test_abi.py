"""The C-ABI library loads, exports every declared entry point, and rejects bad
arguments before touching the device (no GPU needed)."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nsfem_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|const char \*)\s+(nsfem_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from nsfem_amd import _lib
    lib = _lib.load()
    decl = declared_symbols()
    assert len(decl) >= 60
    for name in decl:
        assert hasattr(lib, name), name
    # the Python binding covers exactly the declared ABI
    assert sorted(_lib.exported_symbols()) == decl


def test_version_and_error_string():
    from nsfem_amd import _lib
    lib = _lib.load()
    a, b = ctypes.c_int(), ctypes.c_int()
    assert lib.nsfem_version(ctypes.byref(a), ctypes.byref(b)) == 0
    assert (a.value, b.value) == (1, 0)
    assert lib.nsfem_version(None, None) == _lib.NSFEM_EINVAL
    assert b"NULL" in lib.nsfem_last_error()


@pytest.mark.parametrize("name,interval,expect", [
    ("dss", 1, "dss"),
    ("cholesky", 1, "Invalid solver name"),
    ("gs", 0, "Invalid MSE check interval"),
])
def test_solver_factory_rejects(name, interval, expect):
    from nsfem_amd import _lib, cu
    rs = np.array([0, 1], np.int32)
    col = np.array([0], np.int32)
    val = np.array([2.0])
    with pytest.raises(_lib.NsfemInvalidArgument, match=expect):
        cu.createSolver(name, 1, (1, rs, col, val), 10, 1e-6, interval)


def test_solver_factory_rejects_bad_channels_and_mixed_f32():
    from nsfem_amd import _lib, cu
    rs = np.array([0, 1], np.int32)
    col = np.array([0], np.int32)
    with pytest.raises(_lib.NsfemInvalidArgument, match="channels"):
        cu.createSolver("gs", 0, (1, rs, col, np.array([2.0])), 10, 1e-6, 1)
    with pytest.raises(_lib.NsfemInvalidArgument, match="FP64"):
        cu.createSolver("pcg_mixed", 1, (1, rs, col, np.array([2.0], np.float32)), 10, 1e-6, 1)


def test_csr_validation():
    from nsfem_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    rs = np.array([0, 2], np.int32)
    col = np.array([0, 5], np.int32)  # column out of range
    val = np.array([1.0, 2.0])
    rc = lib.nsfem_csr_create(1, 2, 2, rs.ctypes.data, col.ctypes.data, val.ctypes.data, 1, ctypes.byref(h))
    assert rc == _lib.NSFEM_EINVAL
    rc = lib.nsfem_csr_create(1, 2, 2, rs.ctypes.data, col.ctypes.data, val.ctypes.data, 7, ctypes.byref(h))
    assert rc == _lib.NSFEM_EINVAL


def test_zero_size_allocation_rejected():
    from nsfem_amd import _lib
    lib = _lib.load()
    p = ctypes.c_void_p()
    assert lib.nsfem_malloc(0, ctypes.byref(p)) == _lib.NSFEM_EINVAL
    assert b"zero" in lib.nsfem_last_error()


def test_mesh_errors():
    from nsfem_amd import _lib, fem
    with pytest.raises(_lib.NsfemError, match="Failed to open"):
        fem.TriangleMesh.load_gmsh("/nonexistent.msh")
    # a mesh without the DFG border groups cannot build a Chorin context
    xy = np.array([[0, 0], [1, 0], [0, 1]], float)
    m = fem.TriangleMesh.from_arrays(xy, [[0, 1, 2]], np.zeros((0, 3), np.int32), ["Interior"])
    with pytest.raises(_lib.NsfemInvalidArgument, match="No Left border"):
        fem.ChorinContext(m, fem.DfgConditions())
