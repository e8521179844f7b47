"""The allocation-free dot product (nsfem_dot_noalloc) through the C ABI
against the allocating one (nsfem_dot): the same kernels, so the same
result bit for bit, on two calls in a row (the second smaller, so it reads
fewer of the shared scratch's partial sums than the first left behind), and
against an FP64 host sum to rounding."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    from nsfem_amd import _lib
    p = ctypes.c_void_p()
    _lib.call("nsfem_malloc", ctypes.c_size_t(a.nbytes), ctypes.byref(p))
    _lib.call("nsfem_upload", p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes))
    return p


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dot_noalloc_matches_dot(dtype):
    from nsfem_amd import _lib
    dt = _lib.F64 if dtype == np.float64 else _lib.F32
    rng = np.random.default_rng(11)
    for n in (3_000_001, 70_001):
        x = rng.uniform(-1, 1, n).astype(dtype)
        y = rng.uniform(-1, 1, n).astype(dtype)
        dx, dy = _dev(x), _dev(y)
        try:
            a, b = ctypes.c_double(), ctypes.c_double()
            _lib.call("nsfem_dot_noalloc", dt, n, dx, dy, ctypes.byref(a))
            _lib.call("nsfem_dot", dt, n, dx, dy, ctypes.byref(b))
            assert a.value == b.value, (n, a.value, b.value)
            ref = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
            assert abs(a.value - ref) <= 1e-12 * float(np.abs(x.astype(np.float64)) @ np.abs(y.astype(np.float64)))
        finally:
            _lib.call("nsfem_free", dx)
            _lib.call("nsfem_free", dy)


def test_dot_noalloc_rejects_bad_arguments():
    from nsfem_amd import _lib
    out = ctypes.c_double(7.0)
    with pytest.raises(_lib.NsfemInvalidArgument):
        _lib.call("nsfem_dot_noalloc", _lib.F64, -1, None, None, ctypes.byref(out))
    with pytest.raises(_lib.NsfemInvalidArgument):
        _lib.call("nsfem_dot_noalloc", 5, 10, None, None, ctypes.byref(out))
    _lib.call("nsfem_dot_noalloc", _lib.F64, 0, None, None, ctypes.byref(out))
    assert out.value == 0.0
