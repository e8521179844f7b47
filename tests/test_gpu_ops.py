"""GPU parity of the generic sparse / vector kernels against numpy / scipy."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _rand_csr(rng, rows, cols, density, empty_rows=()):
    m = sp.random(rows, cols, density=density, random_state=np.random.RandomState(rng.integers(1 << 31)), format="csr")
    m = m.tolil()
    for r in empty_rows:
        m.rows[r] = []
        m.data[r] = []
    m = m.tocsr()
    m.sort_indices()
    return m


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 2e-6)])
@pytest.mark.parametrize("shape,density", [((1, 1), 1.0), ((37, 53), 0.2), ((5000, 4000), 0.002), ((3000, 3000), 0.02)])
def test_spmv_spmm(dtype, tol, shape, density):
    from nsfem_amd import cu
    rng = np.random.default_rng(11)
    m = _rand_csr(rng, *shape, density, empty_rows=[0] if shape[0] > 1 else [])
    A = cu.csrF.from_scipy(m, dtype)
    x = rng.standard_normal(shape[1] * 3).astype(dtype)
    xv = cu.vec(x)
    for nch in (1, 2, 3):
        y = cu.vec(shape[0] * nch, dtype)
        A.mul(xv, y, numCh=nch)
        got = y.download()
        ref = np.concatenate([m @ x[c * shape[1]:(c + 1) * shape[1]].astype(np.float64) for c in range(nch)])
        scale = max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(got - ref)) <= tol * scale * max(1, m.getnnz(1).max())


@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
def test_spmv_lane_widths_agree(lanes):
    from nsfem_amd import cu
    rng = np.random.default_rng(12)
    m = _rand_csr(rng, 2000, 2000, 0.01)
    A = cu.csrF.from_scipy(m, np.float64)
    x = cu.vec(rng.standard_normal(2000))
    y = cu.vec(2000, np.float64)
    A.mul(x, y, lanes=lanes)
    assert np.allclose(y.download(), m @ x.download(), rtol=1e-13, atol=1e-13)


def test_spmv_alpha_beta_and_reference_wrappers():
    from nsfem_amd import cu
    rng = np.random.default_rng(13)
    m = _rand_csr(rng, 300, 200, 0.05)
    A = cu.csrF.from_scipy(m)
    op = cu.spmv(A)
    x = rng.standard_normal(200)
    op.x.upload(x)
    op.compute()
    assert np.allclose(op.b.download(), m @ x, rtol=1e-13, atol=1e-14)
    y0 = rng.standard_normal(300)
    y = cu.vec(y0)
    A.mul(op.x, y, alpha=-2.0, beta=0.5)
    assert np.allclose(y.download(), -2.0 * (m @ x) + 0.5 * y0, rtol=1e-13, atol=1e-13)
    mm = cu.spmm(A, 2)
    X = rng.standard_normal(400)
    mm.x.upload(X)
    mm.compute()
    assert np.allclose(mm.b.download(), np.concatenate([m @ X[:200], m @ X[200:]]), rtol=1e-13, atol=1e-14)
    # values are mutable in place (csrF.values)
    A.values.upload(2 * m.data)
    op.compute()
    assert np.allclose(op.b.download(), 2 * (m @ x), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_blas1(dtype):
    from nsfem_amd import cu
    rng = np.random.default_rng(14)
    n = 100003
    a = rng.standard_normal(n).astype(dtype)
    b = rng.standard_normal(n).astype(dtype)
    va, vb = cu.vec(a), cu.vec(b)
    cu.saxpy(n, va, vb, -0.25)
    ref = (b + dtype(-0.25) * a)
    assert np.allclose(vb.download(), ref, rtol=1e-6 if dtype == np.float32 else 1e-15)
    cu.scale(n, vb, 3.0)
    ref = ref * 3
    assert np.allclose(vb.download(), ref, rtol=1e-6 if dtype == np.float32 else 1e-15)
    d = cu.dot(n, va, va)
    assert abs(d - float(a.astype(np.float64) @ a)) <= 1e-12 * d
    assert abs(cu.nrm2(n, va) - np.sqrt(d)) <= 1e-12 * np.sqrt(d)
    c = cu.vec(n, dtype)
    cu.geam(n, 2.0, va, -1.0, vb, c)
    assert np.allclose(c.download(), 2 * a - vb.download(), rtol=1e-6, atol=1e-6)
    # empty inputs
    assert cu.dot(0, va, va) == 0.0


def test_scatter_gather_and_copies():
    from nsfem_amd import cu
    rng = np.random.default_rng(15)
    dense = rng.standard_normal(1000)
    idx = rng.choice(1000, 137, replace=False).astype(np.int32)
    vals = rng.standard_normal(137)
    vd, vi, vv = cu.vec(dense), cu.vec(idx), cu.vec(vals)
    out = cu.vec(137, np.float64)
    cu.gather(vd, vi, out)
    assert np.array_equal(out.download(), dense[idx])
    cu.scatter(vi, vv, vd)
    ref = dense.copy()
    ref[idx] = vals
    assert np.array_equal(vd.download(), ref)
    other = cu.vec(1000, np.float64)
    vd.copyTo(other)
    assert np.array_equal(other.download(), ref)
    other.memsetZero()
    assert not other.download().any()


def test_copy_bandwidth_hook():
    """The achievable-bandwidth hook copies, times and frees: a plausible
    rate for a 64 MiB copy (above 100 GB/s, below 20 TB/s read + write), and
    bad sizes are rejected."""
    import ctypes
    from nsfem_amd import _lib
    v = ctypes.c_double()
    _lib.call("nsfem_copy_bandwidth", ctypes.c_size_t(64 << 20), 3, ctypes.byref(v))
    assert 1e11 < v.value < 2e13
    with pytest.raises(Exception):
        _lib.call("nsfem_copy_bandwidth", ctypes.c_size_t(8), 3, ctypes.byref(v))
