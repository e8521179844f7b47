"""Python mirror of the reference's CudaUtils API (namespace cu), implemented
over the C ABI of libnsfem_hip.so.

  cu::vec<T>          include/cu/vec.h:16-387         -> vec
  cu::csrF            src/csrF.cu:5-40                -> csrF
  cu::spmv / cu::spmm src/spmv.cu:5-63, spmm.cu:5-70  -> spmv / spmm
  cu::saxpy / scale   src/blas.cu:51-67               -> saxpy / scale
  cu::AbstractSolver  include/cu/solvers/abstractSolver.h:8-32
  cu::createSolver    src/solvers/solverFactory.cu:209-237

Errors keep the reference's split: invalid arguments raise ValueError
(NsfemInvalidArgument), everything else a RuntimeError subclass."""

import ctypes

import numpy as np

from . import _lib
from ._lib import F32, F64, call

_NP = {F32: np.float32, F64: np.float64}


def dtype_id(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return F32
    if dt == np.float64:
        return F64
    raise ValueError(f"unsupported dtype {dt} (float32 or float64)")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class vec:
    """A device buffer of n elements (owning, or a view that never frees)."""

    def __init__(self, n_or_host, dtype=np.float64, _ptr_value=None, _owner=None):
        if _ptr_value is not None:
            self.dtype = np.dtype(dtype)
            self.n = int(n_or_host)
            self._p = ctypes.c_void_p(_ptr_value)
            self._owning = False
            self._owner = _owner
            return
        host = None
        if isinstance(n_or_host, (int, np.integer)):
            n = int(n_or_host)
            self.dtype = np.dtype(dtype)
        else:
            host = np.ascontiguousarray(n_or_host)
            self.dtype = host.dtype
            n = host.size
        if n == 0:
            raise _lib.NsfemInvalidArgument(_lib.NSFEM_EINVAL, "Size can't be zero")
        self.n = n
        self._p = ctypes.c_void_p()
        self._owning = True
        call("nsfem_malloc", n * self.dtype.itemsize, ctypes.byref(self._p))
        if host is not None:
            self.upload(host)

    @classmethod
    def view(cls, ptr, n, dtype, owner=None):
        """Non-owning view of externally managed device memory (vec.h:65-71)."""
        return cls(n, dtype, _ptr_value=ptr if isinstance(ptr, int) else ptr.value, _owner=owner)

    @property
    def ptr(self):
        return self._p

    def size(self):
        return self.n

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.size != self.n:
            raise _lib.NsfemInvalidArgument(_lib.NSFEM_EINVAL, f"Size mismatch - got {host.size}, expected {self.n}")
        call("nsfem_upload", self._p, _ptr(host), host.nbytes)

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        call("nsfem_download", _ptr(out), self._p, out.nbytes)
        return out

    def copyTo(self, dst):
        if dst.n != self.n or dst.dtype != self.dtype:
            raise _lib.NsfemInvalidArgument(_lib.NSFEM_EINVAL, "vec copy: size mismatch")
        call("nsfem_copy", dst._p, self._p, self.n * self.dtype.itemsize)

    def memsetZero(self):
        call("nsfem_memset_zero", self._p, self.n * self.dtype.itemsize)

    def free(self):
        if self._owning and self._p:
            call("nsfem_free", self._p)
        self._p = ctypes.c_void_p()
        self._owning = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class csrF:
    """Device CSR built from host arrays (deep copy); values mutable in place."""

    def __init__(self, rows, cols, rowStart, column, values, dtype=None):
        values = np.ascontiguousarray(values, dtype=dtype if dtype is not None else np.asarray(values).dtype)
        self.dtype = values.dtype
        rs = np.ascontiguousarray(rowStart, dtype=np.int32)
        col = np.ascontiguousarray(column, dtype=np.int32)
        self.rows, self.cols, self.nnz = int(rows), int(cols), int(col.size)
        self._h = ctypes.c_void_p()
        call("nsfem_csr_create", self.rows, self.cols, self.nnz, _ptr(rs), _ptr(col), _ptr(values), dtype_id(self.dtype),
             ctypes.byref(self._h))
        p = ctypes.c_void_p()
        call("nsfem_csr_values", self._h, ctypes.byref(p))
        self.values = vec.view(p.value, max(1, self.nnz), self.dtype, owner=self)

    @classmethod
    def from_scipy(cls, m, dtype=np.float64):
        m = m.tocsr()
        m.sort_indices()
        return cls(m.shape[0], m.shape[1], m.indptr, m.indices, m.data.astype(dtype))

    @property
    def handle(self):
        return self._h

    def mul(self, x, y, numCh=1, alpha=1.0, beta=0.0, lanes=0, ldx=None, ldy=None):
        call("nsfem_csr_mul", self._h, numCh, x.ptr, ldx if ldx is not None else self.cols, y.ptr,
             ldy if ldy is not None else self.rows, alpha, beta, lanes)

    def __del__(self):
        try:
            if self._h:
                call("nsfem_csr_destroy", self._h)
                self._h = ctypes.c_void_p()
        except Exception:
            pass


class spmv:
    """b = M x with default operands x (cols) and b (rows)."""

    def __init__(self, m: csrF):
        self.mat = m
        self.x = vec(m.cols, m.dtype)
        self.b = vec(m.rows, m.dtype)

    def compute(self, x=None, b=None):
        x = self.x if x is None else x
        b = self.b if b is None else b
        call("nsfem_spmv", self.mat.handle, x.ptr, b.ptr)


class spmm:
    """B = M X, X and B column-major with numCh columns."""

    def __init__(self, m: csrF, numCh):
        if numCh < 1:
            raise _lib.NsfemInvalidArgument(_lib.NSFEM_EINVAL, "SPMM channels must be at least 1")
        self.mat = m
        self.numCh = numCh
        self.x = vec(m.cols * numCh, m.dtype)
        self.b = vec(m.rows * numCh, m.dtype)

    def compute(self, x=None, b=None):
        x = self.x if x is None else x
        b = self.b if b is None else b
        call("nsfem_spmm", self.mat.handle, self.numCh, x.ptr, b.ptr)


def saxpy(n, src, dst, alpha):
    """dst += alpha * src"""
    call("nsfem_axpy", dtype_id(dst.dtype), n, alpha, src.ptr, dst.ptr)


def scale(n, dst, alpha):
    call("nsfem_scale", dtype_id(dst.dtype), n, alpha, dst.ptr)


def geam(n, alpha, a, beta, b, c):
    call("nsfem_geam", dtype_id(c.dtype), n, alpha, a.ptr, beta, b.ptr, c.ptr)


def dot(n, x, y):
    out = ctypes.c_double()
    call("nsfem_dot", dtype_id(x.dtype), n, x.ptr, y.ptr, ctypes.byref(out))
    return out.value


def nrm2(n, x):
    out = ctypes.c_double()
    call("nsfem_nrm2", dtype_id(x.dtype), n, x.ptr, ctypes.byref(out))
    return out.value


def scatter(idx, values, dense):
    call("nsfem_scatter", dtype_id(dense.dtype), idx.n, idx.ptr, values.ptr, dense.ptr)


def gather(dense, idx, values):
    call("nsfem_gather", dtype_id(dense.dtype), idx.n, dense.ptr, idx.ptr, values.ptr)


class AbstractSolver:
    """Solver handle with the reference's AbstractSolver interface."""

    def __init__(self, handle, num_ch, rows, io_dtype):
        self._h = handle
        self._numCh = num_ch
        self.rows = rows
        self.dtype = _NP[io_dtype]
        r, s = ctypes.c_void_p(), ctypes.c_void_p()
        call("nsfem_solver_rhs", self._h, ctypes.byref(r))
        call("nsfem_solver_sol", self._h, ctypes.byref(s))
        self._rhs = vec.view(r.value, rows * num_ch, self.dtype, owner=self)
        self._sol = vec.view(s.value, rows * num_ch, self.dtype, owner=self)

    def getNumCh(self):
        return self._numCh

    def getLastMse(self, ch=-1):
        out = ctypes.c_double()
        call("nsfem_solver_last_mse", self._h, ch, ctypes.byref(out))
        return out.value

    def getLastRelRes(self, ch=0, system=0):
        out = ctypes.c_double()
        call("nsfem_solver_last_relres", self._h, system, ch, ctypes.byref(out))
        return out.value

    def getLastIterations(self, system=None):
        out = ctypes.c_int()
        if system is None:
            call("nsfem_solver_last_iterations", self._h, ctypes.byref(out))
        else:
            call("nsfem_solver_last_iterations_system", self._h, system, ctypes.byref(out))
        return out.value

    def getLastOuterIterations(self):
        """pcg_mixed with mixedRestart: refinement steps of the last solve."""
        out = ctypes.c_int()
        call("nsfem_solver_last_outer_iterations", self._h, ctypes.byref(out))
        return out.value

    def getRhs(self):
        return self._rhs

    def getSol(self):
        return self._sol

    def solve(self):
        call("nsfem_solver_solve", self._h)

    def setMaxIters(self, n):
        call("nsfem_solver_set_max_iters", self._h, n)

    def setTargetMse(self, mse):
        call("nsfem_solver_set_target_mse", self._h, mse)

    def setMseCheckInterval(self, k):
        call("nsfem_solver_set_mse_check_interval", self._h, k)

    def setOption(self, key, value):
        call("nsfem_solver_set_option", self._h, key.encode(), float(value))

    def timeKernel(self, which, reps=50):
        """(average ms per launch, algorithmic bytes per launch, launches per
        PCG iteration) of one of the iteration's kernels."""
        ms, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        call("nsfem_solver_time_kernel", self._h, which.encode(), int(reps), ctypes.byref(ms), ctypes.byref(b), ctypes.byref(n))
        return ms.value, b.value, n.value

    def reshape(self, lanes=0, rowsPerBlock=0):
        call("nsfem_solver_reshape", self._h, int(lanes), int(rowsPerBlock))

    def shape(self):
        """(lanes per row, rows per workgroup) of the solver kernels."""
        a, b = ctypes.c_int(), ctypes.c_int()
        call("nsfem_solver_shape", self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def reshapeSweeps(self, lanes=0, rowsPerBlock=0):
        call("nsfem_solver_reshape_sweeps", self._h, int(lanes), int(rowsPerBlock))

    def kindRows(self):
        """Rows per workgroup of the (direction, update, rho) kernels."""
        a, b, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call("nsfem_solver_kind_rows", self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(d))
        return a.value, b.value, d.value

    def sweepShape(self):
        """(lanes per row, rows per workgroup) of the colour sweeps."""
        a, b = ctypes.c_int(), ctypes.c_int()
        call("nsfem_solver_sweep_shape", self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def autotune(self, reps=10):
        """Picks the fastest (lanes, rows per workgroup) by timing one
        graph-replayed PCG iteration, then the sweeps' own shape
        (sweepShape()); overwrites internal work vectors."""
        a, b = ctypes.c_int(), ctypes.c_int()
        call("nsfem_solver_autotune", self._h, int(reps), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def numColors(self, system=0):
        out = ctypes.c_int()
        call("nsfem_solver_num_colors", self._h, system, ctypes.byref(out))
        return out.value

    def coloring(self, system=0, n=None):
        k = self.numColors(system)
        n = self.rows if n is None else n
        col = np.empty(n, dtype=np.int32)
        ps = np.empty(k + 1, dtype=np.int32)
        call("nsfem_solver_coloring", self._h, system, _ptr(col), _ptr(ps))
        return col, ps

    def __del__(self):
        try:
            if self._h:
                call("nsfem_solver_destroy", self._h)
                self._h = None
        except Exception:
            pass


def _csr_arrays(m):
    """(rows, rowStart, column, values) from a scipy matrix or a tuple."""
    if hasattr(m, "tocsr"):
        m = m.tocsr()
        m.sort_indices()
        return m.shape[0], m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data
    rows, rowStart, column, values = m
    return rows, np.asarray(rowStart, np.int32), np.asarray(column, np.int32), np.asarray(values)


def createSolver(name, numCh, m, maxIters, targetMse, mseCheckInterval, dtype=None):
    """cu::createSolver: name "gs" | "pcg" | "pcg_mixed" | "bpcg" | "bpcg_mixed" | "amg_pcg" | "amg_pcg_mixed"
    ("dss" is rejected).
    m: scipy sparse matrix or (rows, rowStart, column, values).  The value
    dtype (float32 / float64) sets the compute precision."""
    rows, rs, col, vals = _csr_arrays(m)
    vals = np.ascontiguousarray(vals, dtype=dtype if dtype is not None else vals.dtype)
    rs = np.ascontiguousarray(rs)
    col = np.ascontiguousarray(col)
    h = ctypes.c_void_p()
    call("nsfem_solver_create", name.encode(), numCh, rows, col.size, _ptr(rs), _ptr(col), _ptr(vals), dtype_id(vals.dtype),
         maxIters, targetMse, mseCheckInterval, ctypes.byref(h))
    io = ctypes.c_int()
    call("nsfem_solver_io_dtype", h, ctypes.byref(io))
    return AbstractSolver(h, numCh, rows, io.value)


def createBatchSolver(name, numCh, matrices, maxIters, targetMse, mseCheckInterval, dtype=np.float64):
    """Independent systems solved together (block-diagonal, per-system convergence)."""
    arrs = [_csr_arrays(m) for m in matrices]
    rows = np.array([a[0] for a in arrs], np.int32)
    nnz = np.array([a[2].size for a in arrs], np.int32)
    keep = [(np.ascontiguousarray(a[1]), np.ascontiguousarray(a[2]), np.ascontiguousarray(a[3], np.float64)) for a in arrs]
    k = len(arrs)
    rsp = (ctypes.c_void_p * k)(*[x[0].ctypes.data for x in keep])
    cp = (ctypes.c_void_p * k)(*[x[1].ctypes.data for x in keep])
    vp = (ctypes.c_void_p * k)(*[x[2].ctypes.data for x in keep])
    h = ctypes.c_void_p()
    call("nsfem_solver_create_batch", name.encode(), numCh, k, _ptr(rows), _ptr(nnz), ctypes.cast(rsp, ctypes.c_void_p),
         ctypes.cast(cp, ctypes.c_void_p), ctypes.cast(vp, ctypes.c_void_p), dtype_id(dtype), maxIters, targetMse,
         mseCheckInterval, ctypes.byref(h))
    io = ctypes.c_int()
    call("nsfem_solver_io_dtype", h, ctypes.byref(io))
    total = int(rows.sum())
    s = AbstractSolver(h, numCh, total, io.value)
    s.system_rows = rows
    s.system_offsets = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
    return s


def mcgs_coloring(rows, rowStart, column):
    """Smallest-last greedy colouring (host): (coloring, partitionStart)."""
    rs = np.ascontiguousarray(rowStart, np.int32)
    col = np.ascontiguousarray(column, np.int32)
    coloring = np.empty(rows, np.int32)
    ps = np.empty(rows + 1, np.int32)
    k = ctypes.c_int()
    call("nsfem_mcgs_coloring", rows, _ptr(rs), _ptr(col), _ptr(coloring), _ptr(ps), ctypes.byref(k))
    return coloring, ps[: k.value + 1].copy()
