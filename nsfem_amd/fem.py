"""Meshes, the Chorin context and the GPU time loop (Python side of the C ABI).

  mesh::parseTriangleGmsh / createMesh   cpp/libs/Mesh/src/io.cpp:65-214, concreteMesh.cpp:78-293
  buildChorinContext                     cpp/execs/NavierStokes/src/buildContext.cpp:17-142
  solveNsChorinCuda                      cpp/execs/NavierStokes/src/chorinCuda.cu:154-379
  ChorinCudaConfig / SolverConfig        include/NavierStokes/chorinCudaConfig.h:6-17
  DfgConditions                          include/NavierStokes/dfgCondtions.h:4-13
"""

import ctypes
import dataclasses
import math
import time

import numpy as np

from . import _lib
from ._lib import (F32, F64, PRESSURE_START_EXTRAPOLATE, PRESSURE_START_LAST, TAU_CLAMP, TAU_REJECT,  # noqa: F401
                   TAU_UNCHECKED, VELOCITY_START_EXTRAPOLATE, VELOCITY_START_LAST, VELOCITY_START_ZERO, call)

_NP = {F32: np.float32, F64: np.float64}


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class TriangleMesh:
    """Corner-only triangle mesh with border elements and named groups."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def load_gmsh(cls, path):
        h = ctypes.c_void_p()
        call("nsfem_mesh_load_gmsh", str(path).encode(), ctypes.byref(h))
        return cls(h)

    @classmethod
    def from_arrays(cls, xy, triangles, border_edges, groups):
        xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
        tri = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        be = np.ascontiguousarray(border_edges, np.int32).reshape(-1, 3)
        names = (ctypes.c_char_p * len(groups))(*[g.encode() for g in groups])
        h = ctypes.c_void_p()
        call("nsfem_mesh_from_arrays", xy.shape[0], _ptr(xy), tri.shape[0], _ptr(tri), be.shape[0], _ptr(be), len(groups),
             ctypes.cast(names, ctypes.c_void_p), ctypes.byref(h))
        return cls(h)

    @property
    def handle(self):
        return self._h

    def info(self):
        n, t, b, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call("nsfem_mesh_info", self._h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(b), ctypes.byref(g))
        return n.value, t.value, b.value, g.value

    def arrays(self):
        n, t, b, g = self.info()
        xy = np.empty((n, 2), np.float64)
        tri = np.empty((t, 3), np.int32)
        border = np.empty((b, 3), np.int32)
        call("nsfem_mesh_arrays", self._h, _ptr(xy), _ptr(tri), _ptr(border))
        names = []
        buf = ctypes.create_string_buffer(256)
        for k in range(g):
            call("nsfem_mesh_group_name", self._h, k, buf, 256)
            names.append(buf.value.decode())
        return xy, tri, border, names

    def __del__(self):
        try:
            if self._h:
                call("nsfem_mesh_destroy", self._h)
                self._h = None
        except Exception:
            pass


@dataclasses.dataclass
class DfgConditions:
    viscosity: float = 1e-3
    peakVelocity: float = 1.0
    channelHeight: float = 0.41
    profileScale: float = 1.0

    def left_velocity(self, y):
        h = self.channelHeight
        return self.profileScale * self.peakVelocity * y * (h - y) / (h * h)


MATRICES = {
    "velocityMass": _lib.MAT_VELOCITY_MASS,
    "velocityStiffness": _lib.MAT_VELOCITY_STIFFNESS,
    "pressureStiffness": _lib.MAT_PRESSURE_STIFFNESS,
    "pressureStiffnessInternal": _lib.MAT_PRESSURE_STIFFNESS_INTERNAL,
    "velocityPressureDiv": _lib.MAT_VELOCITY_PRESSURE_DIV,
    "pressureVelocityDiv": _lib.MAT_PRESSURE_VELOCITY_DIV,
}


class ChorinContext:
    """Host-side set-up of one mesh: P2/P1 numbering, Dirichlet lists, FP64 matrices."""

    def __init__(self, mesh: TriangleMesh, cond: DfgConditions, threads=0, gpu_assembly=False, _handle=None):
        self.cond = cond
        if _handle is None:
            c = _lib.DfgConditions(cond.viscosity, cond.peakVelocity, cond.channelHeight, cond.profileScale)
            self._h = ctypes.c_void_p()
            flags = _lib.NSFEM_CONTEXT_GPU_ASSEMBLY if gpu_assembly else 0
            call("nsfem_context_build_ex", mesh.handle, ctypes.byref(c), threads, flags, ctypes.byref(self._h))
        else:
            self._h = _handle
        self._init_sizes()

    @classmethod
    def batch(cls, meshes, conds, threads=0, gpu_assembly=False):
        """Independent instances as one block-diagonal context (velocity
        [ux of all; uy of all]); the Chorin solvers keep one system per instance."""
        if len(meshes) != len(conds) or not meshes:
            raise ValueError("need one condition per mesh")
        arr = (_lib.DfgConditions * len(conds))(*[
            _lib.DfgConditions(c.viscosity, c.peakVelocity, c.channelHeight, c.profileScale) for c in conds])
        hs = (ctypes.c_void_p * len(meshes))(*[m.handle.value for m in meshes])
        h = ctypes.c_void_p()
        flags = _lib.NSFEM_CONTEXT_GPU_ASSEMBLY if gpu_assembly else 0
        call("nsfem_context_build_batch", ctypes.cast(hs, ctypes.c_void_p), arr, len(meshes), threads, flags, ctypes.byref(h))
        return cls(None, None, _handle=h)

    def num_systems(self):
        out = ctypes.c_int()
        call("nsfem_context_num_systems", self._h, ctypes.byref(out))
        return out.value

    def system_offsets(self):
        k = self.num_systems()
        vo = np.empty(k, np.int32)
        po = np.empty(k, np.int32)
        call("nsfem_context_system_offsets", self._h, _ptr(vo), _ptr(po))
        return vo, po

    def _init_sizes(self):
        N, V, nd, ni, E = (ctypes.c_int() for _ in range(5))
        call("nsfem_context_sizes", self._h, ctypes.byref(N), ctypes.byref(V), ctypes.byref(nd), ctypes.byref(ni), ctypes.byref(E))
        self.numVelocityNodes = N.value
        self.numPressureNodes = V.value
        self.numDirichletVelocity = nd.value
        self.numInternalPressure = ni.value
        self.numElements = E.value

    @property
    def handle(self):
        return self._h

    def matrix(self, name):
        """(rows, cols, rowStart, column, values) copies of a context matrix."""
        rows, cols, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rs, col, val = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        call("nsfem_context_matrix", self._h, MATRICES[name], ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz),
             ctypes.byref(rs), ctypes.byref(col), ctypes.byref(val))
        r, n = rows.value, nnz.value
        rowStart = np.ctypeslib.as_array(ctypes.cast(rs, ctypes.POINTER(ctypes.c_int)), (r + 1,)).copy()
        column = np.ctypeslib.as_array(ctypes.cast(col, ctypes.POINTER(ctypes.c_int)), (max(n, 1),))[:n].copy()
        values = np.ctypeslib.as_array(ctypes.cast(val, ctypes.POINTER(ctypes.c_double)), (max(n, 1),))[:n].copy()
        return r, cols.value, rowStart, column, values

    def scipy(self, name):
        import scipy.sparse as sp
        r, c, rs, col, val = self.matrix(name)
        return sp.csr_matrix((val, col, rs), shape=(r, c))

    def dirichlet(self, which):
        """which: 'vx', 'vy' or 'p' -> (ids, values)."""
        k = {"vx": 0, "vy": 1, "p": 2}[which]
        cnt = ctypes.c_int()
        call("nsfem_context_dirichlet", self._h, k, ctypes.byref(cnt), None, None)
        ids = np.empty(cnt.value, np.int32)
        vals = np.empty(cnt.value, np.float64)
        call("nsfem_context_dirichlet", self._h, k, ctypes.byref(cnt), _ptr(ids), _ptr(vals))
        return ids, vals

    def internal_pressure(self):
        ids = np.empty(self.numInternalPressure, np.int32)
        call("nsfem_context_internal_pressure", self._h, _ptr(ids))
        return ids

    def velocity_mesh(self):
        xy = np.empty((self.numVelocityNodes, 2), np.float64)
        el = np.empty((self.numElements, 6), np.int32)
        call("nsfem_context_velocity_mesh", self._h, _ptr(xy), _ptr(el))
        return xy, el

    def pressure_mesh(self):
        xy = np.empty((self.numPressureNodes, 2), np.float64)
        el = np.empty((self.numElements, 3), np.int32)
        call("nsfem_context_pressure_mesh", self._h, _ptr(xy), _ptr(el))
        return xy, el

    def fast_convection_map(self, threads=0):
        """The FastConvection integration map as a scipy CSR matrix (nnz_C x 2N)."""
        import scipy.sparse as sp
        rows, cols, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rs, col, val = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        call("nsfem_fast_convection_map", self._h, threads, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz),
             ctypes.byref(rs), ctypes.byref(col), ctypes.byref(val))
        try:
            r, n = rows.value, nnz.value
            a = np.ctypeslib.as_array(ctypes.cast(rs, ctypes.POINTER(ctypes.c_int)), (r + 1,)).copy()
            b = np.ctypeslib.as_array(ctypes.cast(col, ctypes.POINTER(ctypes.c_int)), (max(n, 1),))[:n].copy()
            v = np.ctypeslib.as_array(ctypes.cast(val, ctypes.POINTER(ctypes.c_double)), (max(n, 1),))[:n].copy()
        finally:
            for p in (rs, col, val):
                _lib.load().nsfem_host_free(p)
        return sp.csr_matrix((v, b, a), shape=(r, cols.value))

    def __del__(self):
        try:
            if self._h:
                call("nsfem_context_destroy", self._h)
                self._h = None
        except Exception:
            pass


@dataclasses.dataclass
class SolverConfig:
    method: str = "gs"
    maxIterations: int = 200
    targetMse: float = 1e-6
    mseCheckInterval: int = 1
    lanes: int = 0         # lanes per matrix row (1, 2, 4, 8, 16); 0 picks from the mean row length
    rowsPerBlock: int = 0  # rows per workgroup of the solver kernels; 0 is the default
    sweepLanes: int = 0    # the colour sweeps' own shape; 0, 0: as lanes, rowsPerBlock
    sweepRowsPerBlock: int = 0


@dataclasses.dataclass
class ChorinConfig:
    velocitySolver: SolverConfig = dataclasses.field(default_factory=SolverConfig)
    pressureSolver: SolverConfig = dataclasses.field(default_factory=SolverConfig)
    dtype: int = F64
    innerTol: float = 1e-4
    eagerLaunches: int = 0  # 1: no graph replay of solver iterations (for tracing tools)
    timeStepPolicy: int = 0  # TAU_REJECT: an unstable step is an error; TAU_CLAMP, TAU_UNCHECKED
    stepGraph: int = 0  # 1: each step is one captured graph with device-side convergence (pcg solvers)
    pressureStart: int = 0  # PRESSURE_START_LAST: warm start from p_n; PRESSURE_START_EXTRAPOLATE: 2 p_n - p_n-1
    velocityStart: int = 0  # VELOCITY_START_ZERO (reference), VELOCITY_START_LAST, VELOCITY_START_EXTRAPOLATE


class Chorin:
    """GPU Chorin projection stepping over a ChorinContext.

    instances > 1 (2, 4, 8 or 16, single-mesh context): an ensemble of
    independent flows on the one mesh sharing every matrix, instance b's
    inflow scaled by inflow_scale[b]; the state arrays then hold the
    instances one after another (velocity B x 2N, pressure B x V)."""

    def __init__(self, ctx: ChorinContext, cfg: ChorinConfig, instances=1, inflow_scale=None):
        self.ctx = ctx
        self.cfg = cfg
        self._m1 = cfg.velocitySolver.method.encode()
        self._m2 = cfg.pressureSolver.method.encode()
        v, p = cfg.velocitySolver, cfg.pressureSolver
        c = _lib.ChorinConfigC(_lib.SolverConfig(self._m1, v.maxIterations, v.targetMse, v.mseCheckInterval, v.lanes,
                                                 v.rowsPerBlock, v.sweepLanes, v.sweepRowsPerBlock),
                               _lib.SolverConfig(self._m2, p.maxIterations, p.targetMse, p.mseCheckInterval, p.lanes,
                                                 p.rowsPerBlock, p.sweepLanes, p.sweepRowsPerBlock),
                               cfg.dtype, cfg.innerTol, cfg.eagerLaunches, cfg.timeStepPolicy, cfg.stepGraph,
                               cfg.pressureStart, cfg.velocityStart)
        self._h = ctypes.c_void_p()
        self.instances = int(instances)
        if inflow_scale is None:
            scale = None
        else:
            scale = np.ascontiguousarray(inflow_scale, np.float64)
            if scale.size != self.instances:
                raise ValueError("inflow_scale needs one value per instance")
            scale = scale.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        call("nsfem_chorin_create_ensemble", ctx.handle, ctypes.byref(c), self.instances, scale, ctypes.byref(self._h))
        self.dtype = _NP[cfg.dtype]
        self.N = ctx.numVelocityNodes
        self.V = ctx.numPressureNodes

    def set_time_step(self, tau):
        """Under TAU_REJECT a step above the stability limit raises
        NsfemError (the previous step stays); TAU_CLAMP takes the limit."""
        call("nsfem_chorin_set_time_step", self._h, float(tau))

    def step_graph_info(self):
        """dict(builds, build_seconds, rollbacks, budgets) of the captured-step mode."""
        b, r = ctypes.c_int(), ctypes.c_int()
        sec = ctypes.c_double()
        bud = (ctypes.c_int * 3)()
        call("nsfem_chorin_step_graph_info", self._h, ctypes.byref(b), ctypes.byref(sec), ctypes.byref(r), bud)
        return {"builds": b.value, "build_seconds": sec.value, "rollbacks": r.value, "budgets": list(bud)}

    def set_step_graph_budgets(self, tentative, pressure, final):
        call("nsfem_chorin_set_step_graph_budgets", self._h, int(tentative), int(pressure), int(final))

    def time_step_info(self):
        """dict(tauRequested, tauUsed, lambdaNu = lambda_max(M^-1 nu K),
        tauMax = 1 / lambdaNu); runs the estimate on first use."""
        v = [ctypes.c_double() for _ in range(4)]
        call("nsfem_chorin_time_step_info", self._h, *[ctypes.byref(x) for x in v])
        return dict(zip(("tauRequested", "tauUsed", "lambdaNu", "tauMax"), (x.value for x in v)))

    def step(self):
        st = _lib.StepStats()
        call("nsfem_chorin_step", self._h, ctypes.byref(st))
        return st.as_dict()

    def download(self):
        u = np.empty(2 * self.N * self.instances, self.dtype)
        p = np.empty(self.V * self.instances, self.dtype)
        call("nsfem_chorin_download", self._h, _ptr(u), _ptr(p))
        return u, p

    def download_async(self):
        """Starts copying the state after the last step to host memory while
        later steps run; download_wait() returns it."""
        call("nsfem_chorin_download_async", self._h)

    def download_wait(self):
        u = np.empty(2 * self.N * self.instances, self.dtype)
        p = np.empty(self.V * self.instances, self.dtype)
        call("nsfem_chorin_download_wait", self._h, _ptr(u), _ptr(p))
        return u, p

    def upload_velocity(self, u):
        u = np.ascontiguousarray(u, self.dtype)
        if u.size != 2 * self.N * self.instances:
            raise ValueError("velocity must have 2N entries per instance")
        call("nsfem_chorin_upload_velocity", self._h, _ptr(u))

    def reset(self):
        call("nsfem_chorin_reset", self._h)

    def state_pointers(self):
        u, p = ctypes.c_void_p(), ctypes.c_void_p()
        call("nsfem_chorin_state", self._h, ctypes.byref(u), ctypes.byref(p))
        return u.value, p.value

    def time_kernel(self, which, reps=50):
        """(average ms per launch, algorithmic bytes per launch, launches per
        solver iteration -- 1 for the once-per-step kernels) of a step kernel."""
        ms, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        call("nsfem_chorin_time_kernel", self._h, which.encode(), int(reps), ctypes.byref(ms), ctypes.byref(b), ctypes.byref(n))
        return ms.value, b.value, n.value

    def shape(self):
        """(velocity lanes, velocity rows per workgroup, pressure lanes,
        pressure rows per workgroup, the same four for the colour sweeps, then
        velocity direction / update / rho rows, pressure direction / update /
        rho rows, the velocity and pressure direction kernels' lanes, then
        whether each solver fuses its reductions) of the solver kernels."""
        out = (ctypes.c_int * 18)()
        call("nsfem_chorin_shape", self._h, out)
        return tuple(out)

    def autotune(self, reps=10):
        """Autotunes both solvers' launch shapes (the state is kept); returns
        shape()."""
        out = (ctypes.c_int * 18)()
        call("nsfem_chorin_autotune", self._h, int(reps), out)
        return tuple(out)

    def set_solver_option(self, key, value, which="both"):
        """Solver.setOption for the step's solvers: which = "velocity",
        "pressure" or "both"."""
        w = {"velocity": 0, "pressure": 1, "both": 2}[which]
        call("nsfem_chorin_set_solver_option", self._h, w, key.encode(), float(value))

    def assemble_convection(self, u_host):
        """Returns (A, C) values in the velocity pattern for a host velocity."""
        from .cu import vec
        nnz = self.ctx.matrix("velocityStiffness")[3].size
        u = vec(np.ascontiguousarray(u_host, self.dtype))
        a = vec(nnz, self.dtype)
        c = vec(nnz, self.dtype)
        call("nsfem_chorin_assemble_convection", self._h, u.ptr, a.ptr, c.ptr)
        return a.download(), c.download()

    def __del__(self):
        try:
            if self._h:
                call("nsfem_chorin_destroy", self._h)
                self._h = None
        except Exception:
            pass


@dataclasses.dataclass
class TimeStepSolution:
    time: float
    velocity: np.ndarray
    pressure: np.ndarray


def solve_ns_chorin(mesh: TriangleMesh, cond: DfgConditions, tau0: float, max_t: float, cfg: ChorinConfig,
                    frame_step: int = 1, verbose: bool = False, autotune: int = 10):
    """solveNsChorinCuda: numTimeSteps = ceil(maxT / tau0), tau = maxT / numTimeSteps,
    numTimeSteps + 1 steps from u = Dirichlet values (under TAU_CLAMP tau0 is
    first lowered to the stability limit); every frame_step-th state
    is kept (the frames the reference renders, navierStokes.cpp:109), each
    copied to the host while the following steps run.  autotune > 0: the
    solvers' launch shapes are tuned first (that many replays per candidate)."""
    ctx = ChorinContext(mesh, cond)
    ch = Chorin(ctx, cfg)
    if autotune > 0:
        ch.autotune(autotune)
    if cfg.timeStepPolicy == TAU_CLAMP:
        # lowered to the stability limit first, so the run still ends at max_t
        tau0 = min(tau0, ch.time_step_info()["tauMax"])
    n_steps = int(math.ceil(max_t / tau0))
    tau = max_t / n_steps
    ch.set_time_step(tau)
    out = []
    stats = []
    pending = None  # time of the state being downloaded
    for it in range(n_steps + 1):
        t0 = time.perf_counter()
        st = ch.step()
        st["wall_ms"] = 1e3 * (time.perf_counter() - t0)
        stats.append(st)
        if it % frame_step == 0:
            if pending is not None:
                out.append(TimeStepSolution(pending, *ch.download_wait()))
            ch.download_async()
            pending = (it + 1) * tau
        if verbose:
            print(f"{it} / {n_steps}: {st['wall_ms']:.3f} ms  tent iters {st['itersTent']} pressure iters "
                  f"{st['itersPressure']} final iters {st['itersFinal']}", flush=True)
    if pending is not None:
        out.append(TimeStepSolution(pending, *ch.download_wait()))
    return out, stats


def estimate_lambda_max(ctx: ChorinContext, iters=40, seed=1, graphs=True):
    """Largest eigenvalue of M^-1 K (velocity mass / stiffness) by power
    iteration on the GPU; bounds the forward-Euler diffusion step,
    tau * nu * lambda_max <= 2."""
    from . import cu
    rng = np.random.default_rng(seed)
    K = cu.csrF(*ctx.matrix("velocityStiffness")[:2], *ctx.matrix("velocityStiffness")[2:])
    M = cu.csrF(*ctx.matrix("velocityMass")[:2], *ctx.matrix("velocityMass")[2:])
    solver = cu.createSolver("pcg", 1, ctx.scipy("velocityMass"), 500, 1e-10, 1)
    if not graphs:
        solver.setOption("graphs", 0)
    n = ctx.numVelocityNodes
    x = cu.vec(rng.standard_normal(n))
    y = cu.vec(n, np.float64)
    mx = cu.vec(n, np.float64)
    lam = 0.0
    for _ in range(iters):
        K.mul(x, solver.getRhs())
        solver.getSol().memsetZero()
        solver.solve()
        solver.getSol().copyTo(x)
        K.mul(x, y)
        M.mul(x, mx)
        lam = cu.dot(n, x, y) / cu.dot(n, x, mx)
        cu.scale(n, x, 1.0 / cu.nrm2(n, x))
    return lam
