"""ctypes binding of libnsfem_hip.so (declarations in include/nsfem_hip.h).

The library is loaded from the package directory only; a missing library is
an error (there is no fallback implementation)."""

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnsfem_hip.so")

NSFEM_OK = 0
NSFEM_EINVAL = 1
NSFEM_ERUNTIME = 2
NSFEM_EDEVICE = 3
NSFEM_ENONFINITE = 4

F32 = 0
F64 = 1
TAU_REJECT = 0     # time-step policies (nsfem_chorin_config.timeStepPolicy)
TAU_CLAMP = 1
PRESSURE_START_LAST = 0
PRESSURE_START_EXTRAPOLATE = 1
VELOCITY_START_ZERO = 0
VELOCITY_START_LAST = 1
VELOCITY_START_EXTRAPOLATE = 2
TAU_UNCHECKED = 2
NSFEM_CONTEXT_GPU_ASSEMBLY = 1

MAT_VELOCITY_MASS = 0
MAT_VELOCITY_STIFFNESS = 1
MAT_PRESSURE_STIFFNESS = 2
MAT_PRESSURE_STIFFNESS_INTERNAL = 3
MAT_VELOCITY_PRESSURE_DIV = 4
MAT_PRESSURE_VELOCITY_DIV = 5


class NsfemError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class NsfemInvalidArgument(NsfemError, ValueError):
    pass


class NsfemNonFinite(NsfemError):
    pass


class NsfemDeviceError(NsfemError):
    pass


_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class DfgConditions(ctypes.Structure):
    _fields_ = [("viscosity", ctypes.c_double), ("peakVelocity", ctypes.c_double),
                ("channelHeight", ctypes.c_double), ("profileScale", ctypes.c_double)]


class SolverConfig(ctypes.Structure):
    _fields_ = [("method", ctypes.c_char_p), ("maxIterations", ctypes.c_int),
                ("targetMse", ctypes.c_double), ("mseCheckInterval", ctypes.c_int),
                ("lanes", ctypes.c_int), ("rowsPerBlock", ctypes.c_int),
                ("sweepLanes", ctypes.c_int), ("sweepRowsPerBlock", ctypes.c_int)]


class ChorinConfigC(ctypes.Structure):
    _fields_ = [("velocitySolver", SolverConfig), ("pressureSolver", SolverConfig), ("dtype", ctypes.c_int),
                ("innerTol", ctypes.c_double), ("eagerLaunches", ctypes.c_int), ("timeStepPolicy", ctypes.c_int),
                ("stepGraph", ctypes.c_int), ("pressureStart", ctypes.c_int), ("velocityStart", ctypes.c_int)]


class StepStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "mseTentX", "mseTentY", "msePressure", "mseFinalX", "mseFinalY", "relTent", "relPressure", "relFinal",
        "msConvection", "msTentative", "msPressure", "msFinal", "msTotal")] + \
        [(n, ctypes.c_int) for n in ("itersTent", "itersPressure", "itersFinal", "outerTent", "outerPressure", "outerFinal")] + \
        [(n, ctypes.c_double) for n in ("tauRequested", "tauUsed")] + [("capturedStep", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# name -> argtypes (restype is int unless listed in _RESTYPES)
_SIGNATURES = {
    "nsfem_last_error": [],
    "nsfem_version": [_c_int_p, _c_int_p],
    "nsfem_synchronize": [],
    "nsfem_copy_bandwidth": [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
    "nsfem_malloc": [ctypes.c_size_t, ctypes.POINTER(_vp)],
    "nsfem_free": [_vp],
    "nsfem_upload": [_vp, _vp, ctypes.c_size_t],
    "nsfem_download": [_vp, _vp, ctypes.c_size_t],
    "nsfem_copy": [_vp, _vp, ctypes.c_size_t],
    "nsfem_memset_zero": [_vp, ctypes.c_size_t],
    "nsfem_csr_create": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.POINTER(_vp)],
    "nsfem_csr_destroy": [_vp],
    "nsfem_csr_values": [_vp, ctypes.POINTER(_vp)],
    "nsfem_csr_shape": [_vp, _c_int_p, _c_int_p, _c_int_p, _c_int_p],
    "nsfem_spmv": [_vp, _vp, _vp],
    "nsfem_spmm": [_vp, ctypes.c_int, _vp, _vp],
    "nsfem_csr_mul": [_vp, ctypes.c_int, _vp, ctypes.c_long, _vp, ctypes.c_long, ctypes.c_double, ctypes.c_double, ctypes.c_int],
    "nsfem_axpy": [ctypes.c_int, ctypes.c_long, ctypes.c_double, _vp, _vp],
    "nsfem_scale": [ctypes.c_int, ctypes.c_long, ctypes.c_double, _vp],
    "nsfem_geam": [ctypes.c_int, ctypes.c_long, ctypes.c_double, _vp, ctypes.c_double, _vp, _vp],
    "nsfem_dot": [ctypes.c_int, ctypes.c_long, _vp, _vp, _c_double_p],
    "nsfem_dot_noalloc": [ctypes.c_int, ctypes.c_long, _vp, _vp, _c_double_p],
    "nsfem_nrm2": [ctypes.c_int, ctypes.c_long, _vp, _c_double_p],
    "nsfem_scatter": [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp],
    "nsfem_gather": [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp],
    "nsfem_convert": [ctypes.c_int, ctypes.c_int, ctypes.c_long, _vp, _vp],
    "nsfem_solver_create": [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                            ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)],
    "nsfem_solver_create_batch": [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)],
    "nsfem_solver_destroy": [_vp],
    "nsfem_solver_num_ch": [_vp, _c_int_p],
    "nsfem_solver_rows": [_vp, _c_int_p],
    "nsfem_solver_io_dtype": [_vp, _c_int_p],
    "nsfem_solver_rhs": [_vp, ctypes.POINTER(_vp)],
    "nsfem_solver_sol": [_vp, ctypes.POINTER(_vp)],
    "nsfem_solver_solve": [_vp],
    "nsfem_solver_last_mse": [_vp, ctypes.c_int, _c_double_p],
    "nsfem_solver_last_relres": [_vp, ctypes.c_int, ctypes.c_int, _c_double_p],
    "nsfem_solver_last_iterations": [_vp, _c_int_p],
    "nsfem_solver_last_iterations_system": [_vp, ctypes.c_int, _c_int_p],
    "nsfem_solver_last_outer_iterations": [_vp, _c_int_p],
    "nsfem_solver_set_max_iters": [_vp, ctypes.c_int],
    "nsfem_solver_set_target_mse": [_vp, ctypes.c_double],
    "nsfem_solver_set_mse_check_interval": [_vp, ctypes.c_int],
    "nsfem_solver_set_option": [_vp, ctypes.c_char_p, ctypes.c_double],
    "nsfem_solver_time_kernel": [_vp, ctypes.c_char_p, ctypes.c_int, _c_double_p, _c_double_p, _c_double_p],
    "nsfem_solver_num_colors": [_vp, ctypes.c_int, _c_int_p],
    "nsfem_solver_reshape": [_vp, ctypes.c_int, ctypes.c_int],
    "nsfem_solver_shape": [_vp, _c_int_p, _c_int_p],
    "nsfem_solver_reshape_sweeps": [_vp, ctypes.c_int, ctypes.c_int],
    "nsfem_solver_sweep_shape": [_vp, _c_int_p, _c_int_p],
    "nsfem_solver_kind_rows": [_vp, _c_int_p, _c_int_p, _c_int_p],
    "nsfem_solver_autotune": [_vp, ctypes.c_int, _c_int_p, _c_int_p],
    "nsfem_solver_coloring": [_vp, ctypes.c_int, _vp, _vp],
    "nsfem_mcgs_coloring": [ctypes.c_int, _vp, _vp, _vp, _vp, _c_int_p],
    "nsfem_mesh_load_gmsh": [ctypes.c_char_p, ctypes.POINTER(_vp)],
    "nsfem_mesh_from_arrays": [ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.POINTER(_vp)],
    "nsfem_mesh_destroy": [_vp],
    "nsfem_mesh_info": [_vp, _c_int_p, _c_int_p, _c_int_p, _c_int_p],
    "nsfem_mesh_arrays": [_vp, _vp, _vp, _vp],
    "nsfem_mesh_group_name": [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int],
    "nsfem_context_build": [_vp, ctypes.POINTER(DfgConditions), ctypes.c_int, ctypes.POINTER(_vp)],
    "nsfem_context_build_ex": [_vp, ctypes.POINTER(DfgConditions), ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)],
    "nsfem_context_build_batch": [_vp, ctypes.POINTER(DfgConditions), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)],
    "nsfem_context_num_systems": [_vp, _c_int_p],
    "nsfem_context_system_offsets": [_vp, _vp, _vp],
    "nsfem_context_destroy": [_vp],
    "nsfem_context_sizes": [_vp, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p],
    "nsfem_context_matrix": [_vp, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                             ctypes.POINTER(_vp)],
    "nsfem_context_dirichlet": [_vp, ctypes.c_int, _c_int_p, _vp, _vp],
    "nsfem_context_internal_pressure": [_vp, _vp],
    "nsfem_context_velocity_mesh": [_vp, _vp, _vp],
    "nsfem_context_pressure_mesh": [_vp, _vp, _vp],
    "nsfem_fast_convection_map": [_vp, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                  ctypes.POINTER(_vp)],
    "nsfem_host_free": [_vp],
    "nsfem_chorin_create": [_vp, ctypes.POINTER(ChorinConfigC), ctypes.POINTER(_vp)],
    "nsfem_chorin_create_ensemble": [_vp, ctypes.POINTER(ChorinConfigC), ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_vp)],
    "nsfem_chorin_instances": [_vp, ctypes.POINTER(ctypes.c_int)],
    "nsfem_chorin_destroy": [_vp],
    "nsfem_chorin_set_time_step": [_vp, ctypes.c_double],
    "nsfem_chorin_time_step_info": [_vp, _c_double_p, _c_double_p, _c_double_p, _c_double_p],
    "nsfem_chorin_set_step_graph_budgets": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "nsfem_chorin_step_graph_info": [_vp, ctypes.POINTER(ctypes.c_int), _c_double_p, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)],
    "nsfem_chorin_step": [_vp, ctypes.POINTER(StepStats)],
    "nsfem_chorin_state": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    "nsfem_chorin_download": [_vp, _vp, _vp],
    "nsfem_chorin_upload_velocity": [_vp, _vp],
    "nsfem_chorin_reset": [_vp],
    "nsfem_chorin_assemble_convection": [_vp, _vp, _vp, _vp],
    "nsfem_chorin_shape": [_vp, _vp],
    "nsfem_chorin_autotune": [_vp, ctypes.c_int, _vp],
    "nsfem_chorin_download_async": [_vp],
    "nsfem_chorin_download_wait": [_vp, _vp, _vp],
    "nsfem_chorin_set_solver_option": [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_double],
    "nsfem_chorin_time_kernel": [_vp, ctypes.c_char_p, ctypes.c_int, _c_double_p, _c_double_p, _c_double_p],
}
_RESTYPES = {"nsfem_last_error": ctypes.c_char_p, "nsfem_host_free": None}

_lock = threading.Lock()
_lib = None


def library_path():
    return LIB_PATH


def load():
    """Loads the HIP library (raises if it is not built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m nsfem_amd.build` "
                              "(the HIP path has no fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = lib
        return lib


def exported_symbols():
    return sorted(_SIGNATURES)


def check(rc):
    if rc == NSFEM_OK:
        return
    msg = load().nsfem_last_error().decode(errors="replace")
    cls = {NSFEM_EINVAL: NsfemInvalidArgument, NSFEM_ENONFINITE: NsfemNonFinite,
           NSFEM_EDEVICE: NsfemDeviceError}.get(rc, NsfemError)
    raise cls(rc, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))
