"""ctypes declarations of the C ABI in include/opentm.h.

Loading fails loudly when libopentm.so is missing: there is no CPU
fallback for any engine call.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopentm.so")

OTM_OK = 0
OTM_E_ARG = -1
OTM_E_HIP = -2
OTM_E_NOCONV = -3
OTM_E_UNSUPPORTED = -4
OTM_E_NUMERIC = -5
OTM_E_NODEV = -6
OTM_E_STATE = -7
OTM_E_OOM = -8
OTM_HOST = 0
OTM_DEVICE = 1
OTM_ANY = 2
PROF_CLASSES = ["fine_gs", "fine_apply", "fine_resid", "coarse_gs", "coarse_resid", "transfer",
                "coarsest", "vector", "energy", "setup"]

# Every function the header declares, with (restype, argtypes).
_c = ctypes
_p = _c.c_void_p
_i = _c.c_int
_d = _c.c_double
_ip = _c.POINTER(_c.c_int)
_dp = _c.POINTER(_c.c_double)


class OtmConfig(ctypes.Structure):
    _fields_ = [
        ("reso", _c.c_int * 3),
        ("heat_ratio", _c.c_double * 2),
        ("fem_rel_thres", _c.c_double),
        ("max_cg_iter", _c.c_int),
        ("pre_sweeps", _c.c_int),
        ("post_sweeps", _c.c_int),
        ("min_coarse", _c.c_int),
        ("device", _c.c_int),
        ("verbose", _c.c_int),
        ("coarse_sweeps", _c.c_int),
        ("fp64_refine", _c.c_int),
        ("fp16_stencils", _c.c_int),
        ("exact_smoother", _c.c_int),
        ("graph_vcycle", _c.c_int),
        ("fine_tile", _c.c_int),
    ]


class OtmInitParams(ctypes.Structure):
    _fields_ = [
        ("initway", _c.c_int),
        ("seed", _c.c_int),
        ("volratio", _c.c_double),
        ("n_period", _c.c_int),
        ("sym", _c.c_int),
        ("coeffs", _c.c_void_p),
        ("ncoeffs", _c.c_int),
    ]


class OtmDensityParams(ctypes.Structure):
    _fields_ = [
        ("filter_kind", _c.c_int),
        ("filter_radius", _c.c_double),
        ("filter_periodic", _c.c_int),
        ("penal", _c.c_double),
    ]


class OtmLogParams(ctypes.Structure):
    _fields_ = [
        ("prefix", _c.c_char_p),
        ("logrho", _c.c_int),
        ("logc", _c.c_int),
        ("logsens", _c.c_int),
        ("logobj", _c.c_int),
    ]


class OtmCustomParams(ctypes.Structure):
    _fields_ = [
        ("base", OtmConfig),
        ("model", _c.c_int),
        ("initway", _c.c_int),
        ("seed", _c.c_int),
        ("max_iter", _c.c_int),
        ("finthres", _c.c_double),
        ("volratio", _c.c_double),
        ("target", _c.c_double * 6),
        ("sym", _c.c_int),
        ("log", OtmLogParams),
    ]


class OtmRequest(ctypes.Structure):
    _fields_ = [
        ("cfg", OtmConfig),
        ("kappa", _c.c_void_p),
        ("dC", _c.c_double * 9),
        ("want_sens", _c.c_int),
    ]


class OtmResult(ctypes.Structure):
    _fields_ = [
        ("status", _c.c_int),
        ("C", _c.c_double * 9),
        ("iters", _c.c_int * 3),
        ("relres", _c.c_double * 3),
        ("queue_ms", _c.c_double),
        ("latency_ms", _c.c_double),
    ]


class OtmServerCounters(ctypes.Structure):
    _fields_ = [
        ("completed", _c.c_longlong),
        ("steps", _c.c_longlong),
        ("contexts_created", _c.c_longlong),
        ("contexts_evicted", _c.c_longlong),
        ("active_peak", _c.c_int),
        ("queued", _c.c_int),
        ("resident_bytes", _c.c_double),
        ("budget_bytes", _c.c_double),
    ]


OTM_TUNE_MAX = 16


class OtmTuneReport(ctypes.Structure):
    _fields_ = [
        ("n", _c.c_int),
        ("fine_tile", _c.c_int * OTM_TUNE_MAX),
        ("coarse_sweeps", _c.c_int * OTM_TUNE_MAX),
        ("ms", _c.c_double * OTM_TUNE_MAX),
        ("iters", _c.c_int * OTM_TUNE_MAX),
        ("ok", _c.c_int * OTM_TUNE_MAX),
        ("chosen", _c.c_int),
    ]


class OtmCustomReport(ctypes.Structure):
    _fields_ = [
        ("iterations", _c.c_int),
        ("converged", _c.c_int),
        ("objective", _c.c_double),
        ("volume", _c.c_double),
        ("C", _c.c_double * 9),
        ("unconverged_solves", _c.c_int),
    ]


SIGNATURES = {
    "otm_config_init": (None, [_c.POINTER(OtmConfig)]),
    "otm_last_error": (_c.c_char_p, []),
    "otm_version": (_i, []),
    "otm_create": (_i, [_c.POINTER(OtmConfig), _c.POINTER(_p)]),
    "otm_destroy": (_i, [_p]),
    "otm_levels": (_i, [_p, _ip]),
    "otm_level_reso": (_i, [_p, _i, _ip]),
    "otm_update": (_i, [_p, _p, _i, _i]),
    "otm_heat_matrix": (_i, [_p, _dp]),
    "otm_solve_stats": (_i, [_p, _ip, _dp]),
    "otm_fine_tile_shape": (_i, [_i, _ip]),
    "otm_fine_tile": (_i, [_p, _ip]),
    "otm_autotune": (_i, [_p, _i, _c.POINTER(OtmTuneReport)]),
    "otm_sensitivity": (_i, [_p, _dp, _p, _i, _i]),
    "otm_element_energy": (_i, [_p, _p, _i]),
    "otm_get_chi": (_i, [_p, _p, _i]),
    "otm_set_chi": (_i, [_p, _p, _i]),
    "otm_reset_displacement": (_i, [_p]),
    "otm_apply_level": (_i, [_p, _i, _p, _p, _i]),
    "otm_apply_fine_pcg": (_i, [_p, _p, _p, _p, _i]),
    "otm_vcycle": (_i, [_p, _p, _p, _i]),
    "otm_get_stencil": (_i, [_p, _i, _p, _i]),
    "otm_get_kappa": (_i, [_p, _p, _i]),
    "otm_solve": (_i, [_p, _d, _i, _dp]),
    "otm_mg_cycles": (_i, [_p, _i, _dp]),
    "otm_transfer": (_i, [_p, _i, _i, _p, _p, _i]),
    "otm_smooth_level": (_i, [_p, _i, _p, _p, _i, _i, _i]),
    "otm_coarsest_solve": (_i, [_p, _p, _p, _i]),
    "otm_conv_filter": (_i, [_ip, _i, _d, _i, _i, _p, _p, _i]),
    "otm_conv_filter_noalloc": (_i, [_ip, _i, _d, _i, _i, _p, _p, _p]),
    "otm_density_params_init": (None, [_c.POINTER(OtmDensityParams)]),
    "otm_set_design": (_i, [_p, _p, _i]),
    "otm_get_design": (_i, [_p, _p, _i]),
    "otm_eval": (_i, [_p, _c.POINTER(OtmDensityParams), _dp]),
    "otm_apply_design": (_i, [_p, _c.POINTER(OtmDensityParams)]),
    "otm_backward": (_i, [_p, _dp, _p, _i]),
    "otm_oc_create": (_i, [_i, _d, _d, _d, _i, _c.POINTER(_p)]),
    "otm_oc_destroy": (_i, [_p]),
    "otm_oc_set_bisection": (_i, [_p, _i, _d]),
    "otm_oc_filter_sens": (_i, [_p, _p, _p, _i, _ip, _d, _i]),
    "otm_oc_update": (_i, [_p, _p, _p, _d, _i, _ip, _dp]),
    "otm_oc_step": (_i, [_p, _p, _d, _d, _ip, _dp]),
    "otm_mma_create": (_i, [_i, _i, _d, _d, _d, _d, _i, _c.POINTER(_p)]),
    "otm_mma_destroy": (_i, [_p]),
    "otm_mma_set_bounds": (_i, [_p, _d, _d]),
    "otm_mma_set_bounds_array": (_i, [_p, _p, _p, _i]),
    "otm_mma_set_move": (_i, [_p, _d]),
    "otm_mma_update": (_i, [_p, _i, _p, _p, _dp, _c.POINTER(_p), _i, _ip]),
    "otm_mma_profile_enable": (_i, [_p, _i]),
    "otm_mma_profile_read": (_i, [_p, _i, _c.POINTER(_c.c_longlong), _dp, _dp]),
    "otm_init_density": (_i, [_p, _i, _i, _d]),
    "otm_init_params_init": (None, [_c.POINTER(OtmInitParams)]),
    "otm_init_design": (_i, [_p, _c.POINTER(OtmInitParams), _p]),
    "otm_symmetrize": (_i, [_c.POINTER(_i), _i, _p, _i]),
    "otm_write_density": (_i, [_c.c_char_p, _c.POINTER(_i), _p, _i]),
    "otm_read_density": (_i, [_c.c_char_p, _c.POINTER(_i), _p, _i]),
    "otm_symmetrize_design": (_i, [_p, _i]),
    "otm_design_stats": (_i, [_p, _dp, _dp]),
    "otm_custom_params_init": (None, [_c.POINTER(OtmCustomParams)]),
    "otm_log_params_init": (None, [_c.POINTER(OtmLogParams)]),
    "otm_log_iteration": (_i, [_p, _c.POINTER(OtmLogParams), _i, _dp, _d, _p, _i]),
    "otm_field_reduce": (_i, [_p, _c.c_longlong, _i, _d, _i, _dp, _p, _d]),
    "otm_run_custom": (_i, [_c.POINTER(OtmCustomParams), _p, _p, _c.POINTER(OtmCustomReport)]),
    "otm_profile_enable": (_i, [_p, _i]),
    "otm_profile_read": (_i, [_p, _i, _c.POINTER(_c.c_longlong), _dp, _dp]),
    "otm_device_alloc": (_i, [_i, _c.c_size_t, _c.POINTER(_p)]),
    "otm_device_free": (_i, [_p]),
    "otm_memcpy": (_i, [_p, _p, _c.c_size_t, _i, _i]),
    "otm_synchronize": (_i, [_p]),
    "otm_batch_create": (_i, [_i, _c.POINTER(_p)]),
    "otm_batch_destroy": (_i, [_p]),
    "otm_batch_update": (_i, [_p, _c.POINTER(_p), _i, _c.POINTER(_p), _i, _i]),
    "otm_batch_heat_matrix": (_i, [_p, _c.POINTER(_p), _i, _c.POINTER(_c.c_double), _c.POINTER(_i)]),
    "otm_batch_sensitivity": (_i, [_p, _c.POINTER(_p), _i, _c.POINTER(_c.c_double), _c.POINTER(_p), _i, _i]),
    "otm_batch_stats": (_i, [_p, _c.POINTER(_c.c_longlong), _c.POINTER(_i)]),
    "otm_request_init": (None, [_c.POINTER(OtmRequest)]),
    "otm_server_create": (_i, [_i, _c.c_double, _i, _c.POINTER(_p)]),
    "otm_server_destroy": (_i, [_p]),
    "otm_server_submit": (_i, [_p, _c.POINTER(OtmRequest), _c.POINTER(_c.c_longlong)]),
    "otm_server_wait": (_i, [_p, _c.c_longlong, _c.POINTER(OtmResult), _p]),
    "otm_server_stats": (_i, [_p, _c.POINTER(OtmServerCounters)]),
}

_lib = None


class OtmError(RuntimeError):
    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed with status {code}: {msg}")
        self.code = code


# Engine handles still open, closed in order by an exit hook while the HIP
# runtime is alive; after it ran, finalisers make no engine calls (the
# runtime's own teardown may already have started).
_live = []
_closed_at_exit = False


def track(obj):
    import weakref
    _live.append(weakref.ref(obj))


def shutting_down():
    return _closed_at_exit


def _close_all():
    global _closed_at_exit
    # servers and batches hold contexts: close them first
    objs = [r() for r in _live]
    objs = [o for o in objs if o is not None]
    order = {"Server": 0, "Batch": 1, "OCOptimizer": 2, "MMAOptimizer": 3, "Homogenization_H": 4,
             "DeviceBuffer": 5}
    objs.sort(key=lambda o: order.get(type(o).__name__, 6))
    for o in objs:
        try:
            o.close()
        except Exception:
            pass
    _live.clear()
    _closed_at_exit = True


import atexit as _atexit  # noqa: E402

_atexit.register(_close_all)


def load():
    """Load libopentm.so (built by opentm_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP engine first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status, func):
    if status != OTM_OK:
        msg = load().otm_last_error()
        raise OtmError(status, func, msg.decode() if msg else "")
