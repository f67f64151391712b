"""ctypes binding of libcostuum_amd.so (the C ABI in include/costuum_amd.h).

The library must be built (costuum_amd.build) before use; there is no
fallback implementation: a missing library raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcostuum_amd.so")

CTM_OK = 0
CTM_ERR_ARGUMENT = -1
CTM_ERR_HIP = -2
CTM_ERR_STATE = -3
CTM_ERR_MEMORY = -4
CTM_CONVERGED = 0
CTM_NOT_CONVERGED = 1
CTM_PRECISION_FP64 = 0
CTM_PRECISION_MIXED = 1

# Every entry point declared in include/costuum_amd.h.
EXPORTS = (
    "ctm_create", "ctm_destroy", "ctm_last_error", "ctm_stream",
    "ctm_synchronize", "ctm_tmatrix_batch", "ctm_tmatrix_fixed",
    "ctm_tmatrix_size", "ctm_tmatrix_get", "ctm_batch_orders",
    "ctm_profile_enable", "ctm_profile_reset", "ctm_profile_read",
    "ctm_observables", "ctm_orientation_average", "ctm_average_get",
    "ctm_query", "ctm_m0_sums", "ctm_gauss_legendre", "ctm_rv_over_rs",
    "ctm_adaptive_quadrature", "ctm_orientation_coefficients",
    "ctm_weighted_reduce", "ctm_tmatrix_tangents", "ctm_tangent_get",
    "ctm_observables_select", "ctm_vjp", "ctm_observables_noalloc",
    "ctm_tmatrix_set", "ctm_grain_memory_estimate", "ctm_precision_stats",
    "ctm_observables_derivative", "ctm_plan_create", "ctm_plan_run",
    "ctm_plan_info_get", "ctm_plan_destroy", "ctm_plan_tune",
    "ctm_plan_tuning_set",
)

# double f(double x, void *arg): the host callback type (ctm_real_fn)
REAL_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_void_p)


class CtmQuery(ctypes.Structure):
    _fields_ = [("grain", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("theta_in", ctypes.c_double), ("phi_in", ctypes.c_double),
                ("theta_out", ctypes.c_double), ("phi_out", ctypes.c_double)]


QUERY_OUT = 40

KERNELS = ("gauss_legendre", "wigner", "tables", "assemble", "solve",
           "converge", "sums", "obs_x", "obs_sum", "average",
           "assemble_small", "assemble_wide", "assemble_pack", "refine",
           "assemble_window")


class CtmObservableSpec(ctypes.Structure):
    _fields_ = [("n_theta", ctypes.c_int32),
                ("theta", ctypes.c_void_p),
                ("do_extinction", ctypes.c_int32),
                ("do_absorption", ctypes.c_int32),
                ("account_for_scattering", ctypes.c_int32),
                ("ngauss", ctypes.c_int32),
                ("n_scatter_theta", ctypes.c_int32),
                ("scatter_gl", ctypes.c_int32)]


class CtmKernelStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64),
                ("milliseconds", ctypes.c_double),
                ("flops", ctypes.c_double),
                ("bytes", ctypes.c_double)]


class CtmConvergence(ctypes.Structure):
    _fields_ = [("min_order", ctypes.c_uint32),
                ("max_order", ctypes.c_uint32),
                ("glfac", ctypes.c_uint32),
                ("precision", ctypes.c_uint32),
                ("tolerance", ctypes.c_double)]


class CtmPlanOutputs(ctypes.Structure):
    _fields_ = [("nmax", ctypes.c_void_p), ("status", ctypes.c_void_p),
                ("qsca", ctypes.c_void_p), ("qext", ctypes.c_void_p),
                ("obs_qext", ctypes.c_void_p),
                ("obs_qextpol", ctypes.c_void_p),
                ("obs_qextcpol", ctypes.c_void_p),
                ("obs_qabs", ctypes.c_void_p),
                ("obs_qabspol", ctypes.c_void_p)]


class CtmPlanInfo(ctypes.Structure):
    _fields_ = [("runs", ctypes.c_int64), ("reruns", ctypes.c_int64),
                ("mixed_blocks", ctypes.c_int64),
                ("mixed_fallback", ctypes.c_int64),
                ("pool_bytes", ctypes.c_int64),
                ("allm_items", ctypes.c_int32),
                ("window_items", ctypes.c_int32 * 3)]


class CtmPlanTuning(ctypes.Structure):
    _fields_ = [("window_scale", ctypes.c_double),
                ("window_add", ctypes.c_double),
                ("micro_ratio", ctypes.c_double),
                ("ms_default", ctypes.c_double),
                ("ms_chosen", ctypes.c_double),
                ("trials", ctypes.c_int32),
                ("runs_per_trial", ctypes.c_int32),
                ("big_first", ctypes.c_int32),
                ("obs_lean", ctypes.c_int32)]


_lib = None


def lib():
    """Loads the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libcostuum_amd.so is not built; run python -m costuum_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.c_void_p  # device or host double pointers
    L.ctm_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    L.ctm_create.restype = ctypes.c_int
    L.ctm_destroy.argtypes = [vp]
    L.ctm_destroy.restype = ctypes.c_int
    L.ctm_last_error.argtypes = [vp]
    L.ctm_last_error.restype = ctypes.c_char_p
    L.ctm_stream.argtypes = [vp]
    L.ctm_stream.restype = vp
    L.ctm_synchronize.argtypes = [vp]
    L.ctm_synchronize.restype = ctypes.c_int
    L.ctm_tmatrix_batch.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp, dp,
                                    ctypes.POINTER(CtmConvergence), dp, dp,
                                    dp, dp]
    L.ctm_tmatrix_batch.restype = ctypes.c_int
    L.ctm_tmatrix_fixed.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp, dp,
                                    dp, dp, ctypes.c_uint32, dp, dp]
    L.ctm_tmatrix_fixed.restype = ctypes.c_int
    L.ctm_tmatrix_size.argtypes = [vp, ctypes.c_int]
    L.ctm_tmatrix_size.restype = ctypes.c_size_t
    L.ctm_tmatrix_get.argtypes = [vp, ctypes.c_int, dp, ctypes.c_size_t]
    L.ctm_tmatrix_get.restype = ctypes.c_int
    L.ctm_batch_orders.argtypes = [vp, dp, dp, dp]
    L.ctm_batch_orders.restype = ctypes.c_int
    L.ctm_profile_enable.argtypes = [vp, ctypes.c_int]
    L.ctm_profile_enable.restype = ctypes.c_int
    L.ctm_profile_reset.argtypes = [vp]
    L.ctm_profile_reset.restype = ctypes.c_int
    L.ctm_profile_read.argtypes = [vp, ctypes.c_int,
                                   ctypes.POINTER(CtmKernelStats)]
    L.ctm_profile_read.restype = ctypes.c_int
    L.ctm_observables.argtypes = [vp, ctypes.POINTER(CtmObservableSpec),
                                  dp, dp, dp, dp, dp, dp]
    L.ctm_observables.restype = ctypes.c_int
    L.ctm_observables_noalloc.argtypes = [vp,
                                          ctypes.POINTER(CtmObservableSpec),
                                          dp, dp, dp, dp, dp, dp]
    L.ctm_observables_noalloc.restype = ctypes.c_int
    L.ctm_orientation_average.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                          dp, dp]
    L.ctm_orientation_average.restype = ctypes.c_int
    L.ctm_average_get.argtypes = [vp, ctypes.c_int, dp, ctypes.c_size_t]
    L.ctm_tmatrix_set.argtypes = [vp, ctypes.c_int, dp, ctypes.c_size_t]
    L.ctm_tmatrix_set.restype = ctypes.c_int
    L.ctm_grain_memory_estimate.argtypes = [ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_int32]
    L.ctm_grain_memory_estimate.restype = ctypes.c_double
    L.ctm_precision_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64)]
    L.ctm_precision_stats.restype = ctypes.c_int
    L.ctm_observables_derivative.argtypes = [
        vp, ctypes.c_int, ctypes.POINTER(CtmObservableSpec), dp, dp, dp, dp,
        dp, dp]
    L.ctm_observables_derivative.restype = ctypes.c_int
    L.ctm_average_get.restype = ctypes.c_int
    L.ctm_query.argtypes = [vp, ctypes.c_int, ctypes.POINTER(CtmQuery), dp]
    L.ctm_query.restype = ctypes.c_int
    L.ctm_m0_sums.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp, dp, dp, dp,
                              dp, dp]
    L.ctm_m0_sums.restype = ctypes.c_int
    u32 = ctypes.c_uint32
    L.ctm_tmatrix_tangents.argtypes = [vp, dp, dp]
    L.ctm_tmatrix_tangents.restype = ctypes.c_int
    L.ctm_tangent_get.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp,
                                  ctypes.c_size_t]
    L.ctm_tangent_get.restype = ctypes.c_int
    L.ctm_observables_select.argtypes = [vp, ctypes.c_int]
    L.ctm_observables_select.restype = ctypes.c_int
    L.ctm_vjp.argtypes = [vp, dp, dp, dp]
    L.ctm_vjp.restype = ctypes.c_int
    L.ctm_weighted_reduce.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int64, dp, dp, ctypes.c_int,
                                      dp]
    L.ctm_weighted_reduce.restype = ctypes.c_int
    L.ctm_gauss_legendre.argtypes = [u32, ctypes.c_double, ctypes.c_double,
                                     dp, dp]
    L.ctm_gauss_legendre.restype = ctypes.c_int
    L.ctm_rv_over_rs.argtypes = [ctypes.c_double]
    L.ctm_rv_over_rs.restype = ctypes.c_double
    L.ctm_adaptive_quadrature.argtypes = [REAL_FN, vp, ctypes.c_double,
                                          ctypes.c_double, u32, u32,
                                          ctypes.c_double, ctypes.c_double,
                                          dp]
    L.ctm_adaptive_quadrature.restype = ctypes.c_int
    L.ctm_orientation_coefficients.argtypes = [REAL_FN, vp, u32,
                                               ctypes.c_double,
                                               ctypes.c_double, dp]
    L.ctm_orientation_coefficients.restype = ctypes.c_int
    L.ctm_plan_create.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp, dp,
                                  ctypes.POINTER(CtmConvergence),
                                  ctypes.POINTER(CtmObservableSpec),
                                  ctypes.POINTER(CtmPlanOutputs),
                                  ctypes.POINTER(vp)]
    L.ctm_plan_create.restype = ctypes.c_int
    L.ctm_plan_run.argtypes = [vp]
    L.ctm_plan_run.restype = ctypes.c_int
    L.ctm_plan_info_get.argtypes = [vp, ctypes.POINTER(CtmPlanInfo)]
    L.ctm_plan_info_get.restype = ctypes.c_int
    L.ctm_plan_tune.argtypes = [vp, ctypes.c_int,
                                ctypes.POINTER(CtmPlanTuning)]
    L.ctm_plan_tune.restype = ctypes.c_int
    L.ctm_plan_tuning_set.argtypes = [vp, ctypes.POINTER(CtmPlanTuning)]
    L.ctm_plan_tuning_set.restype = ctypes.c_int
    L.ctm_plan_destroy.argtypes = [vp]
    L.ctm_plan_destroy.restype = ctypes.c_int
    _lib = L
    return L
