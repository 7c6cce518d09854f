"""ctypes bindings of libimfcs_hip.so (include/imfcs.h) and libimfcs_sim.so.

Loading fails loudly when a library has not been built: there is no
Python or CPU fallback for any kernel.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_HERE, "libimfcs_hip.so")
SIM_LIB_PATH = os.path.join(_HERE, "libimfcs_sim.so")

MODEL_GAUSS_2D, MODEL_ACF_1D, MODEL_LINEAR_1D = 1, 2, 3
ESTIMATOR_LSE, ESTIMATOR_MLE = 0, 1
STATE_CONVERGED, STATE_MAX_ITERATION, STATE_SINGULAR_HESSIAN = 0, 1, 2

_i32, _f64, _f32 = ctypes.c_int32, ctypes.c_double, ctypes.c_float


class ACFParameters(ctypes.Structure):
    """imfcs_acf_params == GpufitImFCS.ACFParameters (GpufitImFCS.java:453-486)."""
    _fields_ = [("width", _i32), ("height", _i32), ("win_star", _i32), ("hin_star", _i32),
                ("w_temp", _i32), ("h_temp", _i32), ("pixbinX", _i32), ("pixbinY", _i32),
                ("binningX", _i32), ("binningY", _i32), ("firstframe", _i32), ("lastframe", _i32),
                ("framediff", _i32), ("cfXDistance", _i32), ("cfYDistance", _i32),
                ("correlatorp", _f64), ("correlatorq", _f64), ("frametime", _f64), ("background", _i32),
                ("mtab1", _f64), ("mtabchanumminus1", _f64), ("sampchanumminus1", _f64), ("chanum", _i32),
                ("isNBcalculation", _i32), ("bleachcorr_gpu", _i32), ("bleachcorr_order", _i32),
                ("nopit", _i32), ("ave", _i32)]


class FitModelGeom(ctypes.Structure):
    """imfcs_fit_model."""
    _fields_ = [("ax", _f64), ("ay", _f64), ("s", _f64), ("sz", _f64), ("rx", _f64), ("ry", _f64),
                ("fitobsvol", _f64), ("q2", _f64), ("q3", _f64), ("paramfit", _i32 * 11),
                ("paraminit", _f64 * 11)]


class PipelineConfig(ctypes.Structure):
    """imfcs_config."""
    _fields_ = [("width", _i32), ("height", _i32), ("frames", _i32), ("firstframe", _i32),
                ("lastframe", _i32), ("binning_x", _i32), ("binning_y", _i32), ("pixbin_x", _i32),
                ("pixbin_y", _i32), ("cf_x", _i32), ("cf_y", _i32), ("background", _i32),
                ("correlator_p", _i32), ("correlator_q", _i32), ("frametime", _f64), ("poly_order", _i32),
                ("roi_x", _i32), ("roi_y", _i32), ("roi_w", _i32), ("roi_h", _i32), ("do_acf", _i32),
                ("do_ccf", _i32), ("do_fit", _i32), ("fit_max_iterations", _i32), ("fit_tolerance", _f32),
                ("fit_chi2_fp64", _i32), ("gls", _i32), ("acf_model", FitModelGeom), ("ccf_model", FitModelGeom),
                ("bleach_mode", _i32), ("sliding_window_length", _i32)]

class SimDomain(ctypes.Structure):
    """imfcs_sim_domain: one periodic box of the GPU simulator."""
    _fields_ = [("x0_um", _f64), ("y0_um", _f64), ("w_um", _f64), ("h_um", _f64), ("n_emitters", _i32),
                ("D_um2s", _f64 * 4), ("D2_um2s", _f64), ("frac2", _f64), ("counts_lo", _f64),
                ("counts_hi", _f64)]


BLEACH_POLYNOMIAL, BLEACH_SLIDING_WINDOW, BLEACH_LIN_SEGMENT, BLEACH_SINGLE_EXP, BLEACH_DOUBLE_EXP = 0, 1, 2, 3, 4


# every C-ABI entry point of include/imfcs.h with its ctypes signature
_P, _VP, _SZ = ctypes.POINTER, ctypes.c_void_p, ctypes.c_size_t
SIGNATURES = {
    "imfcs_fit": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _P(_f32), _P(_f32), ctypes.c_int, _P(_f32), _f32,
                                 ctypes.c_int, ctypes.c_int, _P(ctypes.c_int), ctypes.c_int, _SZ, _VP, _P(_f32),
                                 _P(ctypes.c_int), _P(_f32), _P(ctypes.c_int)]),
    "imfcs_get_last_error": (ctypes.c_char_p, []),
    "imfcs_device_available": (ctypes.c_int, []),
    "imfcs_get_versions": (ctypes.c_int, [_P(ctypes.c_int), _P(ctypes.c_int)]),
    "imfcs_reset_gpu": (ctypes.c_int, []),
    "imfcs_reset_device": (ctypes.c_int, []),
    "imfcs_calc_data_bleach_correction": (ctypes.c_int, [_P(_f32), _P(_f32), _P(ACFParameters)]),
    "imfcs_is_binning_memory_sufficient": (ctypes.c_int, [_P(ACFParameters)]),
    "imfcs_calc_binning": (ctypes.c_int, [_P(_f32), _P(_f32), _P(ACFParameters)]),
    "imfcs_is_acf_memory_sufficient": (ctypes.c_int, [_P(ACFParameters)]),
    "imfcs_calc_acf": (ctypes.c_int, [_P(_f32), _P(_f64), _P(_f64), _P(_f64), _P(_f64), _P(_f64), _P(_f64),
                                      _P(_f64), _P(ctypes.c_int), _P(ACFParameters)]),
    "imfcs_plan": (ctypes.c_int, [_P(PipelineConfig), _P(ctypes.c_int), _P(ctypes.c_int), _P(ctypes.c_int)]),
    "imfcs_create": (_VP, [ctypes.c_int]),
    "imfcs_destroy": (None, [_VP]),
    "imfcs_device_alloc": (_VP, [_VP, _SZ]),
    "imfcs_device_free": (ctypes.c_int, [_VP, _VP]),
    "imfcs_host_alloc": (_VP, [_VP, _SZ]),
    "imfcs_host_free": (ctypes.c_int, [_VP, _VP]),
    "imfcs_memcpy_h2d": (ctypes.c_int, [_VP, _VP, _VP, _SZ]),
    "imfcs_memcpy_d2h": (ctypes.c_int, [_VP, _VP, _VP, _SZ]),
    "imfcs_memcpy_d2d": (ctypes.c_int, [_VP, _VP, _VP, _SZ]),
    "imfcs_memcpy_d2d_words": (ctypes.c_int, [_VP, _VP, _VP, _SZ]),
    "imfcs_synchronize": (ctypes.c_int, [_VP]),
    "imfcs_run": (ctypes.c_int, [_VP, _VP, _P(PipelineConfig)]),
    "imfcs_run_noalloc": (ctypes.c_int, [_VP, _VP, _P(PipelineConfig), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "imfcs_batch_create": (_VP, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "imfcs_batch_destroy": (None, [_VP]),
    "imfcs_batch_run": (ctypes.c_int, [_VP, _P(_VP), _P(PipelineConfig)]),
    "imfcs_batch_synchronize": (ctypes.c_int, [_VP]),
    "imfcs_batch_context": (_VP, [_VP, ctypes.c_int]),
    "imfcs_batch_info": (ctypes.c_int, [_VP, _P(ctypes.c_int), _P(ctypes.c_int), _P(ctypes.c_int), _P(_f64)]),
    "imfcs_get_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int)]),
    "imfcs_get_slot_count": (ctypes.c_int, [_VP, _P(ctypes.c_int)]),
    "imfcs_get_run_flags": (ctypes.c_int, [_VP, _P(ctypes.c_int)]),
    "imfcs_run_multi": (ctypes.c_int, [_P(_VP), ctypes.c_int, _VP, _P(PipelineConfig)]),
    "imfcs_refit": (ctypes.c_int, [_VP, _P(PipelineConfig)]),
    "imfcs_get_results": (ctypes.c_int, [_VP, _P(_f64), _P(_f64), _P(_i32), _P(_i32), _P(_f64), _P(_f64),
                                         _P(_i32), _P(_i32), _P(_f64)]),
    "imfcs_get_stage_ms": (ctypes.c_int, [_VP, _P(_f32)]),
    "imfcs_get_correlator_kernel_ms": (ctypes.c_int, [_VP, _P(_f32)]),
    "imfcs_get_gls_covariance": (ctypes.c_int, [_VP, ctypes.c_int, _P(_f64)]),
    "imfcs_mlp_create": (_VP, [_VP, ctypes.c_int, _P(FitModelGeom), _P(_f64), ctypes.c_int, _P(_f64),
                               ctypes.c_uint64]),
    "imfcs_mlp_destroy": (None, [_VP]),
    "imfcs_mlp_train": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _f32, _P(_f32)]),
    "imfcs_mlp_param_count": (ctypes.c_int, []),
    "imfcs_mlp_get_params": (ctypes.c_int, [_VP, _P(_f32)]),
    "imfcs_mlp_target_scale": (ctypes.c_int, [_VP, _P(_f32), _P(_f32)]),
    "imfcs_mlp_gradient": (ctypes.c_int, [_VP, _P(_f32), _P(_f32), ctypes.c_int, _P(_f32), _P(_f32), _P(_f32)]),
    "imfcs_mlp_curves": (ctypes.c_int, [_VP, ctypes.c_int, _P(_f32), _P(_f32)]),
    "imfcs_mlp_predict": (ctypes.c_int, [_VP, _VP, ctypes.c_int, _VP]),
    "imfcs_set_fit_seeder": (ctypes.c_int, [_VP, _VP, _VP]),
    "imfcs_simulate_2d_device": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P(SimDomain),
                                                ctypes.c_int, _f64, _f64, _f64, ctypes.c_int, _f64, ctypes.c_uint64,
                                                _VP]),
}

JNI_EXPORTS = ["fit", "getLastError", "isCudaAvailableInt", "getCudaVersionAsArray", "resetGPU",
               "calcDataBleachCorrection", "isBinningMemorySufficient", "calcBinning", "isACFmemorySufficient",
               "calcACF", "resetDevice"]

_lib = None
_sim = None


def lib():
    """The HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise RuntimeError("libimfcs_hip.so is not built; run imaging_fcs_1_52_amd/build.py "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(HIP_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def sim_lib():
    global _sim
    if _sim is None:
        if not os.path.exists(SIM_LIB_PATH):
            raise RuntimeError("libimfcs_sim.so is not built; run imaging_fcs_1_52_amd/build.py")
        L = ctypes.CDLL(SIM_LIB_PATH)
        L.imfcs_simulate_2d.restype = ctypes.c_int
        L.imfcs_simulate_2d.argtypes = ([_VP] + [ctypes.c_int] * 5 + [_f64] * 4 +
                                        [ctypes.c_int, _f64, _VP, ctypes.c_uint64, ctypes.c_int])
        _sim = L
    return _sim
