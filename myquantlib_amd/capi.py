"""ctypes mirror of include/myql_capi.h and the loader of libmyql_hip.so.

The structs here must stay field-for-field identical to the C header; the
package's tests check the sizes against the compiled library.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmyql_hip.so")

# status codes
OK, E_INVALID_ARG, E_NV_FELLER, E_HIP, E_OOM, E_UNSUPPORTED = range(6)
GREEK_NONE, GREEK_ESSENTIAL, GREEK_FULL = range(3)
SCHEME_NULL, SCHEME_NV, SCHEME_NCI, SCHEME_EXACT = range(4)
JUMP_NONE, JUMP_MERTON, JUMP_KOU = range(3)
INSTR_EUROPEAN, INSTR_GEO_ASIAN, INSTR_ARITH_ASIAN, INSTR_LOOKBACK, INSTR_BARRIER = range(5)
PAYOFF_VANILLA, PAYOFF_CASH, PAYOFF_ASSET = range(3)
CALL, PUT = range(2)
BARRIER_UP, BARRIER_DOWN = range(2)
BARRIER_IN, BARRIER_OUT = range(2)
REBATE_NONE, REBATE_AT_HIT, REBATE_AT_MATURITY = range(3)
FP64, FP32 = range(2)
STREAMS_JUMP, STREAMS_SEEDED, STREAMS_SOBOL = range(3)
FLAG_REFERENCE_NCI_ROWS = 1
FLAG_AUTOTUNE = 2
MLMC_GILES, MLMC_ML2R = range(2)


class HestonParams(C.Structure):
    _fields_ = [("kappa", C.c_double), ("theta", C.c_double), ("sigma", C.c_double),
                ("rho", C.c_double), ("v0", C.c_double)]


class Model(C.Structure):
    _fields_ = [("n_factors", C.c_int32), ("jump_kind", C.c_int32), ("vol", C.c_double),
                ("heston", HestonParams * 2), ("jump", C.c_double * 4)]


class Instrument(C.Structure):
    _fields_ = [("kind", C.c_int32), ("payoff", C.c_int32), ("option_type", C.c_int32),
                ("fixed_strike", C.c_int32), ("barrier_dir", C.c_int32),
                ("barrier_action", C.c_int32), ("rebate_timing", C.c_int32),
                ("is_strip", C.c_int32), ("maturity", C.c_double), ("barrier", C.c_double),
                ("rebate", C.c_double), ("rebate_rate", C.c_double), ("payoff_eps", C.c_double),
                ("past_avg_log", C.c_double), ("past_avg_S", C.c_double),
                ("t_elapsed", C.c_double), ("hist_min_S", C.c_double), ("hist_max_S", C.c_double),
                ("strikes", C.POINTER(C.c_double)), ("n_strikes", C.c_int64)]


class MCConfig(C.Structure):
    _fields_ = [("num_paths", C.c_uint64), ("time_steps", C.c_uint64), ("seed", C.c_uint64),
                ("fd_bump", C.c_double), ("vol_bump", C.c_double), ("t_bump", C.c_double),
                ("r_bump", C.c_double), ("precision", C.c_int32), ("stream_layout", C.c_int32),
                ("num_streams", C.c_uint64), ("flags", C.c_uint32), ("qmc_replicates", C.c_uint32)]


_PD = C.POINTER(C.c_double)


# The result arrays (double *) are declared as plain addresses so that a
# batch of thousands of results can be pointed at numpy buffers with integer
# arithmetic instead of one ctypes pointer object per field.
_VP = C.c_void_p


class MCResult(C.Structure):
    _fields_ = [("price", _VP), ("price_std_err", _VP), ("delta", _VP), ("delta_std_err", _VP),
                ("gamma", _VP), ("gamma_std_err", _VP), ("vega", _VP * 2),
                ("vega_std_err", _VP * 2), ("theta", _VP), ("theta_std_err", _VP), ("rho", _VP),
                ("rho_std_err", _VP), ("vanna", _VP * 2), ("vanna_std_err", _VP * 2),
                ("charm", _VP), ("charm_std_err", _VP), ("total_steps", C.c_uint64),
                ("device_ms", C.c_double)]


class MCJob(C.Structure):
    _fields_ = [("model", C.POINTER(Model)), ("schemes", C.c_int32 * 2),
                ("greek_mode", C.c_int32), ("instrument", C.POINTER(Instrument)),
                ("config", C.POINTER(MCConfig)), ("S0", C.c_double), ("r", C.c_double),
                ("q", C.c_double), ("result", C.POINTER(MCResult))]


class FourierConfig(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("start_bound_guess", C.c_double)]


class FourierJob(C.Structure):
    _fields_ = [("model", C.POINTER(Model)), ("instrument", C.POINTER(Instrument)),
                ("greek_mode", C.c_int32), ("config", C.POINTER(FourierConfig)),
                ("S0", C.c_double), ("r", C.c_double), ("q", C.c_double),
                ("result", C.POINTER(MCResult))]


class MLMCConfig(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("L_min", C.c_uint64), ("L_max", C.c_uint64),
                ("N0", C.c_uint64), ("base_steps", C.c_uint64), ("alpha", C.c_double),
                ("beta", C.c_double), ("adaptive_rates", C.c_int32), ("seed", C.c_uint64),
                ("fd_bump", C.c_double), ("precision", C.c_int32),
                ("stream_layout", C.c_int32), ("num_streams", C.c_uint64), ("flags", C.c_uint32),
                ("estimator", C.c_int32)]


class MLMCResult(C.Structure):
    _fields_ = [("price", C.c_double), ("price_std_err", C.c_double),
                ("total_steps", C.c_uint64), ("n_levels", C.c_int32),
                ("paths_per_level", C.POINTER(C.c_uint64)), ("mean_diffs", _PD),
                ("var_diffs", _PD), ("alpha_estimated", C.c_double),
                ("beta_estimated", C.c_double), ("device_ms", C.c_double),
                ("delta", C.c_double), ("delta_std_err", C.c_double),
                ("gamma", C.c_double), ("gamma_std_err", C.c_double),
                ("level_weights", _PD)]


CV_NPARAM, CV_MAX_POINTS, CV_MAX_JOBS, CV_NREG = 12864, 16, 16, 4


class CVConfig(C.Structure):
    _fields_ = [("train_paths", C.c_uint64), ("train_seed", C.c_uint64), ("steps", C.c_uint32),
                ("batch", C.c_uint32), ("lr", C.c_double), ("points", C.c_uint32),
                ("init_seed", C.c_uint32)]


class CVNet(C.Structure):
    _fields_ = [("params", C.c_float * CV_NPARAM), ("norm", C.c_float * 6), ("n_jobs", C.c_int32),
                ("n_strikes", C.c_int32), ("maturity", C.c_double * CV_MAX_JOBS),
                ("coef", C.c_double * (CV_MAX_JOBS * CV_NREG * 64)), ("final_loss", C.c_double),
                ("train_ms", C.c_double)]


class CVResult(C.Structure):
    _fields_ = [("plain_price", _PD), ("plain_std_err", _PD), ("var_ratio", _PD),
                ("used_cv", C.POINTER(C.c_int32))]


# Every extern "C" symbol declared in include/myql_capi.h.
EXPORTED = [
    "myql_create", "myql_destroy", "myql_last_error", "myql_tuning_report", "myql_mc_calculate",
    "myql_mc_calculate_noalloc", "myql_mc_calculate_batch", "myql_mlmc_calculate", "myql_fourier_calculate_batch", "myql_cv_train",
    "myql_mc_calculate_batch_cv", "myql_debug_stream_draws", "myql_debug_normcdfinv",
    "myql_debug_poissinv", "myql_debug_poisson_counts", "myql_debug_qmc_points", "myql_debug_chi2_quantile", "myql_debug_path_payoffs", "myql_debug_set_launch_shape",
    "myql_debug_cv_input_grad",
    "myql_debug_cv_param_grad",
    "myql_debug_cv_jump_expectation",
]


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libmyql_hip.so and declare the argument types of every entry point.

    Raises OSError when the library has not been built: there is no fallback.
    """
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `python __graft_entry__.py` "
                      "(the HIP engine has no CPU fallback)")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.myql_create.restype = vp
    lib.myql_create.argtypes = [C.c_int]
    lib.myql_destroy.restype = None
    lib.myql_destroy.argtypes = [vp]
    lib.myql_last_error.restype = C.c_char_p
    lib.myql_last_error.argtypes = [vp]
    lib.myql_fourier_calculate_batch.restype = C.c_int
    lib.myql_fourier_calculate_batch.argtypes = [vp, C.POINTER(FourierJob), C.c_int64]
    lib.myql_tuning_report.restype = C.c_int
    lib.myql_tuning_report.argtypes = [vp, C.c_char_p, C.c_int64]
    lib.myql_mc_calculate.restype = C.c_int
    lib.myql_mc_calculate.argtypes = [vp, C.POINTER(Model), C.POINTER(C.c_int32),
                                      C.POINTER(Instrument), C.c_int32, C.POINTER(MCConfig),
                                      C.c_double, C.c_double, C.c_double, C.POINTER(MCResult)]
    lib.myql_mc_calculate_noalloc.restype = C.c_int
    lib.myql_mc_calculate_noalloc.argtypes = lib.myql_mc_calculate.argtypes
    lib.myql_mc_calculate_batch.restype = C.c_int
    lib.myql_mc_calculate_batch.argtypes = [vp, C.POINTER(MCJob), C.c_int64]
    lib.myql_mlmc_calculate.restype = C.c_int
    lib.myql_mlmc_calculate.argtypes = [vp, C.POINTER(Model), C.POINTER(C.c_int32),
                                        C.POINTER(Instrument), C.c_int32, C.POINTER(MLMCConfig),
                                        C.c_double, C.c_double, C.c_double,
                                        C.POINTER(MLMCResult)]
    lib.myql_debug_stream_draws.restype = C.c_int
    lib.myql_debug_stream_draws.argtypes = [vp, C.c_uint64, C.c_int32, C.c_uint64, C.c_uint64,
                                            C.POINTER(C.c_uint64)]
    lib.myql_debug_normcdfinv.restype = C.c_int
    lib.myql_debug_normcdfinv.argtypes = [vp, _PD, C.c_int64, _PD, C.POINTER(C.c_float)]
    lib.myql_debug_poissinv.restype = C.c_int
    lib.myql_debug_poissinv.argtypes = [vp, _PD, _PD, C.c_int64, _PD]
    lib.myql_debug_qmc_points.restype = C.c_int
    lib.myql_debug_qmc_points.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_int32, C.c_uint64, C.c_uint64,
                                          C.POINTER(C.c_uint32)]
    lib.myql_debug_poisson_counts.restype = C.c_int
    lib.myql_debug_poisson_counts.argtypes = [vp, _PD, C.c_int32, _PD, C.c_int64, C.POINTER(C.c_int32)]
    lib.myql_debug_chi2_quantile.restype = C.c_int
    lib.myql_debug_chi2_quantile.argtypes = [vp, _PD, _PD, C.c_int64, _PD]
    lib.myql_debug_path_payoffs.restype = C.c_int
    lib.myql_debug_path_payoffs.argtypes = [vp, C.POINTER(Instrument), C.c_int32, C.c_double, C.c_double, _PD,
                                            C.c_int64, C.c_int64, _PD, _PD, _PD]
    lib.myql_debug_set_launch_shape.restype = C.c_int
    lib.myql_debug_set_launch_shape.argtypes = [vp] + [C.c_int32] * 8
    pf = C.POINTER(C.c_float)
    lib.myql_cv_train.restype = C.c_int
    lib.myql_cv_train.argtypes = [vp, C.POINTER(MCJob), C.c_int64, C.POINTER(CVConfig),
                                  C.POINTER(CVNet)]
    lib.myql_mc_calculate_batch_cv.restype = C.c_int
    lib.myql_mc_calculate_batch_cv.argtypes = [vp, C.POINTER(MCJob), C.c_int64,
                                               C.POINTER(CVConfig), C.POINTER(CVNet),
                                               C.POINTER(CVResult)]
    lib.myql_debug_cv_input_grad.restype = C.c_int
    lib.myql_debug_cv_input_grad.argtypes = [vp, pf, pf, C.c_int64, pf]
    lib.myql_debug_cv_param_grad.restype = C.c_int
    lib.myql_debug_cv_param_grad.argtypes = [vp, pf, pf, pf, C.c_int64, pf, _PD]
    lib.myql_debug_cv_jump_expectation.restype = C.c_int
    lib.myql_debug_cv_jump_expectation.argtypes = [vp, vp, vp, C.c_int32, C.c_double, _PD, C.c_int64, _PD, _PD, _PD,
                                                   C.POINTER(C.c_int32)]
    return lib
