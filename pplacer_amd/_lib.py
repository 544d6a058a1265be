"""Loads libpplacer_hip.so and declares its C ABI (include/pplacer_hip.h).

The product path has no CPU fallback: if the library is missing, importing
any entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PPLACER_HIP_LIB names another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("PPLACER_HIP_LIB") or os.path.join(HERE, "libpplacer_hip.so")

PP_OK, PP_EINVAL, PP_EDEVICE, PP_ENOMEM, PP_ESTATE, PP_EQUERY = range(6)
PP_FP64, PP_FP32 = 0, 1

Q_STATUS = {0: "ok", 1: "no informative sites", 2: "unknown character", 3: "empty results",
            4: "FindStartFailure", 5: "InvalidStartValues", 6: "assert prox_bl >= 0",
            7: "query too long for the GPU path", 8: "ExceededMaxIter",
            9: "GSL error"}
PP_PRIOR_FLAT_EXP, PP_PRIOR_UNIFORM, PP_PRIOR_INFORMATIVE = 0, 1, 2


class PPError(RuntimeError):
    pass


class Placement(C.Structure):
    _fields_ = [("loc", C.c_int), ("ml_ratio", C.c_double), ("log_like", C.c_double),
                ("dist_bl", C.c_double), ("pend_bl", C.c_double), ("post_prob", C.c_double),
                ("marginal_prob", C.c_double)]


class Prefs(C.Structure):
    _fields_ = [("start_pend", C.c_double), ("max_pend", C.c_double),
                ("initial_tolerance", C.c_double), ("strike_box", C.c_double),
                ("keep_factor", C.c_double), ("max_strikes", C.c_int),
                ("max_pitches", C.c_int), ("keep_at_most", C.c_int), ("calc_pp", C.c_int),
                ("pp_rel_err", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("t_ranking_ms", C.c_double), ("t_ml_ms", C.c_double),
                ("n_logdots", C.c_int64), ("n_like_calls", C.c_int64), ("n_pitched", C.c_int64),
                ("t_stage1_ms", C.c_double), ("t_pitch_ms", C.c_double),
                ("t_refine_ms", C.c_double), ("n_site_pairs", C.c_int64),
                ("n_site_probes_pend", C.c_int64), ("n_site_probes_dist", C.c_int64),
                ("n_launches", C.c_int64), ("t_table_ms", C.c_double),
                ("n_site_probes_pend_pitch", C.c_int64), ("n_site_probes_dist_pitch", C.c_int64),
                ("n_pitch_launches", C.c_int64), ("t_pp_ms", C.c_double),
                ("t_pitch_kernel_ms", C.c_double),
                ("n_fp64_replaced", C.c_int64), ("t_fp64_replace_ms", C.c_double)]


class LaunchConfig(C.Structure):
    _fields_ = [("pitch_streams", C.c_int * 2), ("refine_streams", C.c_int * 2),
                ("chunk", C.c_int), ("graph", C.c_int)]


PP_MAX_BUCKET_STREAMS, PP_MAX_CHUNK = 8, 262144

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u16p = C.POINTER(C.c_uint16)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p
_i, _d = C.c_int, C.c_double

SIGNATURES = {
    "pp_last_error": (C.c_char_p, []),
    "pp_gemmish": (_i, [_dp, _dp, _dp, _i, _i]),
    "pp_dediagonalize": (_i, [_dp, _dp, _dp, _dp, _i]),
    "pp_ten_log_like3": (_i, [_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _dp]),
    "pp_ten_pairwise_prod": (_i, [_dp, _dp, _dp, _i, _i, _i]),
    "pp_ten_statd_pairwise_prod": (_i, [_dp, _dp, _dp, _dp, _i, _i, _i]),
    "pp_ten_masked_logdot": (_i, [_dp, _dp, _u16p, _dp, _i, _i, _i, _dp]),
    "pp_ten_bounded_logdot": (_i, [_dp, _dp, _i, _i, _dp, _i, _i, _i, _dp]),
    "pp_ten_masked_logdot_noalloc": (_i, [C.c_void_p, C.c_void_p, C.c_void_p, _i, _i, _i, C.c_void_p,
                                          C.c_void_p]),
    # test hook of the FP64 engine's wave-parallel sequential sum (not in the header)
    "pp_test_sequential_sum": (_i, [_dp, _i, _dp]),
    "pp_test_cr_math": (_i, [C.c_longlong, _dp, _dp, _dp]),
    "pp_test_cr_math_host": (_i, [C.c_longlong, _dp, _dp, _dp]),
    "pp_mat_log_like3": (_i, [_dp, _dp, _dp, _dp, _i, _i, _dp]),
    "pp_mat_pairwise_prod": (_i, [_dp, _dp, _dp, _i, _i]),
    "pp_mat_statd_pairwise_prod": (_i, [_dp, _dp, _dp, _dp, _i, _i]),
    "pp_mat_masked_logdot": (_i, [_dp, _dp, C.POINTER(C.c_uint16), _i, _i, _dp]),
    "pp_mat_bounded_logdot": (_i, [_dp, _dp, _i, _i, _i, _i, _dp]),
    "pp_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "pp_ctx_destroy": (_i, [_vp]),
    "pp_model_build": (_i, [_vp, _i, _i, _dp, _dp, _dp]),
    "pp_model_get": (_i, [_vp, _dp, _dp, _dp]),
    "pp_model_build_cat": (_i, [_vp, _i, _i, _dp, _dp, _dp]),
    "pp_site_categories": (_i, [_vp, _i, C.POINTER(C.c_int)]),
    "pp_discrete_gamma": (_i, [_i, _d, _dp]),
    "pp_emper_freq": (_i, [_i, C.c_char_p, _i, _i, _dp]),
    "pp_premask": (_i, [_i, C.c_char_p, _i, C.c_char_p, _i, _i, _i, _u8p, _u8p, _ip]),
    "pp_tree_set": (_i, [_vp, _i, _ip, _ip, _dp, _i, C.c_char_p]),
    "pp_tree_upload": (_i, [_vp, _i, _dp, _i, _dp, _i64p, _dp, _i64p, _dp, _i64p]),
    "pp_tree_download": (_i, [_vp, _dp, _i64p, _dp, _i64p, _dp, _i64p]),
    "pp_prefs_default": (None, [C.POINTER(Prefs)]),
    "pp_stage1_scores": (_i, [_vp, _i, C.c_char_p, _d, _i, _dp]),
    "pp_rank_topk": (_i, [_vp, _i, _i, _dp, _i, _ip]),
    "pp_optimize_batch": (_i, [_vp, _i, C.c_char_p, _ip, _ip, _dp, _dp, _d, _d, _i, _dp, _dp,
                               _dp, _ip, _ip]),
    "pp_marginal_batch": (_i, [_vp, _i, C.c_char_p, _ip, _ip, _dp, _dp, _d, _d, _dp, _ip, _ip]),
    "pp_set_prior": (_i, [_vp, _i, _i, _dp]),
    "pp_tree_length": (_i, [_i, _ip, _ip, _dp, _dp]),
    "pp_prior_mean_map": (_i, [_i, _ip, _ip, _dp, _d, _dp]),
    "pp_place_batch": (_i, [_vp, _i, C.c_char_p, C.POINTER(Prefs), _i, C.POINTER(Placement),
                            _ip, _ip, C.POINTER(Stats)]),
    "pp_queries_upload": (_i, [_vp, _i, C.c_char_p]),
    "pp_place_resident": (_i, [_vp, C.POINTER(Prefs), _i, C.POINTER(Stats)]),
    "pp_results_download": (_i, [_vp, C.POINTER(Placement), _ip, _ip]),
    "pp_synchronize": (_i, [_vp]),
    "pp_launch_config_get": (_i, [_vp, C.POINTER(LaunchConfig)]),
    "pp_launch_config_set": (_i, [_vp, C.POINTER(LaunchConfig)]),
    # device GLVs (Glvm.Model's Glv of Gmix_model)
    "pp_glv_make": (_i, [_vp, _i, _i, _i, C.POINTER(_vp)]),
    "pp_glv_free": (_i, [_vp]),
    "pp_glv_dims": (_i, [_vp, _ip, _ip, _ip]),
    "pp_glv_mimic": (_i, [_vp, C.POINTER(_vp)]),
    "pp_glv_copy": (_i, [_vp, C.POINTER(_vp)]),
    "pp_glv_memcpy": (_i, [_vp, _vp]),
    "pp_glv_set_unit": (_i, [_vp]),
    "pp_glv_fp_classify": (_i, [_vp, _ip]),
    "pp_glv_set": (_i, [_vp, _dp, _i64p]),
    "pp_glv_get": (_i, [_vp, _dp, _i64p]),
    "pp_glv_mask_into": (_i, [_u8p, _i, _vp, _vp]),
    "pp_glv_prep_constant_rate_from_lv_arr": (_i, [_vp, _dp]),
    "pp_glv_perhaps_pull_exponent": (_i, [_i, _vp]),
    "pp_glv_masked_logdot": (_i, [_vp, _vp, _u16p, _dp]),
    "pp_glv_bounded_logdot": (_i, [_vp, _vp, _i, _i, _dp]),
    "pp_glv_logdot": (_i, [_vp, _vp, _dp]),
    "pp_glv_pairwise_prod": (_i, [_vp, _vp, _vp]),
    "pp_glv_listwise_prod": (_i, [_vp, C.POINTER(_vp), _i]),
    "pp_glv_summarize_post": (_i, [_vp, _dp]),
    "pp_gmix_log_like3": (_i, [_vp, _vp, _vp, _vp, _dp]),
    "pp_gmix_site_log_like_arr3": (_i, [_vp, _vp, _vp, _vp, _dp]),
    "pp_gmix_slow_log_like3": (_i, [_vp, _vp, _vp, _vp, _dp]),
    "pp_gmix_evolve_into": (_i, [_vp, _vp, _vp, _d]),
    "pp_gmix_statd_pairwise_prod": (_i, [_vp, _vp, _vp, _vp]),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PPError("libpplacer_hip.so is missing at %s: run `python -m pplacer_amd.build`"
                          % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != PP_OK:
        msg = lib().pp_last_error().decode(errors="replace")
        raise PPError("%s failed (code %d): %s" % (what or "libpplacer_hip", rc, msg))
