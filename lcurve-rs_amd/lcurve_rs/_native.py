"""ctypes binding of liblcurve_hip.so (declarations in include/lcurve_hip.h).

The library is loaded from the package directory's parent
(``lcurve-rs_amd/liblcurve_hip.so``) or from ``$LCURVE_HIP_LIB``. A missing
library raises ImportError: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LCURVE_HIP_LIB", os.path.join(os.path.dirname(_HERE), "liblcurve_hip.so"))

LCH_OK = 0
LCH_ERR_INVALID = -1
LCH_ERR_PARSE = -2
LCH_ERR_SETUP = -3
LCH_ERR_DEVICE = -4
LCH_ERR_RUNTIME = -5
LCH_ERR_UNSUPPORTED = -6
LCH_ERR_NOMEM = -7
LCH_BATCH_KEEP_MAPS = 1

D = ctypes.POINTER(ctypes.c_double)
F = ctypes.POINTER(ctypes.c_float)
I = ctypes.POINTER(ctypes.c_int)
VP = ctypes.c_void_p
CP = ctypes.c_char_p
INT = ctypes.c_int


class LcResultC(ctypes.Structure):
    _fields_ = [("wdwarf", ctypes.c_double), ("chisq", ctypes.c_double), ("wnok", ctypes.c_double),
                ("logg1", ctypes.c_double), ("logg2", ctypes.c_double), ("rv1", ctypes.c_double),
                ("rv2", ctypes.c_double), ("sfac", ctypes.c_double * 5)]


class BatchStatsC(ctypes.Structure):
    _fields_ = [("n_sets", ctypes.c_longlong), ("n_valid", ctypes.c_longlong),
                ("n_times", ctypes.c_longlong), ("n_lanes", ctypes.c_longlong),
                ("n_items", ctypes.c_longlong), ("visits", ctypes.c_longlong),
                ("record_bytes", ctypes.c_longlong), ("pairs_evaluated", ctypes.c_longlong),
                ("pairs_by_moments", ctypes.c_longlong),
                ("setup_seconds", ctypes.c_double),
                ("upload_seconds", ctypes.c_double), ("chunk", ctypes.c_int),
                ("persistent", ctypes.c_int), ("pairs_limb", ctypes.c_longlong),
                ("limb_chunks", ctypes.c_longlong), ("moment_evals", ctypes.c_longlong),
                ("limb_items", ctypes.c_longlong), ("limb_slots", ctypes.c_longlong)]


class DPackStatsC(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in (
        "n_sets", "n_valid", "n_segs", "n_elements", "n_pairs", "n_patches", "n_tiles", "n_chunks", "n_codes",
        "n_cand", "n_items", "n_items_plain", "n_items_ecl", "rec_bytes", "input_bytes", "work_bytes", "visits",
        "pairs_evaluated", "pairs_by_moments", "moment_evals", "pairs_limb", "n_limb_chunks", "n_limb_items",
        "split_pairs", "split_limb_lanes")] + \
        [(k, ctypes.c_double) for k in ("setup_core_s", "layout_core_s", "build_wall_s", "concat_s", "upload_s")] + \
        [("chunk", ctypes.c_int), ("limb", ctypes.c_int), ("threads", ctypes.c_int)]


LL = ctypes.POINTER(ctypes.c_longlong)
LCH_NGRID = 11
GRID_NAMES = ("star1f", "star1c", "star2f", "star2c", "disc", "edge", "spot",
              "star1f_cal", "star1c_cal", "star2f_cal", "star2c_cal")


class GridSoaC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_longlong), ("posn", D), ("dirn", D), ("area", F), ("gravity", F),
                ("flux", F), ("ecl_off", LL), ("ingress", D), ("egress", D)]


class FluxParamsC(ctypes.Structure):
    _fields_ = [("iangle", ctypes.c_double), ("q", ctypes.c_double), ("limb1", ctypes.c_int),
                ("limb2", ctypes.c_int), ("ldc1", ctypes.c_double * 4), ("ldc2", ctypes.c_double * 4),
                ("mucrit1", ctypes.c_double), ("mucrit2", ctypes.c_double),
                ("lin_limb_disc", ctypes.c_double), ("quad_limb_disc", ctypes.c_double),
                ("beam_factor1", ctypes.c_double), ("beam_factor2", ctypes.c_double),
                ("spin1", ctypes.c_double), ("spin2", ctypes.c_double),
                ("velocity_scale", ctypes.c_double), ("glens1", ctypes.c_int), ("rlens1", ctypes.c_double),
                ("phase1", ctypes.c_double), ("phase2", ctypes.c_double), ("gint_mode", ctypes.c_int),
                ("scale11", ctypes.c_double), ("scale12", ctypes.c_double),
                ("scale21", ctypes.c_double), ("scale22", ctypes.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "lch_version": (INT, []),
    "lch_last_error": (CP, [VP]),
    "lch_device_count": (INT, []),
    "lch_ctx_create": (INT, [INT, ctypes.POINTER(VP)]),
    "lch_ctx_destroy": (None, [VP]),
    "lch_ctx_set_limb": (INT, [VP, INT]),
    "lch_ctx_set_limb_check": (INT, [VP, INT]),
    "lch_ctx_set_tuning": (INT, [VP, INT, INT]),
    "lch_ctx_sync": (INT, [VP]),
    "lch_model_load_file": (INT, [CP, ctypes.POINTER(VP)]),
    "lch_model_load_text": (INT, [CP, ctypes.POINTER(VP)]),
    "lch_model_copy": (INT, [VP, ctypes.POINTER(VP)]),
    "lch_model_free": (None, [VP]),
    "lch_model_set_param": (INT, [VP, CP, ctypes.c_double]),
    "lch_model_get_param": (INT, [VP, CP, D]),
    "lch_model_get_pparam": (INT, [VP, CP, D, D, D, I, I]),
    "lch_model_set_pparam": (INT, [VP, CP, D, D, D, I]),
    "lch_param_count": (INT, []),
    "lch_param_name": (CP, [INT]),
    "lch_data_count": (INT, [CP, I]),
    "lch_data_read": (INT, [CP, INT, D, D, D, D, D, I]),
    "lch_light_curve": (INT, [VP, VP, D, D, D, D, D, I, INT, INT, INT, D, ctypes.POINTER(LcResultC)]),
    "lch_chisq_batch": (INT, [VP, VP, D, D, D, D, D, I, INT, ctypes.POINTER(CP), INT, D, INT, INT, D]),
    "lch_batch_build": (INT, [VP, VP, D, D, D, D, D, I, INT, ctypes.POINTER(CP), INT, D, INT,
                              D, D, D, INT, ctypes.POINTER(VP)]),
    "lch_batch_build_models": (INT, [VP, ctypes.POINTER(VP), INT, D, D, D, D, D, I, INT, D, D, D, INT,
                                     ctypes.POINTER(VP)]),
    "lch_batch_free": (None, [VP]),
    "lch_batch_set_obs": (INT, [VP, VP, D, D, D]),
    "lch_batch_launch": (INT, [VP, VP, INT]),
    "lch_batch_run": (INT, [VP, VP, INT, D, D, D, D, D]),
    "lch_batch_time": (INT, [VP, VP, INT, INT, D, D]),
    "lch_batch_autotune": (INT, [VP, VP, INT]),
    "lch_batch_stats_get": (INT, [VP, ctypes.POINTER(BatchStatsC)]),
    "lch_dpack_build": (INT, [VP, ctypes.POINTER(VP), INT, D, D, D, D, D, I, INT, D, D, D, INT, INT,
                              ctypes.POINTER(VP)]),
    "lch_dpack_free": (None, [VP]),
    "lch_dpack_upload": (INT, [VP, VP]),
    "lch_dpack_set_slots": (INT, [VP, VP, INT]),
    "lch_dpack_launch": (INT, [VP, VP, INT]),
    "lch_dpack_run": (INT, [VP, VP, INT, D, D, D, D, D]),
    "lch_dpack_time": (INT, [VP, VP, INT, INT, D, D, D]),
    "lch_dpack_set_obs": (INT, [VP, VP, D, D, D]),
    "lch_dpack_stats_get": (INT, [VP, ctypes.POINTER(DPackStatsC)]),
    "lch_dpack_check": (INT, [VP, VP, INT, LL]),
    "lch_batch_sfac": (INT, [VP, D]),
    "lch_selfcheck_patch_codes": (INT, [VP, D, D, I, INT, D, ctypes.POINTER(ctypes.c_longlong),
                                        ctypes.POINTER(ctypes.c_longlong)]),
    "lch_selfcheck_eclipse_prescreen": (INT, [VP, ctypes.POINTER(ctypes.c_longlong), D]),
    "lch_mlp_create": (INT, [VP, I, INT, ctypes.c_ulonglong, INT, ctypes.POINTER(VP)]),
    "lch_mlp_free": (None, [VP]),
    "lch_mlp_param_count": (ctypes.c_longlong, [VP]),
    "lch_mlp_get_params": (INT, [VP, F]),
    "lch_mlp_set_params": (INT, [VP, F]),
    "lch_mlp_set_data": (INT, [VP, F, F, INT]),
    "lch_mlp_train_epoch": (INT, [VP, I, INT, INT, ctypes.c_float, D]),
    "lch_mlp_loss_grad": (INT, [VP, F, F, INT, D, F]),
    "lch_mlp_adam_step": (INT, [VP, ctypes.c_float]),
    "lch_mlp_predict": (INT, [VP, F, INT, F]),
    "lch_mlp_time_step": (INT, [VP, INT, INT, D, D, D]),
    "lch_surr_lik_n_features": (INT, [INT, INT]),
    "lch_model_iscale": (INT, [VP, I]),
    "lch_surr_lik_create": (INT, [VP, INT, D, D, INT, D, INT, ctypes.c_double, D, D, D, ctypes.POINTER(VP)]),
    "lch_surr_lik_eval": (INT, [VP, D, INT, D]),
    "lch_surr_lik_free": (None, [VP]),
    "lch_batch_set_info": (INT, [VP, INT, I, D, D, D, D, CP, INT]),
    "lch_batch_visibility": (INT, [VP, VP, INT, INT, D, INT, ctypes.POINTER(ctypes.c_ubyte), I]),
    "lch_model_upload": (INT, [VP, ctypes.POINTER(GridSoaC), ctypes.POINTER(FluxParamsC), ctypes.POINTER(VP)]),
    "lch_grids_free": (None, [VP]),
    "lch_flux_eval": (INT, [VP, VP, D, D, I, D, ctypes.c_double, D, INT]),
    "lch_flux_eval_static": (INT, [VP, VP, D, D, I, D, ctypes.c_double, D, INT]),
    "lch_model_prepare": (INT, [VP, D, D, I, INT, ctypes.POINTER(VP)]),
    "lch_setup_grid": (INT, [VP, INT, ctypes.POINTER(GridSoaC)]),
    "lch_setup_params": (INT, [VP, ctypes.POINTER(FluxParamsC)]),
    "lch_setup_times": (INT, [VP, INT, D, D, I, D, D]),
    "lch_setup_free": (None, [VP]),
}

_lib = None


def lib():
    """The loaded library (raises ImportError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"liblcurve_hip.so not found at {LIB_PATH}; build it with "
                              f"`python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error(ctx=None):
    """Message of the last failing call of this thread, or of the last
    failing call made with context handle `ctx`."""
    msg = lib().lch_last_error(ctx)
    return msg.decode(errors="replace") if msg else ""


class LcurveError(RuntimeError):
    """A failing engine call; ``code`` is the LCH_ERR_* status."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def check(rc):
    if rc != LCH_OK:
        msg = last_error()
        if rc in (LCH_ERR_INVALID, LCH_ERR_PARSE, LCH_ERR_SETUP, LCH_ERR_UNSUPPORTED):
            raise ValueError(msg)
        raise LcurveError(rc, msg)
    return rc
