"""ctypes binding of libfastbart_hip.so (the C-ABI declared in include/fastbart_hip.h).

The library is loaded from the package directory. If it is missing it is built once with hipcc; if it cannot be built
or loaded, the import fails: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastbart_hip.so")

NODES = 64
STATS_REF16 = 0
STATS_FX64 = 1


class FbhError(RuntimeError):
    pass


class ChainIn(C.Structure):
    _fields_ = [("n", C.c_int), ("p", C.c_int), ("q", C.c_void_p), ("y", C.c_void_p), ("wt", C.c_void_p),
                ("numcut", C.c_void_p), ("ybin", C.c_void_p), ("tau", C.c_float), ("lambda_", C.c_float),
                ("nu", C.c_float), ("sigest", C.c_float), ("offset", C.c_float), ("seed", C.c_uint)]


class Opts(C.Structure):
    _fields_ = [("m", C.c_int), ("nskip", C.c_int), ("ndpost", C.c_int), ("probit", C.c_int),
                ("keep_trees", C.c_int), ("pgtab", C.c_void_p), ("stats_mode", C.c_int), ("node_cap", C.c_int),
                ("device", C.c_int), ("sweeps_per_launch", C.c_int), ("block_threads", C.c_int),
                ("tile_rows", C.c_int), ("row_split", C.c_int), ("packing", C.c_int), ("graph", C.c_int),
                ("chains_per_cu", C.c_int), ("resident", C.c_int), ("probit_redraw", C.c_int)]


TUNE_MAX = 12


class Tune(C.Structure):
    _fields_ = [("packing", C.c_int), ("tile_rows", C.c_int), ("resident", C.c_int), ("chains_per_cu", C.c_int),
                ("ncand", C.c_int), ("cand_packing", C.c_int * TUNE_MAX), ("cand_tile_rows", C.c_int * TUNE_MAX),
                ("cand_resident", C.c_int * TUNE_MAX), ("cand_chains_per_cu", C.c_int * TUNE_MAX),
                ("ms_per_sweep", C.c_double * TUNE_MAX), ("sweeps", C.c_int), ("from_cache", C.c_int)]


class PrepOpts(C.Structure):
    _fields_ = [("ntree", C.c_int), ("numcut", C.c_int), ("rm_const", C.c_int), ("k", C.c_double),
                ("power", C.c_double), ("base", C.c_double), ("sigdf", C.c_double), ("sigquant", C.c_double),
                ("sigest", C.c_double), ("sigmaf", C.c_double), ("lambda_", C.c_double), ("fmean", C.c_double),
                ("binary_offset", C.c_double)]


class Inputs(C.Structure):
    _fields_ = [("n", C.c_int), ("p", C.c_int), ("q", C.POINTER(C.c_uint8)), ("numcut", C.POINTER(C.c_int)),
                ("rm_const", C.POINTER(C.c_int)), ("cutpoints", C.POINTER(C.c_double)),
                ("cut_off", C.POINTER(C.c_int)), ("y", C.POINTER(C.c_float)), ("wt", C.POINTER(C.c_float)),
                ("ybin", C.POINTER(C.c_int)), ("pgtab", C.c_float * 64), ("tau", C.c_float), ("lambda_", C.c_float),
                ("nu", C.c_float), ("sigest", C.c_float), ("offset", C.c_float), ("tau_d", C.c_double),
                ("lambda_d", C.c_double), ("nu_d", C.c_double), ("sigest_d", C.c_double), ("offset_d", C.c_double),
                ("fmean", C.c_double), ("lm_rank", C.c_int), ("priv_", C.c_void_p)]


class ChainOut(C.Structure):
    _fields_ = [("sigma", C.c_void_p), ("trmean", C.c_void_p), ("trees", C.c_void_p), ("stats", C.c_void_p)]


class TraceRec(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("it", "j", "mode", "nx", "v", "c", "nl", "nr", "accepted", "ngood",
                                          "nnogs", "cap")] + \
               [(k, C.c_double) for k in ("pr", "lalpha", "lu", "sl0", "sl1", "sr0", "sr1")] + \
               [("rng_after", C.c_uint32 * 4)]


# every symbol include/fastbart_hip.h declares (checked by the CPU tests)
EXPORTS = ["bart_fit_cpu", "bart_wfit", "bart_set_rng", "fbh_last_status", "fbh_last_error", "fbh_batch_create",
           "fbh_batch_run", "fbh_batch_sync", "fbh_batch_iter", "fbh_batch_fetch", "fbh_batch_state",
           "fbh_batch_yhat", "fbh_batch_timing", "fbh_batch_destroy", "fbh_fit_batch", "fbh_tree_text", "fbh_yhat",
           "fbh_compile_trees", "fbh_trees_info", "fbh_trees_arrays", "fbh_trees_free", "fbh_predict_draw",
           "fbh_predict_draws", "fbh_rows_upload", "fbh_rows_free", "fbh_predict_rows", "fbh_predict_timing",
           "fbh_predict_rows_device", "fbh_predict_rows_device_noalloc", "fbh_predict_last_launch", "fbh_trees_rank_form", "fbh_device_alloc", "fbh_device_free",
           "fbh_device_copy", "fbh_device_sync", "fbh_batch_row_split", "fbh_batch_configure", "fbh_batch_config",
           "fbh_batch_set_order", "fbh_batch_probit_redraw", "fbh_batch_probit_info", "fbh_debug_jump",
           "fbh_batch_autotune", "fbh_batch_autotune_cached",
           "fbh_debug_math", "fbh_debug_libm_host", "fbh_debug_copy", "fbh_batch_trace", "fbh_version",
           "fbh_mlp_create", "fbh_mlp_nparams", "fbh_mlp_params", "fbh_mlp_forward", "fbh_mlp_grad", "fbh_mlp_train",
           "fbh_mlp_timing", "fbh_mlp_destroy", "fbh_batch_learned_setup", "fbh_batch_run_learned",
           "fbh_batch_learned_pairs", "fbh_batch_learned_train", "fbh_batch_learned_timing",
           "fbh_batch_learned_target", "fbh_batch_learned_target_stats", "fbh_debug_split_conditional",
           "fbh_batch_learned_pairs_copy", "fbh_batch_learned_holdout", "fbh_batch_learned_eval",
           "fbh_prep_defaults",
           "fbh_wbart_inputs", "fbh_pbart_inputs", "fbh_inputs_free", "fbh_qnorm", "fbh_qchisq"]

_lock = threading.Lock()
_lib = None


def _declare(L):
    vp, i, d, f = C.c_void_p, C.c_int, C.c_double, C.c_float
    L.bart_fit_cpu.argtypes = [vp] * 8 + [i] * 5 + [f] * 4 + [C.c_uint, i, vp, i, vp, f]
    L.bart_fit_cpu.restype = None
    L.bart_wfit.argtypes = [vp] * 8 + [i] * 5 + [f] * 4 + [C.c_uint, i, vp]
    L.bart_wfit.restype = None
    L.bart_set_rng.argtypes = [vp] * 4
    L.bart_set_rng.restype = None
    L.fbh_last_status.restype = i
    L.fbh_last_error.restype = C.c_char_p
    L.fbh_version.restype = C.c_char_p
    L.fbh_batch_create.argtypes = [C.POINTER(ChainIn), i, C.POINTER(Opts), C.POINTER(vp)]
    L.fbh_batch_run.argtypes = [vp, i]
    L.fbh_batch_sync.argtypes = [vp]
    L.fbh_batch_iter.argtypes = [vp]
    L.fbh_batch_fetch.argtypes = [vp, i, C.POINTER(ChainOut)]
    L.fbh_batch_state.argtypes = [vp, i, vp, vp, vp, vp, vp, vp]
    L.fbh_batch_yhat.argtypes = [vp, i, vp]
    L.fbh_batch_timing.argtypes = [vp, C.POINTER(d), C.POINTER(i), C.POINTER(d)]
    L.fbh_batch_destroy.argtypes = [vp]
    L.fbh_batch_destroy.restype = None
    L.fbh_batch_trace.argtypes = [vp, i, vp, i]
    L.fbh_fit_batch.argtypes = [C.POINTER(ChainIn), i, C.POINTER(Opts), C.POINTER(ChainOut)]
    L.fbh_tree_text.argtypes = [vp, i, i, i, vp, C.c_size_t, vp]
    L.fbh_tree_text.restype = C.c_size_t
    L.fbh_yhat.argtypes = [vp, vp, i, i, i, i, i, vp]
    L.fbh_compile_trees.argtypes = [C.c_char_p, C.c_size_t, vp, vp, i, i, C.POINTER(vp)]
    L.fbh_trees_info.argtypes = [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i)]
    L.fbh_trees_arrays.argtypes = [vp, vp, vp, vp, vp, vp]
    L.fbh_trees_free.argtypes = [vp]
    L.fbh_trees_free.restype = None
    L.fbh_predict_draw.argtypes = [vp, vp, i, i, i, d, vp]
    L.fbh_predict_draws.argtypes = [vp, vp, i, i, vp, i, d, vp]
    L.fbh_rows_upload.argtypes = [vp, i, i, i, C.POINTER(vp)]
    L.fbh_rows_free.argtypes = [vp]
    L.fbh_rows_free.restype = None
    L.fbh_predict_rows.argtypes = [vp, vp, vp, i, d, vp]
    L.fbh_predict_timing.argtypes = [vp, C.POINTER(d), C.POINTER(i)]
    L.fbh_predict_rows_device.argtypes = [vp, vp, vp, i, d, vp, vp]
    L.fbh_predict_rows_device_noalloc.argtypes = [vp, vp, vp, i, d, vp, vp]
    L.fbh_predict_last_launch.argtypes = [vp, C.POINTER(i), C.POINTER(i)]
    L.fbh_trees_rank_form.argtypes = [vp]
    L.fbh_device_alloc.argtypes = [C.c_size_t, i, C.POINTER(vp)]
    L.fbh_device_free.argtypes = [vp, i]
    L.fbh_device_free.restype = None
    L.fbh_device_copy.argtypes = [vp, vp, C.c_size_t, i, i]
    L.fbh_device_sync.argtypes = [i]
    L.fbh_batch_row_split.argtypes = [vp]
    L.fbh_batch_configure.argtypes = [vp, i, i, i, i]
    L.fbh_batch_set_order.argtypes = [vp, vp]
    L.fbh_batch_probit_redraw.argtypes = [vp, i, i]
    L.fbh_batch_probit_info.argtypes = [vp, C.POINTER(i), C.POINTER(i), C.POINTER(C.c_ulonglong)]
    L.fbh_debug_jump.argtypes = [vp, C.c_ulonglong, vp]
    L.fbh_batch_config.argtypes = [vp] + [C.POINTER(i)] * 5
    L.fbh_batch_autotune.argtypes = [vp, i, C.POINTER(Tune)]
    L.fbh_batch_autotune_cached.argtypes = [vp, i, C.c_char_p, C.POINTER(Tune)]
    L.fbh_debug_math.argtypes = [i, vp, vp, C.c_long, vp, vp, i]
    L.fbh_debug_libm_host.argtypes = [i, C.c_long, C.c_long, vp]
    L.fbh_debug_copy.argtypes = [C.c_long, i, C.POINTER(d)]
    L.fbh_mlp_create.argtypes = [vp, i, C.c_uint, i, C.POINTER(vp)]
    L.fbh_mlp_nparams.argtypes = [vp]
    L.fbh_mlp_nparams.restype = C.c_long
    L.fbh_mlp_params.argtypes = [vp, vp, vp]
    L.fbh_mlp_forward.argtypes = [vp, vp, i, vp]
    L.fbh_mlp_grad.argtypes = [vp, vp, vp, vp, i, vp, vp]
    L.fbh_mlp_train.argtypes = [vp, vp, vp, vp, i, i, i, C.c_float, C.c_uint, vp]
    L.fbh_mlp_timing.argtypes = [vp, C.POINTER(d), C.POINTER(i)]
    L.fbh_mlp_destroy.argtypes = [vp]
    L.fbh_mlp_destroy.restype = None
    L.fbh_batch_learned_setup.argtypes = [vp, vp, C.c_float, C.c_long]
    L.fbh_batch_run_learned.argtypes = [vp, i]
    L.fbh_batch_learned_pairs.argtypes = [vp]
    L.fbh_batch_learned_pairs.restype = C.c_long
    L.fbh_batch_learned_train.argtypes = [vp, i, i, C.c_float, C.c_uint, vp]
    L.fbh_batch_learned_timing.argtypes = [vp, C.POINTER(d), C.POINTER(i)]
    L.fbh_batch_learned_target.argtypes = [vp, i, C.c_float, C.c_ulonglong]
    L.fbh_batch_learned_holdout.argtypes = [vp, C.c_long]
    L.fbh_batch_learned_eval.argtypes = [vp, C.c_long, C.c_long, C.POINTER(d), C.POINTER(d), C.POINTER(d)]
    L.fbh_batch_learned_pairs_copy.argtypes = [vp, C.c_long, C.c_long, vp, vp, vp, vp]
    L.fbh_batch_learned_target_stats.argtypes = [vp, C.POINTER(d), C.POINTER(d)]
    L.fbh_debug_split_conditional.argtypes = [i, i, vp, vp, vp, vp, vp, i, i, i, vp, d, d, vp, i, vp, vp,
                                              C.POINTER(d)]
    L.fbh_prep_defaults.argtypes = [i, C.POINTER(PrepOpts)]
    L.fbh_prep_defaults.restype = None
    L.fbh_wbart_inputs.argtypes = [vp, i, i, vp, vp, vp, vp, C.POINTER(PrepOpts), C.POINTER(C.POINTER(Inputs))]
    L.fbh_pbart_inputs.argtypes = [vp, i, i, vp, vp, vp, C.POINTER(PrepOpts), C.POINTER(C.POINTER(Inputs))]
    L.fbh_inputs_free.argtypes = [C.POINTER(Inputs)]
    L.fbh_inputs_free.restype = None
    L.fbh_qnorm.argtypes = [d]
    L.fbh_qnorm.restype = d
    L.fbh_qchisq.argtypes = [d, d]
    L.fbh_qchisq.restype = d
    return L


def lib():
    """The loaded library (built on first use if absent)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                from . import build as _b
                _b.build()
            _lib = _declare(C.CDLL(LIB_PATH))
        return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().fbh_last_error().decode(errors="replace")
        raise FbhError(f"{what or 'fastbart_hip'} failed ({rc}): {msg}")


def ptr(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)
