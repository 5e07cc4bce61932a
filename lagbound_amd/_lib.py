"""ctypes binding of liblbhip.so (include/lbhip.h). The product path has no CPU fallback: if the library is missing
or no HIP device is present, every entry point raises."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("LBHIP_LIB") or os.path.join(HERE, "liblbhip.so")   # LBHIP_LIB: a build variant to load

LB_F32, LB_BF16, LB_F64 = 0, 1, 2
LB_NPARTS = 12
LB_NKERNELS = 8

# every symbol include/lbhip.h declares
EXPORTS = ["lb_create", "lb_destroy", "lb_last_error", "lb_set_graph", "lb_set_levels", "lb_set_level_sig",
           "lb_set_vcs", "lb_set_vcs_sig", "lb_set_triangles", "lb_set_lam", "lb_set_V", "lb_get_V", "lb_set_row_order", "lb_polish_begin", "lb_polish_trial",
           "lb_polish_accept", "lb_polish_get",
           "lb_eval_fgrad", "lb_get_G", "lb_get_multipliers", "lb_phr_update", "lb_sweeps", "lb_separate_triangles", "lb_separate_vcs", "lb_set_vcs_age", "lb_get_vcs_state",
           "lb_set_weights", "lb_sync", "lb_time_eval",
           "lb_kernel_name", "lb_set_tuning", "lb_version", "lb_reduce_exact", "lb_set_option", "lb_certify_dense", "lb_batch_set",
           "lb_batch_set_levels", "lb_batch_eval", "lb_batch_rgd", "lb_batch_get_sig", "lb_capture_eval", "lb_replay_eval",
           "lb_blk_alloc", "lb_blk_set", "lb_blk_get", "lb_blk_set_diag", "lb_blk_op", "lb_blk_elem", "lb_blk_gram",
           "lb_blk_comb", "lb_gram_noalloc", "lb_cert_core"]

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p


class LBError(RuntimeError):
    pass


_lib = None


def load():
    """Load liblbhip.so (raises LBError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise LBError(f"{LIBPATH} is missing: build it with `python -m lagbound_amd.build` (hipcc, gfx950)")
    L = C.CDLL(LIBPATH)
    ctx = C.c_void_p
    sig = {
        "lb_create": ([C.POINTER(ctx), C.c_int, C.c_int, C.c_int], C.c_int),
        "lb_destroy": ([ctx], None),
        "lb_last_error": ([ctx], C.c_char_p),
        "lb_set_graph": ([ctx, C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_double, C.c_double], C.c_int),
        "lb_set_levels": ([ctx, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double], C.c_int),
        "lb_set_level_sig": ([ctx, _f64p], C.c_int),
        "lb_set_vcs": ([ctx, C.c_int, _i32p, _i32p, _f64p, _i64p, _i32p, _f64p, C.c_double], C.c_int),
        "lb_set_vcs_sig": ([ctx, _f64p], C.c_int),
        "lb_set_triangles": ([ctx, C.c_int64, _i32p, _f64p], C.c_int),
        "lb_set_lam": ([ctx, _f64p], C.c_int),
        "lb_set_V": ([ctx, _vp, C.c_int], C.c_int),
        "lb_get_V": ([ctx, _vp, C.c_int], C.c_int),
        "lb_eval_fgrad": ([ctx, C.c_double, C.c_double, _f64p], C.c_int),
        "lb_get_G": ([ctx, _vp, C.c_int], C.c_int),
        "lb_set_row_order": ([ctx, _i64p], C.c_int),
        "lb_polish_begin": ([ctx, _f64p, _f64p], C.c_int),
        "lb_polish_trial": ([ctx, C.c_double, _f64p], C.c_int),
        "lb_polish_accept": ([ctx, _f64p], C.c_int),
        "lb_polish_get": ([ctx, _f64p, _f64p], C.c_int),
        "lb_get_multipliers": ([ctx, _f64p, _f64p, _f64p], C.c_int),
        "lb_phr_update": ([ctx], C.c_int),
        "lb_sweeps": ([ctx, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _i64p, C.c_int], C.c_int),
        "lb_separate_vcs": ([ctx, C.c_int, C.c_double, C.c_int, C.c_int64, _i64p, _i32p, _f64p, _i64p, _f64p], C.c_int),
        "lb_set_vcs_age": ([ctx, C.c_void_p], C.c_int),
        "lb_get_vcs_state": ([ctx, _f64p, C.c_void_p], C.c_int),
        "lb_set_weights": ([ctx, C.c_double, C.c_double], C.c_int),
        "lb_sync": ([ctx], C.c_int),
        "lb_time_eval": ([ctx, C.c_double, C.c_double, C.c_int, _f64p, _f64p], C.c_int),
        "lb_kernel_name": ([C.c_int], C.c_char_p),
        "lb_set_tuning": ([ctx, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "lb_separate_triangles": ([ctx, C.c_int, C.c_double, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                                   C.c_int64, C.c_int, C.c_int, _i64p, _f64p, _i64p, _f64p], C.c_int),
        "lb_version": ([], C.c_char_p),
        "lb_set_option": ([ctx, C.c_char_p, C.c_int64], C.c_int),
        "lb_certify_dense": ([ctx, _f64p, _f64p], C.c_int),
        "lb_batch_set": ([ctx, C.c_int, _i64p, _f64p, _f64p], C.c_int),
        "lb_batch_set_levels": ([ctx, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double], C.c_int),
        "lb_batch_eval": ([ctx, C.c_double, C.c_double, _f64p], C.c_int),
        "lb_batch_rgd": ([ctx, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _f64p],
                         C.c_int),
        "lb_batch_get_sig": ([ctx, _f64p], C.c_int),
        "lb_capture_eval": ([ctx, C.c_double, C.c_double], C.c_int),
        "lb_replay_eval": ([ctx, C.c_int, _f64p], C.c_int),
        "lb_reduce_exact": ([ctx, C.c_int64, C.c_int64, _i64p, C.POINTER(C.c_uint64), _i64p, _i64p, _i64p, _f64p,
                             _i64p], C.c_int),
        "lb_blk_alloc": ([ctx, C.c_int, C.c_int], C.c_int),
        "lb_blk_set": ([ctx, C.c_int, _f64p], C.c_int),
        "lb_blk_get": ([ctx, C.c_int, _f64p], C.c_int),
        "lb_blk_set_diag": ([ctx, C.c_int, _f64p], C.c_int),
        "lb_blk_op": ([ctx, C.c_int, C.c_int, C.c_int], C.c_int),
        "lb_blk_elem": ([ctx, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f64p], C.c_int),
        "lb_blk_gram": ([ctx, C.c_int, _i32p, C.c_int, _i32p, C.c_int, _f64p], C.c_int),
        "lb_blk_comb": ([ctx, C.c_int, C.c_int, _i32p, _f64p, C.c_int], C.c_int),
        "lb_cert_core": ([C.c_int, C.c_int64, C.c_int64, _i32p, _i32p, _f64p, C.c_int64, _i64p, _i32p, _f64p, _f64p,
                          C.c_int, _f64p], C.c_int),
        "lb_gram_noalloc": ([C.c_int, _vp, C.c_int64, C.c_int, C.c_int, C.POINTER(_vp), C.c_int, C.POINTER(_vp),
                             _vp, _vp], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else C.cast(None, t)
