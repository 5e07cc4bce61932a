"""ctypes binding of liblbhost.so (include/lbhost.h): the search's host bookkeeping between device calls (working-set
merges, triangle-pool filtering) and the stable argsorts of its set-up. Raises if the library has not been built."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liblbhost.so")
EXPORTS = ["lbh_ws_merge", "lbh_take_f64", "lbh_take_u8", "lbh_new_triangles", "lbh_argsort_bounded",
           "lbh_argsort_i64", "lbh_tie_order", "lbh_row_gather", "lbh_coo_sum", "lbh_sym_from_lower"]

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with `python -m lagbound_amd.build`")
        L = C.CDLL(LIBPATH)
        L.lbh_ws_merge.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _u8p, C.c_int64, _i64p, _i64p, _i64p, _i32p,
                                   _i64p]
        L.lbh_ws_merge.restype = C.c_int64
        L.lbh_take_f64.argtypes = [C.c_int64, _i64p, _f64p, C.c_int64, _f64p]
        L.lbh_take_u8.argtypes = [C.c_int64, _i64p, _u8p, C.c_int64, _u8p]
        L.lbh_new_triangles.argtypes = [C.c_int64, C.c_int64, _i64p, C.c_int64, _i64p, _i64p]
        L.lbh_new_triangles.restype = C.c_int64
        L.lbh_argsort_bounded.argtypes = [C.c_int64, _i64p, C.c_int64, _i64p]
        L.lbh_argsort_i64.argtypes = [C.c_int64, _i64p, _i64p]
        L.lbh_tie_order.argtypes = [C.c_int64, _f64p, _i64p, _i64p, _i64p]
        L.lbh_row_gather.argtypes = [C.c_int64, _i64p, _i64p, _i64p, _i64p]
        L.lbh_row_gather.restype = C.c_int64
        L.lbh_coo_sum.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p, _f64p, _f64p]
        L.lbh_coo_sum.restype = C.c_int64
        L.lbh_sym_from_lower.argtypes = [C.c_int64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p]
        L.lbh_sym_from_lower.restype = C.c_int64
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def ws_merge(n, P, ptr, col, keep, ni, nR):
    """(ptr, col, src) of the merged working set (lbh_ws_merge)."""
    ptr = np.ascontiguousarray(ptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    keep = np.ascontiguousarray(keep, np.uint8)
    ni = np.ascontiguousarray(ni, np.int64)
    nR = np.ascontiguousarray(nR, np.int64)
    tot = int(keep.sum()) + len(ni)
    po = np.empty(n + 1, np.int64)
    co = np.empty(tot, np.int32)
    so = np.empty(tot, np.int64)
    w = load().lbh_ws_merge(int(n), int(P), _p(ptr, _i64p), _p(col, _i32p), _p(keep, _u8p), len(ni), _p(ni, _i64p),
                            _p(nR, _i64p), _p(po, _i64p), _p(co, _i32p), _p(so, _i64p))
    if w != tot:
        raise ValueError("lbh_ws_merge: a row or cluster out of range")
    return po, co, so


def take(a, src, dtype):
    """out[j] = a[src[j]] for src[j] >= 0, 0 for a new entry (lbh_take_f64 / lbh_take_u8)."""
    src = np.ascontiguousarray(src, np.int64)
    a = np.ascontiguousarray(a, dtype)
    out = np.empty(len(src), dtype)
    if dtype == np.uint8:
        rc = load().lbh_take_u8(len(src), _p(src, _i64p), _p(a, _u8p), len(a), _p(out, _u8p))
    else:
        rc = load().lbh_take_f64(len(src), _p(src, _i64p), _p(a, _f64p), len(a), _p(out, _f64p))
    if rc:
        raise ValueError("lbh_take: an index past the source")
    return out


def new_triangles(n, cand, pool):
    """Indices of the first occurrence of every distinct candidate triangle absent from the pool, ascending key."""
    cand = np.ascontiguousarray(cand, np.int64).reshape(-1, 3)
    pool = np.ascontiguousarray(pool, np.int64).reshape(-1, 3)
    out = np.empty(len(cand), np.int64)
    w = load().lbh_new_triangles(int(n), len(cand), _p(cand, _i64p), len(pool), _p(pool, _i64p), _p(out, _i64p))
    if w < 0:
        raise ValueError("lbh_new_triangles: bad arguments")
    return out[:w]


def argsort_bounded(keys, kmax):
    """np.argsort(keys, kind="stable") for integer keys in [0, kmax), by a counting sort (lbh_argsort_bounded)."""
    keys = np.ascontiguousarray(keys, np.int64)
    out = np.empty(len(keys), np.int64)
    if load().lbh_argsort_bounded(len(keys), _p(keys, _i64p), int(max(kmax, 1)), _p(out, _i64p)):
        raise ValueError("lbh_argsort_bounded: a key outside [0, kmax)")
    return out


def argsort_i64(keys):
    """np.argsort(keys, kind="stable") for int64 keys, by a radix sort (lbh_argsort_i64)."""
    keys = np.ascontiguousarray(keys, np.int64)
    out = np.empty(len(keys), np.int64)
    if load().lbh_argsort_i64(len(keys), _p(keys, _i64p), _p(out, _i64p)):
        raise ValueError("lbh_argsort_i64: bad arguments")
    return out


def tie_order(sc, i, R):
    """np.lexsort((R, i, -sc)) for candidates already in descending score order (lbh_tie_order: only the runs of
    equal scores are sorted); the full lexsort if sc is not descending."""
    sc = np.ascontiguousarray(sc, np.float64)
    i = np.ascontiguousarray(i, np.int64)
    R = np.ascontiguousarray(R, np.int64)
    out = np.empty(len(sc), np.int64)
    rc = load().lbh_tie_order(len(sc), _p(sc, _f64p), _p(i, _i64p), _p(R, _i64p), _p(out, _i64p))
    if rc < 0:
        raise ValueError("lbh_tie_order: bad arguments")
    return out if rc == 0 else np.lexsort((R, i, -sc))


def row_gather(ptr, order):
    """(nptr, idx): the rows of a CSR taken in `order`, entries kept in row order (lbh_row_gather)."""
    ptr = np.ascontiguousarray(ptr, np.int64)
    order = np.ascontiguousarray(order, np.int64)
    n = len(order)
    if len(ptr) != n + 1:
        raise ValueError("row_gather: ptr must have len(order) + 1 entries")
    nptr = np.empty(n + 1, np.int64)
    idx = np.empty(int(ptr[-1] - ptr[0]) if n else 0, np.int64)
    w = load().lbh_row_gather(n, _p(ptr, _i64p), _p(order, _i64p), _p(nptr, _i64p), _p(idx, _i64p))
    if w < 0:
        raise ValueError("lbh_row_gather: bad arguments")
    return nptr, idx[:w]


def coo_sum(N, r, c, v=None, sums=True, abs_sums=False, counts=False):
    """(indptr, indices, sum, abs_sum, count) of COO triplets as a CSR with sorted rows and columns, duplicates combined
    in input order (lbh_coo_sum); the value arrays not asked for are None."""
    r = np.ascontiguousarray(r, np.int64)
    c = np.ascontiguousarray(c, np.int64)
    m = len(r)
    if len(c) != m or (v is not None and len(v) != m):
        raise ValueError("coo_sum: r, c and v must have one length")
    vp = None if v is None else np.ascontiguousarray(v, np.float64)
    indptr = np.empty(int(N) + 1, np.int64)
    idx = np.empty(m, np.int64)
    sv = np.empty(m) if sums else None
    sa = np.empty(m) if abs_sums else None
    sn = np.empty(m) if counts else None
    w = load().lbh_coo_sum(int(N), m, _p(r, _i64p), _p(c, _i64p), None if vp is None else _p(vp, _f64p),
                           _p(indptr, _i64p), _p(idx, _i64p), None if sv is None else _p(sv, _f64p),
                           None if sa is None else _p(sa, _f64p), None if sn is None else _p(sn, _f64p))
    if w < 0:
        raise ValueError("lbh_coo_sum: an index out of range")
    cut = (lambda a: None if a is None else a[:w])
    return indptr, idx[:w], cut(sv), cut(sa), cut(sn)


def sym_from_lower(N, indptr, indices, data):
    """(indptr, indices, data) of the full symmetric CSR of a lower-triangular CSR with sorted rows, by a bucket
    transpose (lbh_sym_from_lower)."""
    ptr = np.ascontiguousarray(indptr, np.int64)
    idx = np.ascontiguousarray(indices, np.int64)
    val = np.ascontiguousarray(data, np.float64)
    nnz = int(ptr[-1]) if len(ptr) else 0
    po = np.empty(int(N) + 1, np.int64)
    io = np.empty(2 * nnz, np.int64)
    vo = np.empty(2 * nnz)
    w = load().lbh_sym_from_lower(int(N), _p(ptr, _i64p), _p(idx, _i64p), _p(val, _f64p), _p(po, _i64p),
                                  _p(io, _i64p), _p(vo, _f64p))
    if w < 0:
        raise ValueError("lbh_sym_from_lower: not a lower-triangular CSR with ascending columns")
    return po, io[:w], vo[:w]
