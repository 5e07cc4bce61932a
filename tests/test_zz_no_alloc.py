"""lb_gram_noalloc (the Gram matrix of the device eigensolver without any allocation: partial sums in a static
device array) against lb_blk_gram on the same blocks, through the C ABI: identical results (the same kernels and the
same fixed summation order), on two calls in a row, the second with a smaller input, and sizes beyond the array
refused. The test allocates its device memory through the HIP runtime the library is linked against; the entry
point under test allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from lagbound_amd import _lib
from lagbound_amd import instances as I

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipFree.argtypes = [C.c_void_p]
        for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipDeviceSynchronize):
            f.restype = C.c_int
        self.h = h
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def empty(self, count):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), 8 * count) == 0
        self.live.append(p)
        return p

    def get(self, p, count):
        out = np.empty(count, np.float64)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, 8 * count, D2H) == 0
        return out

    def close(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _with_context(L, n, r, As, Bs, w):
    """lb_blk_gram on a context holding the same blocks (the existing entry point)."""
    ctx = C.c_void_p()
    assert L.lb_create(C.byref(ctx), 0, _lib.LB_F64, r) == 0
    try:
        inst = I.build_instance(n, np.arange(n - 1, dtype=np.int64) * n + np.arange(1, n))   # a path on n rows
        rc = L.lb_set_graph(ctx, n, inst.rowptr.ctypes.data_as(_lib._i64p),
                            np.ascontiguousarray(inst.col, np.int32).ctypes.data_as(_lib._i32p), None,
                            inst.d.ctypes.data_as(_lib._f64p), inst.L, 0.0)
        assert rc == 0, L.lb_last_error(ctx)
        assert L.lb_blk_alloc(ctx, len(As) + len(Bs), 1) == 0
        for k, X in enumerate(As + Bs):
            X = np.ascontiguousarray(X)
            assert L.lb_blk_set(ctx, k, X.ctypes.data_as(_lib._f64p)) == 0
        diag = -1
        if w is not None:
            assert L.lb_blk_set_diag(ctx, 0, np.ascontiguousarray(w).ctypes.data_as(_lib._f64p)) == 0
            diag = 0
        a = np.arange(len(As), dtype=np.int32)
        b = np.arange(len(As), len(As) + len(Bs), dtype=np.int32)
        out = np.empty((len(As) * r, len(Bs) * r))
        assert L.lb_blk_gram(ctx, len(As), a.ctypes.data_as(_lib._i32p), len(Bs), b.ctypes.data_as(_lib._i32p), diag,
                             out.ctypes.data_as(_lib._f64p)) == 0, L.lb_last_error(ctx)
        return out
    finally:
        L.lb_destroy(ctx)


def _no_alloc(L, hip, n, r, As, Bs, w):
    pa = (C.c_void_p * len(As))(*[hip.put(X) for X in As])
    pb = (C.c_void_p * len(Bs))(*[hip.put(X) for X in Bs])
    pw = hip.put(w) if w is not None else None
    m = len(As) * r * len(Bs) * r
    out = hip.empty(m)
    rc = L.lb_gram_noalloc(0, None, n, r, len(As), pa, len(Bs), pb, pw, out)
    assert rc == 0, rc
    return hip.get(out, m).reshape(len(As) * r, len(Bs) * r)


def test_gram_without_allocation_matches_context_gram():
    L = _lib.load()
    hip = _Hip()
    rng = np.random.default_rng(3)
    try:
        # first call: 3 x 2 blocks of rank 16 with a weight; second call: smaller n, 2 x 1 blocks, no weight
        for n, r, na, nb, weighted in ((5000, 16, 3, 2, True), (1200, 16, 2, 1, False)):
            As = [rng.standard_normal((n, r)) for _ in range(na)]
            Bs = [rng.standard_normal((n, r)) for _ in range(nb)]
            w = rng.random(n) + 0.5 if weighted else None
            print(f"[no-alloc] n={n} r={r} na={na} nb={nb} weighted={weighted}", flush=True)
            got = _no_alloc(L, hip, n, r, As, Bs, w)
            ref = _with_context(L, n, r, As, Bs, w)
            assert np.array_equal(got, ref)
            SA, SB = np.hstack(As), np.hstack(Bs)
            dense = SA.T @ ((w[:, None] if weighted else 1.0) * SB)
            assert np.abs(got - dense).max() <= 1e-12 * n
        # beyond the static array (na r = 6 x 32 = 192 > 96): refused before any device work
        blk = [hip.put(np.zeros((64, 32)))] * 6
        pa = (C.c_void_p * 6)(*blk)
        assert L.lb_gram_noalloc(0, None, 64, 32, 6, pa, 1, pa, None, hip.empty(1)) == -2
        assert L.lb_gram_noalloc(0, None, 64, 12, 1, pa, 1, pa, None, hip.empty(1)) == -1
    finally:
        hip.close()
