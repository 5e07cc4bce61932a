"""Host-side handle of one device context of liblbhip (include/lbhip.h).

`Engine` uploads an `Instance` (graph, pair levels, VCS working sets, triangle pool) once, keeps V and every
multiplier on the device, and exposes the search's device operations: the f+grad evaluation (the reference's
`augmented` + `CV`, alm_vcs.py:1496-1533, 1299-1312), the PHR multiplier update (alm_vcs.py:1704-1711,1723) and the
batched row-step sweeps. There is no CPU fallback: a missing library or device raises `LBError`.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as LB
from ._lib import LBError, ptr, _f64p, _i32p, _i64p
from .vertex_order import Relabeled

DTYPES = {"f32": LB.LB_F32, "bf16": LB.LB_BF16, "f64": LB.LB_F64}
PART_NAMES = ["f", "quad", "zz_over_L", "level", "tri", "vcs", "primal", "C0", "level_nviol", "tri_gmin",
              "vcs_nviol", "lam_sum"]


class _TimedLib:
    """The library with the wall time of every entry-point call added to `owner.lib_s` (most entry points end with a
    stream synchronisation, so this is the context's device time plus launch overhead)."""

    def __init__(self, lib, owner):
        self._lib = lib
        self._owner = owner

    def __getattr__(self, name):
        fn = getattr(self._lib, name)
        owner = self._owner

        def call(*args):
            t = time.perf_counter()
            try:
                return fn(*args)
            finally:
                owner.lib_s += time.perf_counter() - t
        return call


class DeviceEngine:
    """One device context, holding the instance in the numbering it was given."""

    def __init__(self, r: int = 16, dtype: str = "f32", device: int = 0):
        self.lib_s = 0.0
        self.n_pp = self.n_t = self.n_pv = 0          # sizes of what was uploaded (host buffers of the getters)
        self.L = _TimedLib(LB.load(), self)
        self.r = int(r)
        self.dtype = dtype
        self.ctx = C.c_void_p()
        rc = self.L.lb_create(C.byref(self.ctx), int(device), DTYPES[dtype], self.r)
        if rc != 0:
            raise LBError(f"lb_create(device={device}, dtype={dtype}, r={r}) failed with status {rc} "
                          "(no HIP device, or unsupported rank/dtype)")
        self.inst = None

    def close(self):
        if self.ctx:
            self.L.lb_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.lb_last_error(self.ctx)
            raise LBError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    # ------------------------------------------------------------------------------------------------ upload
    def set_tuning(self, heavy_edges=128, heavy_tris=128, seg_len=1024, row_grid=0):
        """Hub thresholds and segment length (take effect at the next set_graph / set_triangles); row_grid 0 keeps
        the current grid."""
        self._chk(self.L.lb_set_tuning(self.ctx, heavy_edges, heavy_tris, seg_len, row_grid), "lb_set_tuning")

    def set_option(self, name: str, value: int):
        self._chk(self.L.lb_set_option(self.ctx, name.encode(), int(value)), "lb_set_option")

    def set_instance(self, inst):
        self.batch_G = 0
        n = inst.n
        self._keep = []
        rowptr = np.ascontiguousarray(inst.rowptr, np.int64)
        col = np.ascontiguousarray(inst.col, np.int32)
        w = None if inst.w is None else np.ascontiguousarray(inst.w, np.float64)
        d = np.ascontiguousarray(inst.d, np.float64)
        self._chk(self.L.lb_set_graph(self.ctx, n, ptr(rowptr, _i64p), ptr(col, _i32p), ptr(w, _f64p), ptr(d, _f64p),
                                      float(inst.L), float(inst.C0)), "lb_set_graph")
        self.set_levels(inst)
        self.set_vcs(inst)
        self.set_triangles(inst.TI, inst.lam)
        self.inst = inst

    def set_levels(self, inst):
        lv = inst.levels
        self.n_pp = sum(l.P * l.P for l in lv)
        if not lv:
            self._chk(self.L.lb_set_levels(self.ctx, 0, None, None, None, None, None, 1.0), "lb_set_levels")
            return
        labs = np.ascontiguousarray(np.concatenate([l.lab for l in lv]), np.int32)
        P = np.array([l.P for l in lv], np.int32)
        sig = np.ascontiguousarray(np.concatenate([l.sig.ravel() for l in lv]), np.float64)
        rq = np.ascontiguousarray(np.concatenate([l.rq for l in lv]), np.float64)
        dsq = np.ascontiguousarray(np.concatenate([l.dsq for l in lv]), np.float64)
        self._chk(self.L.lb_set_levels(self.ctx, len(lv), ptr(labs, _i32p), ptr(P, _i32p), ptr(sig, _f64p),
                                       ptr(rq, _f64p), ptr(dsq, _f64p), float(inst.pairlvl_weight)), "lb_set_levels")

    def set_level_sig(self, sigs):
        sig = np.ascontiguousarray(np.concatenate([np.asarray(s, np.float64).ravel() for s in sigs]))
        self._chk(self.L.lb_set_level_sig(self.ctx, ptr(sig, _f64p)), "lb_set_level_sig")

    def set_vcs(self, inst):
        vc = inst.vcs
        self.n_pv = sum(v.nnz for v in vc)
        if not vc:
            self._chk(self.L.lb_set_vcs(self.ctx, 0, None, None, None, None, None, None, 1.0), "lb_set_vcs")
            return
        labs = np.ascontiguousarray(np.concatenate([v.lab for v in vc]), np.int32)
        P = np.array([v.P for v in vc], np.int32)
        rq = np.ascontiguousarray(np.concatenate([v.rq for v in vc]), np.float64)
        vptr = np.ascontiguousarray(np.concatenate([v.ptr for v in vc]), np.int64)
        col = np.ascontiguousarray(np.concatenate([v.col for v in vc]) if sum(v.nnz for v in vc) else np.zeros(1), np.int32)
        sig = np.ascontiguousarray(np.concatenate([v.sig for v in vc]) if sum(v.nnz for v in vc) else np.zeros(1), np.float64)
        self._chk(self.L.lb_set_vcs(self.ctx, len(vc), ptr(labs, _i32p), ptr(P, _i32p), ptr(rq, _f64p),
                                    ptr(vptr, _i64p), ptr(col, _i32p), ptr(sig, _f64p), float(inst.wv)), "lb_set_vcs")

    def set_vcs_sig(self, sigs):
        sig = np.ascontiguousarray(np.concatenate([np.asarray(s, np.float64) for s in sigs]))
        self._chk(self.L.lb_set_vcs_sig(self.ctx, ptr(sig, _f64p)), "lb_set_vcs_sig")

    def set_triangles(self, TI, lam):
        TI = np.ascontiguousarray(TI, np.int32).reshape(-1, 3)
        lam = np.ascontiguousarray(lam, np.float64)
        self._chk(self.L.lb_set_triangles(self.ctx, len(TI), ptr(TI, _i32p), ptr(lam, _f64p)), "lb_set_triangles")
        self.n_t = len(TI)

    def set_lam(self, lam):
        lam = np.ascontiguousarray(lam, np.float64)
        self._chk(self.L.lb_set_lam(self.ctx, ptr(lam, _f64p)), "lb_set_lam")

    def set_V(self, V):
        V = np.ascontiguousarray(V, np.float64)
        assert V.shape == (self.inst.n, self.r), (V.shape, self.inst.n, self.r)
        self._chk(self.L.lb_set_V(self.ctx, V.ctypes.data_as(C.c_void_p), LB.LB_F64), "lb_set_V")

    def set_row_order(self, order):
        """V / G transfers in the caller's numbering: device row k <-> caller row order[k] (None: identity)."""
        if order is None:
            self._chk(self.L.lb_set_row_order(self.ctx, None), "lb_set_row_order")
            return
        order = np.ascontiguousarray(order, np.int64)
        self._chk(self.L.lb_set_row_order(self.ctx, ptr(order, _i64p)), "lb_set_row_order")

    # ------------------------------------------------------------------------------------------------ RGD polish
    def polish_begin(self, V):
        """Start a device polish at V (caller order, n x r): returns (sum_i g_i, sum Rr.R) of the start."""
        V = np.ascontiguousarray(V, np.float64)
        assert V.shape == (self.inst.n, self.r), (V.shape, self.inst.n, self.r)
        out = np.zeros(2)
        self._chk(self.L.lb_polish_begin(self.ctx, ptr(V, _f64p), ptr(out, _f64p)), "lb_polish_begin")
        return float(out[0]), float(out[1])

    def polish_trial(self, alpha: float) -> float:
        """Evaluate the candidate normalise(V + alpha R): returns its sum_i g_i."""
        out = np.zeros(1)
        self._chk(self.L.lb_polish_trial(self.ctx, float(alpha), ptr(out, _f64p)), "lb_polish_trial")
        return float(out[0])

    def polish_accept(self):
        """Take the last candidate: returns (sum Rr.R, sy, ss, yy) of the new direction."""
        out = np.zeros(4)
        self._chk(self.L.lb_polish_accept(self.ctx, ptr(out, _f64p)), "lb_polish_accept")
        return tuple(float(x) for x in out)

    def polish_get(self):
        """(V, g) of the current iterate, caller order."""
        V = np.empty((self.inst.n, self.r), np.float64)
        g = np.empty(self.inst.n, np.float64)
        self._chk(self.L.lb_polish_get(self.ctx, ptr(V, _f64p), ptr(g, _f64p)), "lb_polish_get")
        return V, g

    def get_V(self) -> np.ndarray:
        out = np.empty((self.inst.n, self.r), np.float64)
        self._chk(self.L.lb_get_V(self.ctx, out.ctypes.data_as(C.c_void_p), LB.LB_F64), "lb_get_V")
        return out

    # ------------------------------------------------------------------------------------------------ eigensolver blocks
    # n x r FP64 blocks and n-vectors kept on the device for the pencil LOBPCG (lb_blk_*; FP64 contexts)
    def blk_alloc(self, nblk, ndiag):
        self._chk(self.L.lb_blk_alloc(self.ctx, int(nblk), int(ndiag)), "lb_blk_alloc")

    def blk_set(self, k, X):
        X = np.ascontiguousarray(X, np.float64)
        assert X.shape == (self.inst.n, self.r), X.shape
        self._chk(self.L.lb_blk_set(self.ctx, int(k), ptr(X, _f64p)), "lb_blk_set")

    def blk_get(self, k):
        out = np.empty((self.inst.n, self.r), np.float64)
        self._chk(self.L.lb_blk_get(self.ctx, int(k), ptr(out, _f64p)), "lb_blk_get")
        return out

    def blk_set_diag(self, k, w):
        w = np.ascontiguousarray(w, np.float64)
        assert w.shape == (self.inst.n,), w.shape
        self._chk(self.L.lb_blk_set_diag(self.ctx, int(k), ptr(w, _f64p)), "lb_blk_set_diag")

    def blk_op(self, dst, src, diag):
        self._chk(self.L.lb_blk_op(self.ctx, int(dst), int(src), int(diag)), "lb_blk_op")

    def blk_elem(self, kind, dst, src, other=-1, diag=-1, theta=None):
        th = None if theta is None else np.ascontiguousarray(theta, np.float64)
        self._chk(self.L.lb_blk_elem(self.ctx, int(kind), int(dst), int(src), int(other), int(diag), ptr(th, _f64p)),
                  "lb_blk_elem")

    def blk_gram(self, a, b, diag=-1):
        a = np.ascontiguousarray(a, np.int32)
        b = np.ascontiguousarray(b, np.int32)
        out = np.empty((len(a) * self.r, len(b) * self.r), np.float64)
        self._chk(self.L.lb_blk_gram(self.ctx, len(a), ptr(a, _i32p), len(b), ptr(b, _i32p), int(diag), ptr(out, _f64p)),
                  "lb_blk_gram")
        return out

    def blk_comb(self, dst, src, coef, accumulate=False):
        src = np.ascontiguousarray(src, np.int32)
        coef = np.ascontiguousarray(coef, np.float64)
        assert coef.shape == (len(src) * self.r, self.r), coef.shape
        self._chk(self.L.lb_blk_comb(self.ctx, int(dst), len(src), ptr(src, _i32p), ptr(coef, _f64p), int(accumulate)),
                  "lb_blk_comb")

    # ------------------------------------------------------------------------------------------------ compute
    def eval(self, rho: float, rt: float | None = None, fetch: bool = True):
        rt = rho if rt is None else rt
        if not fetch:
            self._chk(self.L.lb_eval_fgrad(self.ctx, float(rho), float(rt), None), "lb_eval_fgrad")
            return None
        parts = np.zeros(LB.LB_NPARTS)
        self._chk(self.L.lb_eval_fgrad(self.ctx, float(rho), float(rt), ptr(parts, _f64p)), "lb_eval_fgrad")
        return parts

    def get_G(self) -> np.ndarray:
        out = np.empty((self.inst.n, self.r), np.float64)
        self._chk(self.L.lb_get_G(self.ctx, out.ctypes.data_as(C.c_void_p), LB.LB_F64), "lb_get_G")
        return out

    def get_multipliers(self):
        """(pp, pt, pv) of the last evaluation. The buffers are sized by what this engine uploaded (n_pp, n_t, n_pv),
        which is what the library writes: the C call takes no sizes."""
        pp = np.zeros(max(1, self.n_pp))
        pt = np.zeros(max(1, self.n_t))
        pv = np.zeros(max(1, self.n_pv))
        self._chk(self.L.lb_get_multipliers(self.ctx, ptr(pp, _f64p), ptr(pt, _f64p), ptr(pv, _f64p)),
                  "lb_get_multipliers")
        return pp[:self.n_pp], pt[:self.n_t], pv[:self.n_pv]

    def phr_update(self):
        self._chk(self.L.lb_phr_update(self.ctx), "lb_phr_update")

    def sweeps(self, orders, nbatch: int, rho: float, rt: float | None = None, damp: float = 0.0, kap2: float = 1.0):
        """len(orders) sweeps of batched row steps; orders: (nsw, n) permutations (int64)."""
        rt = rho if rt is None else rt
        orders = np.ascontiguousarray(np.atleast_2d(orders), np.int64)
        assert orders.shape[1] == self.inst.n
        self._chk(self.L.lb_sweeps(self.ctx, len(orders), float(rho), float(rt), float(damp), float(kap2),
                                   ptr(orders, _i64p), int(nbatch)), "lb_sweeps")

    def separate_triangles(self, fam=0, thresh=0.3, maxh=100, maxk=60, hub2=8, cap2=200, vtol=1e-3, cap=1000,
                           nchunk=512, nb=2000):
        """Triangle separation on the resident V (fam 0 wedges, fam 1 one-edge arms). Returns (TI int64 m x 3, val,
        vmin)."""
        out = np.zeros((max(1, cap), 3), np.int64)
        val = np.zeros(max(1, cap))
        cnt = np.zeros(1, np.int64)
        vmin = np.zeros(1)
        self._chk(self.L.lb_separate_triangles(self.ctx, int(fam), float(thresh), int(maxh), int(maxk), int(hub2),
                                               int(cap2), float(vtol), int(cap), int(nchunk), int(nb), ptr(out, _i64p),
                                               ptr(val, _f64p), ptr(cnt, _i64p), ptr(vmin, _f64p)),
                  "lb_separate_triangles")
        m = int(cnt[0])
        return out[:m].copy(), val[:m].copy(), float(vmin[0])

    def separate_vcs(self, m=0, nadd=1000, tol=1e-4, mode="pen"):
        """VCS separation of level m. Returns (i int64, R int32, score, stats dict)."""
        oi = np.zeros(max(1, nadd), np.int64)
        oR = np.zeros(max(1, nadd), np.int32)
        os_ = np.zeros(max(1, nadd))
        cnt = np.zeros(1, np.int64)
        st = np.zeros(4)
        md = {"pen": 0, "mean": 1, "mass": 2}[mode]
        self._chk(self.L.lb_separate_vcs(self.ctx, int(m), float(tol), md, int(nadd), ptr(oi, _i64p), ptr(oR, _i32p),
                                         ptr(os_, _f64p), ptr(cnt, _i64p), ptr(st, _f64p)), "lb_separate_vcs")
        k = int(cnt[0])
        return oi[:k].copy(), oR[:k].copy(), os_[:k].copy(), dict(nviol=st[0], mass=st[1], nviol_ws=st[2], mass_ws=st[3])

    def get_vcs_state(self):
        """(sig, age) of every VCS entry (concatenated over the levels)."""
        tot = self.n_pv
        sig = np.zeros(max(1, tot))
        age = np.zeros(max(1, tot), np.uint8)
        self._chk(self.L.lb_get_vcs_state(self.ctx, ptr(sig, _f64p), age.ctypes.data_as(C.c_void_p)), "lb_get_vcs_state")
        return sig[:tot], age[:tot]

    def set_vcs_age(self, age):
        age = np.ascontiguousarray(age, np.uint8)
        self._chk(self.L.lb_set_vcs_age(self.ctx, age.ctypes.data_as(C.c_void_p)), "lb_set_vcs_age")

    def set_weights(self, pairlvl_weight, vc_weight=1.0):
        self._chk(self.L.lb_set_weights(self.ctx, float(pairlvl_weight), float(vc_weight) * float(pairlvl_weight)),
                  "lb_set_weights")

    # ------------------------------------------------------------------------------------------------ batch mode
    def set_batch(self, row_off, L, C0):
        row_off = np.ascontiguousarray(row_off, np.int64)
        L = np.ascontiguousarray(L, np.float64)
        C0 = np.ascontiguousarray(C0, np.float64)
        self._chk(self.L.lb_batch_set(self.ctx, len(row_off) - 1, ptr(row_off, _i64p), ptr(L, _f64p), ptr(C0, _f64p)),
                  "lb_batch_set")
        self.batch_G = len(row_off) - 1

    def set_batch_levels(self, lab, P, sig, rq, dsq, pairlvl_weight=1.0):
        lab = np.ascontiguousarray(lab, np.int32)
        P = np.ascontiguousarray(P, np.int32)
        sig = np.ascontiguousarray(sig, np.float64)
        rq = np.ascontiguousarray(rq, np.float64)
        dsq = np.ascontiguousarray(dsq, np.float64)
        self._chk(self.L.lb_batch_set_levels(self.ctx, ptr(lab, _i32p), ptr(P, _i32p), ptr(sig, _f64p), ptr(rq, _f64p),
                                             ptr(dsq, _f64p), float(pairlvl_weight)), "lb_batch_set_levels")

    def batch_eval(self, rho, rt=None, fetch=True):
        """Evaluate every graph of the batch; returns parts (G x LB_NPARTS) or None (asynchronous)."""
        rt = rho if rt is None else rt
        if not fetch:
            self._chk(self.L.lb_batch_eval(self.ctx, float(rho), float(rt), None), "lb_batch_eval")
            return None
        parts = np.zeros((self.batch_G, LB.LB_NPARTS))
        self._chk(self.L.lb_batch_eval(self.ctx, float(rho), float(rt), ptr(parts, _f64p)), "lb_batch_eval")
        return parts

    def batch_rgd(self, iters, rho, rt=None, a0=0.3, amin=1e-4, amax=1.0, reset=False):
        """`iters` captured iterations of [batch evaluation, per-graph Riemannian Barzilai-Borwein ascent step];
        returns the per-graph state (G x 4: step, have-previous, <s,s>, <s,y>)."""
        rt = rho if rt is None else rt
        st = np.zeros((self.batch_G, 4))
        self._chk(self.L.lb_batch_rgd(self.ctx, int(iters), float(rho), float(rt), float(a0), float(amin),
                                      float(amax), int(bool(reset)), ptr(st, _f64p)), "lb_batch_rgd")
        return st

    def batch_get_sig(self, nsig):
        """The batch pair-level multipliers (concatenated P_g x P_g, nsig values)."""
        sig = np.zeros(max(1, int(nsig)))
        self._chk(self.L.lb_batch_get_sig(self.ctx, ptr(sig, _f64p)), "lb_batch_get_sig")
        return sig[:int(nsig)]

    def capture_eval(self, rho, rt=None):
        """Record the evaluation chain at (rho, rt) into a HIP graph (single graph or batch)."""
        rt = rho if rt is None else rt
        self._chk(self.L.lb_capture_eval(self.ctx, float(rho), float(rt)), "lb_capture_eval")

    def replay_eval(self, count=1, fetch=True):
        if not fetch:
            self._chk(self.L.lb_replay_eval(self.ctx, int(count), None), "lb_replay_eval")
            return None
        G = getattr(self, "batch_G", 0)
        parts = np.zeros((G, LB.LB_NPARTS) if G else LB.LB_NPARTS)
        self._chk(self.L.lb_replay_eval(self.ctx, int(count), ptr(parts, _f64p)), "lb_replay_eval")
        return parts

    def certify_dense(self, dg):
        """lb_certify_dense at diagonal dg: returns (info, sum R_ij^2, sum of squared off-diagonal magnitudes)."""
        dg = np.ascontiguousarray(dg, np.float64)
        assert len(dg) == self.inst.n
        out = np.zeros(3)
        self._chk(self.L.lb_certify_dense(self.ctx, ptr(dg, _f64p), ptr(out, _f64p)), "lb_certify_dense")
        return out

    def reduce_exact(self, n, keys, rhash=None):
        """Exact reduction of the graph with sorted unique keys u*n+v (u < v) on this context's device
        (reduce_exact.py:12-62). Returns dict(group_of, N, vol, keys, weights, leaf_unions, twin_unions,
        inside_edges)."""
        keys = np.ascontiguousarray(keys, np.int64)
        m = len(keys)
        g = np.zeros(max(1, n), np.int64)
        vol = np.zeros(max(1, n), np.int64)
        kk = np.zeros(max(1, m), np.int64)
        w = np.zeros(max(1, m))
        st = np.zeros(5, np.int64)
        rh = None if rhash is None else np.ascontiguousarray(rhash, np.uint64)
        self._chk(self.L.lb_reduce_exact(self.ctx, int(n), m, ptr(keys, _i64p),
                                         None if rh is None else rh.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         ptr(g, _i64p), ptr(vol, _i64p), ptr(kk, _i64p), ptr(w, _f64p),
                                         ptr(st, _i64p)), "lb_reduce_exact")
        N, nk = int(st[0]), int(st[1])
        return dict(group_of=g[:n].copy(), N=N, vol=vol[:N].copy(), keys=kk[:nk].copy(), weights=w[:nk].copy(),
                    leaf_unions=int(st[2]), twin_unions=int(st[3]), inside_edges=int(st[4]))

    def sync(self):
        self._chk(self.L.lb_sync(self.ctx), "lb_sync")

    def time_eval(self, rho: float, rt: float | None = None, iters: int = 20, per_kernel: bool = True):
        rt = rho if rt is None else rt
        ms = np.zeros(1)
        km = np.zeros(LB.LB_NKERNELS)
        self._chk(self.L.lb_time_eval(self.ctx, float(rho), float(rt), int(iters), ptr(ms, _f64p),
                                      ptr(km, _f64p) if per_kernel else None), "lb_time_eval")
        names = [self.L.lb_kernel_name(k).decode() for k in range(LB.LB_NKERNELS)]
        return float(ms[0]), dict(zip(names, km.tolist())) if per_kernel else None


DEFAULT_ORDER = "degree"


class Engine(Relabeled):
    """The device handle used by the search: a `DeviceEngine` holding the instance with its vertices numbered by
    descending degree (lagbound_amd.vertex_order), every argument and result in the caller's numbering.
    order="given" uploads the instance as given (batch mode packs graphs by their given numbering)."""

    def __init__(self, r: int = 16, dtype: str = "f32", device: int = 0, order: str | None = None):
        make = lambda: DeviceEngine(r=r, dtype=dtype, device=device)  # noqa: E731
        super().__init__(make(), make, order or DEFAULT_ORDER)


def cert_core(device, m, ai, aj, av, cptr, cidx, cval, cd, method=0):
    """The dense core of the lifted sparse certificate on the device (lb_cert_core): returns (info, sum_j (sum_i
    |R_ij|)^2) of the Cholesky factor R of the core's Schur complement (info 0: factored, k > 0: leading minor k not
    positive). method 0: rocSOLVER below 2^31 entries, the library's blocked kernels above; 1: rocSOLVER; 2: the
    blocked kernels (vector trailing updates); 3: the blocked kernels with matrix-core trailing updates."""
    L = LB.load()
    ai, aj = np.ascontiguousarray(ai, np.int32), np.ascontiguousarray(aj, np.int32)
    av, cd, cval = (np.ascontiguousarray(x, np.float64) for x in (av, cd, cval))
    cptr, cidx = np.ascontiguousarray(cptr, np.int64), np.ascontiguousarray(cidx, np.int32)
    out = np.zeros(2)
    rc = L.lb_cert_core(int(device), int(m), len(av), ptr(ai, _i32p), ptr(aj, _i32p), ptr(av, _f64p), len(cd),
                        ptr(cptr, _i64p), ptr(cidx, _i32p), ptr(cval, _f64p), ptr(cd, _f64p), int(method),
                        ptr(out, _f64p))
    if rc == -2:
        raise LBError(f"lb_cert_core: not enough free device memory for a dense core of {m} rows "
                      f"({8 * m * m / 2 ** 30:.1f} GiB)")
    if rc:
        raise LBError(f"lb_cert_core: status {rc}")
    return int(out[0]), float(out[1])
