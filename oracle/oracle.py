"""FP64 CPU oracle for the ALM search -- TEST INFRASTRUCTURE ONLY.

Three independent restatements of one f+grad evaluation (alm_vcs.py:1496-1533 `augmented`, alm_vcs.py:1299-1312
`CV`; est_dual_vcs.py:227-246 at rho = 0):
  * ``eval_c``      -- the C restatement in lb_oracle.c (reference loop structure, OpenMP), also the CPU baseline;
  * ``eval_numpy``  -- a vectorised numpy/scipy restatement (scipy CSR products, bincounts);
  * ``eval_dense``  -- brute force from the mathematical definitions on a dense X = V V^T (small n only):
                       h_QR = sum_{i in Q, j in R, i != j} d_i d_j X_ij, tau_t = X_ab + X_bc - X_ac,
                       g_iR = d_i sum_{j in R, j != i} d_j X_ij, C(y) assembled entry by entry.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this module; the product path never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblb_oracle.so")

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class OrInst(C.Structure):
    _fields_ = [("n", C.c_int64), ("r", C.c_int32), ("rowptr", _i64p), ("col", _i32p), ("w", _f64p), ("d", _f64p),
                ("L", C.c_double), ("C0", C.c_double),
                ("nlev", C.c_int32), ("labs", _i32p), ("P", _i32p), ("sig", _f64p), ("rq", _f64p), ("dsq", _f64p),
                ("pw", C.c_double),
                ("nvcs", C.c_int32), ("vlabs", _i32p), ("vP", _i32p), ("vrq", _f64p), ("vptr", _i64p),
                ("vcol", _i32p), ("vsig", _f64p), ("wv", C.c_double),
                ("T", C.c_int64), ("TI", _i32p), ("lam", _f64p)]


def build(force: bool = False):
    """Compile liblb_oracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(os.path.join(HERE, "lb_oracle.c")):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB)
        L.or_eval.argtypes = [C.POINTER(OrInst), _f64p, C.c_double, C.c_double, _f64p, _f64p, _f64p, _f64p, _f64p,
                              _i64p, _i64p]
        L.or_eval.restype = C.c_int
        L.or_num_threads.restype = C.c_int
        L.or_set_threads.argtypes = [C.c_int]
        L.or_incidence.argtypes = [C.c_int64, C.c_int64, _i32p, _i64p, _i64p]
        L.or_sweeps.argtypes = [C.POINTER(OrInst), _f64p, C.c_int, C.c_double, C.c_double, C.c_double, _i64p, C.c_int,
                                C.c_double, _i64p, _i64p]
        L.or_sweeps.restype = C.c_int
        L.or_separate.argtypes = [C.c_int, C.c_int64, _i64p, _i32p, C.c_int, _f64p, C.c_double, C.c_int64, C.c_int64,
                                  C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int, C.c_int, _i64p, _f64p, _f64p,
                                  _i64p]
        L.or_separate.restype = C.c_int64
        L.or_reduce.argtypes = [C.c_int64, C.c_int64, _i64p, C.POINTER(C.c_uint64), _i64p, _i64p, _i64p, _f64p, _i64p]
        L.or_reduce.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else C.cast(None, t)


class _Packed:
    """Contiguous arrays of an Instance in the oracle's C layout; keeps them alive while the struct is used."""

    def __init__(self, inst, r):
        n = inst.n
        self.rowptr = np.ascontiguousarray(inst.rowptr, np.int64)
        self.col = np.ascontiguousarray(inst.col, np.int32)
        self.w = None if inst.w is None else np.ascontiguousarray(inst.w, np.float64)
        self.d = np.ascontiguousarray(inst.d, np.float64)
        lv = inst.levels
        self.labs = np.ascontiguousarray(np.concatenate([l.lab for l in lv]) if lv else np.zeros(1), np.int32)
        self.P = np.array([l.P for l in lv] or [0], np.int32)
        self.sig = np.ascontiguousarray(np.concatenate([l.sig.ravel() for l in lv]) if lv else np.zeros(1))
        self.rq = np.ascontiguousarray(np.concatenate([l.rq for l in lv]) if lv else np.zeros(1))
        self.dsq = np.ascontiguousarray(np.concatenate([l.dsq for l in lv]) if lv else np.zeros(1))
        vc = inst.vcs
        self.vlabs = np.ascontiguousarray(np.concatenate([v.lab for v in vc]) if vc else np.zeros(1), np.int32)
        self.vP = np.array([v.P for v in vc] or [0], np.int32)
        self.vrq = np.ascontiguousarray(np.concatenate([v.rq for v in vc]) if vc else np.zeros(1))
        self.vptr = np.ascontiguousarray(np.concatenate([v.ptr for v in vc]) if vc else np.zeros(1), np.int64)
        self.vcol = np.ascontiguousarray(np.concatenate([v.col for v in vc]) if vc and sum(v.nnz for v in vc) else np.zeros(1), np.int32)
        self.vsig = np.ascontiguousarray(np.concatenate([v.sig for v in vc]) if vc and sum(v.nnz for v in vc) else np.zeros(1))
        self.TI = np.ascontiguousarray(inst.TI if inst.T else np.zeros((1, 3)), np.int32)
        if inst.T and (self.TI.min() < 0 or self.TI.max() >= n):
            raise ValueError("triangle vertex out of range")
        self.lam = np.ascontiguousarray(inst.lam if inst.T else np.zeros(1), np.float64)
        self.s = OrInst(n, r, _p(self.rowptr, _i64p), _p(self.col, _i32p), _p(self.w, _f64p), _p(self.d, _f64p),
                        float(inst.L), float(inst.C0),
                        len(lv), _p(self.labs, _i32p), _p(self.P, _i32p), _p(self.sig, _f64p), _p(self.rq, _f64p),
                        _p(self.dsq, _f64p), float(inst.pairlvl_weight),
                        len(vc), _p(self.vlabs, _i32p), _p(self.vP, _i32p), _p(self.vrq, _f64p), _p(self.vptr, _i64p),
                        _p(self.vcol, _i32p), _p(self.vsig, _f64p), float(inst.wv),
                        inst.T, _p(self.TI, _i32p), _p(self.lam, _f64p))


PART_NAMES = ["f", "quad", "zz_over_L", "level", "tri", "vcs", "primal", "C0", "level_nviol", "tri_gmin",
              "vcs_nviol", "lam_sum"]


def eval_c(inst, V, rho, rt=None, threads: int | None = None, incidence=None):
    """C oracle evaluation. Returns (parts[12], G, pp (concat P_l^2), pt (T), pv (concat nnz)). `incidence` =
    (cnt, codes) from incidence_c() skips rebuilding the triangle incidence."""
    L_ = lib()
    if threads:
        L_.or_set_threads(int(threads))
    V = np.ascontiguousarray(V, np.float64)
    n, r = V.shape
    assert n == inst.n
    rt = rho if rt is None else rt
    pk = _Packed(inst, r)
    parts = np.zeros(12)
    G = np.zeros((n, r))
    pp = np.zeros(max(1, sum(l.P * l.P for l in inst.levels)))
    pt = np.zeros(max(1, inst.T))
    pv = np.zeros(max(1, sum(v.nnz for v in inst.vcs)))
    ip, ic = (None, None) if incidence is None else incidence
    rc = L_.or_eval(C.byref(pk.s), _p(V, _f64p), float(rho), float(rt), _p(parts, _f64p), _p(G, _f64p),
                    _p(pp, _f64p), _p(pt, _f64p), _p(pv, _f64p), _p(ip, _i64p), _p(ic, _i64p))
    assert rc == 0
    return parts, G, pp[:sum(l.P * l.P for l in inst.levels)], pt[:inst.T], pv[:sum(v.nnz for v in inst.vcs)]


def sweeps_c(inst, V, orders, nbatch, rho, rt=None, damp=0.0, kap2=1.0, threads=None, incidence=None):
    """C oracle sweeps (batch-snapshot semantics). Returns the new V (the input is not modified)."""
    L_ = lib()
    if threads:
        L_.or_set_threads(int(threads))
    V = np.array(V, np.float64, order="C", copy=True)
    n, r = V.shape
    rt = rho if rt is None else rt
    orders = np.ascontiguousarray(np.atleast_2d(orders), np.int64)
    pk = _Packed(inst, r)
    ip, ic = (None, None) if incidence is None else incidence
    rc = L_.or_sweeps(C.byref(pk.s), _p(V, _f64p), len(orders), float(rho), float(rt), float(damp), _p(orders, _i64p),
                      int(nbatch), float(kap2), _p(ip, _i64p), _p(ic, _i64p))
    assert rc == 0
    return V


def separate_c(inst, V, fam=0, thresh=0.3, maxh=100, maxk=60, hub2=8, cap2=200, vtol=1e-3, cap=1000,
               nchunk=512, nb=2000, threads=None, with_hist=False):
    """Triangle separation (fam 0: wedges, separate_ref; fam 1: one-edge arm, separate1_ref). Returns
    (TI (m x 3 int64, (min, centre, max)), val, vmin[, hist (nchunk x nb)])."""
    L_ = lib()
    if threads:
        L_.or_set_threads(int(threads))
    V = np.ascontiguousarray(V, np.float64)
    n, k = V.shape
    out = np.zeros((max(1, cap), 3), np.int64)
    val = np.zeros(max(1, cap))
    vmin = np.zeros(1)
    hist = np.zeros((nchunk, nb), np.int64) if with_hist else None
    rowptr = np.ascontiguousarray(inst.rowptr, np.int64)
    col = np.ascontiguousarray(inst.col, np.int32)
    m = L_.or_separate(int(fam), n, _p(rowptr, _i64p), _p(col, _i32p), k, _p(V, _f64p), float(thresh), int(maxh),
                       int(maxk), int(hub2), int(cap2), float(vtol), int(cap), int(nchunk), int(nb), _p(out, _i64p),
                       _p(val, _f64p), _p(vmin, _f64p), _p(hist, _i64p))
    res = (out[:m].copy(), val[:m].copy(), float(vmin[0]))
    return res + (hist,) if with_hist else res


def incidence_c(n, TI):
    L_ = lib()
    TI = np.ascontiguousarray(TI, np.int32)
    cnt = np.zeros(n + 1, np.int64)
    codes = np.zeros(max(1, 3 * len(TI)), np.int64)
    L_.or_incidence(n, len(TI), _p(TI, _i32p), _p(cnt, _i64p), _p(codes, _i64p))
    return cnt, codes[:3 * len(TI)]


# ----------------------------------------------------------------------------------------------- numpy restatement
def _A(inst):
    data = np.ones(inst.nnz) if inst.w is None else inst.w
    return sp.csr_matrix((data, inst.col, inst.rowptr), shape=(inst.n, inst.n))


def _W(lab, d, P):
    n = len(lab)
    return sp.csr_matrix((d, (lab, np.arange(n))), shape=(P, n))


def eval_numpy(inst, V, rho, rt=None):
    """Vectorised restatement of augmented()/CV() (rho > 0) and est value()/CV() (rho = 0)."""
    rt = rho if rt is None else rt
    V = np.asarray(V, np.float64)
    n, k = V.shape
    d = inst.d
    L = inst.L
    A = _A(inst)
    AV = A @ V
    z = d @ V
    zz = float(z @ z)
    parts = np.zeros(12)
    parts[2] = zz / L
    parts[7] = inst.C0
    parts[11] = float(inst.lam.sum()) if inst.T else 0.0
    G = AV - np.outer(d, z) / L
    pps, levtot, levv = [], 0.0, 0
    for lv in inst.levels:
        Y = _W(lv.lab, d, lv.P) @ V
        if rho > 0:
            H = Y @ Y.T
            H[np.diag_indices(lv.P)] -= lv.dsq
            h = 0.5 * (H + H.T)
            rr = rho * np.outer(lv.rq, lv.rq) * inst.pairlvl_weight
            pp = np.maximum(0.0, lv.sig - rr * h)
            levtot += float(((pp * pp - lv.sig * lv.sig) / (2 * rr)).sum())
            levv += int((h < -1e-9).sum())
        else:
            pp = lv.sig.copy()
        pps.append(pp)
        G += d[:, None] * (pp @ Y)[lv.lab] - (np.diag(pp)[lv.lab] * d ** 2)[:, None] * V
    pt = inst.lam.copy() if inst.T else np.zeros(0)
    tritot, gmin = 0.0, np.inf
    if inst.T:
        a, b, c = (inst.TI[:, q].astype(np.int64) for q in range(3))
        Va, Vb, Vc = V[a], V[b], V[c]
        tau = np.einsum("ij,ij->i", Va, Vb) + np.einsum("ij,ij->i", Vb, Vc) - np.einsum("ij,ij->i", Va, Vc)
        g = 1.0 - tau
        if rho > 0:
            pt = np.maximum(0.0, inst.lam - rt * g)
            tritot = float(((pt * pt - inst.lam * inst.lam) / (2 * rt)).sum())
            gmin = float(g.min())
        h = 0.5 * pt[:, None]
        np.add.at(G, a, h * (Vc - Vb))
        np.add.at(G, b, -h * (Va + Vc))
        np.add.at(G, c, h * (Va - Vb))
    pvs, vcstot, vcsv = [], 0.0, 0
    for vl in inst.vcs:
        Y = _W(vl.lab, d, vl.P) @ V
        rows = np.repeat(np.arange(n), np.diff(vl.ptr))
        R = vl.col.astype(np.int64)
        s = np.einsum("ij,ij->i", V[rows], Y[R])
        di = d[rows]
        g = di * s
        own = R == vl.lab[rows]
        g[own] = di[own] * (s[own] - di[own])
        if rho > 0:
            rr = rho * (inst.wv * vl.rq[R] / (di * di))
            pv = np.maximum(0.0, vl.sig - rr * g)
            vcstot += float(((pv * pv - vl.sig * vl.sig) / (2 * rr)).sum())
        else:
            pv = vl.sig.copy()
        vcsv += int((g < -1e-9).sum())
        pvs.append(pv)
        S = sp.csr_matrix((pv, vl.col, vl.ptr), shape=(n, vl.P))
        ZX = S.T @ (d[:, None] * V)
        ownv = np.zeros(n)
        ownv[rows[own]] = pv[own]
        G += 0.5 * d[:, None] * (S @ Y + ZX[vl.lab]) - (ownv * d ** 2)[:, None] * V
    qa = float(np.einsum("ij,ij->", AV, V))
    if rho > 0:
        parts[1] = qa
        parts[0] = qa - zz / L - levtot - tritot - vcstot + inst.C0
    else:
        parts[1] = float(np.einsum("ij,ij->", G, V))
        parts[0] = inst.C0 + parts[11] + parts[1]
    parts[3], parts[4], parts[5] = levtot, tritot, vcstot
    parts[6] = qa - zz / L + inst.C0
    parts[8], parts[9], parts[10] = levv, gmin, vcsv
    pp_all = np.concatenate([p.ravel() for p in pps]) if pps else np.zeros(0)
    pv_all = np.concatenate(pvs) if pvs else np.zeros(0)
    return parts, G, pp_all, pt, pv_all


# ----------------------------------------------------------------------------------------------- dense brute force
def dense_C(inst, pp_list, pt, pv_list):
    """C(y) = B + sum_l S_l + sum_m S_vc,m + T(pt), assembled entry by entry from the definitions."""
    n = inst.n
    d = inst.d
    Cm = inst.dense_A() - np.outer(d, d) / inst.L
    for lv, pp in zip(inst.levels, pp_list):
        for i in range(n):
            for j in range(n):
                if i != j:
                    Cm[i, j] += d[i] * d[j] * pp[lv.lab[i], lv.lab[j]]
    for vl, pv in zip(inst.vcs, pv_list):
        sig = np.zeros((n, vl.P))
        for i in range(n):
            for e in range(vl.ptr[i], vl.ptr[i + 1]):
                sig[i, vl.col[e]] = pv[e]
        for i in range(n):
            for j in range(n):
                if i != j:
                    Cm[i, j] += d[i] * d[j] * 0.5 * (sig[i, vl.lab[j]] + sig[j, vl.lab[i]])
    for t in range(inst.T):
        a, b, c = (int(x) for x in inst.TI[t])
        w = pt[t]
        Cm[a, b] -= 0.5 * w; Cm[b, a] -= 0.5 * w
        Cm[b, c] -= 0.5 * w; Cm[c, b] -= 0.5 * w
        Cm[a, c] += 0.5 * w; Cm[c, a] += 0.5 * w
    return Cm


def eval_dense(inst, V, rho, rt=None):
    """Brute-force f and G from X = V V^T (n up to a few dozen)."""
    rt = rho if rt is None else rt
    V = np.asarray(V, np.float64)
    n = inst.n
    d = inst.d
    X = V @ V.T
    B = inst.dense_A() - np.outer(d, d) / inst.L
    obj = float((B * X).sum())
    pp_list, levtot = [], 0.0
    for lv in inst.levels:
        P = lv.P
        h = np.zeros((P, P))
        for i in range(n):
            for j in range(n):
                if i != j:
                    h[lv.lab[i], lv.lab[j]] += d[i] * d[j] * X[i, j]
        if rho > 0:
            rr = rho * np.outer(lv.rq, lv.rq) * inst.pairlvl_weight
            pp = np.maximum(0.0, lv.sig - rr * h)
            levtot += float(((pp ** 2 - lv.sig ** 2) / (2 * rr)).sum())
        else:
            pp = lv.sig.copy()
        pp_list.append(pp)
    pt = inst.lam.copy() if inst.T else np.zeros(0)
    tritot = 0.0
    if inst.T and rho > 0:
        tau = np.array([X[a, b] + X[b, c] - X[a, c] for a, b, c in inst.TI])
        pt = np.maximum(0.0, inst.lam - rt * (1.0 - tau))
        tritot = float(((pt ** 2 - inst.lam ** 2) / (2 * rt)).sum())
    pv_list, vcstot = [], 0.0
    for vl in inst.vcs:
        pv = np.zeros(vl.nnz)
        for i in range(n):
            for e in range(vl.ptr[i], vl.ptr[i + 1]):
                R = vl.col[e]
                g = d[i] * sum(d[j] * X[i, j] for j in range(n) if j != i and vl.lab[j] == R)
                if rho > 0:
                    rr = rho * inst.wv * vl.rq[R] / d[i] ** 2
                    pv[e] = max(0.0, vl.sig[e] - rr * g)
                    vcstot += (pv[e] ** 2 - vl.sig[e] ** 2) / (2 * rr)
                else:
                    pv[e] = vl.sig[e]
        pv_list.append(pv)
    Cm = dense_C(inst, pp_list, pt, pv_list)
    G = Cm @ V
    if rho > 0:
        f = obj - levtot - tritot - vcstot + inst.C0
    else:
        f = inst.C0 + (float(inst.lam.sum()) if inst.T else 0.0) + float((Cm * X).sum())
    return f, G, pp_list, pt, pv_list


# ----------------------------------------------------------------------------------------------- numpy sweeps
def sweeps_numpy(inst, V, orders, nbatch, rho, rt=None, damp=0.0, kap2=1.0, hook=None):
    """Vectorised restatement of the batched row steps with per-batch snapshot semantics (sweeps, alm_vcs.py:1360-1436;
    batch_step6 255-351; pair_all 378-415; kq_level 417-428; vcs_state / vcs_pre / vcs_post vcsops.py:127-213,
    308-323). `hook(V, z, Ys, vst)`, if given, sees the running aggregates after every batch (z, the pair levels'
    cluster sums, and per VCS level its cluster sums "Y", stored multipliers "st", "Z" and "cnt"), for tests that
    recompute them from V."""
    rt = rho if rt is None else rt
    V = np.array(V, np.float64, copy=True)
    n, k = V.shape
    d = inst.d
    L = inst.L
    A = _A(inst)
    rows_t = None
    if inst.T:
        a_, b_, c_ = (inst.TI[:, q].astype(np.int64) for q in range(3))
    for order in np.atleast_2d(orders):
        z = d @ V
        Ys = [_W(lv.lab, d, lv.P) @ V for lv in inst.levels]
        Kqs = []
        for lv, Y in zip(inst.levels, Ys):
            if rho > 0:
                H = Y @ Y.T
                H[np.diag_indices(lv.P)] -= lv.dsq
                r = np.outer(lv.rq, lv.rq) * inst.pairlvl_weight
                act = (lv.sig - rho * r * H > 0) * r
                y2 = (Y * Y).sum(1)
                Kqs.append(2 * rho * (act @ y2 + np.diag(act) * y2))
            else:
                Kqs.append(np.zeros(lv.P))
        vst = []
        for vl in inst.vcs:
            Y = _W(vl.lab, d, vl.P) @ V
            rows = np.repeat(np.arange(n), np.diff(vl.ptr))
            R = vl.col.astype(np.int64)
            own = R == vl.lab[rows]
            st = _vcs_eff(vl, V, Y, d, rows, R, own, rho, inst.wv)
            Z = np.zeros_like(Y)
            np.add.at(Z, R, (st * d[rows])[:, None] * V[rows])
            cnt = np.bincount(R, weights=(st > 0).astype(float), minlength=vl.P)
            vst.append(dict(Y=Y, rows=rows, R=R, own=own, st=st, Z=Z, cnt=cnt))
        for rows_b in np.array_split(np.asarray(order, np.int64), nbatch):
            if len(rows_b) == 0:
                continue
            Vs = V.copy()
            vi = Vs[rows_b]
            di = d[rows_b]
            g = A[rows_b] @ Vs - (di / L)[:, None] * (z[None, :] - di[:, None] * vi)
            kap = np.zeros(len(rows_b))
            for lv, Y, Kq in zip(inst.levels, Ys, Kqs):
                if rho > 0:
                    H = Y @ Y.T
                    H[np.diag_indices(lv.P)] -= lv.dsq
                    pp = np.maximum(0.0, lv.sig - rho * np.outer(lv.rq, lv.rq) * inst.pairlvl_weight * H)
                else:
                    pp = lv.sig
                M = pp @ Y
                Q = lv.lab[rows_b]
                g += di[:, None] * (M[Q] - (np.diag(pp)[Q] * di)[:, None] * vi)
                kap += di * di * Kq[Q]
            for vl, s_ in zip(inst.vcs, vst):
                Y = s_["Y"]
                sel = np.isin(s_["rows"], rows_b)
                er = s_["rows"][sel]
                eR = s_["R"][sel]
                eown = s_["own"][sel]
                pos = {int(i): t for t, i in enumerate(rows_b)}
                ix = np.array([pos[int(i)] for i in er], np.int64)
                s = np.einsum("ij,ij->i", Vs[er], Y[eR])
                dd = d[er]
                gg = np.where(eown, dd * (s - dd), dd * s)
                if rho > 0:
                    rr = inst.wv * vl.rq[eR] / (dd * dd)
                    p = vl.sig[sel] - rho * rr * gg
                    y2 = (Y[eR] ** 2).sum(1)
                    y2 = np.where(eown, y2 - 2 * dd * s + dd * dd, y2)
                    kp = np.where(p > 0, rho * rr * dd * dd * y2, 0.0)
                    p = np.maximum(p, 0.0)
                    np.add.at(kap, ix, kp)
                else:
                    p = vl.sig[sel]
                acc = np.zeros((len(rows_b), k))
                np.add.at(acc, ix, p[:, None] * Y[eR])
                pown = np.zeros(len(rows_b))
                sz = np.zeros(len(rows_b))
                pown[ix[eown]] = p[eown]
                sz[ix[eown]] = s_["st"][sel][eown]
                Ri = vl.lab[rows_b]
                g += 0.5 * di[:, None] * (acc - (pown * di)[:, None] * vi) + 0.5 * di[:, None] * (s_["Z"][Ri] - (sz * di)[:, None] * vi)
                if rho > 0:
                    kap += kap2 * rho * di * di * inst.wv * vl.rq[Ri] * (s_["cnt"][Ri] - (sz > 0))
            if inst.T:
                pos = np.full(n, -1, np.int64)
                pos[rows_b] = np.arange(len(rows_b))
                tau = (np.einsum("ij,ij->i", Vs[a_], Vs[b_]) + np.einsum("ij,ij->i", Vs[b_], Vs[c_])
                       - np.einsum("ij,ij->i", Vs[a_], Vs[c_]))
                pt = inst.lam - rt * (1.0 - tau)
                act = pt > 0
                for role, (own_, qv) in enumerate(((a_, Vs[b_] - Vs[c_]), (b_, Vs[a_] + Vs[c_]), (c_, Vs[b_] - Vs[a_]))):
                    m = act & (pos[own_] >= 0)
                    np.add.at(g, pos[own_[m]], -0.5 * pt[m][:, None] * qv[m])
                    np.add.at(kap, pos[own_[m]], rt * (qv[m] ** 2).sum(1))
            gn = np.linalg.norm(g, axis=1)
            w = 0.5 * kap + damp * gn
            q = g + w[:, None] * vi
            nv = np.linalg.norm(q, axis=1)
            ok = nv > 0
            vn = vi.copy()
            vn[ok] = q[ok] / nv[ok][:, None]
            DD = di[:, None] * (vn - vi)
            V[rows_b] = vn
            z = z + DD.sum(0)
            for lv, Y in zip(inst.levels, Ys):
                np.add.at(Y, lv.lab[rows_b], DD)
            for vl, s_ in zip(inst.vcs, vst):
                np.add.at(s_["Y"], vl.lab[rows_b], DD)
                sel = np.flatnonzero(np.isin(s_["rows"], rows_b))
                er = s_["rows"][sel]
                sn = _vcs_eff(vl, V, s_["Y"], d, er, s_["R"][sel], s_["own"][sel], rho, inst.wv, sig=vl.sig[sel])
                so = s_["st"][sel]
                posr = np.full(n, -1, np.int64)
                posr[rows_b] = np.arange(len(rows_b))
                vold = vi[posr[er]]
                np.add.at(s_["Z"], s_["R"][sel], d[er][:, None] * (sn[:, None] * V[er] - so[:, None] * vold))
                np.add.at(s_["cnt"], s_["R"][sel], (sn > 0).astype(float) - (so > 0).astype(float))
                s_["st"][sel] = sn
            if hook is not None:
                hook(V, z, Ys, vst)
    return V


def _vcs_eff(vl, V, Y, d, rows, R, own, rho, wv, sig=None):
    sig = vl.sig if sig is None else sig
    s = np.einsum("ij,ij->i", V[rows], Y[R])
    dd = d[rows]
    if rho <= 0:
        return sig.copy()
    g = np.where(own, dd * (s - dd), dd * s)
    with np.errstate(divide="ignore", invalid="ignore"):
        p = np.maximum(0.0, sig - rho * (wv * vl.rq[R] / (dd * dd)) * g)
    return np.where(dd > 0, p, sig)


# ----------------------------------------------------------------------------------------------- VCS separation
def vcs_separate_numpy(vl, V, d, wv, nadd, tol=1e-4, mode="pen"):
    """vcsops.separate (vcsops.py:395-436) with the full candidate set: for every row i and cluster R not in the
    working set, s = V_i . Y_R (minus d_i and vol minus d_i for the own cluster), violated if s / vol < -tol, score
    pen = wv rq_R s^2 | mean = -s/vol | mass = -d_i s; returns the nadd best (score desc, ties by (i, R)) and the
    statistics (#violated, violation mass, both over all entries and over the working set)."""
    n, k = V.shape
    P = vl.P
    Y = _W(vl.lab, d, P) @ V
    vol = np.maximum(np.bincount(vl.lab, weights=d, minlength=P), 1.0)
    G = V @ Y.T
    S = G.copy()
    VR = np.broadcast_to(vol, (n, P)).copy()
    S[np.arange(n), vl.lab] -= d
    VR[np.arange(n), vl.lab] -= d
    valid = VR > 0
    with np.errstate(divide="ignore", invalid="ignore"):
        xb = np.where(valid, S / np.where(valid, VR, 1.0), 0.0)
    viol = valid & (xb < -tol)
    inws = np.zeros((n, P), bool)
    rows = np.repeat(np.arange(n), np.diff(vl.ptr))
    inws[rows, vl.col] = True
    mass = -d[:, None] * S
    stats = dict(nviol=float(viol.sum()), mass=float(mass[viol].sum()), nviol_ws=float((viol & inws).sum()),
                 mass_ws=float(mass[viol & inws].sum()))
    cand = viol & ~inws
    if mode == "pen":
        sc = wv * vl.rq[None, :] * S * S
    elif mode == "mean":
        sc = -xb
    else:
        sc = mass
    cand &= sc > 0
    ii, RR = np.nonzero(cand)
    ss = sc[ii, RR]
    o = np.lexsort((RR, ii, -ss))[:nadd]
    return ii[o].astype(np.int64), RR[o].astype(np.int32), ss[o], stats


def vcs_update_ws(vl, keep, ni, nR):
    """update_ws / _merge (vcsops.py:439-464): drop entries with keep False, add (ni, nR) with sig 0 and age 0.
    Returns (ptr, col, sig, age, row_old_index_or_-1)."""
    n = len(vl.lab)
    rows_old = np.repeat(np.arange(n), np.diff(vl.ptr))
    ko = np.flatnonzero(keep)
    r = np.r_[rows_old[ko], np.asarray(ni, np.int64)]
    c = np.r_[vl.col[ko].astype(np.int64), np.asarray(nR, np.int64)]
    src = np.r_[ko, np.full(len(ni), -1)]
    o = np.lexsort((c, r))
    r, c, src = r[o], c[o], src[o]
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, r + 1, 1)
    ptr = np.cumsum(ptr)
    return ptr, c.astype(np.int32), src


# ----------------------------------------------------------------------------------------------- exact reduction
def reduce_hash(n):
    """The twin-filter hash weights of reduce_exact.py:31 (numpy PCG64, seed 11)."""
    rng = np.random.default_rng(11)
    return rng.integers(1, 2 ** 62, size=n, dtype=np.int64).astype(np.uint64)


def reduce_c(n, keys, rhash=None):
    """C restatement of reduce_exact.py:12-62. Returns dict(group_of, N, vol, keys, weights, leaf_unions,
    twin_unions, inside_edges)."""
    L_ = lib()
    keys = np.ascontiguousarray(keys, np.int64)
    m = len(keys)
    rh = np.ascontiguousarray(reduce_hash(n) if rhash is None else rhash, np.uint64)
    g = np.zeros(max(1, n), np.int64)
    vol = np.zeros(max(1, n), np.int64)
    kk = np.zeros(max(1, m), np.int64)
    w = np.zeros(max(1, m))
    st = np.zeros(5, np.int64)
    rc = L_.or_reduce(n, m, _p(keys, _i64p), rh.ctypes.data_as(C.POINTER(C.c_uint64)), _p(g, _i64p), _p(vol, _i64p),
                      _p(kk, _i64p), _p(w, _f64p), _p(st, _i64p))
    assert rc == 0
    N, nk = int(st[0]), int(st[1])
    return dict(group_of=g[:n], N=N, vol=vol[:N], keys=kk[:nk], weights=w[:nk], leaf_unions=int(st[2]),
                twin_unions=int(st[3]), inside_edges=int(st[4]))


def reduce_numpy(n, keys):
    """Independent restatement of the reduction's result: the groups are the connected components of the eligible
    edges (leaf edges; adjacent true twins with d^2 < 2m, d > 1, tested on the closed neighbourhoods as sets,
    without any hash), numbered by their smallest vertex in increasing order (scipy connected_components)."""
    import scipy.sparse.csgraph as cg
    keys = np.asarray(keys, np.int64)
    m = len(keys)
    u, v = keys // n, keys % n
    d = np.bincount(np.r_[u, v], minlength=n)
    nb = [set() for _ in range(n)]
    for a, b in zip(u.tolist(), v.tolist()):
        nb[a].add(b)
        nb[b].add(a)
    leaf = (d[u] == 1) | (d[v] == 1)
    twin = np.array([d[a] == d[b] and d[a] > 1 and float(d[a]) ** 2 < 2 * m and (nb[a] | {a}) == (nb[b] | {b})
                     for a, b in zip(u.tolist(), v.tolist())], bool) if m else np.zeros(0, bool)
    el = leaf | twin
    G = sp.csr_matrix((np.ones(int(el.sum())), (u[el], v[el])), shape=(n, n))
    nc, lab = cg.connected_components(G, directed=False)
    first = np.full(nc, n)
    np.minimum.at(first, lab, np.arange(n))
    order = np.argsort(first)
    rank = np.empty(nc, np.int64)
    rank[order] = np.arange(nc)
    grp = rank[lab]
    vol = np.bincount(grp, weights=d, minlength=nc).astype(np.int64)
    gu, gv = grp[u], grp[v]
    inside = gu == gv
    kk, w = np.unique(np.minimum(gu, gv)[~inside] * nc + np.maximum(gu, gv)[~inside], return_counts=True)
    return dict(group_of=grp, N=nc, vol=vol, keys=kk.astype(np.int64), weights=w.astype(np.float64),
                leaf_unions=int(leaf.sum()), twin_unions=int(twin.sum()), inside_edges=int(inside.sum()))
