"""Lifted sparse FP64 certificate of a candidate bound (the reference's certify stage, certify_lifted_vc.py, and the
PSD part of its verifier, verify_cert.py).

At fixed multipliers y >= 0 and a diagonal g, weak duality bounds L Q* by C0 + sum(lam) + sum(g) + sum(sdiag) once
Diag(g) + Diag(sdiag) - C(y) is shown positive semidefinite. The vertex-space matrix is dense (the rank-one d d^T / L
and the cluster terms couple every pair of vertices), so the proof runs on a LIFTED matrix Mt instead
(certify_lifted_vc.py:1-35, 255-387): besides the vertex vectors v_i it has auxiliary coordinates a_p = s_p / vol_p
for the nodes p of binary partial-sum trees over each clustering (s_p = sum of its children's s, s_i = d_i v_i), so that
for every partition (indeed every Gram matrix of unit vectors) the lifted Gram Xt is PSD, |Xt_kl| <= 1, and every link
e_p = vol_p e_p - c_1 e_(child 1) - c_2 e_(child 2) has e_p^T Xt e_p = 0. Then
    Mt = Diag(g (+) 0) - Ct + sum_p mu_p e_p e_p^T + Diag(sdiag),
with Ct the lifted image of C(y) (pair-level terms between cluster nodes, VCS terms between a vertex and a cluster node,
the rank-one term on the root, vertex-vertex terms as in C), satisfies <Mt, Xt> = sum(g) - <C, X> + sum sdiag_k Xt_kk,
and Mt PSD gives <C, X> <= sum(g) + sum(sdiag). With rounding: if fl(Mt) = L D L^T - E (D > 0), the bound grows by
sum |E_kl| plus the assembly error sum |fl(Mt) - Mt|, both priced entry by entry.

This module builds the trees and the terms (numpy), assembles fl(Mt) with its price, and factors it: a minimum-degree
order and the sparse LDL^T on the host (liblbcert.so), the dense core's Schur complement and Cholesky factor on the
device (lb_cert_core), or on the host for a small core when no device is given. The total is summed in exact
rational arithmetic and rounded up, as in lagbound_amd.bound.certify_dense.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from fractions import Fraction

import numpy as np
import scipy.sparse as sp

from .bound import check_multipliers, exact_sum, frac_up, gamma
from . import _host as HL

HERE = os.path.dirname(os.path.abspath(__file__))
CERTLIB = os.path.join(HERE, "liblbcert.so")
U64 = 2.0 ** -53
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_lib = None


def _cert_lib():
    global _lib
    if _lib is None:
        if not os.path.exists(CERTLIB):
            raise RuntimeError(f"{CERTLIB} is missing: build it with `python -m lagbound_amd.build`")
        L = C.CDLL(CERTLIB)
        L.lbc_factor.argtypes = [C.c_int64, _i64p, _i32p, _f64p, C.c_int64, C.c_int64, _i64p, C.POINTER(C.c_void_p)]
        L.lbc_md_order.argtypes = [C.c_int64, _i64p, _i32p, C.c_int64, C.c_int64, _i64p, _i64p]
        L.lbc_info.argtypes = [C.c_void_p, _i64p, _f64p]
        L.lbc_order.argtypes = [C.c_void_p, _i64p]
        L.lbc_core.argtypes = [C.c_void_p, _i32p, _i32p, _f64p, _i64p, _i32p, _f64p, _f64p]
        L.lbc_free.argtypes = [C.c_void_p]
        L.lbc_free.restype = None
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


# ------------------------------------------------------------------------------------------------ lifted trees
def _pair_up(grp, node, vol, key, nxt, links):
    """One round of the partial-sum trees of all groups at once: within each group (items sorted by key), items
    2t and 2t + 1 become the children of a new node (vol = sum, key = the smaller); an odd last item carries over."""
    o = np.lexsort((key, grp))
    grp, node, vol, key = grp[o], node[o], vol[o], key[o]
    start = np.r_[0, np.flatnonzero(np.diff(grp)) + 1]
    size = np.diff(np.r_[start, len(grp)])
    rank = np.arange(len(grp)) - np.repeat(start, size)
    last = np.repeat(start + size - 1, size)
    first = (rank % 2 == 0) & (np.arange(len(grp)) < last)          # left child of a pair
    a = np.flatnonzero(first)
    b = a + 1
    new = nxt + np.arange(len(a), dtype=np.int64)
    links.append((new, vol[a] + vol[b], np.c_[node[a], node[b]], np.c_[vol[a], vol[b]]))
    single = np.ones(len(grp), bool)
    single[a] = False
    single[b] = False
    keep = np.flatnonzero(single)
    return (np.r_[grp[a], grp[keep]], np.r_[new, node[keep]], np.r_[vol[a] + vol[b], vol[keep]],
            np.r_[np.minimum(key[a], key[b]), key[keep]], nxt + len(a))


def _trees(groups, node, vol, key, nxt, links):
    """Partial-sum trees over the items of every group; returns (group ids, top node, top volume, top key) per group
    and the next free node id."""
    while True:
        counts = np.bincount(groups)
        if len(groups) == 0 or counts.max() <= 1:
            break
        groups, node, vol, key, nxt = _pair_up(groups, node, vol, key, nxt, links)
    return groups, node, vol, key, nxt


class Lifted:
    """The lifted matrix of an instance at a diagonal g: N coordinates (n vertices first), the links (parent, vol,
    children (2), child coefficients (2)), the cluster -> node maps of every level and VCS level, the root, the
    shift weights w, and the lower-triangle COO terms of Mt without the shift."""

    def __init__(self, inst, g, key=None, mu_c=1e4, sa=0.01):
        n, d, L = inst.n, np.asarray(inst.d, np.float64), float(inst.L)
        if key is None:
            key = np.arange(n)
        self.n = n
        live = d > 0
        # distinct clusterings (a pair level and a VCS level with the same labels share one tree)
        labs = []

        def find(lab):
            lab = np.asarray(lab, np.int64)
            for c, x in enumerate(labs):
                if len(x) == len(lab) and np.array_equal(x, lab):
                    return c
            labs.append(lab)
            return len(labs) - 1

        lv_c = [find(lv.lab) for lv in inst.levels]
        vc_c = [find(v.lab) for v in inst.vcs]
        links = []
        nxt = n
        cl_nodes = []                   # per clustering: (node, coef) per cluster (node -1: no volume)
        for lab in labs:
            P = int(lab.max()) + 1 if len(lab) else 0
            idx = np.flatnonzero(live)
            grp, top, tvol, tkey, nxt = _trees(lab[idx], idx.astype(np.int64), d[idx], np.asarray(key)[idx], nxt,
                                               links)
            nodeP = np.full(P, -1, np.int64)
            coefP = np.zeros(P)
            keyP = np.zeros(P, np.int64)
            nodeP[grp] = top
            coefP[grp] = tvol
            keyP[grp] = tkey
            cl_nodes.append((nodeP, coefP, keyP))
        # the root: a tree over the cluster nodes of the clustering with the fewest clusters, or over all vertices
        if labs:
            c = int(np.argmin([len(x[0]) for x in cl_nodes]))
            nodeP, coefP, keyP = cl_nodes[c]
            ok = nodeP >= 0
            items = (np.zeros(int(ok.sum()), np.int64), nodeP[ok], coefP[ok], keyP[ok])
        else:
            idx = np.flatnonzero(live)
            items = (np.zeros(len(idx), np.int64), idx.astype(np.int64), d[idx], np.asarray(key)[idx])
        _, top, tvol, _, nxt = _trees(*items, nxt, links)
        self.root, self.rootvol = int(top[0]), float(tvol[0])
        if self.rootvol != L:
            raise ValueError(f"the root's volume {self.rootvol} is not L = {L}")
        self.N = nxt
        if links:
            self.link_p = np.concatenate([x[0] for x in links])
            self.link_vol = np.concatenate([x[1] for x in links])
            self.link_ch = np.concatenate([x[2] for x in links])
            self.link_cf = np.concatenate([x[3] for x in links])
        else:
            self.link_p = np.zeros(0, np.int64)
            self.link_vol = np.zeros(0)
            self.link_ch = np.zeros((0, 2), np.int64)
            self.link_cf = np.zeros((0, 2))
        self.lv_nodes = [cl_nodes[c][:2] for c in lv_c]
        self.vc_nodes = [cl_nodes[c][:2] for c in vc_c]
        self.mu_c = float(mu_c)
        self.w = np.r_[np.maximum(d, 1.0), np.full(self.N - n, float(sa))]
        self._terms(inst, np.asarray(g, np.float64))

    def _terms(self, inst, g):
        n, d, N = self.n, np.asarray(inst.d, np.float64), self.N
        R, Cc, V = [], [], []

        def put(r, c, v):
            r = np.asarray(r, np.int64)
            c = np.asarray(c, np.int64)
            v = np.broadcast_to(np.asarray(v, np.float64), r.shape)
            R.append(np.maximum(r, c))
            Cc.append(np.minimum(r, c))
            V.append(np.array(v))

        ar = np.arange(n)
        put(ar, ar, g)                                              # Diag(g) on the vertices
        put([self.root], [self.root], [self.rootvol * (self.rootvol / inst.L)])   # d d^T / L = L |a_root|^2
        if len(self.link_p):                                        # mu_p e_p e_p^T
            mu = self.mu_c / self.link_vol
            ids = np.c_[self.link_p, self.link_ch]
            co = np.c_[self.link_vol, -self.link_cf]
            for a in range(3):
                for b in range(a, 3):
                    put(ids[:, a], ids[:, b], mu * co[:, a] * co[:, b])
        for lv, (node, coef) in zip(inst.levels, self.lv_nodes):    # pair levels
            sig = np.asarray(lv.sig, np.float64)
            isaux = node >= n
            Q, Rr = np.nonzero(np.triu(sig))
            ok = (node[Q] >= 0) & (node[Rr] >= 0) & ~((Q == Rr) & ~isaux[Q])
            Q, Rr = Q[ok], Rr[ok]
            put(node[Q], node[Rr], -sig[Q, Rr] * coef[Q] * coef[Rr])
            lab = np.asarray(lv.lab, np.int64)
            sq = np.diag(sig)[lab]
            nz = (sq > 0) & (d > 0) & isaux[lab]
            put(ar[nz], ar[nz], sq[nz] * d[nz] * d[nz])
        for vl, (node, coef) in zip(inst.vcs, self.vc_nodes):       # VCS entries
            rows = np.repeat(np.arange(n), np.diff(vl.ptr))
            col = np.asarray(vl.col, np.int64)
            s = np.asarray(vl.sig, np.float64)
            lab = np.asarray(vl.lab, np.int64)
            own = col == lab[rows]
            keep = (s > 0) & (node[col] >= 0) & (d[rows] > 0) & ~(own & (node[col] < n))
            rows, col, s, own = rows[keep], col[keep], s[keep], own[keep]
            put(rows, node[col], -0.5 * s * d[rows] * coef[col])
            put(rows[own], rows[own], s[own] * d[rows[own]] * d[rows[own]])
        u = np.repeat(np.arange(n), np.diff(inst.rowptr))           # vertex-vertex: A and triangles (Mt = -C)
        wv = np.ones(inst.nnz) if inst.w is None else np.asarray(inst.w, np.float64)
        up = u < inst.col
        put(u[up], inst.col[up], -wv[up])
        if inst.T:
            TI = np.asarray(inst.TI, np.int64)
            lam = np.asarray(inst.lam, np.float64)
            h = 0.5 * lam
            nz = lam > 0
            i, j, k = TI[nz, 0], TI[nz, 1], TI[nz, 2]
            put(i, j, h[nz])
            put(j, k, h[nz])
            put(i, k, -h[nz])
        self.R = np.concatenate(R)
        self.C = np.concatenate(Cc)
        self.V = np.concatenate(V)

    def assemble(self, sdiag):
        """fl(Mt) with the diagonal shift sdiag (N values, used exactly as given): the lower triangle as CSR, and the
        assembly price sum_kl gamma(h_kl + 5) sum |terms_kl| over the full matrix (off-diagonal entries twice), h_kl
        the number of terms of entry kl: each term is a product of at most three floats and each entry a sum of h
        terms, so |fl(Mt)_kl - Mt_kl| <= gamma(h_kl + 5) sum |terms_kl| whatever the summation order."""
        N = self.N
        r = np.r_[self.R, np.arange(N)]
        c = np.r_[self.C, np.arange(N)]
        v = np.r_[self.V, np.asarray(sdiag, np.float64)]
        if not np.isfinite(v).all():
            raise ValueError("non-finite term in the lifted matrix")
        it = np.int32 if N < 2 ** 31 - 1 else np.int64
        # one pass (liblbhost): the entries' sums, the sums of their terms' absolute values and their term counts
        indptr, indices, sm, sab, cnt = HL.coo_sum(N, r, c, v, abs_sums=True, counts=True)
        M = sp.csr_matrix((sm, indices.astype(it), indptr.astype(it)), shape=(N, N))
        rows = np.repeat(np.arange(N), np.diff(indptr))
        mult = np.where(rows == indices, 1.0, 2.0)
        hmax = float(cnt.max())
        gam = (cnt + 5.0) * U64 / (1.0 - (cnt + 5.0) * U64)
        price = float(np.sum(mult * gam * sab)) * (1.0 + 4.0 * len(sab) * U64) * 1.0001
        return M, price, hmax


def full_symmetric(Ml):
    """The full symmetric CSR (both triangles and the diagonal) of a lower-triangle CSR with sorted rows, without
    arithmetic (a bucket transpose in liblbhost: no sort)."""
    N = Ml.shape[0]
    indptr, indices, data = HL.sym_from_lower(N, Ml.indptr, Ml.indices, Ml.data)
    it = np.int32 if N < 2 ** 31 - 1 else np.int64
    return sp.csr_matrix((data, indices.astype(it), indptr.astype(it)), shape=Ml.shape)


def vertex_key(inst, dense_deg=1000, core_deg=256):
    """Tree locality key of the vertices: their position in a minimum-degree order of the vertex graph (edges and
    triangle pairs), so that vertices eliminated close together share tree nodes (the reference orders them by a
    nested-dissection position, certify_lifted_vc.py:259-260)."""
    n = inst.n
    u = np.repeat(np.arange(n), np.diff(inst.rowptr))
    r, c = [u], [np.asarray(inst.col, np.int64)]
    if inst.T:
        TI = np.asarray(inst.TI, np.int64)
        r += [TI[:, 0], TI[:, 2]]
        c += [TI[:, 2], TI[:, 0]]
    r, c = np.concatenate(r), np.concatenate(c)
    # the symmetric pattern of the pairs, rows and columns sorted (liblbhost: one pass, no values)
    ptr, idx64, _, _, _ = HL.coo_sum(n, np.r_[r, c], np.r_[c, r], sums=False)
    order = np.zeros(n, np.int64)
    ns = np.zeros(1, np.int64)
    idx = idx64.astype(np.int32)
    rc = _cert_lib().lbc_md_order(n, _p(ptr, _i64p), _p(idx, _i32p), dense_deg, core_deg, _p(order, _i64p),
                                  _p(ns, _i64p))
    if rc:
        raise RuntimeError("vertex ordering failed")
    key = np.empty(n, np.int64)
    key[order] = np.arange(n)
    return key


# ------------------------------------------------------------------------------------------------ factorization
def _dense_core_host(m, ai, aj, av, cptr, cidx, cval, cd):
    """The core on the host (small cores, no device): S = A22 - L21 D1 L21^T, its Cholesky factor R and the sum of its
    column sums of |R|, squared. Returns (info, sum, 0) with info = 0 (factored) or k > 0."""
    S = np.zeros((m, m))
    np.add.at(S, (ai, aj), av)
    S = S + np.tril(S, -1).T
    nsp = len(cd)
    for k in range(nsp):
        c = cidx[cptr[k]:cptr[k + 1]]
        if len(c):
            v = cval[cptr[k]:cptr[k + 1]]
            S[np.ix_(c, c)] -= cd[k] * np.outer(v, v)
    try:
        Rm = np.linalg.cholesky(S)
    except np.linalg.LinAlgError:
        return 1, 0.0
    cs = np.abs(Rm).sum(axis=0)
    return 0, float(np.dot(cs, cs))


def factor_price(Ml, device=None, dense_deg=1000, core_deg=512, max_core=170_000, order=None, core_method=0,
                 log=None):
    """Prove fl(Mt) = L D L^T + R R^T - E with D > 0 and bound sum |E_kl|. Returns (ok, price, info). The price is
    gamma(h + 4) (sum |A| + sum_k D_k (sum_i |L_ik|)^2) for the sparse part and the Schur complement (h: most products
    in one entry, plus one) and gamma(m + 1) sum_k (sum_i |R_ik|)^2 for the dense core factor R (m rows; Higham,
    Accuracy and Stability of Numerical Algorithms, Thm 10.3 with |E| <= gamma(m + 1) |R| |R^T|), each evaluated with
    an upward allowance for its own rounding."""
    L = _cert_lib()
    F = full_symmetric(Ml)
    N = F.shape[0]
    ptr = F.indptr.astype(np.int64)
    idx = F.indices.astype(np.int32)
    val = F.data.astype(np.float64)
    t0 = time.time()
    h = C.c_void_p()
    rc = L.lbc_factor(N, _p(ptr, _i64p), _p(idx, _i32p), _p(val, _f64p), int(dense_deg), int(core_deg),
                      None if order is None else _p(np.ascontiguousarray(order, np.int64), _i64p), C.byref(h))
    if rc < 0:
        raise RuntimeError("lbc_factor: bad arguments")
    try:
        io = np.zeros(9, np.int64)
        do = np.zeros(3)
        L.lbc_info(h, _p(io, _i64p), _p(do, _f64p))
        info = dict(N=int(N), nnz=int(F.nnz), sparse_pivots=int(io[1]), core=int(io[2]), nnzL=int(io[3]),
                    hmax=int(io[4]), maxcol=int(io[5]), sparse_s=round(time.time() - t0, 3))
        if log:
            log(f"sparse LDL^T: {info['sparse_pivots']} pivots, nnz(L) {info['nnzL']}, core {info['core']} rows "
                f"({info['sparse_s']} s)")
        if rc > 0:
            info.update(failed="sparse pivot not positive", at=int(rc - 1))
            return False, None, info
        m, nsp = int(io[2]), int(io[1])
        sumA, sumLDL = float(do[0]), float(do[1])
        core_sum, core_info = 0.0, 0
        if m:
            if m > max_core:
                info.update(failed=f"dense core of {m} rows above the limit {max_core}")
                return False, None, info
            na, nc = int(io[6]), int(io[7])
            ai = np.empty(na, np.int32)
            aj = np.empty(na, np.int32)
            av = np.empty(na)
            cptr = np.empty(nsp + 1, np.int64)
            cidx = np.empty(nc, np.int32)
            cval = np.empty(nc)
            cd = np.empty(nsp)
            L.lbc_core(h, _p(ai, _i32p), _p(aj, _i32p), _p(av, _f64p), _p(cptr, _i64p), _p(cidx, _i32p),
                       _p(cval, _f64p), _p(cd, _f64p))
            t1 = time.time()
            if device is None:
                core_info, core_sum = _dense_core_host(m, ai, aj, av, cptr, cidx, cval, cd)
            else:
                from .engine import cert_core
                core_info, core_sum = cert_core(device, m, ai, aj, av, cptr, cidx, cval, cd, method=core_method)
            info["core_s"] = round(time.time() - t1, 3)
            info["core_on"] = "host" if device is None else "device"
            if core_info != 0:
                info.update(failed=f"dense core not positive definite (leading minor {core_info})")
                return False, None, info
    finally:
        L.lbc_free(h)
    hmax = info["hmax"]
    nterm = info["nnz"] + info["nnzL"] + N
    p_sparse = gamma(hmax + 4) * (sumA + sumLDL) * (1.0 + gamma(nterm + 8))
    p_core = gamma(m + 1) * core_sum * (1.0 + gamma(m * m + 8)) if m else 0.0
    underflow = 8.0 * (nterm + m * m) * 2.0 ** -1022                  # gradual underflow of the products
    price = (p_sparse + p_core + underflow) * 1.01
    info.update(p_sparse=p_sparse, p_core=p_core)
    return True, price, info


# ------------------------------------------------------------------------------------------------ certificate
def certify_sparse(inst, g, shift, sa=0.01, mu_c=1e4, device=None, key=None, dense_deg=1000, core_deg=512,
                   max_core=170_000, log=None, keep=False):
    """Prove L Q* <= C0 + sum(lam) + sum(g) + sum(sdiag) + price with the lifted matrix at sdiag = shift w (w = max(d,
    1) on the vertices, sa on the auxiliary nodes). Returns dict(passed, UB, ...); keep=True also returns the lifted
    structure, g and sdiag under "proof" (for certificate_record)."""
    check_multipliers(inst)
    t0 = time.time()
    if key is None:
        key = vertex_key(inst)
    lf = Lifted(inst, g, key=key, mu_c=mu_c, sa=sa)
    sdiag = float(shift) * lf.w
    Ml, p_asm, hmax_asm = lf.assemble(sdiag)
    ok, p_fact, info = factor_price(Ml, device=device, dense_deg=dense_deg, core_deg=core_deg, max_core=max_core,
                                    log=log)
    info.update(n=inst.n, N=lf.N, aux=lf.N - inst.n, links=len(lf.link_p), shift=float(shift),
                seconds=round(time.time() - t0, 2))
    if not ok:
        return dict(passed=False, reason=info.get("failed", "factorization failed"), info=info)
    proof = dict(lifted=lf, g=np.asarray(g, np.float64), sdiag=sdiag) if keep else None
    total = (Fraction(inst.C0) + (exact_sum(inst.lam) if inst.T else Fraction(0)) + exact_sum(g) + exact_sum(sdiag)
             + Fraction(p_asm) + Fraction(p_fact))
    UB = frac_up(total / Fraction(inst.L))
    res = dict(passed=True, UB=UB, price=(p_asm + p_fact) / inst.L, price_asm=p_asm / inst.L,
               price_fact=p_fact / inst.L, shift_part=float(exact_sum(sdiag)) / inst.L, info=info)
    if keep:
        res["proof"] = proof
    return res


def certify_estimate_sparse(inst, est, margin=1e-7, device=None, tries=8, log=None, **kw):
    """certify_sparse at the estimate's best round: shift = factor max(0, -lmin) + margin, starting at factor 1 and
    doubling factor and margin while a pivot fails (the lifted pencil's smallest eigenvalue differs from the vertex
    pencil's estimate; the reference likewise raises its shift until the factorization succeeds,
    certify_lifted_vc.py:556-568)."""
    base = max(0.0, -est["lmin"])
    factor, mg = 1.0, margin
    key = vertex_key(inst)
    res = None
    for k in range(tries):
        shift = factor * base + mg
        res = certify_sparse(inst, est["g"], shift, device=device, key=key, log=log, **kw)
        if log:
            log(f"sparse certificate at shift {shift:.4e}: " + (f"UB {res['UB']:.9f}" if res["passed"]
                                                               else res["reason"]))
        if res["passed"]:
            res["factor"], res["tries"] = factor, k + 1
            return res
        if "core" in res.get("reason", "") and "limit" in res.get("reason", ""):
            return res
        factor *= 2.0
        mg = max(mg * 10.0, 1e-6 * max(base, 1e-3))
    return res


# ------------------------------------------------------------------------------------------------ certificate record
def certificate_record(inst, res):
    """Everything an independent verifier needs, as plain arrays (the reference's --dump_cert,
    verify_cert.py:130-177): the graph, volumes, L, C0, the multipliers, g, sdiag, the link stiffness, the lifted
    trees (link parent, volume, children, child coefficients, root), the cluster -> node maps of every level and VCS
    level, and the certified UB. `res` is a passed certify_sparse result made with keep=True."""
    if not res.get("passed") or "proof" not in res:
        raise ValueError("certificate_record needs a passed certify_sparse(..., keep=True) result")
    pf = res["proof"]
    lf = pf["lifted"]
    rec = dict(n=np.int64(inst.n), N=np.int64(lf.N), L=np.float64(inst.L), C0=np.float64(inst.C0),
               d=np.asarray(inst.d, np.float64), rowptr=np.asarray(inst.rowptr, np.int64),
               col=np.asarray(inst.col, np.int64),
               w=np.zeros(0) if inst.w is None else np.asarray(inst.w, np.float64),
               TI=np.asarray(inst.TI, np.int64).reshape(-1, 3) if inst.T else np.zeros((0, 3), np.int64),
               lam=np.asarray(inst.lam, np.float64) if inst.T else np.zeros(0),
               g=pf["g"], sdiag=np.asarray(pf["sdiag"], np.float64), mu_c=np.float64(lf.mu_c),
               link_p=lf.link_p, link_vol=lf.link_vol, link_ch=lf.link_ch, link_cf=lf.link_cf,
               root=np.int64(lf.root), UB=np.float64(res["UB"]),
               nlev=np.int64(len(inst.levels)), nvcs=np.int64(len(inst.vcs)))
    for k, (lv, (node, coef)) in enumerate(zip(inst.levels, lf.lv_nodes)):
        rec[f"lv{k}_lab"] = np.asarray(lv.lab, np.int64)
        rec[f"lv{k}_sig"] = np.asarray(lv.sig, np.float64)
        rec[f"lv{k}_node"], rec[f"lv{k}_coef"] = np.asarray(node, np.int64), np.asarray(coef, np.float64)
    for k, (vl, (node, coef)) in enumerate(zip(inst.vcs, lf.vc_nodes)):
        rec[f"vc{k}_lab"] = np.asarray(vl.lab, np.int64)
        rec[f"vc{k}_ptr"] = np.asarray(vl.ptr, np.int64)
        rec[f"vc{k}_col"] = np.asarray(vl.col, np.int64)
        rec[f"vc{k}_sig"] = np.asarray(vl.sig, np.float64)
        rec[f"vc{k}_node"], rec[f"vc{k}_coef"] = np.asarray(node, np.int64), np.asarray(coef, np.float64)
    return rec


def write_certificate(path, inst, res):
    """certificate_record written as an .npz (plain arrays, loadable without pickles)."""
    np.savez(path, **certificate_record(inst, res))
