"""Independent verifier of a lifted modularity upper-bound certificate (the role of the reference's verify_cert.py:1-50,
verify at verify_cert.py:965-1020).

It imports nothing from the certifier (sparse_cert, bound, alm, engine). From the instance and the certificate record
(sparse_cert.write_certificate: multipliers, g, sdiag, link stiffness, the lifted trees and the cluster -> node maps,
the certifier's UB) it checks, in this order:

 1. the instance: a simple symmetric graph with positive weights, vol >= weighted degree, sum(vol) = L and
    sum(vol) - 2 sum_{i<j} w_ij = C0 (the contraction bookkeeping: C0 counts the edges inside contracted groups twice);
 2. the multipliers: finite and >= 0, pair-level sig symmetric, labels and indices in range, three distinct vertices
    per triangle; sdiag >= 0 and the link stiffness >= 0;
 3. the lifted trees: each auxiliary node is the parent of exactly one link, its children come before it, every child
    coefficient is the child's volume and vol_p is their sum (all volumes integers, so the sums are exact). Then every
    auxiliary vector x_p = s_p / vol_p is a convex combination of vertex rows (||x_p|| <= 1) and every link vanishes on
    the lifted Gram matrix. Expanding the root and every cluster node to its leaves: the root covers each vertex of
    positive volume once, and the node of cluster R covers exactly R's vertices of positive volume, once each, with
    coefficient vol(R);
 4. fl(Mt) assembled here from the constraint definitions (one lower-triangle term per constraint coefficient, summed
    per entry in sorted order) and its price sum_kl gamma(h_kl + 5) sum |terms_kl|;
 5. the lifting identity H^T (Mt - Diag(sdiag)) H = Diag(g) - C(y) (H the lifting map: vertex rows to lifted rows),
    with C(y) applied by a separate vertex-space implementation: on the identity for n <= 1000, on random vectors above;
 6. positive semidefiniteness: an up-looking LDL^T of fl(Mt) (liblbverify) with a per-entry running residual bound, in
    a minimum-degree order (the order is the one thing taken from the certifier's library, liblbcert: a permutation
    cannot make a false proof pass);
 7. UB = (C0 + sum lam + sum g + sum sdiag + price_asm + price_psd) / L in exact rationals, rounded up; each price is
    charged with an allowance for its own floating-point evaluation.

The claim (verify_cert.py:6-21): for every partition X, L Q(X) = C0 + <B, X> <= C0 + sum(lam) + <C(y), X>, since every
constraint family used (triangles lam (1 - X_ij - X_jk + X_ik) >= 0, pair-level sig_QR sum d_i d_j X_ij >= 0, VCS
s_iR d_i sum_{j in R, j != i} d_j X_ij >= 0) holds for every partition; and <C, X> <= sum(g) + sum(sdiag) + prices when
the lifted matrix is PSD up to the priced residual.

check_reduction re-checks a graph reduction from the original graph (verify_cert.py:874-959): the reduced weights,
volumes and C0 from the original edges, the eligibility of every contraction (leaf or twin rule, recomputed), and the
exact identity Q_orig(labels) = (C0 + <B', X'>) / L.

CLI (the reference's `verify_cert.py INSTANCE CERT --method ldl --orig ORIG --labels LAB --out R.json`):
    python -m lagbound_amd.verify [INSTANCE_DIR] CERT.npz [--orig ORIG_DIR --labels LAB.npy] [--out R.json]
exit status 0 = VERIFIED.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from fractions import Fraction

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
VERIFYLIB = os.path.join(HERE, "liblbverify.so")
ORDERLIB = os.path.join(HERE, "liblbcert.so")
UNIT = 2.0 ** -53
_i64p, _i32p, _f64p = C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)


class Rejected(Exception):
    """A check failed: the certificate is not verified."""


def need(cond, msg):
    if not cond:
        raise Rejected(msg)


def _gam(k):
    k = np.asarray(k, np.float64)
    return k * UNIT / (1.0 - k * UNIT)


def rational_sum(a) -> Fraction:
    """Exact sum of float64 values: each is m 2^e with integer m; values with equal e are summed as integers."""
    a = np.asarray(a, np.float64).ravel()
    a = a[a != 0]
    if not len(a):
        return Fraction(0)
    need(np.isfinite(a).all(), "non-finite value in an exact sum")
    m, e = np.frexp(a)
    mi = (m * 2.0 ** 53).astype(np.int64)
    e = e.astype(np.int64) - 53
    tot = Fraction(0)
    for ex in np.unique(e):
        tot += Fraction(int(mi[e == ex].astype(object).sum())) * Fraction(2) ** int(ex)
    return tot


def round_up(q: Fraction) -> float:
    """The smallest float64 >= q."""
    x = float(q)
    while Fraction(x) < q:
        x = float(np.nextafter(x, np.inf))
    return x


def _lib(path, what):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing ({what}): build it with `python -m lagbound_amd.build`")
    return C.CDLL(path)


# ------------------------------------------------------------------------------------------------ inputs
def load_instance_dir(D):
    """Instance directory in the reference's layout (graph.json, keys.npy u*n+v with u < v, optional weights.npy,
    vol.npy and meta.json with L_orig / inside_edges): n, the edge list, volumes, L and C0."""
    n = int(json.load(open(os.path.join(D, "graph.json")))["n"])
    keys = np.load(os.path.join(D, "keys.npy")).astype(np.int64)
    u, v = keys // n, keys % n
    w = np.load(os.path.join(D, "weights.npy")).astype(np.float64) if os.path.exists(os.path.join(D, "weights.npy")) \
        else np.ones(len(keys))
    if os.path.exists(os.path.join(D, "vol.npy")):
        d = np.load(os.path.join(D, "vol.npy")).astype(np.float64)
    else:
        d = np.bincount(np.r_[u, v], weights=np.r_[w, w], minlength=n)
    meta = json.load(open(os.path.join(D, "meta.json"))) if os.path.exists(os.path.join(D, "meta.json")) else {}
    L = float(meta.get("L_orig", d.sum()))
    C0 = 2.0 * float(meta.get("inside_edges", 0))
    return dict(n=n, u=u, v=v, w=w, d=d, L=L, C0=C0)


def load_certificate(path):
    """The certificate record as a dict of arrays (no pickled objects)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def instance_from_record(c):
    """The graph stored in the certificate record (unordered edge list)."""
    n = int(c["n"])
    rp, col = c["rowptr"].astype(np.int64), c["col"].astype(np.int64)
    row = np.repeat(np.arange(n), np.diff(rp))
    w = c["w"].astype(np.float64) if c["w"].size else np.ones(len(col))
    up = row < col
    return dict(n=n, u=row[up], v=col[up], w=w[up], d=c["d"].astype(np.float64), L=float(c["L"]), C0=float(c["C0"]),
                both_directions=(row, col, w))


# ------------------------------------------------------------------------------------------------ 1-3 structure
def check_instance(g):
    n, u, v, w, d = g["n"], g["u"], g["v"], g["w"], g["d"]
    need(len(d) == n and np.isfinite(d).all() and (d >= 0).all(), "volumes must be finite and >= 0")
    need(((u >= 0) & (u < v) & (v < n)).all(), "edges must be u < v < n")
    key = u * n + v
    need(len(np.unique(key)) == len(key), "repeated edge")
    need(np.isfinite(w).all() and (w > 0).all(), "edge weights must be positive")
    if "both_directions" in g:              # the stored CSR must hold every edge in both directions
        row, col, ww = g["both_directions"]
        need(len(row) == 2 * len(u) and not (row == col).any(), "the stored graph is not symmetric or has self-loops")
        lo, hi = np.minimum(row, col), np.maximum(row, col)
        k2 = lo * n + hi
        o1, o2 = np.argsort(k2, kind="stable"), np.argsort(key, kind="stable")
        need(np.array_equal(k2[o1][0::2], key[o2]) and np.array_equal(k2[o1][1::2], key[o2]),
             "the stored graph is not symmetric")
        need(np.array_equal(ww[o1][0::2], w[o2]) and np.array_equal(ww[o1][1::2], w[o2]), "asymmetric edge weights")
    wdeg = np.bincount(np.r_[u, v], weights=np.r_[w, w], minlength=n)
    need((d >= wdeg).all(), "a volume is below the weighted degree")
    need(rational_sum(d) == Fraction(g["L"]), "sum of the volumes differs from L")
    need(rational_sum(d) - 2 * rational_sum(w) == Fraction(g["C0"]), "sum(vol) - 2 sum(w) differs from C0")
    need(all(float(x).is_integer() for x in (g["L"], g["C0"])) and np.all(d == np.round(d)),
         "volumes, L and C0 must be integers")


def check_multipliers(c, n):
    T = c["TI"].reshape(-1, 3).astype(np.int64) if c["TI"].size else np.zeros((0, 3), np.int64)
    lam = c["lam"].astype(np.float64)
    need(len(lam) == len(T), "one lam per triangle")
    need(np.isfinite(lam).all() and (lam >= 0).all(), "triangle multipliers must be finite and >= 0")
    if len(T):
        need(((T >= 0) & (T < n)).all(), "triangle vertex out of range")
        need(((T[:, 0] != T[:, 1]) & (T[:, 1] != T[:, 2]) & (T[:, 0] != T[:, 2])).all(),
             "a triangle repeats a vertex")
    for k in range(int(c["nlev"])):
        lab, sig = c[f"lv{k}_lab"].astype(np.int64), c[f"lv{k}_sig"].astype(np.float64)
        need(len(lab) == n and (lab >= 0).all() and (lab < len(sig)).all(), f"pair level {k}: labels out of range")
        need(sig.ndim == 2 and sig.shape[0] == sig.shape[1] == len(c[f"lv{k}_node"]),
             f"pair level {k}: sig must be square, one row per cluster node")
        need(np.isfinite(sig).all() and (sig >= 0).all() and np.array_equal(sig, sig.T),
             f"pair level {k}: sig must be finite, >= 0 and symmetric")
    for k in range(int(c["nvcs"])):
        lab, ptr = c[f"vc{k}_lab"].astype(np.int64), c[f"vc{k}_ptr"].astype(np.int64)
        col, s = c[f"vc{k}_col"].astype(np.int64), c[f"vc{k}_sig"].astype(np.float64)
        P = len(c[f"vc{k}_node"])
        need(len(lab) == n and (lab >= 0).all() and (lab < P).all(), f"VCS level {k}: labels out of range")
        need(len(ptr) == n + 1 and ptr[0] == 0 and ptr[-1] == len(col) and (np.diff(ptr) >= 0).all(),
             f"VCS level {k}: bad row pointers")
        need(((col >= 0) & (col < P)).all(), f"VCS level {k}: cluster out of range")
        need(len(s) == len(col) and np.isfinite(s).all() and (s >= 0).all(), f"VCS level {k}: sig must be >= 0")
        rows = np.repeat(np.arange(n), np.diff(ptr))
        need(len(np.unique(rows * P + col)) == len(col), f"VCS level {k}: repeated (vertex, cluster) entry")
    sd = c["sdiag"].astype(np.float64)
    need(len(sd) == int(c["N"]) and np.isfinite(sd).all() and (sd >= 0).all(), "sdiag must be N finite values >= 0")
    need(np.isfinite(c["g"]).all() and len(c["g"]) == n, "g must be n finite values")
    need(float(c["mu_c"]) >= 0 and np.isfinite(float(c["mu_c"])), "the link stiffness must be >= 0")


def _expand(nodes, owner, n, c0, c1):
    """(owner, vertex) pairs of the leaves under each node (auxiliary ids >= n expand to their two children)."""
    nodes, owner = nodes.astype(np.int64), owner.astype(np.int64)
    out_o, out_v = [], []
    for _ in range(200):
        aux = nodes >= n
        out_o.append(owner[~aux])
        out_v.append(nodes[~aux])
        if not aux.any():
            return np.concatenate(out_o), np.concatenate(out_v)
        a = nodes[aux] - n
        nodes = np.r_[c0[a], c1[a]]
        owner = np.r_[owner[aux], owner[aux]]
    raise Rejected("lifted tree deeper than 200 levels")


def check_tree(c, d):
    """Structural checks of the lifted trees; returns the volume of every lifted node."""
    n, N = int(c["n"]), int(c["N"])
    lp = c["link_p"].astype(np.int64)
    lv = c["link_vol"].astype(np.float64)
    ch = c["link_ch"].reshape(-1, 2).astype(np.int64)
    cf = c["link_cf"].reshape(-1, 2).astype(np.float64)
    need(len(lp) == N - n and np.array_equal(np.sort(lp), np.arange(n, N)),
         "every auxiliary node must be the parent of exactly one link")
    need(((ch >= 0) & (ch < lp[:, None])).all(), "a link's children must come before its parent (acyclic)")
    vol = np.zeros(N)
    vol[:n] = d
    order = np.argsort(lp)
    lp, lv, ch, cf = lp[order], lv[order], ch[order], cf[order]
    vol[lp] = lv
    need(np.all(lv == np.round(lv)) and np.all(np.abs(lv) < 2.0 ** 52), "link volumes must be integers")
    need(np.array_equal(cf[:, 0], vol[ch[:, 0]]) and np.array_equal(cf[:, 1], vol[ch[:, 1]]),
         "a child coefficient differs from the child's volume")
    need(np.array_equal(lv, cf[:, 0] + cf[:, 1]) and (lv > 0).all(), "a link volume differs from its children's sum")
    c0 = np.full(N - n, -1, np.int64)
    c1 = np.full(N - n, -1, np.int64)
    c0[lp - n], c1[lp - n] = ch[:, 0], ch[:, 1]
    live = np.flatnonzero(d > 0)
    root = int(c["root"])
    need(0 <= root < N and vol[root] == float(c["L"]), "the root's volume must be L")
    _, leaves = _expand(np.array([root]), np.zeros(1), n, c0, c1)
    need(np.array_equal(np.sort(leaves), live), "the root must cover every vertex of positive volume once")
    maps = [(f"lv{k}", c[f"lv{k}_lab"]) for k in range(int(c["nlev"]))] + \
           [(f"vc{k}", c[f"vc{k}_lab"]) for k in range(int(c["nvcs"]))]
    for name, lab in maps:
        lab = lab.astype(np.int64)
        node, coef = c[f"{name}_node"].astype(np.int64), c[f"{name}_coef"].astype(np.float64)
        P = len(node)
        cvol = np.bincount(lab, weights=d, minlength=P)
        has = node >= 0
        need(len(cvol) == P and np.array_equal(coef[has], cvol[has]) and np.all(cvol[~has] == 0),
             f"{name}: a cluster coefficient differs from the cluster volume")
        need((node[has] < N).all() and np.array_equal(vol[node[has]], coef[has]),
             f"{name}: a cluster node's volume differs from its coefficient")
        own, lv_ = _expand(node[has], np.flatnonzero(has), n, c0, c1)
        o = np.argsort(lv_, kind="stable")
        need(np.array_equal(lv_[o], live) and np.array_equal(own[o], lab[live]),
             f"{name}: a cluster node does not cover exactly its cluster's vertices")
    return vol


# ------------------------------------------------------------------------------------------------ 4 assembly
def assemble(c, g, vol):
    """fl(Mt) (lower CSR) and price_asm, from the constraint definitions:
    Mt = Diag(g (+) 0) - lifted C(y) + sum_p mu_p e_p e_p^T + Diag(sdiag), where in lifted coordinates
      B = W - d d^T / L:     -w_ij at (i, j); +L at (root, root)   (d d^T / L = L x_root x_root^T, vol_root = L);
      triangle (i,j,k), lam: +lam/2 at (i,j) and (j,k), -lam/2 at (i,k)       (C holds -lam/2, -lam/2, +lam/2);
      pair level:            -sig_QR vol_Q vol_R at (node_Q, node_R), Q <= R, and +sig_QQ d_i^2 at (i, i) for the
                             i != j restriction (a one-vertex cluster's diagonal pair is empty and skipped);
      VCS entry (i, R, s):   -s d_i vol_R / 2 at (i, node_R), and +s d_i^2 at (i, i) when i is in R (skipped when
                             R = {i});
      link p (children a, b, coefficients ca, cb):  mu_c / vol_p (vol_p e_p - ca e_a - cb e_b)(...)^T."""
    n, N, L = g["n"], int(c["N"]), g["L"]
    d = g["d"]
    rows, cols, vals = [], [], []

    def term(r, cc, v):
        r, cc = np.asarray(r, np.int64), np.asarray(cc, np.int64)
        rows.append(np.maximum(r, cc))
        cols.append(np.minimum(r, cc))
        vals.append(np.broadcast_to(np.asarray(v, np.float64), r.shape).copy())

    ids = np.arange(n)
    term(ids, ids, c["g"])
    term(g["u"], g["v"], -g["w"])
    root = int(c["root"])
    term([root], [root], [L])                                   # L |x_root|^2, vol_root = L (checked)
    if len(c["lam"]):
        T = c["TI"].reshape(-1, 3).astype(np.int64)
        lam = c["lam"].astype(np.float64)
        pos = lam > 0
        T, h = T[pos], 0.5 * lam[pos]
        term(T[:, 0], T[:, 1], h)
        term(T[:, 1], T[:, 2], h)
        term(T[:, 0], T[:, 2], -h)
    for k in range(int(c["nlev"])):
        lab, sig = c[f"lv{k}_lab"].astype(np.int64), c[f"lv{k}_sig"].astype(np.float64)
        node, coef = c[f"lv{k}_node"].astype(np.int64), c[f"lv{k}_coef"].astype(np.float64)
        Q, R = np.nonzero(np.triu(sig > 0))
        keep = (node[Q] >= 0) & (node[R] >= 0) & ~((Q == R) & (node[Q] < n))
        Q, R = Q[keep], R[keep]
        term(node[Q], node[R], -sig[Q, R] * coef[Q] * coef[R])
        sq = np.diag(sig)[lab]
        fix = (sq > 0) & (d > 0) & (node[lab] >= n)
        term(ids[fix], ids[fix], sq[fix] * d[fix] * d[fix])
    for k in range(int(c["nvcs"])):
        lab, ptr = c[f"vc{k}_lab"].astype(np.int64), c[f"vc{k}_ptr"].astype(np.int64)
        col, s = c[f"vc{k}_col"].astype(np.int64), c[f"vc{k}_sig"].astype(np.float64)
        node, coef = c[f"vc{k}_node"].astype(np.int64), c[f"vc{k}_coef"].astype(np.float64)
        i = np.repeat(ids, np.diff(ptr))
        inside = lab[i] == col
        keep = (s > 0) & (d[i] > 0) & (node[col] >= 0) & ~(inside & (node[col] < n))
        i, col, s, inside = i[keep], col[keep], s[keep], inside[keep]
        term(i, node[col], -0.5 * s * d[i] * coef[col])
        term(i[inside], i[inside], s[inside] * d[i[inside]] * d[i[inside]])
    lp = c["link_p"].astype(np.int64)
    if len(lp):
        ch = c["link_ch"].reshape(-1, 2).astype(np.int64)
        cf = c["link_cf"].reshape(-1, 2).astype(np.float64)
        lv = c["link_vol"].astype(np.float64)
        mu = float(c["mu_c"]) / lv
        e_id = [lp, ch[:, 0], ch[:, 1]]
        e_co = [lv, -cf[:, 0], -cf[:, 1]]
        for a in range(3):
            for b in range(a, 3):
                term(e_id[a], e_id[b], mu * e_co[a] * e_co[b])
    term(np.arange(N), np.arange(N), c["sdiag"])
    r, cc, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    need(np.isfinite(v).all(), "non-finite term in the lifted matrix")
    key = r * N + cc
    o = np.argsort(key, kind="stable")
    key, v = key[o], v[o]
    first = np.r_[0, np.flatnonzero(np.diff(key)) + 1]
    val = np.add.reduceat(v, first)
    absum = np.add.reduceat(np.abs(v), first)
    h = np.diff(np.r_[first, len(key)]).astype(np.float64)
    ek = key[first]
    er, ec = ek // N, ek % N
    mult = np.where(er == ec, 1.0, 2.0)
    price = float(np.sum(mult * _gam(h + 5.0) * absum))
    need(len(v) < 1e12, "too many terms for the price allowance")
    price = price * 1.001
    Ml = sp.csr_matrix((val, (er, ec)), shape=(N, N))
    Ml.sort_indices()
    return Ml, price, dict(terms=int(len(v)), entries=int(len(val)), hmax=int(h.max()))


# ------------------------------------------------------------------------------------------------ 5 identity
def lift(c, Z, n):
    """Lifted rows X = H Z: vertex rows as given, x_p = (ca x_a + cb x_b) / vol_p in parent order."""
    N = int(c["N"])
    X = np.zeros((N, Z.shape[1]))
    X[:n] = Z
    lp = c["link_p"].astype(np.int64)
    o = np.argsort(lp)
    ch = c["link_ch"].reshape(-1, 2).astype(np.int64)[o]
    cf = c["link_cf"].reshape(-1, 2).astype(np.float64)[o]
    lv = c["link_vol"].astype(np.float64)[o]
    for p, a, b, ca, cb, vp in zip(lp[o].tolist(), ch[:, 0].tolist(), ch[:, 1].tolist(), cf[:, 0].tolist(),
                                   cf[:, 1].tolist(), lv.tolist()):
        X[p] = (ca * X[a] + cb * X[b]) / vp
    return X


def lift_adjoint(c, Y, n):
    """H^T Y: the adjoint of lift (children receive ca / vol_p times the parent's row, parents in reverse order)."""
    Y = Y.copy()
    lp = c["link_p"].astype(np.int64)
    o = np.argsort(lp)[::-1]
    ch = c["link_ch"].reshape(-1, 2).astype(np.int64)[o]
    cf = c["link_cf"].reshape(-1, 2).astype(np.float64)[o]
    lv = c["link_vol"].astype(np.float64)[o]
    for p, a, b, ca, cb, vp in zip(lp[o].tolist(), ch[:, 0].tolist(), ch[:, 1].tolist(), cf[:, 0].tolist(),
                                   cf[:, 1].tolist(), lv.tolist()):
        Y[a] += (ca / vp) * Y[p]
        Y[b] += (cb / vp) * Y[p]
    return Y[:n]


def apply_C(c, g, Z):
    """C(y) Z in vertex space, from the constraint definitions (no lifting):
    C = W - d d^T / L + T(lam) + sum_levels [sig_(lab i, lab j) d_i d_j]_(i != j)
        + sum_VCS [s_iR d_i d_j / 2 for j in R \\ {i}, symmetrised]."""
    n, d, L = g["n"], g["d"], g["L"]
    u, v, w = g["u"], g["v"], g["w"]
    A = sp.csr_matrix((np.r_[w, w], (np.r_[u, v], np.r_[v, u])), shape=(n, n))
    out = A @ Z - np.outer(d, d @ Z) / L
    if len(c["lam"]):
        T = c["TI"].reshape(-1, 3).astype(np.int64)
        h = 0.5 * c["lam"].astype(np.float64)
        i, j, k = T[:, 0], T[:, 1], T[:, 2]
        Tm = sp.csr_matrix((np.r_[-h, -h, -h, -h, h, h], (np.r_[i, j, j, k, i, k], np.r_[j, i, k, j, k, i])),
                           shape=(n, n))
        out += Tm @ Z
    for k in range(int(c["nlev"])):
        lab, sig = c[f"lv{k}_lab"].astype(np.int64), c[f"lv{k}_sig"].astype(np.float64)
        Y = np.zeros((len(sig), Z.shape[1]))
        np.add.at(Y, lab, d[:, None] * Z)
        out += d[:, None] * (sig @ Y)[lab] - (np.diag(sig)[lab] * d * d)[:, None] * Z
    for k in range(int(c["nvcs"])):
        lab, ptr = c[f"vc{k}_lab"].astype(np.int64), c[f"vc{k}_ptr"].astype(np.int64)
        col, s = c[f"vc{k}_col"].astype(np.int64), c[f"vc{k}_sig"].astype(np.float64)
        P = len(c[f"vc{k}_node"])
        i = np.repeat(np.arange(n), np.diff(ptr))
        inside = lab[i] == col
        Y = np.zeros((P, Z.shape[1]))
        np.add.at(Y, lab, d[:, None] * Z)
        # row i: s_iR d_i / 2 (Y_R - [i in R] d_i Z_i)
        contrib = (0.5 * s * d[i])[:, None] * (Y[col] - np.where(inside, d[i], 0.0)[:, None] * Z[i])
        np.add.at(out, i, contrib)
        # the transposed half: vertex j in R receives s_iR d_i d_j / 2 Z_i for every entry (i, R), i != j
        Wr = np.zeros((P, Z.shape[1]))
        np.add.at(Wr, col, (0.5 * s * d[i])[:, None] * Z[i])
        selfpart = np.zeros((n, Z.shape[1]))
        np.add.at(selfpart, i[inside], (0.5 * s[inside] * d[i[inside]])[:, None] * Z[i[inside]])
        out += d[:, None] * (Wr[lab] - selfpart)
    return out


def identity_check(c, g, Ml, k=8, dense_max=1000, seed=1):
    """max over columns of |H^T M0 H Z - (g Z - C Z)| / scale, M0 = fl(Mt) - Diag(sdiag)."""
    n, N = g["n"], int(c["N"])
    Z = np.eye(n) if n <= dense_max else np.random.default_rng(seed).standard_normal((n, k))
    F = Ml + sp.tril(Ml, -1).T
    M0 = (F - sp.diags(c["sdiag"].astype(np.float64))).tocsr()
    Y1 = lift_adjoint(c, M0 @ lift(c, Z, n), n)
    Y2 = c["g"][:, None] * Z - apply_C(c, g, Z)
    absM = abs(M0)
    sc = lift_adjoint(c, absM @ lift(c, np.abs(Z), n), n) + np.abs(c["g"])[:, None] * np.abs(Z)
    err = float(np.max(np.abs(Y1 - Y2) / (sc.max(axis=0, keepdims=True) + 1e-300)))
    return err, ("identity" if n <= dense_max else f"random k={k}")


# ------------------------------------------------------------------------------------------------ 6 PSD
def md_order(F, dense_deg=1000):
    """Minimum-degree order of the full pattern (liblbcert's quotient-graph ordering, no dense core)."""
    lib = _lib(ORDERLIB, "ordering")
    lib.lbc_md_order.argtypes = [C.c_int64, _i64p, _i32p, C.c_int64, C.c_int64, _i64p, _i64p]
    N = F.shape[0]
    ptr, idx = F.indptr.astype(np.int64), F.indices.astype(np.int32)
    order, ns = np.zeros(N, np.int64), np.zeros(1, np.int64)
    rc = lib.lbc_md_order(N, ptr.ctypes.data_as(_i64p), idx.ctypes.data_as(_i32p), dense_deg, 2 ** 62,
                          order.ctypes.data_as(_i64p), ns.ctypes.data_as(_i64p))
    need(rc == 0, "ordering failed")
    need(np.array_equal(np.sort(order), np.arange(N)), "the ordering is not a permutation")
    return order


def ldl_price(F, order, max_nnz=2e8, want_factor=False):
    """lbv_ldl on the full symmetric CSR F in the given order: (status, price, info[, (D, Lp, Li, Lx)])."""
    lib = _lib(VERIFYLIB, "verifier")
    lib.lbv_ldl.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _i64p, C.c_int64, _f64p, _f64p, _i64p, _i64p, _i32p,
                            _f64p]
    N = F.shape[0]
    ptr, idx = F.indptr.astype(np.int64), F.indices.astype(np.int32)
    val = F.data.astype(np.float64)
    order = np.ascontiguousarray(order, np.int64)
    out, iout = np.zeros(4), np.zeros(4, np.int64)
    args = (N, ptr.ctypes.data_as(_i64p), idx.ctypes.data_as(_i32p), val.ctypes.data_as(_f64p),
            order.ctypes.data_as(_i64p), int(max_nnz))
    fac = None
    if want_factor:
        rc = lib.lbv_ldl(*args, None, out.ctypes.data_as(_f64p), iout.ctypes.data_as(_i64p), None, None, None)
        if rc in (0, 1):
            nz = int(iout[1])
            D, Lp, Li, Lx = np.zeros(N), np.zeros(N + 1, np.int64), np.zeros(nz, np.int32), np.zeros(nz)
            rc = lib.lbv_ldl(*args, D.ctypes.data_as(_f64p), out.ctypes.data_as(_f64p), iout.ctypes.data_as(_i64p),
                             Lp.ctypes.data_as(_i64p), Li.ctypes.data_as(_i32p), Lx.ctypes.data_as(_f64p))
            fac = (D, Lp, Li, Lx)
    else:
        rc = lib.lbv_ldl(*args, None, out.ctypes.data_as(_f64p), iout.ctypes.data_as(_i64p), None, None, None)
    info = dict(nnzL=int(iout[1]), hmax=int(iout[2]), flops=float(out[2]), failed_pivot=int(iout[0]))
    if rc == -1:
        raise RuntimeError("lbv_ldl: bad arguments")
    if rc == -2:
        return "too_large", None, info, None
    if rc == 1:
        return "not_psd", None, info, fac
    nterm = info["nnzL"] * 2 + N
    need(nterm * max(info["hmax"], 1) < 5e11, "too many terms for the PSD price allowance")
    underflow = 4.0 * nterm * (info["hmax"] + 4) * 2.0 ** -1074
    return "ok", (float(out[0]) + underflow) * 1.001, info, fac


# ------------------------------------------------------------------------------------------------ verify
def verify(cert, instance=None, max_nnz=2e8, id_check=True, orig=None, log=None):
    """Verify a certificate record (dict of arrays or an .npz path) against an instance (a directory in the
    reference's layout, a dict as load_instance_dir returns, or None for the graph stored in the record).
    Returns dict(passed, UB, ...); `fail` names the first check that failed. `orig` = (n_orig, keys_orig, group_of,
    labels or None) adds check_reduction."""
    t0 = time.time()
    say = log or (lambda *a: None)
    out = dict(passed=False)
    try:
        c = load_certificate(cert) if isinstance(cert, str) else cert
        g = instance_from_record(c) if instance is None else (
            load_instance_dir(instance) if isinstance(instance, str) else instance)
        n = g["n"]
        need(int(c["n"]) == n and float(c["L"]) == g["L"] and float(c["C0"]) == g["C0"] and
             np.array_equal(c["d"].astype(np.float64), g["d"]), "the record's n / L / C0 / vol differ from the instance")
        out.update(n=n, N=int(c["N"]), cert_UB=float(c["UB"]))
        check_instance(g)
        check_multipliers(c, n)
        vol = check_tree(c, g["d"])
        say(f"instance, multipliers and lifted trees OK (n {n}, N {int(c['N'])})")
        Ml, p_asm, ainfo = assemble(c, g, vol)
        out["assembly"] = ainfo
        if id_check:
            err, mode = identity_check(c, g, Ml)
            out.update(identity_relerr=err, identity_mode=mode)
            say(f"lifting identity: rel err {err:.2e} ({mode})")
            need(err < 1e-9, f"the lifted matrix does not represent Diag(g) - C(y) (rel err {err:.2e})")
        F = (Ml + sp.tril(Ml, -1).T).tocsr()
        F.sort_indices()
        status, p_psd, pinfo, _ = ldl_price(F, md_order(F), max_nnz=max_nnz)
        out["psd"] = pinfo
        need(status != "too_large", f"PSD proof not attempted: nnz(L) {pinfo['nnzL']} above the limit {max_nnz:.0e}")
        need(status == "ok", f"PSD proof failed: pivot {pinfo['failed_pivot']} not positive")
        tot = (Fraction(g["C0"]) + rational_sum(c["lam"]) + rational_sum(c["g"]) + rational_sum(c["sdiag"])
               + Fraction(p_asm) + Fraction(p_psd))
        UB = round_up(tot / Fraction(g["L"]))
        out.update(passed=True, UB=UB, price_asm=p_asm / g["L"], price_psd=p_psd / g["L"])
        say(f"VERIFIED UB {UB:.9f} (certifier {float(c['UB']):.9f}; price_asm {p_asm / g['L']:.2e}, "
            f"price_psd {p_psd / g['L']:.2e})")
        if orig is not None:                   # (n_orig, keys_orig, group_of, labels or None)
            out["reduction"] = check_reduction(*orig, reduced=g)
    except Rejected as e:
        out.update(passed=False, fail=str(e))
        say("NOT VERIFIED:", e)
    out["seconds"] = round(time.time() - t0, 3)
    return out


# ------------------------------------------------------------------------------------------------ reduction
def _find(par, x):
    while par[x] != x:
        par[x] = par[par[x]]
        x = par[x]
    return x


def check_reduction(n0, keys0, group_of, labels=None, reduced=None):
    """Re-check a reduction of the original graph (n0 vertices, keys u*n0+v, u < v) into `reduced` (dict with n, u, v,
    w, d, L, C0 in the reduced numbering) given group_of (original vertex -> reduced vertex):
    - volumes = summed original degrees, C0 = 2 #edges inside groups, weights = original edge counts between groups;
    - every group is joined by eligible pairs of the ORIGINAL graph, recomputed here: an edge (a, b) inside a group
      joins a and b when deg a = 1 or deg b = 1 (leaf rule), or when N(a) \\ {b} = N(b) \\ {a} with deg^2 < 2m
      (adjacent twins); both hold in every optimal partition, so the reduction keeps the optimum;
    - with labels (original or reduced numbering): Q_orig = (C0 + <B', X'>) / L exactly."""
    keys0 = np.asarray(keys0, np.int64)
    u0, v0 = keys0 // n0, keys0 % n0
    need(len(np.unique(keys0)) == len(keys0) and (u0 < v0).all(), "the original graph is not simple")
    grp = np.asarray(group_of, np.int64)
    n = reduced["n"]
    need(len(grp) == n0 and grp.min() == 0 and grp.max() == n - 1, "group_of does not map onto the reduced vertices")
    m0 = len(keys0)
    deg = np.bincount(np.r_[u0, v0], minlength=n0).astype(np.int64)
    L = 2 * m0
    need(reduced["L"] == L, "L differs from 2 m of the original graph")
    need(np.array_equal(np.bincount(grp, weights=deg, minlength=n).astype(np.int64),
                        reduced["d"].astype(np.int64)), "reduced volumes differ from the summed original degrees")
    gu, gv = grp[u0], grp[v0]
    inside = gu == gv
    need(2 * int(inside.sum()) == reduced["C0"], "C0 differs from 2 x the edges inside groups")
    a, b = np.minimum(gu[~inside], gv[~inside]), np.maximum(gu[~inside], gv[~inside])
    kk, cnt = np.unique(a * n + b, return_counts=True)
    rk = reduced["u"] * n + reduced["v"]
    o = np.argsort(rk)
    need(np.array_equal(kk, rk[o]) and np.array_equal(cnt.astype(np.float64), reduced["w"][o]),
         "reduced weights differ from the original edge counts")
    # eligibility witnesses
    A = sp.csr_matrix((np.ones(2 * m0, np.int8), (np.r_[u0, v0], np.r_[v0, u0])), shape=(n0, n0))
    A.sort_indices()
    ptr, col = A.indptr, A.indices
    par = list(range(n0))
    nleaf = ntwin = nbad = 0
    for x, y in zip(u0[inside].tolist(), v0[inside].tolist()):
        rx, ry = _find(par, x), _find(par, y)
        if rx == ry:
            continue
        if deg[x] == 1 or deg[y] == 1:
            nleaf += 1
        elif deg[x] == deg[y] and deg[x] * deg[x] < L and np.array_equal(
                col[ptr[x]:ptr[x + 1]][col[ptr[x]:ptr[x + 1]] != y], col[ptr[y]:ptr[y + 1]][col[ptr[y]:ptr[y + 1]] != x]):
            ntwin += 1
        else:
            nbad += 1
            continue
        par[max(rx, ry)] = min(rx, ry)
    roots = np.array([_find(par, x) for x in range(n0)], np.int64)
    need(len(np.unique(grp * n0 + roots)) == n, "a contracted group is not joined by eligible leaf / twin pairs")
    res = dict(bookkeeping="ok", leaf_unions=nleaf, twin_unions=ntwin, ineligible_inside_edges=nbad)
    if labels is not None:
        lab = np.asarray(labels, np.int64)
        if len(lab) == n:
            lr = lab
        else:
            need(len(lab) == n0, "labels match neither graph")
            _, first = np.unique(grp, return_index=True)
            lr = lab[first]
        lr = np.unique(lr, return_inverse=True)[1]
        lo = lr[grp]
        P = int(lr.max()) + 1
        volc = np.bincount(lo, weights=deg, minlength=P).astype(np.int64)
        Qo = Fraction(2 * int((lo[u0] == lo[v0]).sum()), L) - Fraction(sum(int(x) ** 2 for x in volc), L * L)
        same = lr[reduced["u"]] == lr[reduced["v"]]
        volr = np.bincount(lr, weights=reduced["d"], minlength=P).astype(np.int64)
        BX = Fraction(2 * int(reduced["w"][same].astype(np.int64).sum())) - Fraction(sum(int(x) ** 2 for x in volr), L)
        Qr = (Fraction(reduced["C0"]) + BX) / L
        need(Qo == Qr, f"reduction identity failed: Q_orig {float(Qo)} != reduced {float(Qr)}")
        res.update(Q_orig=float(Qo), identity="exact")
    return res


def check_reduction_dirs(orig_dir, red_dir, labels=None):
    """check_reduction on instance directories in the reference's layout (red_dir holds group_of.npy)."""
    n0 = int(json.load(open(os.path.join(orig_dir, "graph.json")))["n"])
    keys0 = np.load(os.path.join(orig_dir, "keys.npy")).astype(np.int64)
    grp = np.load(os.path.join(red_dir, "group_of.npy")).astype(np.int64)
    return check_reduction(n0, keys0, grp, labels, reduced=load_instance_dir(red_dir))


def main(argv=None):
    pa = argparse.ArgumentParser(description="independent verifier of a lifted modularity upper-bound certificate")
    pa.add_argument("paths", nargs="+", help="[INSTANCE_DIR] CERT.npz")
    pa.add_argument("--orig", default=None, help="original (unreduced) instance directory: check the reduction")
    pa.add_argument("--labels", default=None, help="labels for the reduction identity (original or reduced numbering)")
    pa.add_argument("--max_nnz", type=float, default=2e8)
    pa.add_argument("--out", default="")
    a = pa.parse_args(argv)
    inst_dir, cert = (a.paths[0], a.paths[1]) if len(a.paths) > 1 else (None, a.paths[0])
    r = verify(cert, inst_dir, max_nnz=a.max_nnz, log=lambda *s: print("[verify]", *s, flush=True))
    if a.orig and r["passed"]:
        try:
            r["reduction"] = check_reduction_dirs(a.orig, inst_dir, np.load(a.labels) if a.labels else None)
        except Rejected as e:
            r.update(passed=False, fail=f"reduction: {e}")
    s = json.dumps(r, indent=1, default=float)
    print(s)
    if a.out:
        open(a.out, "w").write(s)
    print("VERIFIED" if r["passed"] else "NOT VERIFIED", f"UB {r['UB']:.9f}" if r["passed"] else r.get("fail"))
    return 0 if r["passed"] else 1


if __name__ == "__main__":
    sys.exit(main())
