"""Instances of the ALM search: the graph in CSR form, cluster levels, the triangle pool and seeded generators.

The layout follows what the solver loads from a (reduced) instance directory:
  * adjacency A: CSR, rows sorted, symmetric, ``rowptr`` int64 (n+1), ``col`` int32, ``w`` float64 edge weights
    (None = unweighted, every stored entry 1) -- alm_vcs.py:1106-1117;
  * d: vertex volumes (degrees, or group volumes of a reduced instance), L = original 2m, C0 = 2 * inside edges;
  * pair levels (aggops.PairLevel, aggops.py:51-68; rq/r set-up alm_vcs.py:1254-1259): labels, P, sig (P x P),
    rq = #Q / vol_Q^2, dsq = sum_{i in Q} d_i^2;
  * sparse vertex-cluster (VCS) levels (vcsops.VCSLevel, vcsops.py:31-70): labels, P, rq, working set as row CSR
    (ptr int64, col int32 sorted per row), sig per entry;
  * triangles TI (T x 3, canonical (min(a, c), centre, max(a, c))) with multipliers lam.

Generators (Chung-Lu, planted-partition SBM, label propagation, wedge triangles) are seeded and written for this
project; the reference ships none.
"""
from __future__ import annotations

import dataclasses
import numpy as np

from . import _host as H


@dataclasses.dataclass
class PairLevelData:
    lab: np.ndarray          # int32 (n,)
    P: int
    sig: np.ndarray          # float64 (P, P), symmetric, >= 0
    rq: np.ndarray           # float64 (P,)
    dsq: np.ndarray          # float64 (P,)


@dataclasses.dataclass
class VCSLevelData:
    lab: np.ndarray          # int32 (n,)
    P: int
    rq: np.ndarray           # float64 (P,)
    ptr: np.ndarray          # int64 (n+1,)
    col: np.ndarray          # int32 (nnz,), sorted inside each row
    sig: np.ndarray          # float64 (nnz,)

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    def own(self) -> np.ndarray:
        """Entry id of (i, R(i)) in the working set, or -1 (vcsops.own_entry, vcsops.py:88-91)."""
        n = len(self.lab)
        rows = np.repeat(np.arange(n), np.diff(self.ptr))
        hit = np.flatnonzero(self.col == self.lab[rows])
        out = np.full(n, -1, np.int64)
        # the working set is sorted and unique inside a row, so at most one hit per row
        out[rows[hit]] = hit
        return out


@dataclasses.dataclass
class Instance:
    n: int
    rowptr: np.ndarray       # int64 (n+1,)
    col: np.ndarray          # int32 (nnz,)
    w: np.ndarray | None     # float64 (nnz,) or None (unweighted)
    d: np.ndarray            # float64 (n,)
    L: float
    C0: float = 0.0
    levels: list = dataclasses.field(default_factory=list)      # [PairLevelData]
    pairlvl_weight: float = 1.0
    vcs: list = dataclasses.field(default_factory=list)         # [VCSLevelData]
    vc_weight: float = 1.0
    TI: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3), np.int32))
    lam: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    @property
    def T(self) -> int:
        return int(len(self.TI))

    @property
    def wv(self) -> float:
        """VC weight entering r_iR = wv rq_R / d_i^2 (alm_vcs.py:1153 wvc())."""
        return float(self.vc_weight) * float(self.pairlvl_weight)

    def dense_A(self) -> np.ndarray:
        A = np.zeros((self.n, self.n))
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        A[rows, self.col] = 1.0 if self.w is None else self.w
        return A


# ----------------------------------------------------------------------------------------------------------- setup
def csr_from_keys(n: int, keys: np.ndarray, weights: np.ndarray | None = None):
    """Symmetric CSR with sorted rows from undirected keys u*n+v (u < v), as alm_vcs.py:1106-1117 builds A."""
    keys = np.asarray(keys, np.int64)
    u = keys // n
    v = keys % n
    # keys sorted (u, v): entries (v, u) first, then (u, v); a stable sort by row alone then leaves every row's
    # columns ascending (below-diagonal ones from the first half, above-diagonal ones from the second)
    korder = H.argsort_i64(keys)
    u = u[korder]
    v = v[korder]
    r = np.r_[v, u]
    c = np.r_[u, v]
    order = H.argsort_bounded(r, n)
    c = c[order]
    rowptr = np.zeros(n + 1, np.int64)
    rowptr[1:] = np.cumsum(np.bincount(r, minlength=n))
    w = None
    if weights is not None:
        wk = np.asarray(weights, np.float64)[korder]
        ww = np.r_[wk, wk][order]
        if not np.all(ww == 1.0):
            w = ww
    return rowptr, c.astype(np.int32), w


def degrees(n: int, keys: np.ndarray) -> np.ndarray:
    u = keys // n
    v = keys % n
    return np.bincount(np.r_[u, v], minlength=n).astype(np.float64)


def cap_clusters(lab: np.ndarray, d: np.ndarray, max_clusters: int) -> np.ndarray:
    """Keep the max_clusters-1 largest-volume clusters, merge the rest into one (alm_vcs.py:1133-1137)."""
    lab = np.unique(lab, return_inverse=True)[1]
    if max_clusters and lab.max() + 1 > max_clusters:
        vol = np.bincount(lab, weights=d)
        keep = np.argsort(-vol)[:max_clusters - 1]
        mp = np.full(len(vol), max_clusters - 1)
        mp[keep] = np.arange(len(keep))
        lab = mp[lab]
    return lab.astype(np.int32)


def pair_level(lab: np.ndarray, d: np.ndarray, sig: np.ndarray | None = None) -> PairLevelData:
    """aggops.PairLevel (aggops.py:56-59) + rq (alm_vcs.py:1254-1256)."""
    lab = np.unique(lab, return_inverse=True)[1].astype(np.int32)
    P = int(lab.max()) + 1
    cnt = np.bincount(lab, minlength=P).astype(float)
    vol = np.maximum(np.bincount(lab, weights=d, minlength=P), 1.0)
    rq = cnt / vol ** 2
    dsq = np.bincount(lab, weights=d * d, minlength=P)
    if sig is None:
        sig = np.zeros((P, P))
    return PairLevelData(lab=lab, P=P, sig=np.ascontiguousarray(sig, np.float64), rq=rq, dsq=dsq)


def vcs_level(lab: np.ndarray, d: np.ndarray, ptr, col, sig=None) -> VCSLevelData:
    """vcsops.VCSLevel (vcsops.py:32-51) with a given working set."""
    lab = np.unique(lab, return_inverse=True)[1].astype(np.int32)
    P = int(lab.max()) + 1
    cnt = np.bincount(lab, minlength=P).astype(float)
    vol = np.maximum(np.bincount(lab, weights=d, minlength=P), 1.0)
    rq = cnt / vol ** 2
    ptr = np.ascontiguousarray(ptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    sig = np.zeros(len(col)) if sig is None else np.ascontiguousarray(sig, np.float64)
    return VCSLevelData(lab=lab, P=P, rq=rq, ptr=ptr, col=col, sig=sig)


def canonical_triangles(TI: np.ndarray, n: int, lam: np.ndarray | None = None):
    """(min(a, c), centre, max(a, c)); duplicate triangles merged with their multipliers summed (alm_vcs.py:1193-1196)."""
    TI = np.asarray(TI, np.int64)
    if lam is None:
        lam = np.zeros(len(TI))
    if not len(TI):
        return np.zeros((0, 3), np.int32), np.zeros(0)
    TI = np.c_[np.minimum(TI[:, 0], TI[:, 2]), TI[:, 1], np.maximum(TI[:, 0], TI[:, 2])]
    kk = (TI[:, 0] * n + TI[:, 1]) * n + TI[:, 2]
    _, first, inv = np.unique(kk, return_index=True, return_inverse=True)
    if len(first) < len(TI):
        lam = np.bincount(inv, weights=lam)
        TI = TI[first]
    return TI.astype(np.int32), np.asarray(lam, np.float64)


def init_V(n: int, k: int, rng: np.random.Generator) -> np.ndarray:
    """Gaussian rows, unit-normalised (alm_vcs.py:1214-1216)."""
    V = rng.standard_normal((n, k))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    return V


# ----------------------------------------------------------------------------------------------------------- graphs
def _finish_graph(n_all: int, a: np.ndarray, b: np.ndarray):
    """graph.py:_from_pairs semantics: drop self-loops, dedupe, drop isolated nodes (ids kept in order)."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    keep = lo != hi
    keys = np.unique(lo[keep] * n_all + hi[keep])
    lo = keys // n_all
    hi = keys % n_all
    used = np.zeros(n_all, bool)
    used[lo] = True
    used[hi] = True
    new_id = np.cumsum(used) - 1
    n = int(used.sum())
    return n, new_id[lo] * n + new_id[hi]


def _search_sorted_queries(cw: np.ndarray, q: np.ndarray) -> np.ndarray:
    """np.searchsorted(cw, q, side="right") for queries q >= 0 and a nondecreasing cw, exactly, through a bucket
    table: tab[k] = searchsorted(cw, k / B) is a lower bound of the answer for every q in [k / B, (k + 1) / B)
    (B a power of two, so q * B and k / B are exact), and a few vectorised steps of `while cw[i] <= q: i += 1` finish
    it. Avoids one cache-missing binary search per query."""
    m = len(cw)
    B = 1 << max(10, min(26, int(np.ceil(np.log2(4 * max(m, 1))))))
    tab = np.searchsorted(cw, np.arange(B, dtype=np.float64) / B, side="right")
    out = tab[np.minimum((q * B).astype(np.int64), B - 1)]
    act = np.flatnonzero(out < m)
    for _ in range(64):
        if len(act) == 0:
            break
        step = cw[out[act]] <= q[act]
        act = act[step]
        out[act] += 1
        act = act[out[act] < m]
    if len(act):
        out[act] = np.searchsorted(cw, q[act], side="right")
    return out


def chung_lu(n: int, gamma: float, kbar: float, seed: int, permute: bool = True):
    """Chung-Lu power-law graph: expected degrees w_i = kbar (g-2)/(g-1) (n/(i+1))^(1/(g-1)) (minimum ~1 for
    g = 2.1), m = n kbar / 2 endpoint pairs drawn independently proportional to w, self-loops and duplicates dropped,
    isolated vertices removed. Vertex ids are randomly permuted (hubs are not clustered at low ids).
    Returns (n, keys)."""
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.float64)
    w = kbar * (gamma - 2.0) / (gamma - 1.0) * (n / (i + 1.0)) ** (1.0 / (gamma - 1.0))
    if permute:
        w = w[rng.permutation(n)]
    cw = np.cumsum(w)
    cw /= cw[-1]
    m = int(round(n * kbar / 2.0))
    a = _search_sorted_queries(cw, rng.random(m))
    b = _search_sorted_queries(cw, rng.random(m))
    a = np.minimum(a, n - 1)
    b = np.minimum(b, n - 1)
    return _finish_graph(n, a, b)


def chung_lu_target(n: int, gamma: float, kbar: float, seed: int):
    """Chung-Lu graph whose vertex count and mean degree AFTER dropping duplicates and isolated vertices are close to
    (n, kbar): one pilot draw measures both losses, the second draw (same seed) compensates them."""
    n0, k0 = chung_lu(n, gamma, kbar, seed)
    rn = n0 / n
    rk = (2.0 * len(k0) / max(n0, 1)) / kbar
    return chung_lu(int(round(n / rn)), gamma, kbar / rk, seed)


def sbm(n: int, blocks: int, p_in: float, p_out: float, seed: int):
    """Planted-partition SBM with `blocks` equal blocks (block sizes differ by at most one).
    Edges per block pair ~ Binomial(#pairs, p), then that many distinct pairs uniformly.
    Returns (n, keys, planted labels of the kept vertices)."""
    rng = np.random.default_rng(seed)
    lab = (np.arange(n) * blocks) // n
    starts = np.searchsorted(lab, np.arange(blocks))
    sizes = np.bincount(lab, minlength=blocks)
    ks = []
    for p in range(blocks):
        for q in range(p, blocks):
            sp_, sq = int(sizes[p]), int(sizes[q])
            npairs = sp_ * (sp_ - 1) // 2 if p == q else sp_ * sq
            prob = p_in if p == q else p_out
            ne = rng.binomial(npairs, prob) if npairs > 0 else 0
            if ne == 0:
                continue
            got = np.zeros(0, np.int64)
            while len(got) < ne:
                need = ne - len(got)
                x = rng.integers(0, sp_, size=2 * need + 16) + starts[p]
                y = rng.integers(0, sq, size=2 * need + 16) + starts[q]
                lo = np.minimum(x, y)
                hi = np.maximum(x, y)
                ok = lo != hi
                kk = np.unique(np.r_[got, lo[ok] * n + hi[ok]])
                got = kk
            got = rng.permutation(got)[:ne] if len(got) > ne else got
            ks.append(got)
    keys = np.unique(np.concatenate(ks)) if ks else np.zeros(0, np.int64)
    lo = keys // n
    hi = keys % n
    used = np.zeros(n, bool)
    used[lo] = True
    used[hi] = True
    new_id = np.cumsum(used) - 1
    n2 = int(used.sum())
    return n2, new_id[lo] * n2 + new_id[hi], lab[used].astype(np.int32)


def label_propagation(rowptr: np.ndarray, col: np.ndarray, n: int, seed: int, P: int | None = None,
                      iters: int = 8, dense: bool | None = None) -> np.ndarray:
    """Seeded synchronous label propagation: every vertex takes the most frequent label among its neighbours, ties
    broken by a seeded random key per label; isolated vertices keep their label. Starts from singleton labels
    (P None) or from uniform random labels in [0, P) (a balanced partition of a graph without planted structure)."""
    rng = np.random.default_rng(seed)
    lab = np.arange(n, dtype=np.int64) if P is None else rng.integers(0, P, size=n).astype(np.int64)
    nk = n if P is None else P
    key = rng.permutation(nk).astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    colv = col.astype(np.int64)
    dense = P is not None and (n * nk <= 400_000_000 if dense is None else dense)
    if dense:
        deg = np.diff(rowptr)
        tie = (nk - 1 - key).astype(np.int64)          # larger for a smaller random key
    for _ in range(iters):
        if dense:
            # per-row label counts as an n x P table; score = count * P + tie picks the max count, then the label of
            # the smallest random key, as the sort below does; rows without neighbours keep their label
            cnt = np.bincount(rows * nk + lab[colv], minlength=n * nk).reshape(n, nk)
            score = cnt * nk + tie[None, :]
            best = np.argmax(score, axis=1)
            new = np.where(deg > 0, best, lab)
            del cnt, score
            if np.array_equal(new, lab):
                break
            lab = new
            continue
        kk = rows * nk + lab[colv]
        uq, cnt = np.unique(kk, return_counts=True)
        r_ = uq // nk
        l_ = uq % nk
        # best label per row: max count, then min random key of the label
        order = np.lexsort((key[l_], -cnt, r_))
        r_o = r_[order]
        first = np.r_[True, r_o[1:] != r_o[:-1]]
        new = lab.copy()
        new[r_o[first]] = l_[order][first]
        if np.array_equal(new, lab):
            break
        lab = new
    return np.unique(lab, return_inverse=True)[1].astype(np.int32)


def wedge_triangles(rowptr: np.ndarray, col: np.ndarray, d: np.ndarray, T: int, seed: int,
                    packed: bool | None = None) -> np.ndarray:
    """T distinct wedge triangles (a, j, c): centre j drawn proportional to degree (>= 2), a != c two distinct
    neighbours of j; canonical (min(a, c), j, max(a, c)); duplicates dropped and redrawn."""
    rng = np.random.default_rng(seed)
    n = len(d)
    deg = np.diff(rowptr)
    cand = np.flatnonzero(deg >= 2)
    if len(cand) == 0 or T == 0:
        return np.zeros((0, 3), np.int32)
    p = deg[cand].astype(np.float64)
    cp = np.cumsum(p)
    cp /= cp[-1]
    # one int64 key (lo n + j) n + hi per triangle while n^3 fits in int64; beyond (n >= 2^21) unique rows
    packed = n ** 3 < 2 ** 63 if packed is None else (packed and n ** 3 < 2 ** 63)
    got = np.zeros(0, np.int64) if packed else np.zeros((0, 3), np.int64)
    rounds = 0
    while len(got) < T and rounds < 50:
        need = int((T - len(got)) * 1.2) + 64
        j = cand[np.minimum(np.searchsorted(cp, rng.random(need), side="right"), len(cand) - 1)]
        dj = deg[j]
        x = rng.integers(0, dj)
        y = rng.integers(0, dj - 1)
        y = y + (y >= x)
        a = col[rowptr[j] + x].astype(np.int64)
        c = col[rowptr[j] + y].astype(np.int64)
        lo = np.minimum(a, c)
        hi = np.maximum(a, c)
        if packed:
            got = np.unique(np.r_[got, (lo * n + j) * n + hi])
        else:
            got = _unique_rows(np.r_[got, np.c_[lo, j, hi]])
        rounds += 1
    if not packed:
        if len(got) > T:
            got = got[np.sort(rng.permutation(len(got))[:T])]
        return got.astype(np.int32)
    if len(got) > T:
        got = rng.permutation(got)[:T]
    got = np.sort(got)
    a = got // (n * n)
    j = (got // n) % n
    c = got % n
    return np.c_[a, j, c].astype(np.int32)


def _unique_rows(M: np.ndarray) -> np.ndarray:
    """Distinct rows of an (m, 3) integer array in lexicographic order."""
    if len(M) == 0:
        return M
    M = M[np.lexsort((M[:, 2], M[:, 1], M[:, 0]))]
    keep = np.r_[True, np.any(M[1:] != M[:-1], axis=1)]
    return M[keep]


def random_ws(lab: np.ndarray, P: int, per_row: int, seed: int):
    """Working set with the own cluster plus per_row-1 further random clusters per row (sorted, unique)."""
    rng = np.random.default_rng(seed)
    n = len(lab)
    extra = rng.integers(0, P, size=(n, max(0, per_row - 1)))
    M = np.c_[lab.astype(np.int64), extra]
    M.sort(axis=1)
    keep = np.ones_like(M, bool)
    keep[:, 1:] = M[:, 1:] != M[:, :-1]
    cnt = keep.sum(1)
    ptr = np.r_[0, np.cumsum(cnt)].astype(np.int64)
    col = M[keep].astype(np.int32)
    return ptr, col


def build_instance(n: int, keys: np.ndarray, *, weights=None, vol=None, L=None, C0=0.0) -> Instance:
    rowptr, col, w = csr_from_keys(n, keys, weights)
    d = degrees(n, keys) if vol is None else np.asarray(vol, np.float64)
    if weights is not None and vol is None:
        d = np.zeros(n)
        rows = np.repeat(np.arange(n), np.diff(rowptr))
        np.add.at(d, rows, 1.0 if w is None else w)
    Lv = float(d.sum()) if L is None else float(L)
    return Instance(n=n, rowptr=rowptr, col=col, w=w, d=d, L=Lv, C0=float(C0))


def metric_instance(n: int = 1_000_000, gamma: float = 2.1, kbar: float = 13.0, T: int | None = None, P: int = 64,
                    ws_per_row: int = 4, seeds=(0, 1, 2, 3), sig_scale: float = 1.0) -> tuple:
    """The f+grad metric instance: Chung-Lu graph (seed 0; n and mean degree calibrated after cleaning), pair level
    from 2 rounds of label propagation from P random labels (capped to P clusters),
    sig = sym|N(0,1)| * sig_scale; VCS level with the same labels, own + 3 random entries per row, sig ~ Exp(1) *
    sig_scale; T = 3.33 n wedge triangles (seed 2), lam: half 0, rest Exp(1) (seed 3). Returns (inst, V) with V
    Gaussian unit rows (seed 1), r = 16 columns."""
    s_graph, s_v, s_tri, s_lam = seeds
    n_, keys = chung_lu_target(n, gamma, kbar, s_graph)
    inst = build_instance(n_, keys)
    lab = label_propagation(inst.rowptr, inst.col, n_, seed=s_graph + 100, P=P, iters=2)
    lab = cap_clusters(lab, inst.d, P)
    rs = np.random.default_rng(s_lam + 1000)
    Pq = int(lab.max()) + 1
    S = np.abs(rs.standard_normal((Pq, Pq)))
    sig = 0.5 * (S + S.T) * sig_scale
    inst.levels.append(pair_level(lab, inst.d, sig))
    ptr, wcol = random_ws(lab, Pq, ws_per_row, s_lam + 2000)
    vs = rs.exponential(1.0, size=len(wcol)) * sig_scale
    inst.vcs.append(vcs_level(lab, inst.d, ptr, wcol, vs))
    if T is None:
        T = int(round(3.33 * n_))
    TI = wedge_triangles(inst.rowptr, inst.col, inst.d, T, s_tri)
    rl = np.random.default_rng(s_lam)
    lam = np.where(rl.random(len(TI)) < 0.5, 0.0, rl.exponential(1.0, size=len(TI)))
    inst.TI = TI
    inst.lam = lam
    V = init_V(n_, 16, np.random.default_rng(s_v))
    return inst, V


def add_twins_and_leaves(n: int, keys: np.ndarray, ntwin: int, nleaf: int, seed: int, clique: int = 3):
    """Plant reduction structure into a graph (keys u*n+v, u < v): `ntwin` groups of `clique` new vertices that are
    true twins of a random low-degree vertex x (adjacent to x, to each other and to every neighbour of x), and
    `nleaf` new degree-1 vertices attached to random vertices. Returns (n', sorted unique keys)."""
    rng = np.random.default_rng(seed)
    u, v = keys // n, keys % n
    nb = [[] for _ in range(n)]
    for a, b in zip(u.tolist(), v.tolist()):
        nb[a].append(b)
        nb[b].append(a)
    deg = np.array([len(x) for x in nb])
    low = np.flatnonzero((deg >= 1) & (deg <= 4))
    new = []
    nxt = n
    for x in rng.choice(low, size=min(ntwin, len(low)), replace=False).tolist() if len(low) else []:
        grp = [x] + list(range(nxt, nxt + clique - 1))
        nxt += clique - 1
        for i, a in enumerate(grp):
            for b in grp[i + 1:]:
                new.append((a, b))
        for y in nb[x]:
            for a in grp[1:]:
                new.append((y, a))
    hosts = rng.integers(0, nxt, size=nleaf)
    for h in hosts.tolist():
        new.append((h, nxt))
        nxt += 1
    n2 = nxt
    a = np.r_[u, [p for p, _ in new]].astype(np.int64)
    b = np.r_[v, [q for _, q in new]].astype(np.int64)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    return n2, np.unique(lo * n2 + hi)


def local_moving(rowptr: np.ndarray, col: np.ndarray, w, d: np.ndarray, L: float, gamma: float, seed: int,
                 passes: int = 20) -> np.ndarray:
    """Seeded greedy local moving at resolution gamma (the node-moving phase of Louvain/Leiden): start from
    singletons, visit the vertices in a random order and move each to the neighbouring cluster with the largest gain
    k_{i,C} - gamma d_i vol_C / L (its own cluster counted without it) while some move gains; repeat up to `passes`
    times. High gamma gives many small dense clusters (the reference's levels are Leiden runs at gamma 16 and 64,
    api.py:208-210). A host loop: meant for graphs of up to a few tens of thousands of vertices."""
    n = len(d)
    rng = np.random.default_rng(seed)
    lab = np.arange(n)
    vol = np.asarray(d, np.float64).copy()
    ww = np.ones(len(col)) if w is None else np.asarray(w, np.float64)
    for _ in range(passes):
        moved = 0
        for i in rng.permutation(n).tolist():
            a, b = rowptr[i], rowptr[i + 1]
            if a == b:
                continue
            nb = col[a:b]
            li = lab[i]
            vol[li] -= d[i]
            k = {}
            for j, x in zip(nb.tolist(), ww[a:b].tolist()):
                if j != i:
                    c = lab[j]
                    k[c] = k.get(c, 0.0) + x
            best, bgain = li, k.get(li, 0.0) - gamma * d[i] * vol[li] / L
            for c, kc in k.items():
                g = kc - gamma * d[i] * vol[c] / L
                if g > bgain + 1e-12:
                    best, bgain = c, g
            lab[i] = best
            vol[best] += d[i]
            if best != li:
                moved += 1
        if moved == 0:
            break
    return np.unique(lab, return_inverse=True)[1].astype(np.int32)
