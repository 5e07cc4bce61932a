"""The batched search's host logic on the CPU: the per-graph triangle pool update (inactive triangles leave first,
each graph's room tri_per_n * n_g - kept is filled with its most violated new candidates, candidates already in the
pool are dropped) against a direct per-graph restatement, with a stand-in engine."""
import numpy as np

from lagbound_amd import batch as BT


class _Eng:
    def __init__(self, cand, val):
        self.cand, self.val, self.set = cand, val, None

    def separate_triangles(self, fam, thresh, maxh, vtol, cap):
        assert fam == 0 and cap >= 10
        return self.cand.copy(), self.val.copy(), 0.0

    def set_triangles(self, TI, lam):
        self.set = (TI.copy(), lam.copy())


class _P:
    def __init__(self, ns):
        self.row_off = np.r_[0, np.cumsum(ns)].astype(np.int64)
        self.G = len(ns)
        self.n = int(self.row_off[-1])


def test_pool_update_per_graph_budget():
    rng = np.random.default_rng(0)
    ns = [10, 25, 7, 40]
    p = _P(ns)
    bs = BT.BatchSearch.__new__(BT.BatchSearch)
    bs.p = p
    bs.m = max(ns)

    def tri(g, k):
        o = p.row_off[g]
        out = set()
        while len(out) < k:
            a, b, c = sorted(rng.choice(ns[g], 3, replace=False))
            out.add((o + a, o + b, o + c))
        return np.array(sorted(out), np.int64)

    pools = [tri(g, 6) for g in range(4)]
    TI = np.vstack(pools)
    lam = rng.choice([0.0, 0.5], len(TI))
    cands = [tri(g, 30) for g in range(4)]
    cand = np.vstack(cands + [TI[:5]])                    # some candidates are already in the pool
    val = rng.random(len(cand))
    tg = BT._graph_of(p.row_off, TI[:, 0])
    key = bs._key(TI, tg)
    o = np.argsort(key)
    bs.TI, bs.lam, bs.tg, bs.key = TI[o].astype(np.int32), lam[o], tg[o].astype(np.int32), key[o]
    TI, lam = TI[o], lam[o]
    bs.eng = _Eng(cand, val)
    tri_per_n = 1
    added = bs._separate(tri_per_n, 0.3, 100)
    TI2, lam2 = bs.eng.set
    # restatement, graph by graph
    keep = lam > 0
    exp_T, exp_l = [TI[keep]], [lam[keep]]
    pool_keys = {tuple(t) for t in TI[keep].tolist()}
    g_of = lambda v: int(np.searchsorted(p.row_off, v, side="right") - 1)
    news = []
    for g in range(4):
        have = sum(1 for t in TI[keep] if g_of(t[0]) == g)
        room = max(0, tri_per_n * ns[g] - have)
        cg = [(float(v), tuple(t)) for t, v in zip(cand.tolist(), val) if g_of(t[0]) == g and tuple(t) not in pool_keys]
        seen, uniq = set(), []
        for v, t in sorted(cg, key=lambda x: (x[1])):            # one per key
            if t not in seen:
                seen.add(t)
                uniq.append((v, t))
        uniq.sort(key=lambda x: -x[0])
        news += [t for _, t in uniq[:room]]
    assert added == len(news)
    got = {tuple(t): l for t, l in zip(TI2.tolist(), lam2.tolist())}
    want = {tuple(t): l for t, l in zip(exp_T[0].tolist(), exp_l[0].tolist())}
    want.update({t: 0.0 for t in news})
    assert got == want and len(TI2) == len(want)
    k2 = bs._key(TI2.astype(np.int64), BT._graph_of(p.row_off, TI2[:, 0]))
    assert np.array_equal(k2, np.sort(k2)) and np.array_equal(bs.key, k2)       # the pool stays key-sorted
    for g in range(4):                                    # the budget holds per graph
        assert sum(1 for t in TI2 if g_of(t[0]) == g) <= max(tri_per_n * ns[g], sum(1 for t in TI[keep] if g_of(t[0]) == g))
