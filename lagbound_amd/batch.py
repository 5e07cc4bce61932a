"""Batches of independent graphs evaluated and searched together on one device context (lb_batch_*): the graphs
are packed into one block-diagonal graph (contiguous row ranges, vertex ids offset), each with its own L, C0, rank-1
term and pair level; one launch of each batch kernel evaluates f and G = C(y)V for every graph, and BatchSearch runs
the ALM search of every graph at once (captured iterations of evaluation + per-graph Riemannian step, PHR updates,
wedge separation with per-graph budgets), after which every graph is estimated and certified on its own.

No reference counterpart: the reference evaluates one graph per process (alm_vcs.py `augmented`, `CV`). Each graph's
numbers are those of the single-graph evaluation of the same instance.
"""
from __future__ import annotations

import numpy as np

from . import instances as I
from .engine import Engine


class Packed:
    """Block-diagonal packing of Instances that carry (at most) one pair level each and a triangle pool."""

    def __init__(self, insts):
        self.insts = insts
        self.G = len(insts)
        ns = np.array([x.n for x in insts], np.int64)
        self.row_off = np.r_[0, np.cumsum(ns)].astype(np.int64)
        self.n = int(self.row_off[-1])
        nnzs = [x.nnz for x in insts]
        self.rowptr = np.zeros(self.n + 1, np.int64)
        cols, ws, ds, labs, Ps, sigs, rqs, dsqs, TIs, lams = [], [], [], [], [], [], [], [], [], []
        weighted = any(x.w is not None for x in insts)
        e_off = 0
        for g, x in enumerate(insts):
            o = self.row_off[g]
            self.rowptr[o + 1:o + x.n + 1] = x.rowptr[1:] + e_off
            e_off += x.nnz
            cols.append(x.col.astype(np.int64) + o)
            if weighted:
                ws.append(np.ones(x.nnz) if x.w is None else x.w)
            ds.append(x.d)
            if len(x.levels) > 1:
                raise ValueError("batch mode carries one pair level per graph")
            if x.vcs:
                raise ValueError("batch mode carries no VCS levels")
            if x.levels:
                lv = x.levels[0]
                labs.append(lv.lab.astype(np.int32))
                Ps.append(lv.P)
                sigs.append(lv.sig.ravel())
                rqs.append(lv.rq)
                dsqs.append(lv.dsq)
            if x.T:
                TIs.append(x.TI.astype(np.int64) + o)
                lams.append(x.lam)
        self.col = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)
        self.w = np.concatenate(ws) if weighted else None
        self.d = np.concatenate(ds)
        self.L = np.array([x.L for x in insts])
        self.C0 = np.array([x.C0 for x in insts])
        self.has_levels = bool(labs)
        if self.has_levels:
            if len(labs) != self.G:
                raise ValueError("either every graph of the batch has a pair level or none has")
            self.lab = np.concatenate(labs)
            self.P = np.array(Ps, np.int32)
            self.sig = np.concatenate(sigs)
            self.rq = np.concatenate(rqs)
            self.dsq = np.concatenate(dsqs)
            self.pw = float(insts[0].pairlvl_weight)
        self.TI = np.concatenate(TIs).astype(np.int32) if TIs else np.zeros((0, 3), np.int32)
        self.lam = np.concatenate(lams) if lams else np.zeros(0)
        self.tri_off = np.r_[0, np.cumsum([x.T for x in insts])].astype(np.int64)

    def rows(self, g):
        return slice(int(self.row_off[g]), int(self.row_off[g + 1]))


class BatchEngine:
    """One device context holding a packed batch."""

    def __init__(self, packed: Packed, r=8, dtype="f32", device=0):
        self.p = packed
        self.eng = Engine(r=r, dtype=dtype, device=device, order="given")
        e = self.eng
        union = I.Instance(n=packed.n, rowptr=packed.rowptr, col=packed.col, w=packed.w, d=packed.d,
                           L=float(packed.L.sum()), C0=0.0)
        e.set_instance(union)
        e.set_batch(packed.row_off, packed.L, packed.C0)
        if packed.has_levels:
            e.set_batch_levels(packed.lab, packed.P, packed.sig, packed.rq, packed.dsq, packed.pw)
        if len(packed.TI):
            e.set_triangles(packed.TI, packed.lam)

    def set_V(self, V):
        self.eng.set_V(V)

    def eval(self, rho, rt=None, fetch=True):
        return self.eng.batch_eval(rho, rt, fetch)

    def get_G(self):
        return self.eng.get_G()

    def sync(self):
        self.eng.sync()

    def close(self):
        self.eng.close()


def c4_batch(count=512, n_lo=200, n_hi=20000, r=8, tri_per_n=10, seed=0, log=None):
    """The batched-uneven workload: `count` SBM graphs, n log-uniform in [n_lo, n_hi], block count uniform in
    {4..40}, p_in uniform in [0.02, 0.2], p_out = p_in / 25; a pair level on the planted blocks with sig = sym
    |N(0,1)| / L; 10 n wedge triangles per graph with lam half 0, half Exp(1) / 4."""
    rng = np.random.default_rng(seed)
    insts = []
    for g in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        B = int(rng.integers(4, 41))
        pin = float(rng.uniform(0.02, 0.2))
        n2, keys, lab = I.sbm(n, min(B, n), pin, pin / 25, seed * 100003 + g)
        inst = I.build_instance(n2, keys)
        lv = I.pair_level(lab, inst.d)
        S = np.abs(rng.standard_normal((lv.P, lv.P))) / inst.L
        lv.sig = 0.5 * (S + S.T)
        inst.levels = [lv]
        inst.TI = I.wedge_triangles(inst.rowptr, inst.col, inst.d, tri_per_n * n2, seed * 100003 + g + 1)
        inst.lam = np.where(rng.random(len(inst.TI)) < 0.5, 0.0, rng.exponential(0.25, len(inst.TI)))
        insts.append(inst)
        if log and (g + 1) % 64 == 0:
            log(f"generated {g + 1}/{count} graphs")
    return insts


# ------------------------------------------------------------------------------------------------ batched search
def search_instances(count=512, n_lo=200, n_hi=20000, seed=0, P=64, lpa_iters=3, log=None):
    """The batched-uneven search inputs: the graphs of c4_batch (SBM, n log-uniform in [n_lo, n_hi], 4..40 blocks,
    p_in uniform in [0.02, 0.2], p_out = p_in / 25), each with a seeded label-propagation pair level (P start
    clusters, multipliers 0) and an empty triangle pool: what a search starts from (the planted blocks are not used)."""
    rng = np.random.default_rng(seed)
    insts = []
    for g in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        B = int(rng.integers(4, 41))
        pin = float(rng.uniform(0.02, 0.2))
        n2, keys, _ = I.sbm(n, min(B, n), pin, pin / 25, seed * 100003 + g)
        inst = I.build_instance(n2, keys)
        lab = I.label_propagation(inst.rowptr, inst.col, inst.n, seed * 100003 + g + 7, P=P, iters=lpa_iters)
        inst.levels = [I.pair_level(lab, inst.d)]
        inst.TI = np.zeros((0, 3), np.int32)
        inst.lam = np.zeros(0)
        insts.append(inst)
        if log and (g + 1) % 128 == 0:
            log(f"generated {g + 1}/{count} graphs")
    return insts


def _graph_of(row_off, v):
    return np.searchsorted(row_off, v, side="right") - 1


class BatchSearch:
    """The ALM search of every graph of a batch at once on one device context (the per-graph chain of
    alm_vcs.py:1696-1797 run for all graphs together): per outer iteration, `inner` captured iterations of
    [batch evaluation, per-graph Riemannian Barzilai-Borwein ascent step] (lb_batch_rgd), an evaluation at the final
    V, the PHR multiplier update (triangles and every graph's pair level, lb_phr_update), and every `tri_every`
    outers the wedge triangle separation over the union with a per-graph pool budget of tri_per_n * n_g (inactive
    triangles leave the pool first, the most violated candidates of each graph fill its room, alm_vcs.py:1735-1762).

    The inner solver differs from the reference's block-coordinate sweeps (which visit one graph's rows in batches):
    a gradient step moves every row of every graph at once, so one launch sequence serves all graphs. The bound's
    validity does not depend on the search, and each graph's bound is certified on its own afterwards (finish)."""

    def __init__(self, insts, r=8, dtype="f32", device=0, seed=0, log=None):
        from .alm import init_partition_V
        self.log = log or (lambda *a: None)
        self.insts = insts
        self.p = Packed(insts)
        self.be = BatchEngine(self.p, r=r, dtype=dtype, device=device)
        self.eng = self.be.eng
        rng = np.random.default_rng(seed)
        V = np.zeros((self.p.n, r))
        for g, x in enumerate(insts):
            Vg = init_partition_V(x, x.levels[0].lab, r, rng, lambda *a: None)
            V[self.p.rows(g)] = Vg / np.linalg.norm(Vg, axis=1, keepdims=True)
        self.be.set_V(V)
        # the pool is kept sorted by a per-graph key (union vertex of the first corner, then the other two corners
        # relative to the graph's first row), with every triangle's graph id beside it, so an update is one merge
        self.m = int(np.diff(self.p.row_off).max()) if self.p.G else 1
        if self.p.n * self.m * self.m >= 2 ** 63:
            raise ValueError("graphs too large for the triangle keys of a batch")
        tg = _graph_of(self.p.row_off, self.p.TI[:, 0]) if len(self.p.TI) else np.zeros(0, np.int64)
        key = self._key(self.p.TI.astype(np.int64), tg)
        o = np.argsort(key, kind="stable")
        self.TI, self.lam = np.ascontiguousarray(self.p.TI[o], np.int32), self.p.lam[o].copy()
        self.tg, self.key = tg[o].astype(np.int32), key[o]
        self.nsig = int(sum(int(P) ** 2 for P in self.p.P))
        self.history = []

    def _key(self, TI, tg):
        off = self.p.row_off[tg]
        return (TI[:, 0].astype(np.int64) * self.m + (TI[:, 1] - off)) * self.m + (TI[:, 2] - off)

    def _separate(self, tri_per_n, thresh, maxh):
        """Pool update: inactive triangles (lam = 0) leave, then each graph's room tri_per_n * n_g - kept is filled
        with its most violated new wedge candidates (one per triangle, none already in the pool)."""
        p = self.p
        keep = np.flatnonzero(self.lam > 0)
        self.TI, self.lam, self.tg, self.key = self.TI[keep], self.lam[keep], self.tg[keep], self.key[keep]
        ns = np.diff(p.row_off)
        room = np.maximum(0, tri_per_n * ns - np.bincount(self.tg, minlength=p.G))
        added = 0
        if room.sum():
            cap = int(1.3 * room.sum()) + 10
            new, val, _ = self.eng.separate_triangles(0, thresh=thresh, maxh=maxh, vtol=1e-3, cap=cap)
            if len(new):
                gn = _graph_of(p.row_off, new[:, 0])
                kn = self._key(new, gn)
                uk, fi = np.unique(kn, return_index=True)          # one candidate per triangle
                pos = np.searchsorted(self.key, uk)
                isnew = np.ones(len(uk), bool)
                ok = pos < len(self.key)
                isnew[ok] = self.key[pos[ok]] != uk[ok]
                fi = fi[isnew]
                new, val, gn, kn = new[fi], val[fi], gn[fi], kn[fi]
                o = np.lexsort((-val, gn))                         # by graph, most violated first
                first = np.searchsorted(gn[o], np.arange(p.G))
                rank = np.arange(len(o)) - first[gn[o]]
                sel = o[rank < room[gn[o]]]
                sel = sel[np.argsort(kn[sel], kind="stable")]
                added = len(sel)
                at = np.searchsorted(self.key, kn[sel])
                self.TI = np.insert(self.TI, at, new[sel].astype(np.int32), axis=0)
                self.lam = np.insert(self.lam, at, 0.0)
                self.tg = np.insert(self.tg, at, gn[sel].astype(np.int32))
                self.key = np.insert(self.key, at, kn[sel])
        self.eng.set_triangles(self.TI, self.lam)
        return added

    def run(self, outer=24, inner=100, rho=0.02, tri_weight=1.0, tri_every=4, tri_per_n=10, thresh=0.3, maxh=100,
            a0=0.3, amin=1e-4, amax=0.3, autotune=False):
        """The outer loop; autotune=True picks the batch kernels' launch shapes (lagbound_amd.tune.autotune_batch)
        once the first triangle pool is in place."""
        import time
        rt = rho * tri_weight
        for it in range(outer):
            t0 = time.perf_counter()
            st = self.eng.batch_rgd(inner, rho, rt, a0=a0, amin=amin, amax=amax, reset=True)
            t1 = time.perf_counter()
            parts = self.eng.batch_eval(rho, rt)
            if len(self.TI):
                _, pt, _ = self.eng.get_multipliers()
                self.lam = np.asarray(pt, np.float64).copy()
            self.eng.phr_update()
            added = 0
            if tri_every and it % tri_every == 0 and it < outer - 1:
                added = self._separate(tri_per_n, thresh, maxh)
                if autotune and it == 0:
                    from . import tune
                    self.tuning, _ = tune.autotune_batch(self.be, rho, reps=5)
                    self.log(f"autotuned batch shapes {self.tuning}")
            t2 = time.perf_counter()
            rec = dict(it=it, inner_s=round(t1 - t0, 4), other_s=round(t2 - t1, 4), T=int(len(self.TI)), added=added,
                       aug_mean=float(np.mean(parts[:, 0] / self.p.L)), step_med=float(np.median(st[:, 0])))
            self.history.append(rec)
            self.log(f"outer {it}: {inner} captured iterations {rec['inner_s']:.3f} s, separation/PHR "
                     f"{rec['other_s']:.3f} s, {rec['T']} triangles (+{added}), mean augmented/L {rec['aug_mean']:.6f}")
        return self.history

    def results(self):
        """Per graph: (Instance with the final multipliers, V rows)."""
        p = self.p
        V = self.be.eng.get_V()
        sig = self.eng.batch_get_sig(self.nsig)
        so = np.r_[0, np.cumsum([int(P) ** 2 for P in p.P])]
        gt = self.tg
        out = []
        for g, x in enumerate(self.insts):
            y = I.Instance(n=x.n, rowptr=x.rowptr, col=x.col, w=x.w, d=x.d, L=x.L, C0=x.C0)
            lv = x.levels[0]
            y.levels = [I.pair_level(lv.lab, x.d, sig[so[g]:so[g + 1]].reshape(lv.P, lv.P))]
            sel = gt == g
            y.TI = (self.TI[sel] - p.row_off[g]).astype(np.int32)
            y.lam = self.lam[sel].copy()
            out.append((y, V[p.rows(g)].astype(np.float64)))
        return out

    def close(self):
        self.be.close()


def finish_graph(inst, V, device=0, rgd_iters=120, eig_maxiter=60, margin=1e-7, dense_max_n=3000):
    """Candidate bound (FP64 assembly, device operator) and FP64 certificate of one graph after the batched search:
    the dense certificate up to dense_max_n vertices, the lifted sparse one above."""
    from . import bound as B
    from . import sparse_cert as SC
    lag = B.DeviceLagrangian(inst, r=V.shape[1], dtype="f64", device=device)
    try:
        est = B.estimate(inst, V, lag=lag, rgd_iters=rgd_iters, eig_maxiter=eig_maxiter, eig_block=8, escape=0,
                         log=lambda *a: None)
    finally:
        lag.close()
    if inst.n <= dense_max_n:
        cert = B.certify_estimate(inst, est, margin=margin, device=device)
    else:
        cert = SC.certify_estimate_sparse(inst, est, margin=margin, device=device)
    lab = inst.levels[0].lab.astype(np.int64)
    u = np.repeat(np.arange(inst.n), np.diff(inst.rowptr))
    vol = np.bincount(lab, weights=inst.d)
    lb = float((inst.C0 + np.sum(lab[u] == lab[inst.col]) - (vol ** 2).sum() / inst.L) / inst.L)
    return dict(n=inst.n, T=int(inst.T), ub=float(est["ub"]), dv=float(est["dv"]), lb=lb,
                certified=bool(cert["passed"]), UB=cert.get("UB"), reason=cert.get("reason"))


def _finish_worker(args):
    inst, V, device = args
    return finish_graph(inst, V, device=device)


def finish_all(pairs, device=0, workers=8):
    """finish_graph for every (instance, V) pair, largest first, on `workers` spawned processes (each with its own
    device contexts); results in input order."""
    import multiprocessing as mp
    order = sorted(range(len(pairs)), key=lambda k: -pairs[k][0].n)
    if workers <= 1:
        res = {k: finish_graph(pairs[k][0], pairs[k][1], device=device) for k in order}
    else:
        ctx = mp.get_context("spawn")
        with ctx.Pool(workers, initializer=_limit_threads) as pool:
            out = pool.map(_finish_worker, [(pairs[k][0], pairs[k][1], device) for k in order], chunksize=1)
        res = dict(zip(order, out))
    return [res[k] for k in range(len(pairs))]


def _limit_threads():
    from threadpoolctl import threadpool_limits
    threadpool_limits(limits=2)


def search_main(a, log):
    """The batched search end to end: inputs, search (timed), per-graph estimate + certificate (timed)."""
    import time
    t = time.perf_counter()
    insts = search_instances(a.count, a.n_lo, a.n_hi, log=log)
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    bs = BatchSearch(insts, r=a.r, dtype="f32", log=log)
    t_setup = time.perf_counter() - t
    t = time.perf_counter()
    hist = bs.run(outer=a.outer, inner=a.inner, rho=a.rho, autotune=bool(a.autotune))
    t_search = time.perf_counter() - t
    pairs = bs.results()
    p = bs.p
    bs.close()
    t = time.perf_counter()
    res = finish_all(pairs, workers=a.workers)
    t_finish = time.perf_counter() - t
    cert = np.array([r["certified"] for r in res])
    UB = np.array([r["UB"] if r["certified"] else np.nan for r in res])
    lb = np.array([r["lb"] for r in res])
    gap = UB - lb
    it_s = float(np.mean([h["inner_s"] for h in hist])) / a.inner
    out = {"graphs": p.G, "n": p.n, "nnz": int(len(p.col)), "r": a.r, "rho": a.rho, "outer": a.outer,
           "tuning": getattr(bs, "tuning", None),
           "inner": a.inner, "T_final": int(hist[-1]["T"]), "generate_s": round(t_gen, 2),
           "setup_s": round(t_setup, 2), "search_s": round(t_search, 3),
           "ms_per_captured_iteration": round(1e3 * it_s, 4),
           "graph_iterations_per_s": round(p.G / it_s, 1), "finish_s": round(t_finish, 2),
           "finish_workers": a.workers, "certified": int(cert.sum()),
           "uncertified": {str(k): res[k]["reason"] for k in np.flatnonzero(~cert)[:8]},
           "certified_gap_UB_minus_lb": {q: round(float(np.nanpercentile(gap, v)), 5)
                                         for q, v in (("p10", 10), ("p50", 50), ("p90", 90))} if cert.any() else None,
           "UB_ge_lb": bool(np.all(UB[cert] >= lb[cert]))}
    log(f"search {t_search:.2f} s ({1e3 * it_s:.3f} ms per captured iteration of all {p.G} graphs), "
        f"finish {t_finish:.1f} s, certified {int(cert.sum())}/{p.G}")
    return out


def main(argv=None):
    import argparse
    import json
    import sys
    import time
    ap = argparse.ArgumentParser(prog="lagbound_amd.batch", description="Time the batched evaluation of many "
                                 "uneven graphs against one context per graph; prints one JSON line.")
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--n_lo", type=int, default=200)
    ap.add_argument("--n_hi", type=int, default=20000)
    ap.add_argument("--r", type=int, default=8)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--rho", type=float, default=0.02)
    ap.add_argument("--single", type=int, default=1, help="also time one context per graph")
    ap.add_argument("--search", type=int, default=0, help="1: the batched search with per-graph certificates instead")
    ap.add_argument("--outer", type=int, default=24)
    ap.add_argument("--inner", type=int, default=100)
    ap.add_argument("--workers", type=int, default=8, help="processes for the per-graph estimate and certificate")
    ap.add_argument("--autotune", type=int, default=1, help="1: pick the batch kernels' launch shapes first")
    a = ap.parse_args(argv)
    log = lambda *s: print("[batch]", *s, file=sys.stderr, flush=True)
    if a.search:
        out = search_main(a, log)
        print(json.dumps(out), flush=True)
        return 0
    t = time.perf_counter()
    insts = c4_batch(a.count, a.n_lo, a.n_hi, a.r, log=log)
    p = Packed(insts)
    log(f"{p.G} graphs, n {p.n}, nnz {len(p.col)}, T {len(p.TI)}, built in {time.perf_counter() - t:.1f}s")
    rng = np.random.default_rng(1)
    V = rng.standard_normal((p.n, a.r))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    be = BatchEngine(p, r=a.r, dtype="f32")
    be.set_V(V)
    tuning = None
    if a.autotune:
        from . import tune
        t = time.perf_counter()
        best, trials = tune.autotune_batch(be, a.rho, log=log)
        tuning = {"best": best, "ms_best": round(min(ms for _, ms in trials), 4),
                  "ms_default": round(next(ms for c, ms in trials if c == dict(batch_light=128, batch_light_grid=8192,
                                                                            batch_heavy_grid=8192)), 4),
                  "candidates": len(trials), "seconds": round(time.perf_counter() - t, 1)}
        log(f"autotuned batch shapes {best}: {tuning['ms_best']} ms (library defaults {tuning['ms_default']} ms)")
    for _ in range(3):
        be.eval(a.rho, fetch=False)
    be.sync()
    t = time.perf_counter()
    for _ in range(a.steps):
        be.eval(a.rho, fetch=False)
    be.sync()
    ms_batch = 1e3 * (time.perf_counter() - t) / a.steps
    # the same evaluation captured once as a HIP graph and replayed (lb_capture_eval / lb_replay_eval)
    be.eng.capture_eval(a.rho)
    be.eng.replay_eval(3, fetch=False)
    be.sync()
    t = time.perf_counter()
    be.eng.replay_eval(a.steps, fetch=False)
    be.sync()
    ms_replay = 1e3 * (time.perf_counter() - t) / a.steps
    be.close()
    out = {"graphs": p.G, "n": p.n, "nnz": int(len(p.col)), "T": int(len(p.TI)), "r": a.r, "rho": a.rho,
           "ms_per_batch_eval": round(ms_batch, 4), "graph_evals_per_s": round(p.G * 1e3 / ms_batch, 1),
           "ms_per_batch_eval_captured": round(ms_replay, 4),
           "alg_bytes": int(p.n * a.r * 8 + len(p.col) * 4 + p.n * 24 + len(p.TI) * 16)}
    out["hbm_GBps"] = round(out["alg_bytes"] / (ms_batch * 1e-3) / 1e9, 1)
    if tuning:
        out["tuning"] = tuning
    log(f"batch: {ms_batch:.3f} ms per evaluation of all {p.G} graphs; captured replays {ms_replay:.3f} ms")
    if a.single:
        from .engine import Engine
        tot = 0.0
        for g, x in enumerate(insts):
            e = Engine(r=a.r, dtype="f32")
            e.set_instance(x)
            e.set_V(V[p.rows(g)])
            e.eval(a.rho, fetch=False)
            e.sync()
            t = time.perf_counter()
            for _ in range(5):
                e.eval(a.rho, fetch=False)
            e.sync()
            tot += (time.perf_counter() - t) / 5
            e.close()
        out["ms_one_context_per_graph"] = round(1e3 * tot, 4)
        out["batch_speedup"] = round(1e3 * tot / ms_batch, 2)
        log(f"one context per graph: {1e3 * tot:.3f} ms for all graphs")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
