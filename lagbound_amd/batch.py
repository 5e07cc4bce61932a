"""Batches of independent graphs evaluated together on one device context (lb_batch_*): the graphs are packed into
one block-diagonal graph (contiguous row ranges, vertex ids offset), each with its own L, C0, rank-1 term and pair
level; one launch of each batch kernel evaluates f and G = C(y)V for every graph.

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
    a = ap.parse_args(argv)
    log = lambda *s: print("[batch]", *s, file=sys.stderr, flush=True)
    t = time.perf_counter()
    insts = c4_batch(a.count, a.n_lo, a.n_hi, a.r, log=log)
    p = Packed(insts)
    log(f"{p.G} graphs, n {p.n}, nnz {len(p.col)}, T {len(p.TI)}, built in {time.perf_counter() - t:.1f}s")
    rng = np.random.default_rng(1)
    V = rng.standard_normal((p.n, a.r))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    be = BatchEngine(p, r=a.r, dtype="f32")
    be.set_V(V)
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
