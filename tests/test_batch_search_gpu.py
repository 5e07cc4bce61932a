"""The batched search (lb_batch_rgd, lagbound_amd.batch.BatchSearch) on the device.

- every captured iteration against a numpy restatement of the step from the same state: the tangent gradient
  R = G - (G . V) V from an eager evaluation, the first step a0 / rmax_g, then Barzilai-Borwein
  <s, s> / |<s, y>| clamped to [amin, amax] / rmax_g, and the retraction, with R and V rounded to FP32 as stored
  (V to 2e-6, steps to 1e-6 relative); eager evaluations and synchronizations run between the replays;
- the batch PHR update copies every graph's pair-level multipliers (lb_batch_get_sig = the last evaluation's pp);
- a 12-graph search (n 200..3000): every graph's bound is certified and at least the modularity of its level
  partition; on four graphs it is within 5% of the single-graph pipeline's (sweeps) certified bound;
- the 512-graph batched-uneven workload (n 200..20,000): the search and every graph's certificate, timed.
"""
import time

import numpy as np
import pytest

from lagbound_amd import batch as BT

pytestmark = pytest.mark.gpu


def _state(bs):
    e = bs.eng
    V = e.get_V().astype(np.float64)
    return V


def _step_numpy(V, G, off, prev, a0, amin, amax):
    """One step of every graph from V with the evaluation's G; prev = (Vp, Rp, alpha) or None."""
    Vn = np.empty_like(V)
    R = (G - np.sum(G * V, axis=1, keepdims=True) * V).astype(np.float32).astype(np.float64)   # stored as FP32
    al = np.zeros(len(off) - 1)
    for g in range(len(off) - 1):
        s_ = slice(off[g], off[g + 1])
        rmax = np.sqrt((R[s_] ** 2).sum(1)).max()
        sc = 1.0 / rmax if rmax > 0 else 0.0
        if prev is None:
            a = a0 * sc
        else:
            Vp, Rp, ap = prev
            s, y = V[s_] - Vp[s_], R[s_] - Rp[s_]
            sy = float((s * y).sum())
            a = float((s * s).sum()) / abs(sy) if sy != 0 else ap[g]
        a = min(max(a, amin * sc), amax * sc)
        al[g] = a
        W = V[s_] + a * R[s_]
        Vn[s_] = (W / np.linalg.norm(W, axis=1, keepdims=True)).astype(np.float32)
    return Vn, R, al


def test_captured_iterations_match_numpy():
    insts = BT.search_instances(8, 200, 2500, seed=3)
    bs = BT.BatchSearch(insts, r=8, dtype="f32", seed=1)
    try:
        # move off the partition start (whose wedges are not violated), then separate triangles, so that the
        # evaluations below carry triangle terms
        bs.eng.batch_rgd(10, 0.02, 0.02, reset=True)
        bs.eng.batch_eval(0.02, 0.02)
        bs.eng.phr_update()
        bs._separate(3, 0.05, 100)
        print(f"[bsearch] {len(bs.TI)} triangles separated", flush=True)
        assert len(bs.TI) > 0
        off = bs.p.row_off
        a0, amin, amax = 0.3, 1e-4, 1.0
        prev = None
        for k in range(4):
            V = _state(bs)
            bs.eng.batch_eval(0.02, 0.02)                      # eager evaluation at the current V
            G = bs.eng.get_G().astype(np.float64)
            Vn, R, al = _step_numpy(V, G, off, prev, a0, amin, amax)
            st = bs.eng.batch_rgd(1, 0.02, 0.02, a0=a0, amin=amin, amax=amax, reset=(k == 0))
            bs.eng.sync()
            got = _state(bs)
            err = np.abs(got - Vn).max()
            print(f"[bsearch] iteration {k}: max |V - numpy| {err:.2e}, steps {st[:3, 0]} vs {al[:3]}", flush=True)
            assert err <= 2e-6, k
            assert np.allclose(st[:, 0], al, rtol=1e-6), k
            prev = (V, R, al)
    finally:
        bs.close()


def test_batch_phr_update_copies_pair_multipliers():
    insts = BT.search_instances(5, 200, 1500, seed=4)
    bs = BT.BatchSearch(insts, r=8, dtype="f32", seed=2)
    try:
        bs.eng.batch_rgd(5, 0.05, 0.05, reset=True)
        bs.eng.batch_eval(0.05, 0.05)
        pp, _, _ = bs.eng.get_multipliers()
        sig0 = bs.eng.batch_get_sig(bs.nsig)
        bs.eng.phr_update()
        sig1 = bs.eng.batch_get_sig(bs.nsig)
        print(f"[bsearch] pair multipliers: {np.count_nonzero(sig1)} nonzero of {bs.nsig}, max {sig1.max():.3e}",
              flush=True)
        assert np.all(sig0 == 0) and np.any(sig1 > 0) and np.all(sig1 >= 0)
        if len(pp) == bs.nsig:
            assert np.array_equal(pp, sig1)
    finally:
        bs.close()


def test_small_batched_search_certifies_every_graph():
    """12 graphs: every bound certified and at least the level partition's modularity; on the first four graphs
    the certified bound is within 5% of the one the single-graph pipeline (block-coordinate sweeps,
    lagbound_amd.serve.solve_request, 24 outers) certifies for the same graph."""
    from lagbound_amd import serve
    insts = BT.search_instances(12, 200, 3000, seed=5)
    bs = BT.BatchSearch(insts, r=8, dtype="f32", seed=0, log=lambda *a: print("[bsearch]", *a, flush=True))
    try:
        hist = bs.run(rho=0.02)
        pairs = bs.results()
    finally:
        bs.close()
    assert hist[-1]["T"] > 0
    res = BT.finish_all(pairs, workers=1)
    for g, r in enumerate(res):
        print(f"[bsearch] graph {g} n {r['n']} T {r['T']}: lb {r['lb']:.5f} ub {r['ub']:.5f} UB {r['UB']} "
              f"certified {r['certified']}", flush=True)
        assert r["certified"], r
        assert r["UB"] >= r["lb"] - 1e-12
    p = serve.Params(k=8, outer=24)
    for g, x in enumerate(BT.search_instances(4, 200, 3000, seed=5)):
        u = np.repeat(np.arange(x.n), np.diff(x.rowptr))
        keys = np.unique(np.minimum(u, x.col).astype(np.int64) * x.n + np.maximum(u, x.col))
        one, _ = serve.solve_request(serve.Request(rid=g, n=x.n, keys=keys, seed=g), p)
        print(f"[bsearch] graph {g}: batched UB {res[g]['UB']:.5f}, single-graph sweeps UB {one['UB']:.5f}",
              flush=True)
        assert one["certified"] and res[g]["UB"] <= 1.05 * one["UB"]


@pytest.mark.timeout(1100)
def test_c4_batched_search_512_graphs():
    t0 = time.time()
    insts = BT.search_instances(512, 200, 20000, seed=0)
    bs = BT.BatchSearch(insts, r=8, dtype="f32", seed=0, log=lambda *a: print("[bsearch]", *a, flush=True))
    try:
        t1 = time.time()
        hist = bs.run(rho=0.02)
        t2 = time.time()
        pairs = bs.results()
    finally:
        bs.close()
    res = BT.finish_all(pairs, workers=8)
    t3 = time.time()
    ok = [r["certified"] for r in res]
    it_ms = 1e3 * np.mean([h["inner_s"] for h in hist]) / 100
    print(f"[bsearch] 512 graphs: inputs {t1 - t0:.1f} s, search {t2 - t1:.2f} s ({it_ms:.3f} ms per captured "
          f"iteration), estimate + certificate {t3 - t2:.1f} s, certified {sum(ok)}/512, final pool "
          f"{hist[-1]['T']} triangles", flush=True)
    bad = [r for r in res if not r["certified"]]
    assert not bad, bad[:3]
    assert all(r["UB"] >= r["lb"] - 1e-12 for r in res)
