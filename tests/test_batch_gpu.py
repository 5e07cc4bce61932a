"""Batch mode on the device: one launch of each batch kernel evaluates many uneven graphs; every graph's f parts and
G rows match the FP64 oracle evaluation of that graph alone (FP32: 1e-5 relative; FP64: 1e-11), and an
empty-level batch works. The variant that also compares the triangle multipliers runs the getter whose host buffer
was once sized by the union instance's triangle count (0) rather than the uploaded count (a host heap overrun, shown
on the CPU by tests/test_engine_host.py and fixed).

The batched-uneven workload itself (512 SBM graphs, n log-uniform in [200, 20000], 10 n triangles each, r = 8, FP32)
is checked graph by graph against the oracle in test_c4_batch_512_graphs."""
import numpy as np
import pytest

import oracle as O
from lagbound_amd.batch import BatchEngine, Packed
from batch_case import make_batch

pytestmark = pytest.mark.gpu

KEYS = [(0, "f"), (1, "quad"), (2, "zz/L"), (3, "level"), (4, "tri"), (6, "primal"), (7, "C0"), (11, "lamsum")]


def _parity(dtype, r, tol, rho, multipliers):
    insts = make_batch(12, seed=1)
    p = Packed(insts)
    print(f"[batch] {dtype} r={r} rho={rho}: G={p.G} n={p.n} nnz={len(p.col)} T={len(p.TI)}", flush=True)
    rng = np.random.default_rng(5)
    V = rng.standard_normal((p.n, r))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    be = BatchEngine(p, r=r, dtype=dtype)
    be.set_V(V)
    parts = be.eval(rho)
    G = be.get_G()
    pt_all = be.eng.get_multipliers()[1] if multipliers else None
    be.close()
    Vs = V.astype(np.float32).astype(np.float64) if dtype == "f32" else V
    for g, x in enumerate(insts):
        sl = p.rows(g)
        po, Go, _, pto, _ = O.eval_c(x, Vs[sl], rho, threads=4)
        scale = max(abs(po[1]), abs(po[2]), abs(po[6]), 1.0)
        for k, name in KEYS:
            assert abs(parts[g, k] - po[k]) <= tol * scale + 1e-12, (g, name, parts[g, k], po[k])
        assert parts[g, 8] == po[8]
        if rho > 0 and x.T:
            assert abs(parts[g, 9] - po[9]) <= 10 * tol
            if multipliers:
                t0, t1 = p.tri_off[g], p.tri_off[g + 1]
                assert np.abs(pt_all[t0:t1] - pto).max() <= tol * max(1.0, np.abs(pto).max()) * 10
        err = np.linalg.norm(G[sl] - Go) / max(np.linalg.norm(Go), 1e-300)
        assert err <= tol, (g, err)


@pytest.mark.parametrize("dtype,r,tol", [("f32", 8, 1e-5), ("f64", 8, 1e-11), ("f32", 16, 1e-5)])
@pytest.mark.parametrize("rho", [0.0, 0.02])
def test_batch_parts_and_G_match_single_graph_oracle(dtype, r, tol, rho):
    _parity(dtype, r, tol, rho, multipliers=False)


@pytest.mark.parametrize("dtype,r,tol", [("f32", 8, 1e-5), ("f64", 8, 1e-11), ("f32", 16, 1e-5)])
@pytest.mark.parametrize("rho", [0.0, 0.02])
def test_batch_matches_single_graph_oracle(dtype, r, tol, rho):
    _parity(dtype, r, tol, rho, multipliers=True)


def test_batch_without_levels():
    insts = make_batch(5, seed=2, with_levels=False)
    p = Packed(insts)
    V = np.random.default_rng(1).standard_normal((p.n, 8))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    be = BatchEngine(p, r=8, dtype="f64")
    be.set_V(V)
    parts = be.eval(0.02)
    G = be.get_G()
    be.close()
    for g, x in enumerate(insts):
        po, Go, *_ = O.eval_c(x, V[p.rows(g)], 0.02, threads=4)
        assert abs(parts[g, 0] - po[0]) <= 1e-11 * max(1.0, abs(po[1]))
        assert np.linalg.norm(G[p.rows(g)] - Go) <= 1e-11 * np.linalg.norm(Go)


@pytest.mark.timeout(900)
def test_c4_batch_512_graphs():
    """The batched-uneven workload at its own size: 512 SBM graphs (sum n ~ 2.2M, 243M CSR entries, 21.5M triangles)
    in one context, r = 8, FP32. Every graph's f parts, G rows and triangle multipliers match the FP64 oracle of that
    graph alone (FP32 bars: 1e-5 relative; pt within 1e-5 rt + 1e-6 max |pt|), at rho = 0 and rho = 0.02."""
    from lagbound_amd.batch import c4_batch
    insts = c4_batch(512, log=lambda s: print("[c4]", s, flush=True))
    p = Packed(insts)
    print(f"[c4] G={p.G} n={p.n} nnz={len(p.col)} T={len(p.TI)}", flush=True)
    rng = np.random.default_rng(1)
    V = rng.standard_normal((p.n, 8))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    be = BatchEngine(p, r=8, dtype="f32")
    be.set_V(V)
    Vs = V.astype(np.float32).astype(np.float64)
    for rho in (0.0, 0.02):
        parts = be.eval(rho)
        G = be.get_G()
        pt_all = be.eng.get_multipliers()[1]
        worst = 0.0
        for g, x in enumerate(insts):
            sl = p.rows(g)
            po, Go, _, pto, _ = O.eval_c(x, Vs[sl], rho, threads=8)
            scale = max(abs(po[1]), abs(po[2]), abs(po[6]), 1.0)
            for k, name in KEYS:
                assert abs(parts[g, k] - po[k]) <= 1e-5 * scale + 1e-12, (g, name, parts[g, k], po[k])
            err = np.linalg.norm(G[sl] - Go) / max(np.linalg.norm(Go), 1e-300)
            assert err <= 1e-5, (g, err)
            worst = max(worst, err)
            if rho > 0 and x.T:
                t0, t1 = p.tri_off[g], p.tri_off[g + 1]
                assert np.abs(pt_all[t0:t1] - pto).max() <= 1e-5 * rho + 1e-6 * max(1.0, np.abs(pto).max()), g
        print(f"[c4] rho={rho}: all {p.G} graphs match; worst G rel err {worst:.2e}", flush=True)
    be.close()


def test_batch_autotuner_shapes_keep_results():
    """lagbound_amd.tune.autotune_batch times every batch launch-shape candidate; under the chosen shapes, and under
    the extremes of the light-row threshold (every row one slot / every row one wavefront), every graph's parts and
    G match the oracle as in the default shapes (FP32, 1e-5)."""
    from lagbound_amd import tune
    insts = make_batch(12, seed=2)
    p = Packed(insts)
    rng = np.random.default_rng(6)
    V = rng.standard_normal((p.n, 8))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    be = BatchEngine(p, r=8, dtype="f32")
    try:
        be.set_V(V)
        best, trials = tune.autotune_batch(be, 0.02, reps=5, log=lambda *a: print("[batch]", *a, flush=True))
        assert len(trials) == len(list(tune.batch_candidates()))
        Vs = V.astype(np.float32).astype(np.float64)
        for cfg in (best, dict(batch_light=1 << 30, batch_light_grid=64, batch_heavy_grid=64),
                    dict(batch_light=1, batch_light_grid=8192, batch_heavy_grid=7)):
            tune.apply_batch(be.eng, cfg)
            parts = be.eval(0.02)
            G = be.get_G()
            for g, x in enumerate(insts):
                sl = p.rows(g)
                po, Go, _, _, _ = O.eval_c(x, Vs[sl], 0.02, threads=4)
                scale = max(abs(po[1]), abs(po[2]), abs(po[6]), 1.0)
                for k, name in KEYS:
                    assert abs(parts[g, k] - po[k]) <= 1e-5 * scale + 1e-12, (cfg, g, name)
                assert np.linalg.norm(G[sl] - Go) <= 1e-5 * max(np.linalg.norm(Go), 1e-300), (cfg, g)
    finally:
        be.close()
