"""Large cluster counts on the device: pair and VCS levels with P >= 1000 clusters, where the per-workgroup cluster
rows no longer fit the kernels' LDS budget (sumP (r + 2) 8 bytes > 48 KB), so the evaluation runs its global-atomic
cluster sums, the unstaged row kernel and the global VCS pass, and the sweeps their non-LDS paths. The production
solver runs with up to 6000 clusters per level (--max_clusters 6000, pipeline.py:116-117).

Same bars as the small-P tests: evaluation f / G within 1e-5 relative in FP32 and 1e-11 in FP64 of the oracle;
sweeps 1e-9 (FP64); the PHR update equals the oracle evaluated at the updated multipliers."""
import numpy as np
import pytest

import oracle as O
from lagbound_amd import instances as I
from test_eval_gpu import _check, _engine

pytestmark = pytest.mark.gpu


def big_cluster_instance(n=12000, P=1200, P2=0, Pv=1000, ws=4, k=16, seed=21, T_per_n=2.0):
    """Chung-Lu graph with a pair level of P random clusters (and optionally a second of P2), a VCS level of Pv
    clusters with `ws` working-set entries per row, and a wedge triangle pool."""
    n_, keys = I.chung_lu(n, 2.1, 13.0, seed)
    inst = I.build_instance(n_, keys)
    rng = np.random.default_rng(seed + 1)

    def labels(p):
        lab = rng.integers(0, p, n_)
        lab[:p] = np.arange(p)                   # every cluster non-empty
        return rng.permutation(lab)

    for p in [P] + ([P2] if P2 else []):
        lab = labels(p)
        S = np.abs(rng.standard_normal((p, p))) * (p * p / inst.L)
        inst.levels.append(I.pair_level(lab, inst.d, 0.5 * (S + S.T)))
    if Pv:
        labv = labels(Pv)
        ptr, col = I.random_ws(labv, Pv, ws, seed + 3)
        inst.vcs.append(I.vcs_level(labv, inst.d, ptr, col, rng.exponential(1.0, len(col)) * (Pv * Pv / inst.L)))
    TI = I.wedge_triangles(inst.rowptr, inst.col, inst.d, int(T_per_n * n_), seed + 5)
    inst.TI = TI
    inst.lam = np.where(rng.random(len(TI)) < 0.5, 0.0, rng.exponential(1.0, len(TI)))
    V = I.init_V(n_, k, rng)
    return inst, V


def _lds_exceeded(inst, r):
    sumP = sum(lv.P for lv in inst.levels) + sum(v.P for v in inst.vcs)
    return sumP * (r + 2) * 8 > 48 * 1024


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("r", [16, 32])
def test_eval_with_large_cluster_counts(dtype, r):
    inst, V = big_cluster_instance(k=r, P2=1500 if r == 32 else 0)
    assert _lds_exceeded(inst, r)
    eng = _engine(inst, V, dtype, r)
    for rho in (0.0, 0.02):
        _check(eng, inst, rho, dtype, f"large P r={r}")
    eng.close()


@pytest.mark.parametrize("r", [16, 32])
def test_sweeps_with_large_cluster_counts(r):
    from lagbound_amd.engine import Engine
    inst, V = big_cluster_instance(n=6000, k=r, P=1100, Pv=1000)
    assert _lds_exceeded(inst, r)
    rng = np.random.default_rng(31)
    orders = np.stack([rng.permutation(inst.n) for _ in range(2)])
    for step_wave in (1, 0):
        eng = Engine(r=r, dtype="f64")
        eng.set_option("step_wave", step_wave)
        eng.set_instance(inst)
        eng.set_V(V)
        V0 = eng.get_V()
        print(f"[large-P sweeps] r={r} step_wave={step_wave} n={inst.n}", flush=True)
        eng.sweeps(orders, 16, 0.02, damp=0.1)
        W = eng.get_V()
        eng.close()
        Wo = O.sweeps_c(inst, V0, orders, 16, 0.02, damp=0.1)
        assert np.abs(W - Wo).max() <= 1e-9


def test_phr_update_with_large_cluster_counts():
    inst, V = big_cluster_instance(n=8000, k=16, P=1300, Pv=1200)
    eng = _engine(inst, V, "f64", 16)
    rho = 0.05
    eng.eval(rho)
    pp, pt, pv = eng.get_multipliers()
    eng.phr_update()
    o = 0
    for lv in inst.levels:
        lv.sig = pp[o:o + lv.P * lv.P].reshape(lv.P, lv.P).copy()
        o += lv.P * lv.P
    inst.lam = pt.copy()
    inst.vcs[0].sig = pv.copy()
    _check(eng, inst, rho, "f64", "large P after phr update")
    eng.close()
