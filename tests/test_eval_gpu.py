"""Parity of the HIP f+grad evaluation (liblbhip through the C ABI) with the FP64 oracle on the same inputs.

Tolerances (stated by the precision contract of the search): FP32 and BF16-stored V: f within 1e-5 relative,
||G - G64||_F / ||G64||_F <= 1e-5, with the oracle evaluated on the stored (for BF16: dequantised) V; FP64 path:
1e-11 relative. The penalty parts (and f) also get their FP64 rounding floor 4u sum (p^2 + s^2) / (2 c) (see
_penalty_floor), which only matters where rho c is tiny against the multipliers. Effective multipliers: pp (FP64 on the device) 1e-9 relative; pt within 1e-5 * rt absolute (tau in
FP32); pv 1e-6 relative to |pv| + |sig| (FP32-stored multipliers).
"""
import numpy as np
import pytest

import oracle as O
from conftest import small_instance, chung_lu_instance

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-5, "bf16": 1e-5, "f64": 1e-11}
U64 = 2.0 ** -53


def _penalty_floor(inst, rho, pp, pt, pv):
    """Absolute rounding floor of the penalty parts, each a sum of (p^2 - s^2) / (2 c) over its entries (level:
    c = rho rq_Q rq_R w, alm_vcs.py:476-496; triangles: c = rho, 498-509; VCS: c = rho wv rq_R / d_i^2,
    vcsops.py:266-288). With c tiny the difference of squares cancels, and any change of rounding (summation order,
    FMA contraction) moves a term by a few ulps of (p^2 + s^2) / (2 c): the oracle built with and without FMA
    contraction differs by 4e-11 relative in f on the 300k-vertex instance. Returns {3: level, 4: tri, 5: vcs}."""
    if rho <= 0.0:
        return {3: 0.0, 4: 0.0, 5: 0.0}
    fl = {3: 0.0, 4: 0.0, 5: 0.0}
    o = 0
    for lv in inst.levels:
        P = lv.P
        rr = rho * np.outer(lv.rq, lv.rq) * inst.pairlvl_weight
        p_ = pp[o:o + P * P].reshape(P, P)
        ok = rr > 0
        fl[3] += float(np.sum((p_[ok] ** 2 + lv.sig[ok] ** 2) / (2 * rr[ok])))
        o += P * P
    if inst.T:
        fl[4] = float(np.sum((pt ** 2 + inst.lam ** 2) / (2 * rho)))
    o = 0
    for v in inst.vcs:
        rows = np.repeat(np.arange(inst.n), np.diff(v.ptr))
        d = inst.d[rows]
        c = rho * inst.wv * v.rq[v.col] / np.where(d > 0, d * d, 1.0)
        ok = (d > 0) & (c > 0)
        p_ = pv[o:o + v.nnz]
        fl[5] += float(np.sum((p_[ok] ** 2 + v.sig[ok] ** 2) / (2 * c[ok])))
        o += v.nnz
    return {k: 4 * U64 * x for k, x in fl.items()}


def _engine(inst, V, dtype, r, tuning=None):
    from lagbound_amd.engine import Engine
    eng = Engine(r=r, dtype=dtype)
    if tuning:
        eng.set_tuning(*tuning)
    eng.set_instance(inst)
    eng.set_V(V)
    return eng


def _check(eng, inst, rho, dtype, label):
    print(f"[parity] {label} dtype={dtype} rho={rho} n={inst.n} T={inst.T}", flush=True)
    parts = eng.eval(rho)
    G = eng.get_G()
    pp, pt, pv = eng.get_multipliers()
    Vs = eng.get_V()                       # the stored V (BF16: dequantised)
    po, Go, ppo, pto, pvo = O.eval_c(inst, Vs, rho)
    tol = TOL[dtype]
    floor = _penalty_floor(inst, rho, ppo, pto, pvo)
    f_err = abs(parts[0] - po[0])
    g_err = np.linalg.norm(G - Go) / max(1e-300, np.linalg.norm(Go))
    assert f_err <= tol * max(1.0, abs(po[0])) + sum(floor.values()), (label, "f", parts[0], po[0], f_err, floor)
    assert g_err <= tol, (label, "G", g_err)
    for k in (1, 2, 3, 4, 5, 6):
        assert abs(parts[k] - po[k]) <= max(tol * max(1.0, abs(po[k])), tol * max(1.0, abs(po[0]))) + floor.get(k, 0.0), \
            (label, k, parts[k], po[k], floor.get(k, 0.0))
    assert parts[8] == po[8] or dtype != "f64"
    assert np.allclose(pp, ppo, rtol=1e-9 if dtype != "f64" else 1e-12, atol=1e-12)
    if inst.T:
        rt = rho
        assert np.abs(pt - pto).max() <= (1e-5 * max(rt, 1e-30) + 1e-6 * np.abs(pto).max()) if dtype != "f64" else 1e-12 * max(1.0, np.abs(pto).max())
    if len(pvo):
        # FP32-stored multipliers: pv = sig - rho r g loses relative accuracy when it nearly cancels, so the bound is
        # relative to the inputs |sig| (and |pv|)
        sig = np.concatenate([v.sig for v in inst.vcs])
        rel = 1e-6 if dtype != "f64" else 1e-12
        assert np.all(np.abs(pv - pvo) <= rel * (np.abs(pvo) + np.abs(sig)) + 1e-300)
    return parts, G


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f64"])
@pytest.mark.parametrize("r", [8, 16, 32])
def test_small_parity(dtype, r):
    inst, V = small_instance(n=60, seed=r, k=r)
    eng = _engine(inst, V, dtype, r)
    for rho in (0.0, 0.3):
        _check(eng, inst, rho, dtype, f"small r={r}")


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_small_weighted_and_segments(dtype):
    """Weighted reduced-instance form (edge multiplicities, C0 > 0) and the hub segment path forced by tiny
    thresholds (every row with more than 2 neighbours / triangle records is split into 3-item segments)."""
    inst, V = small_instance(n=50, seed=3, k=16, weighted=True, ntri=60)
    eng = _engine(inst, V, dtype, 16, tuning=(2, 2, 3))
    for rho in (0.0, 0.5):
        _check(eng, inst, rho, dtype, "weighted+segments")


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f64"])
def test_chung_lu_parity(dtype):
    inst, V = chung_lu_instance(n=30000, seed=1, k=16)
    eng = _engine(inst, V, dtype, 16)
    for rho in (0.0, 0.02):
        _check(eng, inst, rho, dtype, "chung-lu 30k")


def test_empty_pools():
    """No levels, no VCS, no triangles: G = B V and f = <B, X> (+C0)."""
    inst, V = small_instance(n=40, seed=2, k=16, ntri=0, vcs_per_row=0)
    inst.levels = []
    eng = _engine(inst, V, "f32", 16)
    for rho in (0.0, 0.1):
        _check(eng, inst, rho, "f32", "empty pools")


def test_phr_update_on_device():
    """lb_phr_update (sig <- pp, lam <- pt, VCS sig <- pv) equals the oracle evaluated with the updated multipliers."""
    inst, V = chung_lu_instance(n=8000, seed=4, k=16)
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
    _check(eng, inst, rho, "f64", "after phr update")
    _check(eng, inst, 0.0, "f64", "after phr update, rho=0")


def test_repeatable():
    inst, V = chung_lu_instance(n=20000, seed=6, k=16)
    eng = _engine(inst, V, "f32", 16)
    p1 = eng.eval(0.02)
    G1 = eng.get_G()
    p2 = eng.eval(0.02)
    G2 = eng.get_G()
    assert np.linalg.norm(G1 - G2) <= 1e-6 * np.linalg.norm(G1)
    assert abs(p1[0] - p2[0]) <= 1e-9 * abs(p1[0])


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_overlap_matches_one_stream(dtype):
    """The hub segments on the context's second stream (option "overlap", the default) give the same results as the
    whole chain on one stream, with and without a strided segment grid, and match the oracle. Same up to the
    run-to-run rounding of the FP64 LDS atomics in the cluster sums (as test_repeatable): 1e-12 relative in FP64,
    1e-9 (f parts) / 1e-6 (G) in FP32."""
    inst, V = chung_lu_instance(n=20000, seed=7, k=16)
    eng = _engine(inst, V, dtype, 16, tuning=(8, 8, 16, 0))
    out = {}
    for ov, grid in ((1, 0), (0, 0), (1, 3), (0, 3)):
        eng.set_option("overlap", ov)
        eng.set_option("seg_grid", grid)
        for rho in (0.0, 0.02):
            p = eng.eval(rho)
            out[(ov, grid, rho)] = (p, eng.get_G(), eng.get_multipliers()[1])
    tf, tg = (1e-12, 1e-12) if dtype == "f64" else (1e-9, 1e-6)

    def same(a, b):
        fin = np.isfinite(a[0])
        assert np.array_equal(fin, np.isfinite(b[0]))
        assert np.allclose(a[0][fin], b[0][fin], rtol=tf, atol=tf * abs(a[0][0]))
        assert np.linalg.norm(a[1] - b[1]) <= tg * np.linalg.norm(a[1])
        assert np.allclose(a[2], b[2], rtol=tg, atol=tg * max(1.0, np.abs(a[2]).max(initial=0.0)))

    for rho in (0.0, 0.02):
        for key in ((0, 0), (1, 3), (0, 3)):
            same(out[(1, 0, rho)], out[key + (rho,)])
    eng.set_option("overlap", 1)
    eng.set_option("seg_grid", 0)
    for rho in (0.0, 0.02):
        _check(eng, inst, rho, dtype, "overlap")


@pytest.mark.parametrize("which", ["distinct", "second_level", "two_vcs"])
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_vcs_label_sets(dtype, which):
    """The cluster sums of a VCS level whose labels equal a pair level's are copied from that level, not
    accumulated again. Against the oracle: VCS labels different from every pair level ("distinct"), equal to the
    second pair level ("second_level"), and two VCS levels, one a copy of the first pair level and one distinct."""
    from lagbound_amd import instances as I
    inst, V = small_instance(n=60, seed=11, k=16, ntri=40, blocks=4)
    rng = np.random.default_rng(5)
    lab1 = np.asarray(inst.levels[0].lab)
    other = rng.integers(0, 4, inst.n)
    other[:4] = [0, 1, 2, 3]
    sets = {"distinct": [other], "second_level": [np.asarray(inst.levels[1].lab)], "two_vcs": [lab1, other]}[which]
    inst.vcs = []
    for k, lab in enumerate(sets):
        P = int(lab.max()) + 1
        ptr, col = I.random_ws(lab, P, 2, 3 + k)
        inst.vcs.append(I.vcs_level(lab, inst.d, ptr, col, rng.exponential(0.1, len(col))))
    eng = _engine(inst, V, dtype, 16)
    for rho in (0.0, 0.3):
        _check(eng, inst, rho, dtype, f"vcs labels {which}")


def test_metric_instance_parity():
    """The metric instance (n ~ 1M, r = 16, nnz ~ 13M, T ~ 3.4M, pair + VCS P = 64) in FP32 against the oracle."""
    from lagbound_amd import instances as I
    inst, V = I.metric_instance()
    eng = _engine(inst, V, "f32", 16)
    _check(eng, inst, 0.02, "f32", "metric instance")


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f64"])
@pytest.mark.parametrize("r,P", [(16, 64), (8, 16), (32, 32)])
def test_vcs_pass_forms_agree(dtype, r, P):
    """The one-pass VCS kernel (multipliers per entry, Z on the FP64 matrix cores) and the two-pass form (multipliers,
    then Z by LDS adds) give the same multipliers and the same evaluation: pv to 1e-6 relative to |pv| + |sig| (the
    dot V_i . Y_R is summed in another order, so a stored FP32 multiplier can move by one rounding), f and G to 1e-9
    relative (FP32 / BF16 V: G is FP32) and both match the oracle; with either workgroup count and tiles of 16 or
    32 rows."""
    inst, V = chung_lu_instance(n=6000, seed=11, k=r, P=P)
    eng = _engine(inst, V, dtype, r)
    sig = np.concatenate([v.sig for v in inst.vcs])
    out = {}
    for kern, grid, tile in ((0, 0, 16), (1, 0, 16), (1, 1536, 16), (0, 512, 16), (1, 0, 32)):
        eng.set_option("vcs_kernel", kern)
        eng.set_option("vcs_grid", grid)
        eng.set_option("vcs_tile", tile)
        parts = np.array(eng.eval(0.02))
        out[(kern, grid, tile)] = (parts, eng.get_G(), eng.get_multipliers()[2])
    p0, G0, pv0 = out[(0, 0, 16)]
    for key, (p1, G1, pv1) in out.items():
        assert np.max(np.abs(pv1 - pv0) / (np.abs(pv0) + np.abs(sig) + 1e-30)) <= 1e-6, key
        gt = 1e-9 if dtype == "f64" else 2e-6
        assert np.linalg.norm(G1 - G0) <= gt * np.linalg.norm(G0), key
        assert abs(p1[0] - p0[0]) <= 1e-9 * abs(p0[0]) + 1e-9, key
    po, Go, ppo, pto, pvo = O.eval_c(inst, eng.get_V(), 0.02)
    floor = sum(_penalty_floor(inst, 0.02, ppo, pto, pvo).values())
    tol = TOL[dtype]
    for key, (p1, G1, _) in out.items():
        assert abs(p1[0] - po[0]) <= tol * max(1.0, abs(po[0])) + floor, key
        assert np.linalg.norm(G1 - Go) <= tol * np.linalg.norm(Go), key
    eng.close()
