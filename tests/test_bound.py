"""Candidate bound (estimate) and dense FP64 certificate, on the CPU with the oracle's C(y) products.

Pins: exact sums against Fraction arithmetic; the dense C(y) against the oracle's entry-by-entry assembly; the
certificate against brute-force maximum modularity over every partition (weak duality must hold), against the
closed form n (lambda_max(B) + delta) / L at zero multipliers, and against a shift just below the true eigenvalue
(Cholesky must refuse); the estimate's eigenvalue against a dense generalized eigensolver."""
import itertools
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
from lagbound_amd import bound as B
from lagbound_amd import instances as I
from oracle_engine import OracleLagrangian


def _partitions(n):
    """Every set partition of range(n) as a label vector (restricted growth strings)."""
    def rec(i, lab, m):
        if i == n:
            yield list(lab)
            return
        for c in range(m + 1):
            lab.append(c)
            yield from rec(i + 1, lab, max(m, c + 1))
            lab.pop()
    yield from rec(0, [], 0)


def _modularity(inst, lab):
    lab = np.asarray(lab)
    X = (lab[:, None] == lab[None, :]).astype(float)
    Bm = inst.dense_A() - np.outer(inst.d, inst.d) / inst.L
    return (inst.C0 + float((Bm * X).sum())) / inst.L


def _with_multipliers(inst, seed, scale=0.02, ntri=12):
    rng = np.random.default_rng(seed)
    n = inst.n
    lab = rng.integers(0, 3, n)
    lv = I.pair_level(lab, inst.d)
    s = rng.random((lv.P, lv.P)) * scale / inst.L
    lv.sig = 0.5 * (s + s.T)
    inst.levels = [lv]
    vl = I.vcs_level(rng.integers(0, 3, n), inst.d, np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))
    keys = np.unique(rng.integers(0, n, 2 * n) * vl.P + rng.integers(0, vl.P, 2 * n))
    rows, cols = keys // vl.P, keys % vl.P
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, rows + 1, 1)
    vl.ptr, vl.col = np.cumsum(ptr), cols.astype(np.int32)
    vl.sig = rng.random(len(cols)) * scale / inst.L
    inst.vcs = [vl]
    TI = np.sort(np.array([rng.choice(n, 3, replace=False) for _ in range(ntri)]), axis=1)
    perm = rng.integers(0, 3, ntri)                      # centre anywhere: (a, b, c) with tau = X_ab + X_bc - X_ac
    TI = np.array([[t[(p + 1) % 3], t[p], t[(p + 2) % 3]] for t, p in zip(TI, perm)])
    inst.TI = TI.astype(np.int32)
    inst.lam = rng.random(ntri) * 0.3
    return inst


def _tiny(seed=0, n=8):
    rng = np.random.default_rng(seed)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.45]
    pairs += [(i, i + 1) for i in range(n - 1)]
    keys = np.unique([min(i, j) * n + max(i, j) for i, j in pairs]).astype(np.int64)
    return I.build_instance(n, keys)


def _pencil_min(inst, g, s):
    Cm, _ = B.dense_C(inst)
    M = np.diag(g) - Cm
    return float(sl.eigh(M, np.diag(s), eigvals_only=True)[0])


def test_exact_sum_and_round_up():
    rng = np.random.default_rng(1)
    a = rng.standard_normal(500) * 2.0 ** rng.integers(-60, 60, 500)
    ref = sum((Fraction(float(x)) for x in a), Fraction(0))
    assert B.exact_sum(a) == ref
    assert B.exact_sum([]) == 0
    x = Fraction(1, 3)
    u = B.frac_up(x)
    assert Fraction(u) >= x and Fraction(float(np.nextafter(u, -np.inf))) < x
    assert B.frac_up(Fraction(0.5)) == 0.5


def test_dense_C_matches_oracle_assembly():
    inst = _with_multipliers(_tiny(2, n=9), 3)
    Cm, F = B.dense_C(inst)
    Co = O.dense_C(inst, [lv.sig for lv in inst.levels], inst.lam, [vl.sig for vl in inst.vcs])
    assert np.array_equal(Cm, Cm.T)
    assert np.abs(Cm - 0.5 * (Co + Co.T)).max() <= 1e-14
    assert (F >= 0).all() and F.max() < 1e-12


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_certificate_bounds_every_partition(seed):
    inst = _with_multipliers(_tiny(seed, n=8), 10 + seed)
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((inst.n, 4))
    W /= np.linalg.norm(W, axis=1, keepdims=True)
    Cm, _ = B.dense_C(inst)
    g = np.einsum("ij,ij->i", Cm @ W, W)
    s = np.maximum(inst.d, 1.0)
    mu = _pencil_min(inst, g, s)
    res = B.certify_dense(inst, g, (max(0.0, -mu) + 1e-9) * s)
    assert res["passed"], res
    best = max(_modularity(inst, lab) for lab in _partitions(inst.n))
    assert res["UB"] >= best
    expect = (inst.C0 + inst.lam.sum() + g.sum() + (max(0.0, -mu) + 1e-9) * s.sum()) / inst.L
    assert expect - 1e-12 <= res["UB"] <= expect + res["price"] + 1e-12
    assert res["price"] < 1e-10
    # a shift below the true eigenvalue leaves M indefinite: the certificate must refuse it
    bad = B.certify_dense(inst, g, (max(0.0, -mu) - 1e-6) * s)
    assert not bad["passed"]


def test_certificate_zero_multipliers_closed_form():
    inst = _tiny(7, n=9)
    Bm = inst.dense_A() - np.outer(inst.d, inst.d) / inst.L
    lmax = float(np.linalg.eigvalsh(Bm)[-1])
    delta = 1e-6
    res = B.certify_dense(inst, np.full(inst.n, lmax + delta))
    assert res["passed"]
    closed = inst.n * (lmax + delta) / inst.L
    assert closed - 1e-12 <= res["UB"] <= closed + res["price"] + 1e-12
    best = max(_modularity(inst, lab) for lab in _partitions(inst.n))
    assert res["UB"] >= best
    assert not B.certify_dense(inst, np.full(inst.n, lmax - 1e-6))["passed"]


def test_certificate_rejects_invalid_multipliers():
    inst = _with_multipliers(_tiny(4, n=8), 5)
    g = np.ones(inst.n)
    inst.lam[0] = -1e-3
    with pytest.raises(ValueError):
        B.certify_dense(inst, g)
    inst.lam[0] = 0.0
    inst.levels[0].sig[0, 1] += 1e-3
    with pytest.raises(ValueError):
        B.certify_dense(inst, g)


def test_estimate_eigenvalue_and_certificate():
    n_, keys, lab = I.sbm(160, 4, 0.2, 0.02, 3)
    inst = _with_multipliers(I.build_instance(n_, keys), 9, scale=0.01, ntri=40)
    rng = np.random.default_rng(0)
    V = rng.standard_normal((n_, 8))
    lag = OracleLagrangian(inst, 8)
    est = B.estimate(inst, V, lag=lag, rgd_iters=60, eig_maxiter=200, log=lambda *a: None)
    assert est["dv"] >= est["dv0"] - 1e-12
    assert est["ub"] == min(r["ub"] for r in est["escape_rounds"])
    rec = est["s_vol"]
    assert rec["ub"] == est["ub"] and abs(rec["ub"] - rec["dv"] - rec["shift"]) < 1e-15
    mu = _pencil_min(inst, est["g"], est["s"])
    assert abs(rec["theta"][0] - mu) <= 1e-7 * max(1.0, abs(mu))
    assert est["lmin"] <= rec["theta"][0]
    cert = B.certify_estimate(inst, est, margin=1e-7)
    assert cert["passed"], cert
    slack = (1e-7 + abs(est["lmin"] - mu)) * est["s"].sum() / inst.L + cert["price"] + 1e-12
    assert abs(cert["UB"] - est["ub"]) <= slack
    assert cert["UB"] >= _modularity(inst, lab)


def test_checkpoint_roundtrip(tmp_path):
    inst = _with_multipliers(_tiny(5, n=9), 6)
    z = {"nlev": 1, "lab0": inst.levels[0].lab.astype(np.int64), "sig0": inst.levels[0].sig, "nvcs": 1,
         "vcslab0": inst.vcs[0].lab.astype(np.int64), "vcsptr0": inst.vcs[0].ptr, "vcscol0": inst.vcs[0].col,
         "vcssig0": inst.vcs[0].sig.astype(np.float32), "TI": inst.TI.astype(np.int64), "lam": inst.lam}
    np.savez(tmp_path / "c.npz", **z)
    inst2 = B.instance_from_checkpoint(_tiny(5, n=9), np.load(tmp_path / "c.npz"))
    assert np.array_equal(inst2.levels[0].sig, inst.levels[0].sig)
    assert np.array_equal(inst2.TI, inst.TI) and np.array_equal(inst2.lam, inst.lam)
    assert np.array_equal(inst2.vcs[0].sig, inst.vcs[0].sig.astype(np.float32).astype(np.float64))


def test_host_lobpcg_matches_dense_pencil():
    """lobpcg_pencil on host blocks (b = 12 columns, n = 300 > 4 b, so the block iteration runs) reaches the four
    smallest eigenvalues of the dense pencil (Diag(g) - C, Diag(s)); the reported residuals are those of the returned
    pairs; a start block holding a basis of W plus random columns reaches the same values."""
    n_, keys, _ = I.sbm(300, 5, 0.08, 0.01, 7)
    inst = _with_multipliers(I.build_instance(n_, keys), 4, scale=0.01, ntri=60)
    rng = np.random.default_rng(1)
    W = rng.standard_normal((n_, 8))
    W /= np.linalg.norm(W, axis=1, keepdims=True)
    lag = OracleLagrangian(inst, 8)
    _, g, _ = B.value(lag, W)
    s = np.maximum(inst.d, 1.0)
    ref = _pencil_min(inst, g, s)
    Cm, _ = B.dense_C(inst)
    ref4 = sl.eigh(np.diag(g) - Cm, np.diag(s), eigvals_only=True)[:4]
    for X0 in (None, np.linalg.svd(W, full_matrices=False)[0]):
        th, rn, X, trace = B.pencil_min(lag, g, s, 12, X0=X0, rng=np.random.default_rng(2), maxiter=300, atol=1e-9,
                                        on_device=False, log=lambda *a: None)
        assert X.shape == (n_, 12) and trace
        assert abs(th[0] - ref) <= 1e-8 * max(1.0, abs(ref))
        assert np.allclose(th[:4], ref4, rtol=1e-6, atol=1e-8)
        MX = g[:, None] * X - lag.CX(X)
        xs = np.einsum("ij,ij->j", X, s[:, None] * X)
        R = MX - s[:, None] * X * th[None, :]
        assert np.allclose(np.sqrt(np.einsum("ij,ij->j", R, R / s[:, None]) / xs), rn, rtol=1e-6, atol=1e-12)


def test_dense_pencil_for_tiny_graphs():
    """At most 4 b rows: the pencil is solved densely (same values as the generalized eigensolver)."""
    inst = _with_multipliers(_tiny(3, n=9), 2)
    g = np.random.default_rng(5).standard_normal(inst.n)
    s = np.maximum(inst.d, 1.0)
    th, rn, X, _ = B.pencil_min(OracleLagrangian(inst, 8), g, s, 8, log=lambda *a: None)
    assert abs(th[0] - _pencil_min(inst, g, s)) <= 1e-10 * max(1.0, abs(th[0]))
    assert rn.max() <= 1e-8


def test_ascent_increases_value_and_keeps_unit_rows():
    """The nonmonotone ascent never ends below its start value, keeps unit rows, and its g is the row products of
    C W at the returned W."""
    n_, keys, _ = I.sbm(200, 4, 0.1, 0.01, 9)
    inst = _with_multipliers(I.build_instance(n_, keys), 6, scale=0.01, ntri=30)
    lag = OracleLagrangian(inst, 8)
    W0 = np.random.default_rng(3).standard_normal((n_, 8))
    W0 /= np.linalg.norm(W0, axis=1, keepdims=True)
    f0 = B.value(lag, W0)[0]
    W, f, g = B.ascend(lag, W0, iters=60, on_device=False, log=lambda *a: None)
    assert f >= f0
    assert np.abs(np.linalg.norm(W, axis=1) - 1).max() <= 1e-12
    assert np.allclose(g, np.einsum("ij,ij->i", lag.CX(W), W), rtol=1e-12, atol=1e-12)
    assert abs(f - B.value(lag, W)[0]) <= 1e-12 * abs(f)


def test_bb_step_rule():
    """Alternating Barzilai-Borwein steps for the ascent (curvature sy < 0): odd iterations ss / |sy|, even |sy| / yy;
    no curvature: the step doubles; always clipped to [0.05, 50]."""
    assert B.bb_step(1, -2.0, 6.0, 4.0, 1.0) == 3.0
    assert B.bb_step(2, -2.0, 6.0, 4.0, 1.0) == 0.5
    assert B.bb_step(3, 1e-3, 6.0, 4.0, 1.5) == 3.0
    assert B.bb_step(3, 1e-3, 6.0, 4.0, 40.0) == 50.0
    assert B.bb_step(1, -1.0, 1e3, 1.0, 1.0) == 50.0 and B.bb_step(2, -1e-6, 1.0, 1.0, 1.0) == 0.05


def test_escape_directions():
    theta = np.array([-1.0, -0.5, -0.2, 0.1])
    X = np.random.default_rng(0).standard_normal((20, 4))
    U = B.escape_directions(theta, X, ratio=0.3, count=4)
    assert U.shape == (20, 2) and np.allclose(np.abs(U).max(axis=0), 0.5)
    assert B.escape_directions(np.array([0.1, 0.2]), X[:, :2]) is None
