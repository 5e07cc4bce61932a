"""The lifted sparse certificate on the CPU (host factorization, dense core on the host).

Pinned by:
- the lifting identity: for unit rows V, the lifted vectors (v_i, and a_p = s_p / vol_p for every tree node) give
  <Mt - Diag(sdiag), Xt> = sum(g) - <C(y), V V^T> with C(y) assembled independently (bound.dense_C), and every link
  has e_p^T Xt e_p = 0 -- a dropped term, a wrong sign or a wrong tree coefficient breaks one of them;
- the verdict against the lifted matrix's own smallest eigenvalue (the factorization succeeds above it and fails below);
- the dense certificate (vertex space) at the same g and vertex shift: the UBs differ by the auxiliary shift and the
  two rounding prices;
- brute force: every certified UB is at least the best modularity over all partitions of 9-vertex graphs.
"""
import numpy as np
import pytest

from lagbound_amd import bound as B
from lagbound_amd import instances as I
from lagbound_amd import sparse_cert as S
from oracle_engine import OracleLagrangian
from test_bound import _modularity, _partitions, _tiny, _with_multipliers


def _lifted_vectors(lf, V, d):
    """Lifted vectors: vertex rows V, auxiliary rows a_p = (c_1 x_(child 1) + c_2 x_(child 2)) / vol_p (children
    first in creation order)."""
    X = np.zeros((lf.N, V.shape[1]))
    X[:len(V)] = V
    for p, vol, ch, cf in zip(lf.link_p, lf.link_vol, lf.link_ch, lf.link_cf):
        X[p] = (cf[0] * X[ch[0]] + cf[1] * X[ch[1]]) / vol
    return X


def _instance(seed, n=40, ntri=40):
    n_, keys, lab = I.sbm(n, 4, 0.3, 0.05, seed)
    return _with_multipliers(I.build_instance(n_, keys), seed + 3, scale=0.05, ntri=ntri), lab


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lifting_identity(seed):
    inst, _ = _instance(seed)
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(inst.n)
    lf = S.Lifted(inst, g, key=rng.permutation(inst.n))
    sdiag = rng.random(lf.N)
    Ml, price, _ = lf.assemble(sdiag)
    M = S.full_symmetric(Ml).toarray()
    assert np.array_equal(M, M.T)
    Cm, _ = B.dense_C(inst)
    for _ in range(3):
        V = rng.standard_normal((inst.n, 5))
        V /= np.linalg.norm(V, axis=1, keepdims=True)
        Xl = _lifted_vectors(lf, V, inst.d)
        Xt = Xl @ Xl.T
        lhs = float(np.sum((M - np.diag(sdiag)) * Xt))
        rhs = float(g.sum()) - float(np.sum(Cm * (V @ V.T)))
        assert abs(lhs - rhs) <= 1e-9 * (1 + abs(rhs))
        assert np.abs(np.diag(Xt)).max() <= 1 + 1e-12
        for p, vol, ch, cf in zip(lf.link_p, lf.link_vol, lf.link_ch, lf.link_cf):
            e = np.zeros(lf.N)
            e[p], e[ch[0]], e[ch[1]] = vol, -cf[0], -cf[1]
            assert abs(e @ Xt @ e) <= 1e-9 * vol * vol
    # the root holds the whole volume, every level's cluster nodes their cluster's volume
    assert lf.rootvol == inst.L
    for (node, coef), lv in zip(lf.lv_nodes, inst.levels):
        assert np.allclose(coef, np.bincount(lv.lab, weights=inst.d, minlength=lv.P))
    assert price > 0


def test_verdict_follows_the_lifted_eigenvalue():
    """The factorization succeeds at shifts above the lifted matrix's PSD threshold and fails below it."""
    inst, _ = _instance(4, n=50)
    lag = OracleLagrangian(inst, 8)
    V = np.random.default_rng(1).standard_normal((inst.n, 8))
    _, _, g = B.ascend(lag, V / np.linalg.norm(V, axis=1, keepdims=True), iters=30, on_device=False,
                       log=lambda *a: None)
    lf = S.Lifted(inst, g)
    M0 = S.full_symmetric(lf.assemble(np.zeros(lf.N))[0]).toarray()
    w = lf.w
    mu = float(np.linalg.eigvalsh(M0 / np.sqrt(np.outer(w, w)))[0])      # smallest eigenvalue of (M0, Diag(w))
    assert mu < 0
    for core_deg in (10 ** 9, 4):                # all sparse; and a dense core on the host
        hi = S.factor_price(lf.assemble((-mu * 1.001 + 1e-9) * w)[0], core_deg=core_deg)
        lo = S.factor_price(lf.assemble((-mu * 0.98) * w)[0], core_deg=core_deg)
        assert hi[0] and not lo[0], (core_deg, hi[2], lo[2])
        assert hi[1] > 0


def test_sparse_agrees_with_dense_certificate():
    """At the same g and vertex shift t s the lifted and the dense certificates both pass, and their UBs differ by the
    auxiliary nodes' shift (t sa per node) up to the two rounding prices."""
    n_, keys, lab = I.sbm(160, 4, 0.2, 0.02, 3)
    inst = _with_multipliers(I.build_instance(n_, keys), 9, scale=0.01, ntri=40)
    est = B.estimate(inst, np.random.default_rng(0).standard_normal((n_, 8)), lag=OracleLagrangian(inst, 8),
                     rgd_iters=60, eig_maxiter=200, log=lambda *a: None)
    sp_ = S.certify_estimate_sparse(inst, est, margin=1e-7)
    assert sp_["passed"], sp_
    shift = sp_["info"]["shift"]
    dn = B.certify_dense(inst, est["g"], shift * est["s"])
    assert dn["passed"]
    aux = shift * 0.01 * (sp_["info"]["N"] - inst.n) / inst.L
    assert abs(sp_["UB"] - (dn["UB"] + aux)) <= sp_["price"] + dn["price"] + 1e-12
    assert sp_["UB"] >= _modularity(inst, lab)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_certified_bound_is_above_the_optimum(seed):
    """9-vertex graphs: the certified UB is at least max over all partitions of the modularity (21147 partitions)."""
    inst = _with_multipliers(_tiny(seed, n=9), seed + 20, scale=0.05, ntri=10)
    lag = OracleLagrangian(inst, 8)
    est = B.estimate(inst, np.random.default_rng(seed).standard_normal((inst.n, 8)), lag=lag, rgd_iters=40,
                     eig_maxiter=100, log=lambda *a: None)
    res = S.certify_estimate_sparse(inst, est)
    assert res["passed"], res
    best = max(_modularity(inst, lab) for lab in _partitions(inst.n))
    assert res["UB"] >= best


def test_refuses_negative_multipliers():
    inst, _ = _instance(5)
    inst.lam[0] = -1e-3
    with pytest.raises(ValueError):
        S.certify_sparse(inst, np.ones(inst.n), 1.0)
