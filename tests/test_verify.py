"""The independent verifier (lagbound_amd.verify) on the CPU.

Pinned by:
- the certifier's own certificates: every passed certify_sparse record verifies, with a UB within the two rounding
  prices of the certifier's;
- brute force: every verified UB is at least the best modularity over all partitions of 9-vertex graphs;
- planted mistakes: a wrong sign in the verifier's own assembly is caught by the lifting identity, and tampered records
  (sdiag lowered, a link coefficient changed, two cluster nodes swapped, a negative multiplier, an asymmetric edge
  weight, a wrong C0, g lowered until the matrix is indefinite) are all rejected;
- the factorization's running bound against the exact rational residual sum |A - L D L^T| of its own factor;
- the reduction check on the reference's own fixtures (tests/test_verifier.py:40-87 there): eligible leaf and twin
  contractions pass with the same union counts, ineligible ones are rejected, and Q_orig = reduced Q exactly.
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

from lagbound_amd import bound as B
from lagbound_amd import instances as I
from lagbound_amd import sparse_cert as S
from lagbound_amd import verify as Vf
from oracle_engine import OracleLagrangian
from test_bound import _modularity, _partitions, _tiny, _with_multipliers


def _certified(inst, seed=0, rgd=60, margin=1e-7):
    est = B.estimate(inst, np.random.default_rng(seed).standard_normal((inst.n, 8)), lag=OracleLagrangian(inst, 8),
                     rgd_iters=rgd, eig_maxiter=200, log=lambda *a: None)
    base = max(0.0, -est["lmin"])
    for k in range(8):
        res = S.certify_sparse(inst, est["g"], (2.0 ** k) * base + margin * 10 ** k, keep=True)
        if res["passed"]:
            return res
    raise AssertionError("no certificate")


def _sbm_instance(seed, n=120):
    n_, keys, lab = I.sbm(n, 4, 0.25, 0.03, seed)
    return _with_multipliers(I.build_instance(n_, keys), seed + 9, scale=0.01, ntri=40), lab


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_certifier_records_verify(seed, tmp_path):
    inst, lab = _sbm_instance(seed)
    res = _certified(inst, seed)
    path = str(tmp_path / "cert.npz")
    S.write_certificate(path, inst, res)
    out = Vf.verify(path)
    assert out["passed"], out
    assert out["identity_relerr"] < 1e-12
    assert abs(out["UB"] - res["UB"]) <= 2 * (res["price"] + out["price_asm"] + out["price_psd"]) + 1e-12
    assert out["UB"] >= _modularity(inst, lab)
    assert Vf.main([path]) == 0


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_verified_bound_is_above_the_optimum(seed):
    """9-vertex graphs (21147 partitions): the verified UB is at least the best modularity."""
    inst = _with_multipliers(_tiny(seed, n=9), seed + 20, scale=0.05, ntri=10)
    res = _certified(inst, seed, rgd=40)
    out = Vf.verify(S.certificate_record(inst, res))
    assert out["passed"], out
    assert out["UB"] >= max(_modularity(inst, lab) for lab in _partitions(inst.n))


def test_planted_assembly_mistake_is_caught_by_the_identity(monkeypatch):
    """The identity check compares the verifier's lifted assembly with a vertex-space C(y): flipping the sign of
    one family of terms in the assembly must show up there."""
    inst, _ = _sbm_instance(4)
    rec = S.certificate_record(inst, _certified(inst, 4))
    assert Vf.verify(rec)["passed"]
    bad = dict(rec)
    orig_assemble = Vf.assemble

    def wrong(c, g, vol):        # the triangle terms with the wrong sign
        c2 = dict(c)
        c2["lam"] = -c["lam"]
        return orig_assemble(c2, g, vol)

    monkeypatch.setattr(Vf, "assemble", wrong)
    out = Vf.verify(bad)
    assert not out["passed"] and "does not represent" in out["fail"]


def test_tampered_records_are_rejected():
    inst, _ = _sbm_instance(5)
    rec = S.certificate_record(inst, _certified(inst, 5))
    n = inst.n

    def rejected(mod, match):
        c = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in rec.items()}
        mod(c)
        out = Vf.verify(c)
        assert not out["passed"] and match in out["fail"], (match, out.get("fail"))

    rejected(lambda c: c["sdiag"].__setitem__(0, -1.0), "sdiag")
    rejected(lambda c: c["link_cf"].__setitem__((0, 0), c["link_cf"][0, 0] + 1.0), "coefficient")
    rejected(lambda c: c["lam"].__setitem__(0, -1e-3), "multipliers")
    rejected(lambda c: c.__setitem__("C0", np.float64(2.0)), "C0")

    def swap_nodes(c):
        node = c["lv0_node"]
        a, b = np.flatnonzero(node >= n)[:2]
        node[a], node[b] = node[b], node[a]
    rejected(swap_nodes, "lv0")

    def asym(c):
        w = np.ones(len(c["col"]))
        w[0] = 2.0
        c["w"] = w
    rejected(asym, "symmetric")
    # g lowered on every vertex: the lifted matrix is no longer PSD
    rejected(lambda c: c.__setitem__("g", c["g"] - 2.0 * np.maximum(inst.d, 1.0)), "PSD")
    # and the untouched record still verifies
    assert Vf.verify(rec)["passed"]


@pytest.mark.parametrize("seed", [0, 1])
def test_running_bound_covers_the_exact_residual(seed):
    """Random sparse symmetric positive definite matrices (and one with a 1e-6 smallest eigenvalue): the exact rational
    sum |A - L D L^T| of the computed factor is at most the price, D matches the dense Cholesky's squared diagonal."""
    rng = np.random.default_rng(seed)
    N = 40
    for tight in (False, True):
        R = sp.random(N, N, density=0.15, random_state=seed, format="csr")
        A = (R @ R.T).toarray() + np.diag(rng.random(N) * 0.1)
        if tight:
            w, Q = np.linalg.eigh(A)
            w[0] = 1e-6 * w[-1]
            A = (Q * w) @ Q.T
            A = 0.5 * (A + A.T)
            A[np.abs(A) < 1e-3 * np.abs(A).max()] = 0.0
            w2 = np.linalg.eigvalsh(A)
            if w2[0] <= 0:
                A += np.eye(N) * (2 * abs(w2[0]) + 1e-8 * w2[-1])
        F = sp.csr_matrix(A)
        F.sort_indices()
        perm = rng.permutation(N)
        st, price, info, fac = Vf.ldl_price(F, perm, want_factor=True)
        assert st == "ok", info
        D, Lp, Li, Lx = fac
        Lm = np.eye(N)
        for j in range(N):
            for p in range(Lp[j], Lp[j + 1]):
                Lm[Li[p], j] = Lx[p]
        Ap = A[np.ix_(perm, perm)]
        resid = Fraction(0)
        for k in range(N):
            for j in range(N):
                s = Fraction(Ap[k, j])
                for i in range(min(k, j) + 1):
                    if Lm[k, i] != 0 and Lm[j, i] != 0:
                        s -= Fraction(Lm[k, i]) * Fraction(D[i]) * Fraction(Lm[j, i])
                resid += abs(s)
        assert resid <= Fraction(price), (float(resid), price)
        c = np.linalg.cholesky(Ap)
        assert np.allclose(D, np.diag(c) ** 2, rtol=1e-8)


def test_indefinite_matrix_fails():
    A = np.eye(6)
    A[2, 2] = -1e-9
    F = sp.csr_matrix(A)
    st, _, info, _ = Vf.ldl_price(F, np.arange(6))
    assert st == "not_psd" and info["failed_pivot"] == 3


def test_exact_sum_survives_cancellation():
    """The total is summed exactly: -1e16 + 1 (a case where nested rounded sums lose the 1)."""
    assert Vf.rational_sum(np.array([-1e16, 1.0, 1e16])) == Fraction(1)
    assert Fraction(Vf.round_up(Fraction(1, 3))) >= Fraction(1, 3)


# ------------------------------------------------------------------------------------------------ reduction
GRAPH = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (3, 4), (4, 5), (5, 6), (4, 6), (0, 7)]


def _reduced(edges, groups):
    """The contraction of an edge list by `groups`, in the reduced-instance form check_reduction reads."""
    edges = np.sort(np.asarray(edges, np.int64), axis=1)
    n0 = int(edges.max()) + 1
    grp = np.unique(np.asarray(groups), return_inverse=True)[1].astype(np.int64)
    n = int(grp.max()) + 1
    deg = np.bincount(edges.ravel(), minlength=n0)
    gu, gv = grp[edges[:, 0]], grp[edges[:, 1]]
    inside = gu == gv
    kk, w = np.unique(np.minimum(gu, gv)[~inside] * n + np.maximum(gu, gv)[~inside], return_counts=True)
    red = dict(n=n, u=kk // n, v=kk % n, w=w.astype(np.float64),
               d=np.bincount(grp, weights=deg, minlength=n).astype(np.float64), L=float(2 * len(edges)),
               C0=float(2 * inside.sum()))
    return n0, edges[:, 0] * n0 + edges[:, 1], grp, red


def _check(edges, groups, labels=None):
    n0, keys, grp, red = _reduced(edges, groups)
    return Vf.check_reduction(n0, keys, grp, labels, reduced=red)


def test_eligible_contractions_pass():
    res = _check(GRAPH, [0, 1, 1, 2, 3, 4, 4, 0])
    assert res["bookkeeping"] == "ok"
    assert (res["leaf_unions"], res["twin_unions"], res["ineligible_inside_edges"]) == (1, 2, 0)
    assert _check(GRAPH, [0, 1, 1, 2, 3, 4, 5, 6])["twin_unions"] == 1          # a partial contraction


@pytest.mark.parametrize("edges,groups", [
    (GRAPH, [0, 1, 2, 3, 3, 4, 5, 6]),                                  # adjacent, not twins
    ([(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 0, 2]),                   # twins, not adjacent
    ([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (1, 4)], [0, 0, 1, 2, 3]),   # deg^2 >= 2m
    ([(0, 1), (1, 2), (2, 3)], [0, 0, 0, 0]),                           # a path in one group
    ([(0, 1), (2, 3)], [0, 1, 0, 1]),                                   # endpoints of two edges
])
def test_ineligible_contractions_are_rejected(edges, groups):
    with pytest.raises(Vf.Rejected, match="eligible"):
        _check(edges, groups)


def test_reduction_bookkeeping_and_identity():
    import oracle as O
    n0, keys = I.add_twins_and_leaves(*I.sbm(60, 3, 0.3, 0.05, 2)[:2], ntwin=6, nleaf=8, seed=3)
    r = O.reduce_c(n0, keys)
    grp, N = r["group_of"], r["N"]
    red = dict(n=N, u=r["keys"] // N, v=r["keys"] % N, w=r["weights"], d=r["vol"].astype(np.float64),
               L=float(2 * len(keys)), C0=float(2 * r["inside_edges"]))
    res = Vf.check_reduction(n0, keys, grp, None, reduced=red)
    assert res["leaf_unions"] + res["twin_unions"] > 0
    labr = np.random.default_rng(1).integers(0, 4, N)
    res = Vf.check_reduction(n0, keys, grp, labr, reduced=red)
    assert res["identity"] == "exact"
    bad = dict(red, C0=red["C0"] + 2)
    with pytest.raises(Vf.Rejected, match="C0"):
        Vf.check_reduction(n0, keys, grp, None, reduced=bad)
