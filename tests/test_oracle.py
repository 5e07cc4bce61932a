"""Pins the FP64 oracle (oracle/lb_oracle.c + oracle/oracle.py) before it is trusted as the parity checker.

No golden vector of the reference pins f, G or the PHR multipliers (its tests only bracket end-to-end bounds), so the
evaluation is pinned by: (1) a dense brute-force evaluation from the mathematical definitions, (2) central finite
differences (grad f = 2 C(y) V in the tangent space of the unit rows), (3) the quadA identity <BV, V> = quadA -
|d^T V|^2 / L (alm_vcs.py:465), (4) the rho -> 0 limit of the PHR value (augmented -> C0 + sum(lam) + <C(y), X>,
est_dual_vcs.py:245-246), (5) hand-computed single-constraint cases, and (6) agreement of the C loops with an
independent numpy/scipy restatement. Planted mistakes (a dropped rank-one term, a flipped triangle sign) are shown to
fail these checks.
"""
import numpy as np
import pytest

import oracle as O
from conftest import small_instance, chung_lu_instance


def _tangent(V, rng):
    D = rng.standard_normal(V.shape)
    return D - np.einsum("ij,ij->i", D, V)[:, None] * V


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("rho", [0.0, 0.3, 5.0])
@pytest.mark.parametrize("weighted", [False, True])
def test_three_way_small(seed, rho, weighted):
    inst, V = small_instance(n=36, seed=seed, weighted=weighted)
    pc, Gc, ppc, ptc, pvc = O.eval_c(inst, V, rho)
    pn, Gn, ppn, ptn, pvn = O.eval_numpy(inst, V, rho)
    fd, Gd, ppd, ptd, pvd = O.eval_dense(inst, V, rho)
    scale = max(1.0, abs(fd))
    assert abs(pc[0] - fd) <= 1e-11 * scale
    assert abs(pn[0] - fd) <= 1e-11 * scale
    gs = max(1.0, np.abs(Gd).max())
    assert np.abs(Gc - Gd).max() <= 1e-11 * gs
    assert np.abs(Gn - Gd).max() <= 1e-11 * gs
    assert np.allclose(ppc, np.concatenate([p.ravel() for p in ppd]), rtol=1e-12, atol=1e-14)
    assert np.allclose(ptc, ptd, rtol=1e-12, atol=1e-14)
    assert np.allclose(pvc, np.concatenate(pvd), rtol=1e-12, atol=1e-14)
    assert np.allclose(pc, pn, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("rho", [0.0, 0.2, 2.0])
def test_finite_differences(rho):
    inst, V = small_instance(n=30, seed=4, sig_scale=0.05)
    rng = np.random.default_rng(9)
    D = _tangent(V, rng)
    _, G, _, _, _ = O.eval_c(inst, V, rho)
    eps = 1e-5
    fp = O.eval_c(inst, V + eps * D, rho)[0][0]
    fm = O.eval_c(inst, V - eps * D, rho)[0][0]
    fd = (fp - fm) / (2 * eps)
    an = 2.0 * float((G * D).sum())
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


def test_planted_mistakes_are_caught():
    """A dropped rank-one term or a flipped triangle sign in G fails the finite-difference check."""
    inst, V = small_instance(n=30, seed=5, sig_scale=0.05)
    rng = np.random.default_rng(3)
    D = _tangent(V, rng)
    rho = 0.5
    _, G, _, pt, _ = O.eval_c(inst, V, rho)
    eps = 1e-5
    fd = (O.eval_c(inst, V + eps * D, rho)[0][0] - O.eval_c(inst, V - eps * D, rho)[0][0]) / (2 * eps)
    z = inst.d @ V
    G_norank1 = G + np.outer(inst.d, z) / inst.L
    # flipped triangle contribution: subtract twice T(pt) V
    Tv = np.zeros_like(G)
    a, b, c = (inst.TI[:, q].astype(int) for q in range(3))
    h = 0.5 * pt[:, None]
    np.add.at(Tv, a, h * (V[c] - V[b])); np.add.at(Tv, b, -h * (V[a] + V[c])); np.add.at(Tv, c, h * (V[a] - V[b]))
    G_flip = G - 2 * Tv
    for bad in (G_norank1, G_flip):
        an = 2.0 * float((bad * D).sum())
        assert abs(fd - an) > 1e-4 * max(1.0, abs(fd))


def test_quadA_identity():
    inst, V = small_instance(n=40, seed=6)
    p, _, _, _, _ = O.eval_c(inst, V, 0.5)
    B = inst.dense_A() - np.outer(inst.d, inst.d) / inst.L
    assert abs((p[1] - p[2]) - float(np.trace(V.T @ B @ V))) <= 1e-10 * max(1.0, abs(p[1]))
    assert abs(p[6] - (float(np.trace(V.T @ B @ V)) + inst.C0)) <= 1e-10 * max(1.0, abs(p[6]))


def test_rho_zero_limit():
    """augmented(V, rho) -> C0 + sum(lam) + <C(y), V V^T> as rho -> 0 (unit rows)."""
    inst, V = small_instance(n=40, seed=7, sig_scale=0.05)
    f0 = O.eval_c(inst, V, 0.0)[0][0]
    f1 = O.eval_c(inst, V, 1e-7)[0][0]
    assert abs(f1 - f0) <= 1e-5 * max(1.0, abs(f0)), (f0, f1)


def test_single_triangle_closed_form():
    """One triangle on a path a - b - c with known V: tau, pt and the PHR value by hand."""
    from lagbound_amd import instances as I
    n = 3
    inst = I.build_instance(n, np.array([0 * 3 + 1, 1 * 3 + 2], np.int64))
    inst.TI = np.array([[0, 1, 2]], np.int32)
    inst.lam = np.array([0.7])
    th = 0.4
    V = np.array([[1.0, 0.0], [np.cos(th), np.sin(th)], [np.cos(2 * th), np.sin(2 * th)]])
    tau = np.cos(th) + np.cos(th) - np.cos(2 * th)
    rho = 0.3
    g = 1.0 - tau
    p = max(0.0, 0.7 - rho * g)
    parts, G, _, pt, _ = O.eval_c(inst, V, rho)
    assert abs(pt[0] - p) < 1e-15
    assert abs(parts[4] - (p * p - 0.49) / (2 * rho)) < 1e-14
    # row a of T(pt) V = pt/2 (V_c - V_b); rows of B V by hand
    d = inst.d
    BV = inst.dense_A() @ V - np.outer(d, d @ V) / inst.L
    assert np.allclose(G[0], BV[0] + 0.5 * p * (V[2] - V[1]), atol=1e-14)
    assert np.allclose(G[1], BV[1] - 0.5 * p * (V[0] + V[2]), atol=1e-14)
    assert np.allclose(G[2], BV[2] + 0.5 * p * (V[0] - V[1]), atol=1e-14)


def test_c_vs_numpy_medium():
    inst, V = chung_lu_instance(n=5000, seed=2, k=16)
    for rho in (0.0, 0.02):
        pc, Gc, ppc, ptc, pvc = O.eval_c(inst, V, rho)
        pn, Gn, ppn, ptn, pvn = O.eval_numpy(inst, V, rho)
        assert abs(pc[0] - pn[0]) <= 1e-10 * max(1.0, abs(pn[0]))
        assert np.linalg.norm(Gc - Gn) <= 1e-11 * np.linalg.norm(Gn)
        assert np.allclose(ptc, ptn, rtol=1e-12, atol=1e-14)
        assert np.allclose(pvc, pvn, rtol=1e-12, atol=1e-14)
        assert np.allclose(ppc, ppn, rtol=1e-10, atol=1e-14)


def test_golden_fixture():
    """Committed fixture (tests/golden/eval_small.npz, written by tests/golden/make_golden.py from the dense
    brute-force evaluation) reproduces with the C oracle."""
    import os
    from lagbound_amd import instances as I
    path = os.path.join(os.path.dirname(__file__), "golden", "eval_small.npz")
    z = np.load(path)
    inst = I.Instance(n=int(z["n"]), rowptr=z["rowptr"], col=z["col"], w=None, d=z["d"], L=float(z["L"]), C0=float(z["C0"]))
    inst.levels.append(I.PairLevelData(lab=z["lab"], P=int(z["P"]), sig=z["sig"], rq=z["rq"], dsq=z["dsq"]))
    inst.vcs.append(I.VCSLevelData(lab=z["vlab"], P=int(z["vP"]), rq=z["vrq"], ptr=z["vptr"], col=z["vcol"], sig=z["vsig"]))
    inst.TI = z["TI"]
    inst.lam = z["lam"]
    for tag, rho in (("r0", 0.0), ("r1", float(z["rho"]))):
        p, G, pp, pt, pv = O.eval_c(inst, z["V"], rho)
        assert abs(p[0] - float(z[f"f_{tag}"])) <= 1e-11 * max(1.0, abs(float(z[f"f_{tag}"])))
        assert np.abs(G - z[f"G_{tag}"]).max() <= 1e-11 * max(1.0, np.abs(z[f"G_{tag}"]).max())
        assert np.allclose(pt, z[f"pt_{tag}"], rtol=1e-12, atol=1e-14)
        assert np.allclose(pv, z[f"pv_{tag}"], rtol=1e-12, atol=1e-14)
