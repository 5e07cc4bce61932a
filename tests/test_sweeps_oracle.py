"""Pins the oracle's batched row steps (or_sweeps) before it is trusted as the parity checker of lb_sweeps.

(1) The C loops agree with an independent numpy restatement over several sweeps and batch counts (this covers the
incremental bookkeeping of z, Y and the VCS Z / cnt / stZ between batches). (2) Rows stay unit. (3) With one row per
batch the first step is pinned from first principles: its direction g equals the row of the pinned evaluation
G = C(y) V plus d_i^2 / L v_i (the diagonal of B removed), and its curvature kappa equals minus the trace of the
Hessian of f in v_i (finite differences) once the p * Hessian(constraint) terms of the unit-row forms (2 B_ii,
2 d_i^2 pp_QQ, 2 d_i^2 pv_own per coordinate) are removed (f = sigma h - rho r h^2 / 2 on the active set), so that
v_new = (g + kappa/2 v) / |g + kappa/2 v| (damp = 0)."""
import numpy as np
import pytest

import oracle as O
from conftest import small_instance, chung_lu_instance


@pytest.mark.parametrize("rho", [0.0, 0.3])
@pytest.mark.parametrize("nbatch", [1, 3, 8])
def test_c_vs_numpy(rho, nbatch):
    inst, V = small_instance(n=40, seed=2, k=8, sig_scale=0.05)
    rng = np.random.default_rng(1)
    orders = np.stack([rng.permutation(inst.n) for _ in range(3)])
    Wc = O.sweeps_c(inst, V, orders, nbatch, rho, damp=0.1)
    Wn = O.sweeps_numpy(inst, V, orders, nbatch, rho, damp=0.1)
    assert np.abs(Wc - Wn).max() < 1e-10
    assert np.abs(np.linalg.norm(Wc, axis=1) - 1).max() < 1e-12


def test_c_vs_numpy_chung_lu():
    inst, V = chung_lu_instance(n=3000, seed=3, k=16)
    rng = np.random.default_rng(2)
    orders = np.stack([rng.permutation(inst.n) for _ in range(2)])
    Wc = O.sweeps_c(inst, V, orders, 16, 0.02, damp=0.0)
    Wn = O.sweeps_numpy(inst, V, orders, 16, 0.02, damp=0.0)
    assert np.abs(Wc - Wn).max() < 1e-9


def _f(inst, V, rho):
    return O.eval_c(inst, V, rho)[0][0]


@pytest.mark.parametrize("rho", [0.0, 0.4])
def test_first_step_from_first_principles(rho):
    inst, V = small_instance(n=30, seed=8, k=8, sig_scale=0.05)
    n, k = V.shape
    order = np.random.default_rng(5).permutation(n)
    i = int(order[0])
    W = O.sweeps_c(inst, V, order[None, :], n, rho, damp=0.0)
    parts, G, pp, pt, pv = O.eval_c(inst, V, rho)
    d = inst.d
    g = G[i] + d[i] ** 2 / inst.L * V[i]
    # curvature: -trace of the Hessian of f in v_i, minus the second-order parts of the unit-row forms
    eps = 1e-4
    f0 = _f(inst, V, rho)
    tr = 0.0
    for c in range(k):
        Vp = V.copy(); Vp[i, c] += eps
        Vm = V.copy(); Vm[i, c] -= eps
        tr += (_f(inst, Vp, rho) - 2 * f0 + _f(inst, Vm, rho)) / eps ** 2
    extra = -2 * k * d[i] ** 2 / inst.L                       # <B, X>: 2 B_ii per coordinate
    o = 0
    for lv in inst.levels:                                    # h_QQ = |Y_Q|^2 - dsq: +pp_QQ 2 d_i^2 per coordinate
        Q = lv.lab[i]
        extra += 2 * k * d[i] ** 2 * pp[o + Q * lv.P + Q]
        o += lv.P * lv.P
    eo = 0
    for vl in inst.vcs:                                       # own entry g = d_i (v_i.Y_R - d_i): +p_own 2 d_i^2
        for e in range(vl.ptr[i], vl.ptr[i + 1]):
            if vl.col[e] == vl.lab[i]:
                extra += 2 * k * d[i] ** 2 * pv[eo + e]
        eo += vl.nnz
    # the VCS own-entry curvature uses the exact gradient d_i (Y_R - d_i v_i) (vcsops.py:145-148) while the unit-row
    # form that f evaluates has d_i (Y_R + d_i v_i): the two squared norms differ by 4 d_i^3 (v_i . Y_R)
    corr = 0.0
    eo = 0
    for vl in inst.vcs:
        Y = np.zeros((vl.P, k))
        np.add.at(Y, vl.lab, d[:, None] * V)
        R = vl.lab[i]
        for e in range(vl.ptr[i], vl.ptr[i + 1]):
            if vl.col[e] == R and pv[eo + e] > 0:
                corr += rho * inst.wv * vl.rq[R] * 4 * d[i] * float(V[i] @ Y[R])
        eo += vl.nnz
    kappa = (extra - tr - corr) if rho > 0 else 0.0
    q = g + 0.5 * kappa * V[i]
    pred = q / np.linalg.norm(q)
    assert np.abs(W[i] - pred).max() < 2e-6, (W[i], pred)


@pytest.mark.parametrize("rho", [0.0, 0.3])
def test_aggregates_match_V_after_every_batch(rho):
    """The incremental bookkeeping between batches, checked from scratch after EVERY batch of the restatement: z = d^T V,
    each pair level's cluster sums Y = W V, each VCS level's cluster sums, and its Z_R = sum over working-set entries
    of st d_i V_i and cnt_R = #(st > 0) with the stored multipliers st; the stored multiplier of every entry whose row
    moved in the batch equals its effective value at the new V and Y (vcs_post, vcsops.py:159-213). A dropped increment,
    a wrong index or a stale st breaks one of them."""
    inst, V = small_instance(n=45, seed=4, k=8, sig_scale=0.05)
    rng = np.random.default_rng(2)
    orders = np.stack([rng.permutation(inst.n) for _ in range(2)])
    d = inst.d
    seen = {"batches": 0, "Vprev": np.array(V, np.float64)}

    def W(lab, P):
        M = np.zeros((P, inst.n))
        M[lab, np.arange(inst.n)] = d
        return M

    def hook(Vc, z, Ys, vst):
        seen["batches"] += 1
        assert np.allclose(z, d @ Vc, atol=1e-11)
        for lv, Y in zip(inst.levels, Ys):
            assert np.allclose(Y, W(lv.lab, lv.P) @ Vc, atol=1e-11)
        moved = np.flatnonzero(np.any(Vc != seen["Vprev"], axis=1))
        for vl, s_ in zip(inst.vcs, vst):
            assert np.allclose(s_["Y"], W(vl.lab, vl.P) @ Vc, atol=1e-11)
            rows, R, st = s_["rows"], s_["R"], s_["st"]
            Z = np.zeros((vl.P, Vc.shape[1]))
            np.add.at(Z, R, (st * d[rows])[:, None] * Vc[rows])
            assert np.allclose(s_["Z"], Z, atol=1e-11)
            assert np.array_equal(s_["cnt"], np.bincount(R, weights=(st > 0).astype(float), minlength=vl.P))
            sel = np.isin(rows, moved)
            if sel.any():
                fresh = O._vcs_eff(vl, Vc, s_["Y"], d, rows[sel], R[sel], s_["own"][sel], rho, inst.wv,
                                   sig=vl.sig[sel])
                assert np.allclose(st[sel], fresh, atol=1e-12)
        seen["Vprev"] = np.array(Vc, copy=True)

    O.sweeps_numpy(inst, V, orders, 5, rho, damp=0.1, hook=hook)
    assert seen["batches"] == 2 * 5
