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
