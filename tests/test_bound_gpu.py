"""Candidate bound on the device: C(y) X products against the oracle, the estimate against the oracle-operator
estimate, |U_fp32 - U_fp64|, and the full chain solver -> estimate -> dense certificate through the drop-in stages."""
import json
import os

import numpy as np
import pytest

from lagbound_amd import bound as B
from lagbound_amd import instances as I
from lagbound_amd import pipeline
from alm_case import _write
from oracle_engine import OracleLagrangian
from test_bound import _modularity, _with_multipliers

pytestmark = pytest.mark.gpu


def _inst(n=600, seed=1):
    n_, keys, lab = I.sbm(n, 6, 0.06, 0.005, seed)
    return _with_multipliers(I.build_instance(n_, keys), 9, scale=0.01, ntri=200), lab


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_device_CX_matches_oracle(dtype, tol):
    inst, _ = _inst()
    X = np.random.default_rng(0).standard_normal((inst.n, 20))
    print(f"[bound] CX dtype={dtype} n={inst.n} width=20", flush=True)
    lag = B.DeviceLagrangian(inst, r=8, dtype=dtype)
    got = lag.CX(X)
    lag.close()
    ref = OracleLagrangian(inst, 8).CX(X)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= tol, err


def test_estimate_device_matches_oracle_and_fp32():
    inst, lab = _inst()
    V = np.random.default_rng(2).standard_normal((inst.n, 8))
    kw = dict(rgd_iters=40, eig_maxiter=120, log=lambda *a: None)
    print("[bound] estimate f64 device vs oracle", flush=True)
    ref = B.estimate(inst, V, lag=OracleLagrangian(inst, 8), **kw)
    e64 = B.estimate(inst, V, dtype="f64", eig_device=False, **kw)     # the same eigensolver as the reference run
    assert abs(e64["dv"] - ref["dv"]) <= 1e-9 * abs(ref["dv"])
    assert abs(e64["ub"] - ref["ub"]) <= 1e-7 * abs(ref["ub"])
    print("[bound] estimate f32 device (FP32 ascent, FP64 assembly and device pencil)", flush=True)
    e32 = B.estimate(inst, V, dtype="f32", **kw)
    du = abs(e32["ub"] - e64["ub"])
    print(f"[bound] U_f64 {e64['ub']:.9f} U_f32 {e32['ub']:.9f} |dU| {du:.3e}", flush=True)
    assert du <= 1e-4 * abs(e64["ub"])
    for e in (e64, e32):                  # both pencils converged to the stated tolerance (bound units)
        rec = e["s_vol"]
        assert rec["resid"][0] * rec["sum_s"] / inst.L <= 1e-5
    cert = B.certify_estimate(inst, e64)
    assert cert["passed"] and cert["UB"] >= _modularity(inst, lab)


def test_device_blocks_match_numpy():
    """lb_blk_*: Gram matrices (with and without a diagonal weight), combinations, residual / scaling and the operator
    Diag(g) X - C X against numpy and the oracle's C X, through the degree numbering of the context."""
    inst, _ = _inst()
    n, r = inst.n, 8
    rng = np.random.default_rng(4)
    X, Y, Z = (rng.standard_normal((n, r)) for _ in range(3))
    w, gd = rng.random(n) + 0.5, rng.standard_normal(n)
    lag = B.DeviceLagrangian(inst, r=r, dtype="f64")
    e = lag.eng.blk_engine()
    e.blk_alloc(6, 2)
    e.blk_set_diag(0, gd)
    e.blk_set_diag(1, w)
    for k, A in enumerate((X, Y, Z)):
        e.blk_set(k, A)
    assert np.array_equal(e.blk_get(1), Y)
    S = np.c_[X, Y]
    T = np.c_[Y, Z]
    tol = 1e-12 * n
    assert np.abs(e.blk_gram([0, 1], [1, 2]) - S.T @ T).max() <= tol
    assert np.abs(e.blk_gram([0, 1], [1, 2], 1) - S.T @ (w[:, None] * T)).max() <= tol
    C = rng.standard_normal((2 * r, r))
    e.blk_comb(3, [0, 1], C)
    assert np.abs(e.blk_get(3) - S @ C).max() <= 1e-12 * np.abs(S @ C).max()
    e.blk_comb(3, [2], C[:r], accumulate=True)
    assert np.abs(e.blk_get(3) - (S @ C + Z @ C[:r])).max() <= 1e-12 * np.abs(S @ C).max()
    th = rng.standard_normal(r)
    e.blk_elem(2, 4, 0, 1, 1, th)
    assert np.abs(e.blk_get(4) - (Y - w[:, None] * X * th[None, :])).max() <= 1e-12
    e.blk_elem(1, 4, 2, -1, 1)
    assert np.abs(e.blk_get(4) - w[:, None] * Z).max() <= 1e-12
    e.blk_op(5, 0, 0)
    ref = gd[:, None] * X - OracleLagrangian(inst, r).CX(X)
    assert np.abs(e.blk_get(5) - ref).max() <= 1e-11 * np.abs(ref).max()
    lag.close()


def test_device_lobpcg_matches_host():
    """The device-block LOBPCG reaches the host LOBPCG's smallest pencil eigenvalue (both converged to the same
    absolute tolerance in bound units), and an estimate through it certifies."""
    inst, lab = _inst()
    V = np.random.default_rng(2).standard_normal((inst.n, 8))
    W = V / np.linalg.norm(V, axis=1, keepdims=True)
    lag = B.DeviceLagrangian(inst, r=8, dtype="f64")
    dv, g, _ = B.value(lag, W)
    s = np.maximum(inst.d, 1.0)
    sums = float(s.sum())
    kw = dict(maxiter=200, chunk=20, atol=1e-7, log=lambda *a: None)
    th_h, rn_h, _, _ = B.pencil_min(lag, g, s, 8, rng=np.random.default_rng(0), on_device=False, **kw)
    th_d, rn_d, Xe, tr = B.pencil_min(lag, g, s, 8, rng=np.random.default_rng(0), **kw)
    print(f"[bound] lmin host {th_h[0]:.10e} (r {rn_h[0]:.2e}) device {th_d[0]:.10e} (r {rn_d[0]:.2e}) "
          f"its {tr[-1][0] if tr else 0}", flush=True)
    assert Xe.shape == (inst.n, 8)
    assert abs(th_d[0] - th_h[0]) * sums / inst.L <= 1e-6
    assert rn_d[0] * sums / inst.L <= 1e-5
    lag.close()
    est = B.estimate(inst, V, dtype="f64", rgd_iters=40, eig_maxiter=120, log=lambda *a: None)
    ref = B.estimate(inst, V, dtype="f64", rgd_iters=40, eig_maxiter=120, eig_device=False, log=lambda *a: None)
    print(f"[bound] ub device-block LOBPCG {est['ub']:.9f}, host LOBPCG {ref['ub']:.9f}", flush=True)
    assert abs(est["ub"] - ref["ub"]) <= 1e-4 * abs(ref["ub"])
    cert = B.certify_estimate(inst, est)
    assert cert["passed"] and cert["UB"] >= _modularity(inst, lab)


def test_solve_estimate_certify_end_to_end(tmp_path):
    ws = pipeline.Workspace(str(tmp_path))
    D, n, keys = _write(str(tmp_path), n=2000, blocks=10)
    pairs = [os.path.join(D, "planted.npy"), os.path.join(D, "coarse.npy")]
    vc = os.path.join(D, "planted.npy")
    print("[bound] solve n=2000", flush=True)
    ck, w = ws.solve("main", pairs, vc, 0.5, 120.0, 30, 8, init=os.path.join(D, "planted.npy"), k=16,
                     extra_args=["--tri_every", "5", "--vcs_every", "4", "--tri_add", "20000", "--tri1_add", "10000"])
    print("[bound] estimate", flush=True)
    res = ws.estimate("est", ck, 0.5, polish_time=120, eig_time=120, extra_args=["--certify", "1"])
    cert = res["certificate"]
    assert cert["passed"], cert
    from lagbound_amd.alm import load_instance
    inst = load_instance(D)
    q = _modularity(inst, np.load(os.path.join(D, "planted.npy")))
    print(f"[bound] Q(planted) {q:.6f} dv {res['dv']:.6f} ub {res['ub']:.6f} certified {cert['UB']:.9f} "
          f"price {cert['price']:.2e}", flush=True)
    assert cert["UB"] >= q
    assert abs(cert["UB"] - res["ub"]) <= 1e-7 * res["s_vol"]["sum_s"] / inst.L + cert["price"] + 1e-12


@pytest.mark.parametrize("weighted", [False, True])
def test_device_certificate_matches_host(weighted):
    n_, keys, lab = I.sbm(500, 5, 0.08, 0.01, 4)
    w = np.random.default_rng(1).integers(1, 4, len(keys)).astype(np.float64) if weighted else None
    inst = _with_multipliers(I.build_instance(n_, keys, weights=w), 3, scale=0.01, ntri=300)
    Cm, _ = B.dense_C(inst)
    W = np.random.default_rng(2).standard_normal((n_, 8))
    W /= np.linalg.norm(W, axis=1, keepdims=True)
    g = np.einsum("ij,ij->i", Cm @ W, W)
    s = np.maximum(inst.d, 1.0)
    import scipy.linalg as sl
    mu = float(sl.eigh(np.diag(g) - Cm, np.diag(s), eigvals_only=True)[0])
    print(f"[bound] device certificate weighted={weighted} n={n_} mu={mu:.6e}", flush=True)
    t = (max(0.0, -mu) + 1e-7) * s
    host = B.certify_dense(inst, g, t)
    dev = B.certify_dense_device(inst, g, t)
    assert host["passed"] and dev["passed"]
    assert abs(dev["UB"] - host["UB"]) <= 1e-9 * abs(host["UB"])
    assert dev["price"] < 1e-9
    bad = B.certify_dense_device(inst, g, (max(0.0, -mu) - 1e-5) * s)
    assert not bad["passed"]


@pytest.mark.parametrize("dtype,ftol,vtol", [("f64", 1e-10, 1e-7), ("f32", 1e-6, 1e-3)])
def test_device_polish_matches_host_polish(dtype, ftol, vtol):
    """The ascent with its row work on the device (DeviceRows, FP64 iterate) against the same ascent with its row work
    in numpy (HostRows) on the same device operator: the same step and backtracking decisions, the same iterate up to
    summation order."""
    inst, _ = _inst()
    V = np.random.default_rng(3).standard_normal((inst.n, 8))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    print(f"[bound] device polish dtype={dtype} n={inst.n}", flush=True)
    lag = B.DeviceLagrangian(inst, r=8, dtype=dtype)
    try:
        Vh, fh, gh = B.ascend(lag, V, iters=40, on_device=False, log=lambda *a: None)
        Vd, fd, gd = B.ascend(lag, V, iters=40, on_device=True, log=lambda *a: None)
    finally:
        lag.close()
    print(f"[bound] f host {fh:.12f} device {fd:.12f}", flush=True)
    assert abs(fd - fh) <= ftol * abs(fh)
    assert np.abs(Vd - Vh).max() <= vtol
    assert np.abs(np.linalg.norm(Vd, axis=1) - 1).max() <= 1e-12
    assert np.abs(gd - gh).max() <= vtol * max(1.0, np.abs(gh).max())


def test_new_graph_invalidates_blocks_and_polish():
    """lb_set_graph releases the eigensolver block reservation and the polish state, which are sized by n: after a
    second, larger graph the block and polish calls return -1 (usage error) instead of indexing the old buffers."""
    import ctypes as C
    from lagbound_amd import _lib
    L = _lib.load()
    ctx = C.c_void_p()
    assert L.lb_create(C.byref(ctx), 0, _lib.LB_F64, 8) == 0
    try:
        f64p = C.POINTER(C.c_double)
        for k, n in enumerate((200, 900)):
            inst = I.build_instance(*I.chung_lu(n, 2.1, 6.0, 3 + k))
            rp = np.ascontiguousarray(inst.rowptr, np.int64)
            col = np.ascontiguousarray(inst.col, np.int32)
            d = np.ascontiguousarray(inst.d, np.float64)
            rc = L.lb_set_graph(ctx, inst.n, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                col.ctypes.data_as(C.POINTER(C.c_int32)), None, d.ctypes.data_as(f64p), inst.L, 0.0)
            assert rc == 0, L.lb_last_error(ctx)
            if k == 0:
                assert L.lb_blk_alloc(ctx, 3, 1) == 0
                Z0 = np.zeros((inst.n, 8))
                assert L.lb_blk_set(ctx, 0, Z0.ctypes.data_as(f64p)) == 0
                W = np.random.default_rng(0).standard_normal((inst.n, 8))
                W /= np.linalg.norm(W, axis=1, keepdims=True)
                out = np.zeros(2)
                assert L.lb_polish_begin(ctx, W.ctypes.data_as(f64p), out.ctypes.data_as(f64p)) == 0
                assert L.lb_blk_op(ctx, 1, 0, -1) == 0
        assert L.lb_blk_op(ctx, 1, 0, -1) == -1
        assert b"lb_blk_alloc" in L.lb_last_error(ctx)
        g = np.zeros(1)
        assert L.lb_polish_trial(ctx, 0.5, g.ctypes.data_as(f64p)) == -1
        Z1 = np.zeros((inst.n, 8))
        assert L.lb_blk_alloc(ctx, 3, 1) == 0 and L.lb_blk_set(ctx, 0, Z1.ctypes.data_as(f64p)) == 0
        assert L.lb_blk_op(ctx, 1, 0, -1) == 0
        assert L.lb_sync(ctx) == 0
    finally:
        L.lb_destroy(ctx)


@pytest.mark.parametrize("r,extra", [(8, 16), (16, 16)])
def test_device_lobpcg_wide_block_matches_dense(r, extra):
    """A pencil block wider than the context (b = r + 16: 3 or 2 device blocks per LOBPCG block, so the Gram matrices
    and combinations over [X, W, P] span more than the library's 6 blocks per call) reaches the smallest eigenvalue
    of the dense pencil (Diag(g) - C, Diag(s)) from scipy.linalg.eigh at n ~ 2000."""
    import scipy.linalg as sl
    n_, keys, lab = I.sbm(2000, 8, 0.02, 0.002, 11)
    inst = _with_multipliers(I.build_instance(n_, keys), 12, scale=0.01, ntri=800)
    rng = np.random.default_rng(6)
    W = rng.standard_normal((n_, r))
    W /= np.linalg.norm(W, axis=1, keepdims=True)
    lag = B.DeviceLagrangian(inst, r=r, dtype="f64")
    try:
        dv, g, _ = B.value(lag, W)
        s = np.maximum(inst.d, 1.0)
        X0 = np.linalg.svd(W, full_matrices=False)[0]
        th, rn, X, tr = B.pencil_min(lag, g, s, r + extra, X0=X0, rng=np.random.default_rng(0), maxiter=400,
                                     atol=1e-8, log=lambda *a: None)
    finally:
        lag.close()
    Cm, _ = B.dense_C(inst)
    mu = float(sl.eigh(np.diag(g) - Cm, np.diag(s), eigvals_only=True, subset_by_index=[0, 0])[0])
    print(f"[bound] wide block r={r} b={X.shape[1]}: theta_1 {th[0]:.12e} dense {mu:.12e} its {tr[-1][0]}", flush=True)
    assert X.shape == (n_, r + extra)
    assert abs(th[0] - mu) * float(s.sum()) / inst.L <= 1e-7
    assert th[0] - rn[0] <= mu + 1e-12 * abs(mu)
