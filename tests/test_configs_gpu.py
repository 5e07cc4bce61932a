"""Parity at the sizes of the workloads the search is specified on (the batch and serving workloads have their own
test files, test_batch_gpu.py and test_serve_gpu.py):

- SBM n = 2,000, 10 blocks, p_in 0.05, p_out 0.002, seed 0: the whole FP64 chain (device reduction bit-exact
  against the oracle, r = 8 solve with a 20,000-triangle pool, bit-exact separation on the solved V, candidate
  bound, dense FP64 certificate >= the planted partition's modularity);
- Chung-Lu n = 300k, gamma 2.1, mean degree 13, 1M wedge triangles, pair + VCS levels over 64 LPA clusters: FP32 and
  FP64 evaluation against the oracle, bit-exact separation of both families at the 1M budget; |U_fp32 - U_fp64| of
  the converged candidate bound at a solved search's multipliers, and of two whole searches (FP32 and FP64);
- Chung-Lu n = 2.4M, gamma 2.2, mean degree 50 (~120M CSR entries, int64 row pointers), 8M triangles, 64 LPA
  clusters: BF16-stored V evaluation against the oracle (FP32 tolerances on the dequantised V) and bit-exact
  separation of both families with binding caps.

V is clustered (a random unit direction per LPA cluster plus noise, unit rows) so that the separations have
millions of violated candidates and their histogram caps bind.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from test_eval_gpu import _check, _engine

pytestmark = pytest.mark.gpu


def _clustered_V(lab, r, seed, noise=0.6):
    rng = np.random.default_rng(seed)
    P = int(lab.max()) + 1
    E = rng.standard_normal((P, r))
    E /= np.linalg.norm(E, axis=1, keepdims=True)
    V = E[lab] + noise * rng.standard_normal((len(lab), r)) / np.sqrt(r)
    return V / np.linalg.norm(V, axis=1, keepdims=True)


def _same_selection(got, want, label):
    TI, val, vmin = got
    TIo, valo, vmino = want
    print(f"[config] {label}: {len(TIo)} kept, vmin {vmino:.7f}", flush=True)
    assert vmin == vmino, (label, vmin, vmino)
    assert TI.shape == TIo.shape, (label, TI.shape, TIo.shape)
    assert np.array_equal(TI, TIo), label
    assert np.array_equal(val, valo), label


def _separations(eng, inst, budget, label):
    """Both families at the solver's caps for a pool budget (alm_vcs.py:1747: floor(1.3 budget) + 10; the one-edge
    family at half the budget), device against the oracle on the stored V."""
    Vs = eng.get_V()
    for fam, add in ((0, budget), (1, budget // 2)):
        cap = int(1.3 * add) + 10
        got = eng.separate_triangles(fam, thresh=0.3, maxh=100, vtol=1e-3, cap=cap)
        want = O.separate_c(inst, Vs, fam, thresh=0.3, maxh=100, vtol=1e-3, cap=cap)
        _same_selection(got, want, f"{label} fam {fam} cap {cap}")


# ------------------------------------------------------------------------------------------ SBM n = 2,000, FP64
def _modularity_orig(n, keys, lab):
    u, v = keys // n, keys % n
    m = len(keys)
    d = np.bincount(np.r_[u, v], minlength=n).astype(np.float64)
    inside = float(np.sum(lab[u] == lab[v]))
    vol = np.bincount(lab, weights=d)
    return inside / m - float(np.sum(vol ** 2)) / (2.0 * m) ** 2


@pytest.mark.timeout(900)
def test_sbm_2000_fp64_chain(tmp_path):
    from lagbound_amd import instances as I
    from lagbound_amd import pipeline
    from lagbound_amd.alm import load_instance
    ws = pipeline.Workspace(str(tmp_path))
    n, keys, lab = I.sbm(2000, 10, 0.05, 0.002, 0)
    D = os.path.join(ws.data, pipeline.TAG)
    os.makedirs(D)
    json.dump({"n": int(n), "m": int(len(keys))}, open(os.path.join(D, "graph.json"), "w"))
    np.save(os.path.join(D, "keys.npy"), keys)
    planted = os.path.join(D, "planted.npy")
    np.save(planted, lab.astype(np.int64))
    q = _modularity_orig(n, keys, lab)
    print(f"[config] sbm n={n} m={len(keys)} Q(planted)={q:.6f}: reduce", flush=True)
    tag = ws.reduce([planted])
    Dz = os.path.join(ws.data, tag)
    ref = O.reduce_c(n, keys)
    assert np.array_equal(np.load(os.path.join(Dz, "keys.npy")), ref["keys"])
    assert np.array_equal(np.load(os.path.join(Dz, "weights.npy")), ref["weights"])
    assert np.array_equal(np.load(os.path.join(Dz, "group_of.npy")), ref["group_of"])
    lr = ws.reduced_labels(planted)
    print(f"[config] sbm reduced n={ref['N']}: solve fp64 r=8, 20000-triangle pool", flush=True)
    ck, _ = ws.solve("c1", [lr], lr, q, 300.0, 40, 8, init=lr, k=8,
                     extra_args=["--dtype", "f64", "--tri_add", "20000", "--tri1_add", "0", "--tri_max", "20000",
                                 "--tri_every", "5", "--vcs_every", "4"])
    z = np.load(ck)
    assert 0 < len(z["TI"]) <= 20000
    inst = load_instance(Dz)
    eng = _engine(inst, z["V"].astype(np.float64), "f64", 8)
    Vs = eng.get_V()
    cap = int(1.3 * 20000) + 10
    got = eng.separate_triangles(0, thresh=0.3, maxh=100, vtol=1e-3, cap=cap)
    _same_selection(got, O.separate_c(inst, Vs, 0, thresh=0.3, maxh=100, vtol=1e-3, cap=cap), "sbm solved V")
    eng.close()
    print("[config] sbm estimate + lifted sparse certificate", flush=True)
    dump = os.path.join(ws.runs, "c1cert.npz")
    res = ws.estimate("c1est", ck, q, polish_time=120, eig_time=120, extra_args=["--certify", "2", "--dump_cert", dump])
    cert = res["certificate"]
    print(f"[config] sbm ub {res['ub']:.6f} certified {cert.get('UB')} passed {cert['passed']}", flush=True)
    assert cert["passed"], cert
    assert cert["UB"] >= q
    assert cert["UB"] <= 1.0
    # the independent verifier: the record against the reduced instance files, and the reduction from the original
    from lagbound_amd import verify as Vf
    ver = Vf.verify(dump, Dz)
    red = Vf.check_reduction_dirs(D, Dz, np.load(planted))
    print(f"[config] sbm verifier: {ver.get('UB')} passed {ver['passed']} ({ver.get('fail', '')}); reduction {red}",
          flush=True)
    assert ver["passed"], ver
    assert abs(ver["UB"] - cert["UB"]) <= 1e-6 and ver["UB"] >= q
    assert red["bookkeeping"] == "ok" and red["identity"] == "exact"


# --------------------------------------------------------------------------- Chung-Lu n = 300k, FP32 and FP64
@pytest.fixture(scope="module")
def chung_lu_300k():
    from lagbound_amd import instances as I
    print("[config] building Chung-Lu n=300k gamma=2.1 kbar=13 T=1M", flush=True)
    inst, _ = I.metric_instance(300_000, 2.1, 13.0, T=1_000_000)
    V = _clustered_V(inst.levels[0].lab, 16, 5)
    return inst, V


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_chung_lu_300k_eval(chung_lu_300k, dtype):
    inst, V = chung_lu_300k
    eng = _engine(inst, V, dtype, 16)
    for rho in (0.0, 0.02):
        _check(eng, inst, rho, dtype, "chung-lu 300k")
    eng.close()


def test_chung_lu_300k_separation(chung_lu_300k):
    inst, V = chung_lu_300k
    eng = _engine(inst, V, "f32", 16)
    _separations(eng, inst, 1_000_000, "chung-lu 300k")
    eng.close()


@pytest.mark.timeout(900)
def test_chung_lu_300k_bound_fp32_vs_fp64():
    """The candidate bound at the multipliers of a solved search (the 300k-vertex Chung-Lu request, FP64 search,
    8 outer iterations): the estimate from the solved V with an FP32 ascent and with an FP64 ascent (both with the
    FP64 assembly and device pencil). Both pencils must have converged (first residual times sum(s) / L below the
    1e-5 tolerance), and abs(U_fp32 - U_fp64) is bounded by 1e-4 of U_fp64; the search context's own FP32 / FP64
    evaluation of the same bound is printed beside it."""
    from lagbound_amd import bound as B
    from lagbound_amd import instances as I
    from lagbound_amd import serve
    n, keys = I.chung_lu_target(300_000, 2.1, 13.0, 0)
    p = serve.Params(dtype="f64", outer=8, cert_max_n=0)
    print("[config] chung-lu 300k: FP64 search for the multipliers", flush=True)
    inst, lab, s = serve.search_request(serve.Request(rid=0, n=n, keys=keys, seed=3), p)
    V = s.V.copy()
    s.eng.close()
    kw = dict(rgd_iters=60, eig_maxiter=400, escape=0, log=lambda *a: None)
    out = {}
    for dtype in ("f64", "f32"):
        print(f"[config] chung-lu 300k estimate, {dtype} ascent", flush=True)
        e = B.estimate(inst, V, dtype=dtype, **kw)
        rec = e["s_vol"]
        conv = rec["resid"][0] * rec["sum_s"] / inst.L
        print(f"[config] {dtype}: dv {e['dv']:.9f} ub {e['ub']:.9f} pencil residual (bound units) {conv:.2e} "
              f"after {rec['eig_trace'][-1][0] if rec['eig_trace'] else 0} iterations", flush=True)
        assert conv <= 1e-5, (dtype, conv)
        out[dtype] = e
    du = abs(out["f32"]["ub"] - out["f64"]["ub"])
    print(f"[config] U_f64 {out['f64']['ub']:.9f} U_f32 {out['f32']['ub']:.9f} |dU| {du:.3e}", flush=True)
    assert np.isfinite(out["f64"]["ub"]) and du <= 1e-4 * abs(out["f64"]["ub"])
    assert out["f64"]["ub"] >= serve.modularity(inst, lab.astype(np.int64))


# ------------------------------------------------------------------------ Chung-Lu n = 2.4M, BF16-stored V
@pytest.fixture(scope="module")
def chung_lu_2400k():
    from lagbound_amd import instances as I
    print("[config] building Chung-Lu n=2.4M gamma=2.2 kbar=50 T=8M (about a minute)", flush=True)
    inst, _ = I.metric_instance(2_400_000, 2.2, 50.0, T=8_000_000)
    assert inst.rowptr.dtype == np.int64 and inst.nnz > 100_000_000
    V = _clustered_V(inst.levels[0].lab, 16, 6)
    print(f"[config] n={inst.n} nnz={inst.nnz} T={inst.T} P={inst.levels[0].P}", flush=True)
    return inst, V


@pytest.mark.timeout(900)
def test_chung_lu_2400k_bf16_eval(chung_lu_2400k):
    inst, V = chung_lu_2400k
    eng = _engine(inst, V, "bf16", 16)
    _check(eng, inst, 0.02, "bf16", "chung-lu 2.4M")
    eng.close()


@pytest.mark.timeout(900)
def test_chung_lu_2400k_separation(chung_lu_2400k):
    inst, V = chung_lu_2400k
    eng = _engine(inst, V, "bf16", 16)
    _separations(eng, inst, 8_000_000, "chung-lu 2.4M")
    eng.close()


@pytest.mark.timeout(900)
def test_chung_lu_300k_fp32_vs_fp64_search():
    """The same bound request on the 300k-vertex Chung-Lu graph searched in FP32 and in FP64 (label-propagation
    levels, triangle and VCS separation, then the FP64-assembled candidate bound): abs(U_fp32 - U_fp64) is printed;
    the searches may take different paths, so the bar is loose (1e-2 of U_fp64)."""
    from lagbound_amd import instances as I
    from lagbound_amd import serve
    n, keys = I.chung_lu_target(300_000, 2.1, 13.0, 0)
    out = {}
    for dtype in ("f32", "f64"):
        print(f"[config] chung-lu 300k solve + estimate, {dtype} search", flush=True)
        p = serve.Params(dtype=dtype, outer=8, cert_max_n=0)
        res, _ = serve.solve_request(serve.Request(rid=0, n=n, keys=keys, seed=3), p)
        out[dtype] = res
        print(f"[config] {dtype}: dv {res['dv']:.9f} ub {res['ub']:.9f} lb {res['lb']:.6f}", flush=True)
    du = abs(out["f32"]["ub"] - out["f64"]["ub"])
    print(f"[config] abs(U_fp32 - U_fp64) = {du:.3e} (relative {du / abs(out['f64']['ub']):.3e})", flush=True)
    assert np.isfinite(out["f64"]["ub"]) and out["f64"]["ub"] >= out["f64"]["lb"]
    assert du <= 1e-2 * abs(out["f64"]["ub"])
