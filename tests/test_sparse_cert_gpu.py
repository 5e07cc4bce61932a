"""The lifted sparse certificate with its dense core on the device (lb_cert_core, rocSOLVER dpotrf_64):

- the device core equals the host core (same verdict, same rounding sum to 1e-9) on a core of a few thousand rows;
- on a solved 3000-vertex Chung-Lu request the sparse certificate (device core) and the dense FP64 certificate both
  pass at the same vertex shift, and their bounds differ by the auxiliary nodes' shift up to the two rounding prices;
  the independent verifier accepts the sparse certificate's record;
- the 300k-vertex Chung-Lu request (FP32 search, 8 outer iterations, FP64 estimate) is certified, with the time, the
  dense core's size and its device memory printed.
"""
import time

import numpy as np
import pytest

from lagbound_amd import bound as B
from lagbound_amd import instances as I
from lagbound_amd import serve
from lagbound_amd import sparse_cert as S
from lagbound_amd import verify as Vf

pytestmark = pytest.mark.gpu


def test_device_core_matches_host_core():
    """Lifted matrix of a 12k-vertex Chung-Lu graph with its triangle multipliers, the cluster multipliers zeroed and a
    shift of 5 w (positive definite): a dense core of ~5600 rows, factored on the host and on the device."""
    inst, _ = I.metric_instance(12000, 2.1, 13.0, T=30000, P=16)
    for lv in inst.levels:
        lv.sig[:] = 0.0
    for v in inst.vcs:
        v.sig[:] = 0.0
    g = np.zeros(inst.n)
    lf = S.Lifted(inst, g, key=S.vertex_key(inst))
    Ml = lf.assemble(5.0 * lf.w)[0]
    out = {}
    # host; device with rocSOLVER; device with the blocked kernels (vector, then matrix-core trailing updates)
    for dev, meth in ((None, 0), (0, 1), (0, 2), (0, 3)):
        ok, price, info = S.factor_price(Ml, device=dev, core_deg=256, core_method=meth)
        print(f"[cert] core on {'host' if dev is None else 'device'} (method {meth}): ok {ok} core {info['core']} "
              f"p_core {info.get('p_core')} ({info.get('core_s')} s)", flush=True)
        out[meth] = (ok, price, info)
    assert out[0][2]["core"] > 1000
    assert all(out[k][0] for k in out)
    for meth in (1, 2, 3):
        assert out[meth][2]["p_core"] == pytest.approx(out[0][2]["p_core"], rel=1e-9)
        assert out[meth][1] == pytest.approx(out[0][1], rel=1e-9)
    # and below the PSD threshold every method refuses (a shift of 0 leaves -A - T)
    bad = lf.assemble(0.0 * lf.w)[0]
    for meth in (1, 2, 3):
        assert not S.factor_price(bad, device=0, core_deg=256, core_method=meth)[0]


def _solved(n, dtype="f32", outer=8, seed=3):
    nn, keys = I.chung_lu_target(n, 2.1, 13.0, seed)
    p = serve.Params(dtype=dtype, outer=outer, cert_max_n=0)
    inst, lab, s = serve.search_request(serve.Request(rid=0, n=nn, keys=keys, seed=seed), p)
    V = s.V.copy()
    s.eng.close()
    est = B.estimate(inst, V, dtype="f64", rgd_iters=120, eig_maxiter=300, escape=0, log=lambda *a: None)
    return inst, lab, est


def test_sparse_matches_dense_on_a_solved_request():
    inst, lab, est = _solved(3000)
    sp_ = S.certify_estimate_sparse(inst, est, device=0, keep=True, log=lambda *a: print("[cert]", *a, flush=True))
    assert sp_["passed"], sp_
    ver = Vf.verify(S.certificate_record(inst, sp_), log=lambda *a: print("[verify]", *a, flush=True))
    assert ver["passed"], ver
    assert abs(ver["UB"] - sp_["UB"]) <= sp_["price"] + ver["price_asm"] + ver["price_psd"] + 1e-12
    shift = sp_["info"]["shift"]
    dn = B.certify_dense_device(inst, est["g"], shift * est["s"])
    assert dn["passed"], dn
    aux = shift * 0.01 * (sp_["info"]["N"] - inst.n) / inst.L
    print(f"[cert] n {inst.n}: sparse UB {sp_['UB']:.9f} dense UB {dn['UB']:.9f} aux shift {aux:.3e} "
          f"ub {est['ub']:.9f} core {sp_['info']['core']}", flush=True)
    assert abs(sp_["UB"] - (dn["UB"] + aux)) <= sp_["price"] + dn["price"] + 1e-12
    assert sp_["UB"] >= serve.modularity(inst, lab.astype(np.int64))


@pytest.mark.timeout(1100)
def test_chung_lu_300k_certified():
    t0 = time.time()
    inst, lab, est = _solved(300_000)
    t1 = time.time()
    res = S.certify_estimate_sparse(inst, est, device=0, log=lambda *a: print("[cert]", *a, flush=True))
    t2 = time.time()
    info = res["info"]
    m = info.get("core", 0)
    print(f"[cert] 300k: search + estimate {t1 - t0:.1f} s, certificate {t2 - t1:.1f} s; lifted N {info['N']}, "
          f"core {m} rows ({8 * m * m / 2 ** 30:.1f} GiB on the device), ub {est['ub']:.6f} "
          f"UB {res.get('UB')} passed {res['passed']}", flush=True)
    assert res["passed"], res
    assert res["UB"] >= est["ub"] - 1e-9 and res["UB"] >= serve.modularity(inst, lab.astype(np.int64))
