"""The solver's host logic end to end on the CPU: the device engine is replaced by the FP64 oracle (OracleEngine),
so working-set merges, triangle purge/dedupe, checkpoints, resume and the phase switch run exactly as on the GPU.
Same case and checks as the device end-to-end test, at a smaller size."""
import json
import os

import numpy as np
import pytest

from lagbound_amd import alm, pipeline
from alm_case import _write, _check_ckpt
from oracle_engine import OracleEngine


def _run(monkeypatch, ws, name, pairs, vc, outer, **kw):
    extra = [str(x) for x in kw.pop("extra_args", ())]
    args = pipeline.solver_args(name, pairs, vc, 0.5, 120.0, outer, 4, **kw)
    args[args.index("--k") + 1] = "8"
    monkeypatch.setenv("DATA_DIR", ws.data + os.sep)
    monkeypatch.chdir(ws.root)
    assert alm.main(args + extra) == 0
    return os.path.join(ws.runs, name + "_last.npz")


def test_main_then_proximal_host(tmp_path, monkeypatch, capsys):
    monkeypatch.setattr(alm, "Engine", OracleEngine)
    ws = pipeline.Workspace(str(tmp_path))
    D, n, keys = _write(str(tmp_path), n=240, blocks=4)
    pairs = [os.path.join(D, "planted.npy"), os.path.join(D, "coarse.npy")]
    vc = os.path.join(D, "planted.npy")
    ck = _run(monkeypatch, ws, "main", pairs, vc, 8, init=os.path.join(D, "planted.npy"),
              extra_args=["--tri_every", "3", "--vcs_every", "2", "--tri_add", "400", "--tri1_add", "200"])
    out = capsys.readouterr().out
    w = pipeline.parse_weight(out)
    assert w is not None and w > 0
    z = _check_ckpt(ck, n, 8)
    assert len(z["TI"]) > 0
    assert int(z["nvcs"]) == 1 and len(z["vcscol0"]) > 0
    hist = json.load(open(os.path.join(ws.runs, "main.history.json")))
    assert len(hist) == 8 and all(np.isfinite(h["Laug"]) for h in hist)
    ck2 = _run(monkeypatch, ws, "proximal", pairs, vc, 3, resume=ck, pair_weight=w, proximal=True)
    z2 = _check_ckpt(ck2, n, 8)
    assert len(z2["TI"]) <= len(z["TI"])


def test_solve_estimate_certify_host(tmp_path, monkeypatch, capsys):
    """Solver -> estimate stage (reference arguments, result JSON) -> dense certificate, on the CPU operators. The
    certified bound must be at least the planted partition's modularity and within the rounding price plus the
    margin of the estimate."""
    from lagbound_amd import bound as B
    from lagbound_amd import estimate as E
    from oracle_engine import OracleLagrangian
    from test_bound import _modularity
    monkeypatch.setattr(alm, "Engine", OracleEngine)
    monkeypatch.setattr(B, "DeviceLagrangian", lambda inst, r, dtype="f64", device=0: OracleLagrangian(inst, r))
    ws = pipeline.Workspace(str(tmp_path))
    D, n, keys = _write(str(tmp_path), n=240, blocks=4)
    pairs = [os.path.join(D, "planted.npy"), os.path.join(D, "coarse.npy")]
    vc = os.path.join(D, "planted.npy")
    ck = _run(monkeypatch, ws, "main", pairs, vc, 10, init=os.path.join(D, "planted.npy"),
              extra_args=["--tri_every", "3", "--vcs_every", "2", "--tri_add", "400", "--tri1_add", "200"])
    out = os.path.join(ws.runs, "est.json")
    assert E.main(["--tag", "g__rx", "--ckpt", ck, "--lb", "0.3", "--polish", "rgd", "--polish_time", "60",
                   "--scale", "vol", "--eig_time", "60", "--out", out, "--rgd_iters", "40", "--certify", "1",
                   "--saveW", os.path.join(ws.runs, "W.npz")]) == 0
    res = json.load(open(out))
    for key in ("dv0", "dv", "ub", "s_vol", "escape_rounds", "gap", "certificate"):
        assert key in res, key
    assert res["ub"] >= res["dv"] and res["dv"] >= res["dv0"] - 1e-12
    cert = res["certificate"]
    assert cert["passed"]
    lab = np.load(os.path.join(D, "planted.npy"))
    inst = alm.load_instance(D)
    assert cert["UB"] >= _modularity(inst, lab)
    sums = res["s_vol"]["sum_s"]
    assert abs(cert["UB"] - res["ub"]) <= 1e-7 * sums / inst.L + cert["price"] + 1e-12
    W = np.load(os.path.join(ws.runs, "W.npz"))["V"]
    assert W.shape[0] == n and W.dtype == np.float32


def test_pair_weight_rule_closed_form():
    """pair_weight_rule on the 6-cycle (d_i = 2, L = 12) with V alternating +e1 / -e1 and one level of two clusters
    (even / odd vertices): (B V)_i = -2 v_i, so |(B V)_i| = 2; Y_Q = +-6 e1, rq_Q = 3 / 36 = 1/12,
    K_Q = 2 rho w (2 * 36 + 36) / 144 = 1.5 rho w, kappa_i = d_i^2 K = 6 rho w, hence R(w) = kappa / 2 = 3 rho w and
    R(1) = 3 rho; w_mob = c1 / (3 rho), w_step = c2 / (rho n^2 / L) = c2 / (3 rho). With the defaults (0.1, 0.05) the
    step limit binds; with c1 = 0.01 the mobility does. R(1) does not depend on the weight the rule is measured at."""
    from lagbound_amd.alm import pair_weight_rule
    from lagbound_amd import instances as I
    n = 6
    keys = np.array(sorted(min(i, (i + 1) % n) * n + max(i, (i + 1) % n) for i in range(n)), np.int64)
    inst = I.build_instance(n, keys)
    lab = np.arange(n) % 2
    inst.levels = [I.pair_level(lab, inst.d)]
    V = np.zeros((n, 4))
    V[:, 0] = np.where(lab == 0, 1.0, -1.0)
    rho = 0.1
    w, info = pair_weight_rule(inst, V, rho, 1.0)
    assert info["R1"] == pytest.approx(3 * rho, rel=1e-12)
    assert info["w_mob"] == pytest.approx(0.1 / (3 * rho), rel=1e-12)
    assert info["w_step"] == pytest.approx(0.05 / (3 * rho), rel=1e-12)
    assert w == pytest.approx(0.05 / (3 * rho), rel=1e-12) and info["binding"] == "step"
    w2, info2 = pair_weight_rule(inst, V, rho, 7.5, "0.01,")
    assert info2["R1"] == pytest.approx(3 * rho, rel=1e-12) and info2["c2"] == 0.05
    assert w2 == pytest.approx(0.01 / (3 * rho), rel=1e-12) and info2["binding"] == "mobility"
