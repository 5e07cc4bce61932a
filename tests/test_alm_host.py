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
