"""End-to-end device solver runs through the drop-in Workspace.solve (main phase with triangle and VCS separation,
then the proximal phase resumed from its checkpoint). Checks the reference's output contract: the checkpoint keys,
dtypes and invariants (alm_vcs.py:1576-1589), the history file, the parsed pair weight, and that the multipliers of
the checkpoint evaluate (FP64 oracle, rho = 0) to a finite Lagrangian value."""
import json
import os

import numpy as np
import pytest

import oracle as O
from lagbound_amd import instances as I
from lagbound_amd import pipeline
from alm_case import _write, _check_ckpt

pytestmark = pytest.mark.gpu


def test_main_then_proximal(tmp_path):
    ws = pipeline.Workspace(str(tmp_path))
    D, n, keys = _write(str(tmp_path))
    pairs = [os.path.join(D, "planted.npy"), os.path.join(D, "coarse.npy")]
    vc = os.path.join(D, "planted.npy")
    print("[alm] main phase", flush=True)
    ck, w = ws.solve("main", pairs, vc, 0.5, 120.0, 20, 8, init=os.path.join(D, "planted.npy"), k=8,
                     extra_args=["--tri_every", "3", "--vcs_every", "4", "--tri_add", "2000", "--tri1_add", "1000"])
    assert os.path.exists(ck) and w is not None and w > 0
    log = open(os.path.join(ws.runs, "main.log")).read()
    assert "-> pairlvl_weight" in log and "+" in log and "tri" in log
    z = _check_ckpt(ck, n, 8)
    assert len(z["TI"]) > 0
    assert int(z["nvcs"]) == 1 and len(z["vcscol0"]) > 0
    hist = json.load(open(os.path.join(ws.runs, "main.history.json")))
    assert len(hist) == 20 and all(np.isfinite(h["Laug"]) for h in hist)
    print("[alm] proximal phase", flush=True)
    ck2, w2 = ws.solve("proximal", pairs, vc, 0.5, 120.0, 6, 8, resume=ck, pair_weight=w, proximal=True, k=8)
    z2 = _check_ckpt(ck2, n, 8)
    assert len(z2["TI"]) <= len(z["TI"])
    # the checkpoint's multipliers define a valid Lagrangian: evaluate it with the FP64 oracle
    inst = I.build_instance(n, keys)
    for li in range(int(z2["nlev"])):
        inst.levels.append(I.pair_level(z2[f"lab{li}"], inst.d, z2[f"sig{li}"]))
    inst.vcs.append(I.vcs_level(z2["vcslab0"], inst.d, z2["vcsptr0"], z2["vcscol0"], z2["vcssig0"].astype(np.float64)))
    inst.TI, inst.lam = z2["TI"].astype(np.int32), z2["lam"]
    parts = O.eval_c(inst, z2["V"].astype(np.float64), 0.0)[0]
    assert np.isfinite(parts[0])
