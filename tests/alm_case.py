"""Shared case of the solver end-to-end tests: a planted-partition graph written as a reference instance directory,
and the checks of the reference's checkpoint contract (alm_vcs.py:1576-1589)."""
import json
import os

import numpy as np

from lagbound_amd import instances as I


def _write(ws_root, n=600, blocks=6):
    n_, keys, lab = I.sbm(n, blocks, 0.12, 0.01, 5)
    D = os.path.join(ws_root, "data", "g__rx")
    os.makedirs(D, exist_ok=True)
    json.dump({"n": int(n_), "m": int(len(keys))}, open(os.path.join(D, "graph.json"), "w"))
    np.save(os.path.join(D, "keys.npy"), keys)
    np.save(os.path.join(D, "planted.npy"), lab.astype(np.int64))
    coarse = (lab // 2).astype(np.int64)
    np.save(os.path.join(D, "coarse.npy"), coarse)
    return D, n_, keys


def _check_ckpt(path, n, k):
    z = np.load(path)
    for key in ("TI", "lam", "V", "nlev", "mu", "CI", "om", "PI", "psi", "nvc", "nvcs"):
        assert key in z.files, key
    V = z["V"]
    assert V.dtype == np.float32 and V.shape == (n, k)
    assert np.abs(np.linalg.norm(V.astype(np.float64), axis=1) - 1).max() < 1e-5
    TI = z["TI"]
    lam = z["lam"]
    assert TI.shape[1] == 3 and len(TI) == len(lam) and (lam >= 0).all()
    assert (TI[:, 0] < TI[:, 2]).all()
    for li in range(int(z["nlev"])):
        s = z[f"sig{li}"]
        assert np.allclose(s, s.T) and (s >= 0).all()
    for m in range(int(z["nvcs"])):
        ptr, col, sig = z[f"vcsptr{m}"], z[f"vcscol{m}"], z[f"vcssig{m}"]
        assert ptr[-1] == len(col) == len(sig) and sig.dtype == np.float32 and (sig >= 0).all()
    return z
