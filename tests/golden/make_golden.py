"""Writes tests/golden/eval_small.npz: a small SBM instance (one pair level, one VCS level, triangles), a unit-row V
and the dense brute-force evaluation (oracle.eval_dense) at rho = 0 and rho = 0.3. Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from conftest import small_instance  # noqa: E402


def main():
    inst, V = small_instance(n=24, seed=11, levels=1, ntri=16)
    lv, vl = inst.levels[0], inst.vcs[0]
    out = dict(n=inst.n, rowptr=inst.rowptr, col=inst.col, d=inst.d, L=inst.L, C0=inst.C0, lab=lv.lab, P=lv.P,
               sig=lv.sig, rq=lv.rq, dsq=lv.dsq, vlab=vl.lab, vP=vl.P, vrq=vl.rq, vptr=vl.ptr, vcol=vl.col,
               vsig=vl.sig, TI=inst.TI, lam=inst.lam, V=V, rho=0.3)
    for tag, rho in (("r0", 0.0), ("r1", 0.3)):
        f, G, pp, pt, pv = O.eval_dense(inst, V, rho)
        out[f"f_{tag}"] = f
        out[f"G_{tag}"] = G
        out[f"pt_{tag}"] = pt
        out[f"pv_{tag}"] = np.concatenate(pv)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eval_small.npz")
    np.savez(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
