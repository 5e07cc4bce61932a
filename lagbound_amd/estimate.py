"""Drop-in for the reference's estimate stage, `python _core/est_dual_vcs.py ...` (launched by Workspace.estimate,
pipeline.py:156-170): the candidate bound at a solver checkpoint's multipliers, with every C(y) X product on the
device (lagbound_amd.bound).

It reads the instance directory DATA_DIR/<tag>/ and the checkpoint (alm_vcs.py:1576-1589 keys), runs polish (rgd),
the pencil eigenvalue per scale and the rank-escape rounds, and writes the reference's result JSON (--out) and
W (--saveW). `--certify 1` adds the dense FP64 certificate (small instances only: it factors an n x n matrix);
`--certify 2` the lifted sparse FP64 certificate (lagbound_amd.sparse_cert, its dense core on the device), and
`--dump_cert FILE.npz` writes its record for the independent verifier (python -m lagbound_amd.verify), like the
reference's certifier option of the same name.
Checkpoints with terms the device evaluation does not carry (connectivity multipliers mu / om, VC levels, nested
levels) are refused.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import bound as B
from .alm import load_instance


def parse(argv=None):
    p = argparse.ArgumentParser(prog="lagbound_amd.estimate", description=__doc__.split("\n")[0])
    p.add_argument("--tag", required=True)
    p.add_argument("--ckpt", required=True)
    p.add_argument("--polish", default="rgd", help="rgd | none")
    p.add_argument("--polish_time", type=float, default=600)
    p.add_argument("--ptol", type=float, default=2e-5)
    p.add_argument("--rgd_iters", type=int, default=400)
    p.add_argument("--scale", default="vol", help="comma list of pencil weights s: vol | ones | sqrtvol")
    p.add_argument("--eig_extra", type=int, default=16)
    p.add_argument("--eig_maxiter", type=int, default=400)
    p.add_argument("--eig_tol", type=float, default=1e-9)
    p.add_argument("--eig_chunk", type=int, default=20)
    p.add_argument("--eig_atol", type=float, default=1e-5)
    p.add_argument("--eig_time", type=float, default=1200)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--out", default="")
    p.add_argument("--saveW", default="")
    p.add_argument("--initV", default="")
    p.add_argument("--no_w_init", type=int, default=0)
    p.add_argument("--lb", type=float, default=float("nan"))
    p.add_argument("--escape", type=int, default=2)
    p.add_argument("--escape_min", type=float, default=0.002)
    p.add_argument("--escape_p", type=int, default=4)
    p.add_argument("--escape_ratio", type=float, default=0.3)
    p.add_argument("--device", type=int, default=0)
    p.add_argument("--dtype", default="f64", choices=["f64", "f32", "bf16"])
    p.add_argument("--rank", type=int, default=16, choices=[8, 16, 32], help="device block width of C(y) X")
    p.add_argument("--certify", type=int, default=0,
                   help="1: dense FP64 certificate at the best round; 2: lifted sparse FP64 certificate")
    p.add_argument("--dump_cert", default="", help="with --certify 2: write the certificate record (.npz)")
    p.add_argument("--certify_margin", type=float, default=1e-7)
    return p.parse_args(argv)


def load(tag, ckpt):
    D = os.environ.get("DATA_DIR", "data/")
    inst = load_instance(os.path.join(D, tag))
    z = np.load(ckpt)
    if "mu" in z.files and np.any(z["mu"]):
        raise SystemExit("checkpoint has connectivity multipliers (mu); not supported by the device evaluation")
    if "om" in z.files and len(z["om"]) and np.any(z["om"] > 0):
        raise SystemExit("checkpoint has connectivity cuts (om); not supported by the device evaluation")
    if "nvc" in z.files and int(z["nvc"]) > 0:
        raise SystemExit("checkpoint has VC levels; not supported by the device evaluation")
    if any(f.startswith("nest") for f in z.files):
        raise SystemExit("checkpoint has nested levels; not supported by the device evaluation")
    B.instance_from_checkpoint(inst, z)
    V = np.asarray(z["V"], np.float64)
    return inst, V


def main(argv=None):
    a = parse(argv)
    t0 = time.time()

    def log(*s):
        print(f"[{time.time() - t0:8.1f}s]", *s, flush=True)

    inst, V = load(a.tag, a.ckpt)
    if a.initV:
        V = np.asarray(np.load(a.initV)["V"], np.float64)
    log(f"tag {a.tag} ckpt {a.ckpt} n {inst.n} m {inst.nnz // 2} L {inst.L} k {V.shape[1]} levels "
        f"{[lv.P for lv in inst.levels]} vcs levels {[(v.P, v.nnz) for v in inst.vcs]} ntri {inst.T} "
        f"C0/L {inst.C0 / inst.L:.6f}")
    lag = B.DeviceLagrangian(inst, r=a.rank, dtype=a.dtype, device=a.device)
    try:
        res = B.estimate(inst, V, lag=lag, polish=a.polish, rgd_iters=a.rgd_iters, polish_time=a.polish_time,
                         ptol=a.ptol, scale=a.scale, eig_extra=a.eig_extra, eig_maxiter=a.eig_maxiter,
                         eig_tol=a.eig_tol, eig_chunk=a.eig_chunk, eig_atol=a.eig_atol, eig_time=a.eig_time,
                         escape=a.escape, escape_min=a.escape_min, escape_p=a.escape_p,
                         escape_ratio=a.escape_ratio, seed=a.seed, no_w_init=bool(a.no_w_init), log=log)
    finally:
        lag.close()
    W, g, s = res.pop("W"), res.pop("g"), res.pop("s")
    res.update(tag=a.tag, ckpt=a.ckpt, lb=a.lb, dtype=a.dtype)
    if np.isfinite(a.lb):
        res["gap"] = (res["ub"] - a.lb) / a.lb
    if a.certify == 2:
        from . import sparse_cert as SC
        est = dict(res, g=g, s=s)
        cert = SC.certify_estimate_sparse(inst, est, margin=a.certify_margin, device=a.device, keep=True, log=log)
        if cert["passed"] and a.dump_cert:
            SC.write_certificate(a.dump_cert, inst, cert)
        cert.pop("proof", None)
        res["certificate"] = cert
    elif a.certify:
        cert = B.certify_dense(inst, g, (max(0.0, -res["lmin"]) + a.certify_margin) * s)
        res["certificate"] = cert
        log(f"CERTIFICATE {'passed' if cert['passed'] else 'FAILED'}" +
            (f" UB {cert['UB']:.9f} (rounding price {cert['price']:.3e})" if cert["passed"] else f": {cert['reason']}"))
    log(f"RESULT dv {res['dv']:.6f}  UB {res['ub']:.6f}" +
        (f"  gap {res['gap'] * 100:.2f}%" if np.isfinite(a.lb) else "") + f"  total {time.time() - t0:.0f}s")
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(res, fh, indent=1, default=float)
    if a.saveW:
        np.savez(a.saveW, V=W.astype(np.float32), g=g)
    return 0


if __name__ == "__main__":
    sys.exit(main())
