"""Upper bound of the maximum modularity of a small graph, end to end in one process, in the stage order of the
reference's `_sdp_upper_bound` (src/lagbound/api.py:195-298): exact reduction -> aggregate levels -> main solver
phase -> proximal phase resumed from the main checkpoint -> candidate bound -> lifted sparse FP64 certificate ->
independent verification (the reference's `verify=True`, api.py:280-298).

Differences from the reference's run, all on the input side of the search:
- levels come from seeded greedy local moving at the resolutions `levels` on the reduced graph (the reference runs
  Leiden at those resolutions on the original graph and restricts the labels; igraph/leidenalg are not part of
  this build);
- the search runs in one process (the reference launches its stages as subprocesses), so this entry point is for
  graphs of up to a few thousand vertices; larger graphs go through lagbound_amd.pipeline.
"""
from __future__ import annotations

import os
import tempfile
from fractions import Fraction

import numpy as np

from . import alm
from . import bound as B
from . import instances as I
from . import reduce as RD
from . import sparse_cert as SC
from . import verify as VF


def _flags(tag, n, k, nbatch, outer, seed, extra):
    return ["--tag", tag, "--k", str(k), "--nbatch", str(nbatch), "--inner", "1", "--est_every", "0", "--mom", "0",
            "--shift_p", "0", "--time_limit", "1e9", "--seed", str(seed), "--outer", str(outer),
            "--tri_max", "16000000", "--vcs_max", "150000000", "--vcs_add", "10000000"] + extra


def upper_bound_small(n, keys, k=8, outer=150, prox_outer=40, levels=(16.0, 64.0), seed=0, device=0, dtype="f32",
                      rgd_iters=400, eig_maxiter=400, margin=1e-7, log=None):
    """Returns dict(value, status, estimate, reduced_n, stages...). status: "verified" (the lifted sparse FP64
    certificate passed and the independent verifier, with the reduction check against the original graph, accepted
    it), "certified" (the certificate passed, the verifier did not), "solved_by_reduction" (no edge between groups:
    the value is the exact optimum), or "estimate"."""
    log = log or (lambda *a: None)
    keys = np.asarray(keys, np.int64)
    m = len(keys)
    L = 2 * m
    red = RD.reduce_graph(n, keys, device=device)
    N = red["N"]
    out = dict(n=n, m=m, reduced_n=N, reduced_m=len(red["keys"]), leaf_unions=red["leaf_unions"],
               twin_unions=red["twin_unions"])
    if len(red["keys"]) == 0:                      # api._solved_by_reduction (api.py:170-192): the groups are the optimum
        q = Fraction(2 * red["inside_edges"], L) - Fraction(int((red["vol"].astype(object) ** 2).sum()), L * L)
        out.update(value=B.frac_up(q), status="solved_by_reduction", estimate=float(q))
        return out
    inst = I.build_instance(N, red["keys"], weights=red["weights"], vol=red["vol"].astype(np.float64), L=L,
                            C0=2 * red["inside_edges"])
    # levels: greedy local moving at the resolutions of the reference's Leiden levels (pairs at `levels`, node-cluster
    # rows at the last one), the starting partition at resolution 1
    lm = lambda g, s_: I.local_moving(inst.rowptr, inst.col, inst.w, inst.d, inst.L, g, s_)
    labs = [lm(float(g), seed + 7 + i) for i, g in enumerate(levels)]
    init = lm(1.0, seed + 11)
    nbatch = 512 if N >= 512 * 64 else min(N, max(8, N // 64))
    with tempfile.TemporaryDirectory() as tmp:
        main_save = os.path.join(tmp, "main.npz")
        a = alm.parse(_flags("small", N, k, nbatch, outer, seed,
                             ["--tri_add", "600000", "--tri1_add", "300000", "--tri_every", "9", "--pw_rule", "default",
                              "--save", main_save, "--dtype", dtype, "--device", str(device)]))
        s = alm.Solver(a, inst=inst, labels=labs, vcs_labels=[labs[-1]], init_labels=init, log=log, signals=False)
        s.run()
        w = s.W_AUTO if s.W_AUTO is not None else a.pairlvl_weight
        ck = main_save.replace(".npz", "_last.npz")
        inst2 = I.build_instance(N, red["keys"], weights=red["weights"], vol=red["vol"].astype(np.float64), L=L,
                                 C0=2 * red["inside_edges"])
        a2 = alm.parse(_flags("small", N, k, nbatch, prox_outer, seed,
                              ["--resume", ck, "--tri_add", "0", "--tri1_add", "0", "--tri_every", "1000000",
                               "--pairlvl_weight", str(w), "--phase2_at_it", "1", "--phase2_w", str(w),
                               "--phase2_inner", "4", "--phase2_tri_every", "1000000", "--phase2_est_every", "0",
                               "--dtype", dtype, "--device", str(device)]))
        s2 = alm.Solver(a2, inst=inst2, labels=labs, vcs_labels=[labs[-1]], log=log, signals=False)
        s2.run()
    est = B.estimate(inst2, s2.V, rgd_iters=rgd_iters, eig_maxiter=eig_maxiter, device=device, log=log)
    cert = SC.certify_estimate_sparse(inst2, est, margin=margin, device=device, keep=True)
    out.update(estimate=est["ub"], dv=est["dv"], pair_weight=w, ntri=int(inst2.T))
    if cert["passed"]:
        ver = VF.verify(SC.certificate_record(inst2, cert), orig=(n, keys, red["group_of"], None))
        out.update(value=cert["UB"], status="verified" if ver["passed"] else "certified", price=cert["price"],
                   verified_UB=ver.get("UB"), verifier=ver.get("fail", "passed"))
    else:
        out.update(value=est["ub"], status="estimate", reason=cert.get("reason"))
    return out
