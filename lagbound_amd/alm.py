"""Augmented-Lagrangian low-rank SDP search on the GPU: the drop-in for lagbound's solver stage.

`python -m lagbound_amd.alm ...` accepts the flag set that lagbound's pipeline uses for the solver
(pipeline.py:116-128: main and proximal phases), reads the same instance directory (DATA_DIR/<tag>/{graph.json,
keys.npy, weights.npy, vol.npy, meta.json}, alm_vcs.py:1103-1117) and label files, and writes the same outputs: the
NPZ checkpoint <save>_last.npz with the reference's keys (alm_vcs.py:1576-1589), <save>.history.json, and a log on
stdout including the `-> pairlvl_weight X` line that pipeline.py:131-135 parses. Every computation on V runs through
liblbhip on the device (f+grad evaluation, batched row-step sweeps, PHR updates, triangle and VCS separation); the
host keeps the instance set-up, the row orders (numpy default_rng(seed).permutation per sweep, alm_vcs.py:1387) and
the bookkeeping of the triangle pool and VCS working sets. Flags outside the production set are rejected explicitly.

Main loop (alm_vcs.py:1690-1797): sweeps -> augmented -> PHR update (sig <- pp, lam <- pt, VCS sig <- pv, ages) ->
VCS separation every vcs_every outers (and at vcs_first) -> triangle separation every tri_every outers (purge lam =
0, wedge + one-edge-arm families, dedupe against the pool, best tri_add + tri1_add by violation, wedges first) ->
history / checkpoint / stop on time limit, outer count or SIGTERM.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import resource
import signal
import sys
import time

import numpy as np

from . import _host as H
from . import instances as I
from .engine import Engine


def parse(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--tag", required=True)
    p.add_argument("--pair_labels", nargs="*", default=[])
    p.add_argument("--resume", default="")
    p.add_argument("--init_partition", default="")
    p.add_argument("--init_V", default="")
    p.add_argument("--k", type=int, default=32)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--save", default="")
    p.add_argument("--save_last", type=int, default=1)
    p.add_argument("--outer", type=int, default=10 ** 9)
    p.add_argument("--inner", type=int, default=3)
    p.add_argument("--nbatch", type=int, default=64)
    p.add_argument("--damp", type=float, default=0.0)
    p.add_argument("--time_limit", type=float, default=3600)
    p.add_argument("--rho", type=float, default=0.02)
    p.add_argument("--tri_weight", type=float, default=1.0)
    p.add_argument("--tri", type=int, default=1)
    p.add_argument("--tri_add", type=int, default=500000)
    p.add_argument("--tri_every", type=int, default=3)
    p.add_argument("--tri_max", type=int, default=5000000)
    p.add_argument("--tri_thresh", type=float, default=0.3)
    p.add_argument("--tri_maxh", type=int, default=100)
    p.add_argument("--tri1_add", type=int, default=0)
    p.add_argument("--tri1_maxk", type=int, default=60)
    p.add_argument("--tri1_hub", type=int, default=8)
    p.add_argument("--tri1_cap", type=int, default=200)
    p.add_argument("--est_every", type=int, default=100)
    p.add_argument("--lb", type=float, default=float("nan"))
    p.add_argument("--max_clusters", type=int, default=0)
    p.add_argument("--shift_p", type=int, default=8)
    p.add_argument("--mom", type=float, default=0.0)
    p.add_argument("--ckpt_every", type=int, default=0)
    p.add_argument("--pairlvl_weight", type=float, default=1.0)
    p.add_argument("--pw_rule", default="")
    p.add_argument("--phase2_at_it", type=int, default=0)
    p.add_argument("--phase2_inner", type=int, default=300)
    p.add_argument("--phase2_w", type=float, default=0.2)
    p.add_argument("--phase2_tri_every", type=int, default=1)
    p.add_argument("--phase2_est_every", type=int, default=1)
    p.add_argument("--vcs_labels", nargs="*", default=[])
    p.add_argument("--vc_weight", type=float, default=1.0)
    p.add_argument("--vc_kap2", type=float, default=1.0)
    p.add_argument("--vcs_every", type=int, default=10)
    p.add_argument("--vcs_first", type=int, default=2)
    p.add_argument("--vcs_add", type=int, default=1000000)
    p.add_argument("--vcs_max", type=int, default=20000000)
    p.add_argument("--vcs_drop", type=int, default=5)
    p.add_argument("--vcs_tol", type=float, default=1e-4)
    p.add_argument("--vcs_score", default="pen")
    # device options (no reference counterpart)
    p.add_argument("--dtype", default="f32", help="V storage on the device: f32 | bf16 | f64")
    p.add_argument("--device", type=int, default=0)
    p.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                   help="a kernel switch of the device library (lb_set_option), e.g. sweep_graph=1")
    a, unknown = p.parse_known_args(argv)
    if unknown:
        raise SystemExit(f"unsupported flags for the device solver: {' '.join(unknown)}")
    if a.mom > 0:
        raise SystemExit("--mom > 0 is not supported by the device solver (production runs use --mom 0)")
    if a.est_every > 0:
        raise SystemExit("--est_every > 0 (in-solver estimates) is not supported: run the estimate stage on the "
                         "checkpoint (production runs use --est_every 0)")
    if a.k not in (8, 16, 32):
        raise SystemExit("--k must be 8, 16 or 32 on the device")
    if a.vcs_score not in ("pen", "mean", "mass"):
        raise SystemExit("bad --vcs_score")
    for o in a.option:
        name, _, val = o.partition("=")
        if not name or not val.lstrip("-").isdigit():
            raise SystemExit(f"--option {o!r}: expected NAME=INTEGER")
    return a


def _pw_constants(spec):
    """(c1, c2) of the pair-weight rule: 0.1 and 0.05 unless `spec` is "c1,c2" (an empty field keeps its default)."""
    c = [0.1, 0.05]
    if spec not in ("default", "on", "1", "", None):
        for k, tok in enumerate(str(spec).split(",")[:2]):
            if tok:
                c[k] = float(tok)
    return c[0], c[1]


def pair_weight_rule(inst, V, rho, w0, spec="default"):
    """Pair-level weight of the search (the reference's --pw_rule, alm_vcs.py:1663-1688): the smaller of
      w_mob  = c1 / R(1), R(w) = sum_i d_i kappa_i(w) / max(|(B V)_i|, 1e-9) / sum_i d_i, the volume-weighted ratio of
               a row's pair-level curvature kappa_i(w) = d_i^2 sum_levels K_Q(i), K_Q = 2 rho (sum_R r_QR |Y_R|^2 +
               r_QQ |Y_Q|^2), r = w rq rq^T, Y = cluster sums of d V, to its objective gradient (B V)_i; R is linear
               in w, so R(1) = R(w0) / w0;
      w_step = c2 / (rho n^2 / L).
    Returns (w, info)."""
    import scipy.sparse as sp
    c1, c2 = _pw_constants(spec)
    d = inst.d
    A = sp.csr_matrix((np.ones(inst.nnz) if inst.w is None else inst.w, inst.col, inst.rowptr), shape=(inst.n, inst.n))
    grad = np.linalg.norm(A @ V - np.outer(d, d @ V) / inst.L, axis=1)
    curv = np.zeros(inst.n)
    for lv in inst.levels:
        Y = np.zeros((lv.P, V.shape[1]))
        np.add.at(Y, lv.lab, d[:, None] * V)
        ysq = np.einsum("ij,ij->i", Y, Y)
        rr = w0 * np.outer(lv.rq, lv.rq)
        K = 2.0 * rho * (rr @ ysq + np.diag(rr) * ysq)
        curv += d ** 2 * K[lv.lab]
    R_w0 = float(np.dot(d, curv / np.maximum(grad, 1e-9)) / d.sum())
    R1 = R_w0 / w0
    n2L = inst.n * float(inst.n) / inst.L
    w_mob, w_step = c1 / R1, c2 / (rho * n2L)
    w = min(w_mob, w_step)
    return w, dict(R1=R1, R_w0=R_w0, w_mob=w_mob, w_step=w_step, n2L=n2L, c1=c1, c2=c2,
                   binding="mobility" if w_mob <= w_step else "step")


class Solver:
    """The solver stage. From the command line it reads DATA_DIR/<tag> and the label files named by the flags. In
    process, `inst` (graph only), `labels` / `vcs_labels` (label arrays) and `init_labels` replace the files;
    `log` replaces the timestamped print; `signals` False leaves SIGTERM / SIGINT alone (worker threads)."""

    def __init__(self, a, inst=None, labels=None, vcs_labels=None, init_labels=None, log=None, signals=True):
        self.a = a
        self.t0 = time.time()
        if log is not None:
            self.log = log
        if inst is None:
            D = os.environ.get("DATA_DIR", "data/")
            if not D.endswith("/"):
                D += "/"
            inst = load_instance(D + a.tag)
        self.inst = inst
        n = inst.n
        self.log(f"loaded {a.tag}: n {n} nnz {inst.nnz} L {inst.L:g} C0/L {inst.C0 / inst.L:.6f}")
        z0 = np.load(a.resume) if a.resume else None
        # pair levels (alm_vcs.py:1125-1138, 1254-1259)
        if z0 is not None and not a.pair_labels and labels is None:
            labs = [z0[f"lab{li}"] for li in range(int(z0["nlev"]))]
        else:
            labs = []
            for f in (labels if labels is not None else a.pair_labels):
                lab = np.unique(np.load(f) if isinstance(f, str) else np.asarray(f), return_inverse=True)[1]
                if a.max_clusters and lab.max() + 1 > a.max_clusters:
                    vol_ = np.bincount(lab, weights=inst.d)
                    keep_ = np.argsort(-vol_)[:a.max_clusters - 1]
                    self.log(f"level {os.path.basename(f)}: capped to {a.max_clusters} clusters (merged volume share "
                             f"{1 - vol_[keep_].sum() / vol_.sum():.4f})")
                    lab = I.cap_clusters(lab, inst.d, a.max_clusters)
                labs.append(lab)
        inst.pairlvl_weight = float(a.pairlvl_weight)
        inst.vc_weight = float(a.vc_weight)
        inst.levels = [I.pair_level(lab, inst.d) for lab in labs]
        for lv in inst.levels:
            self.log("pair level clusters", lv.P)
        # sparse VC levels (alm_vcs.py:1155-1167): empty working sets, or the checkpoint's
        vfiles = list(vcs_labels) if vcs_labels is not None else list(a.vcs_labels)
        inst.vcs = [I.vcs_level(np.load(f) if isinstance(f, str) else np.asarray(f), inst.d, np.zeros(n + 1, np.int64),
                                np.zeros(0, np.int32)) for f in vfiles]
        self.ages = [np.zeros(0, np.uint8) for _ in inst.vcs]
        if z0 is not None and "nvcs" in z0.files:
            if not inst.vcs:
                inst.vcs = [I.vcs_level(z0[f"vcslab{m}"], inst.d, np.zeros(n + 1, np.int64), np.zeros(0, np.int32))
                            for m in range(int(z0["nvcs"]))]
                self.ages = [np.zeros(0, np.uint8) for _ in inst.vcs]
            for m, vl in enumerate(inst.vcs):
                if m < int(z0["nvcs"]) and np.array_equal(vl.lab, np.unique(z0[f"vcslab{m}"], return_inverse=True)[1]):
                    vl.ptr = z0[f"vcsptr{m}"].astype(np.int64)
                    vl.col = z0[f"vcscol{m}"].astype(np.int32)
                    vl.sig = z0[f"vcssig{m}"].astype(np.float64)
                    self.ages[m] = np.zeros(vl.nnz, np.uint8)
                    self.log(f"resumed vcs level {m}: nnz {vl.nnz} sig sum {vl.sig.sum():.4g}")
        for vl in inst.vcs:
            self.log(f"VCS level: P {vl.P} active-set nnz {vl.nnz}")
        # V, triangle pool (alm_vcs.py:1187-1216)
        self.rng = np.random.default_rng(a.seed)
        TI = np.zeros((0, 3), np.int32)
        lam = np.zeros(0)
        if z0 is not None:
            V = z0["V"].astype(np.float64)
            TI, lam = I.canonical_triangles(z0["TI"].astype(np.int64), n, z0["lam"].astype(np.float64))
            for li, lv in enumerate(inst.levels):
                if f"sig{li}" in z0.files and z0[f"sig{li}"].shape == lv.sig.shape:
                    lv.sig = z0[f"sig{li}"].astype(np.float64)
            z0.close()
        elif init_labels is not None or a.init_partition:
            part = np.asarray(init_labels) if init_labels is not None else np.load(a.init_partition)
            V = init_partition_V(inst, part, a.k, self.rng, self.log)
        elif a.init_V:
            V = np.load(a.init_V).astype(np.float64)
        else:
            V = self.rng.standard_normal((n, a.k))
        V /= np.linalg.norm(V, axis=1, keepdims=True)
        if V.shape[1] != a.k:
            raise SystemExit(f"V has {V.shape[1]} columns, --k {a.k}")
        inst.TI, inst.lam = TI, lam
        self.eng = Engine(r=a.k, dtype=a.dtype, device=a.device)
        for o in a.option:
            name, _, val = o.partition("=")
            self.eng.set_option(name, int(val))
        self.eng.set_instance(inst)
        self.eng.set_V(V)
        self._upload_ages()
        self.W_AUTO = None
        if (a.pw_rule) and inst.levels:
            self.pw_rule(V)
        self.stop = False
        if signals:
            signal.signal(signal.SIGTERM, self._stop)
            signal.signal(signal.SIGINT, self._stop)

    # ------------------------------------------------------------------------------------------------ helpers
    def log(self, *s):
        rss = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 2 ** 10
        print(f"[{time.time() - self.t0:8.0f}s {rss:6.0f}MB]", *s, flush=True)

    def _stop(self, *_):
        self.stop = True

    def _upload_ages(self):
        if self.inst.vcs and sum(v.nnz for v in self.inst.vcs):
            self.eng.set_vcs_age(np.concatenate(self.ages))

    def set_pair_weight(self, w):
        """r_QR = rq_Q rq_R w and wv = vc_weight * w everywhere (set_pair_weight, alm_vcs.py:1270-1275)."""
        self.a.pairlvl_weight = float(w)
        self.inst.pairlvl_weight = float(w)
        self.eng.set_weights(float(w), self.a.vc_weight)

    def pw_rule(self, V):
        """Choose the pair weight by pair_weight_rule at the solver's rho and install it."""
        w, info = pair_weight_rule(self.inst, V, self.a.rho, float(self.a.pairlvl_weight), self.a.pw_rule)
        self.W_AUTO = w
        self.log(f"pair weight: mobility at unit weight {info['R1']:.6g} -> {info['w_mob']:.6g}, step limit "
                 f"{info['w_step']:.6g} (n^2/L {info['n2L']:.4g}), {info['binding']} binds -> pairlvl_weight {w:.6g}")
        self.set_pair_weight(w)

    def sweeps(self, nsw, rho, rt, damp):
        n = self.inst.n
        orders = np.stack([self.rng.permutation(n) for _ in range(nsw)]) if nsw else np.zeros((0, n), np.int64)
        if nsw:
            self.eng.sweeps(orders, self.a.nbatch, rho, rt, damp, self.a.vc_kap2)

    def fetch_multipliers(self):
        pp, pt, pv = self.eng.get_multipliers()
        return pp, pt, pv

    # ------------------------------------------------------------------------------------------------ separation
    def vcs_separation(self):
        a = self.a
        msg = ""
        sig_all, age_all = self.eng.get_vcs_state()
        o = 0
        changed = False
        for m, vl in enumerate(self.inst.vcs):
            vl.sig = sig_all[o:o + vl.nnz].copy()
            self.ages[m] = age_all[o:o + vl.nnz].copy()
            o += vl.nnz
        for m, vl in enumerate(self.inst.vcs):
            ts = time.time()
            nz0 = vl.nnz
            keep = (self.ages[m] < a.vcs_drop) if a.vcs_drop > 0 else np.ones(vl.nnz, bool)
            room = max(0, a.vcs_max - int(keep.sum()))
            ni, nR, sc, stv = self.eng.separate_vcs(m, nadd=max(1, min(a.vcs_add, room)), tol=a.vcs_tol,
                                                    mode=a.vcs_score)
            if room == 0:
                ni, nR = ni[:0], nR[:0]
            ptr, col, src = _merge_ws(vl, keep, ni, nR)
            sig = _take_kept(vl.sig, src, np.float64)
            age = _take_kept(self.ages[m], src, np.uint8)
            vl.ptr, vl.col, vl.sig = ptr, col, sig
            self.ages[m] = age
            changed = True
            L2 = self.inst.L ** 2
            msg += (f" vcs[{m}] nnz {nz0}->{vl.nnz} (-{nz0 - int(keep.sum())} +{len(ni)}) viol {stv['nviol']:.4g} "
                    f"(ws {stv['nviol_ws']:.4g}) mass/L2 {stv['mass'] / L2:.4g} (ws {stv['mass_ws'] / L2:.4g}) "
                    f"[{time.time() - ts:.1f}s]")
        if changed:
            self.eng.set_vcs(self.inst)
            self.eng.set_weights(self.a.pairlvl_weight, self.a.vc_weight)
            self._upload_ages()
        return msg

    def tri_separation(self, lam):
        a = self.a
        n = self.inst.n
        TI = self.inst.TI.astype(np.int64)
        keep = lam > 0
        TI = TI[keep]
        lam = lam[keep]
        room = max(0, a.tri_max - len(TI))
        new = np.zeros((0, 3), np.int64)
        if room:
            cap0 = int(min(a.tri_add, room) * 1.3) + 10
            new, val, _ = self.eng.separate_triangles(0, thresh=a.tri_thresh, maxh=a.tri_maxh, vtol=1e-3, cap=cap0)
            if a.tri1_add:
                cap1 = int(a.tri1_add * 1.3) + 10
                n1, v1, _ = self.eng.separate_triangles(1, thresh=a.tri_thresh, maxh=a.tri_maxh, maxk=a.tri1_maxk,
                                                        hub2=a.tri1_hub, cap2=a.tri1_cap, vtol=1e-3, cap=cap1)
                if len(n1):
                    o1 = np.argsort(-v1, kind="stable")[:a.tri1_add]
                    new = np.vstack([new, n1[o1]])
                    val = np.r_[val + 10.0, v1[o1]]           # keep both budgets (wedges first)
            if len(new):
                new, val = _new_triangles(new, val, TI, n)
                o = np.argsort(-val, kind="stable")[:min(a.tri_add + a.tri1_add, room)]
                new = new[o]
                TI = np.vstack([TI, new]) if len(TI) else new
                lam = np.r_[lam, np.zeros(len(new))]
        self.inst.TI = TI.astype(np.int32)
        self.inst.lam = lam
        self.eng.set_triangles(self.inst.TI, lam)
        return f" +{len(new)}tri"

    # ------------------------------------------------------------------------------------------------ checkpoint
    def save(self, path):
        """NPZ with the reference's keys (alm_vcs.py:1576-1589); written atomically."""
        inst = self.inst
        n = inst.n
        V = self.eng.get_V()
        dd = {"TI": inst.TI.astype(np.int64), "lam": inst.lam, "V": V.astype(np.float32), "nlev": len(inst.levels),
              "mu": np.zeros(n), "CI": np.zeros((0, 3), np.int64), "om": np.zeros(0), "PI": np.zeros((0, 2), np.int64),
              "psi": np.zeros(0)}
        for li, lv in enumerate(inst.levels):
            dd[f"lab{li}"] = lv.lab.astype(np.int64)
            dd[f"sig{li}"] = 0.5 * (lv.sig + lv.sig.T)
            dd[f"dense{li}"] = 1
            dd[f"pair{li}"] = 1
        dd["nvc"] = 0
        dd["nvcs"] = len(inst.vcs)
        if inst.vcs and sum(v.nnz for v in inst.vcs):
            sig_all, _ = self.eng.get_vcs_state()
        o = 0
        for m, vl in enumerate(inst.vcs):
            sig = sig_all[o:o + vl.nnz] if vl.nnz else np.zeros(0)
            o += vl.nnz
            dd[f"vcslab{m}"] = vl.lab.astype(np.int64)
            dd[f"vcsptr{m}"] = vl.ptr
            dd[f"vcscol{m}"] = vl.col
            dd[f"vcssig{m}"] = sig.astype(np.float32)
            dd[f"vcsfull{m}"] = 0
        tmp = path.replace(".npz", ".tmp.npz")
        np.savez(tmp, **dd)
        os.replace(tmp, path)

    def sync_level_sig(self, pp):
        o = 0
        for lv in self.inst.levels:
            lv.sig = pp[o:o + lv.P * lv.P].reshape(lv.P, lv.P).copy()
            o += lv.P * lv.P

    # ------------------------------------------------------------------------------------------------ main loop
    def run(self, close=True):
        """The outer loop. With close=False the device context stays open (V and the final multipliers resident) for
        the caller, who closes it (self.eng.close())."""
        a = self.a
        inst = self.inst
        L = inst.L
        rho = a.rho
        self.log(f"search: {inst.n} vertices, {inst.nnz // 2} edges, rank {a.k}, pair levels "
                 f"{[lv.P for lv in inst.levels]}, {inst.T} triangles, rho {rho}, {a.nbatch} batches, {a.dtype}; "
                 f"C0/L {inst.C0 / L:.6f}")
        parts = self.eng.eval(rho, rho * a.tri_weight)
        self.log(f"initial augmented value {parts[0] / L:.6f}, primal {parts[6] / L:.6f}")
        hist = []
        tsweep = tsep = tvcs = 0.0
        PH2 = False
        last = False
        for it in range(1, a.outer + 1):
            if not PH2 and a.phase2_at_it > 0 and it > a.phase2_at_it:
                PH2 = True
                w2 = a.phase2_w * self.W_AUTO if self.W_AUTO is not None else a.phase2_w
                self.set_pair_weight(w2)
                a.inner = a.phase2_inner
                a.tri_every = a.phase2_tri_every
                a.est_every = a.phase2_est_every
                self.log(f"proximal phase from outer iteration {it}: {a.inner} sweeps per update, pair weight "
                         f"{a.pairlvl_weight:.4g}, triangle separation every {a.tri_every}")
            ts = time.time()
            self.sweeps(a.inner, rho, rho * a.tri_weight, a.damp)
            self.eng.sync()
            tsweep += time.time() - ts
            parts = self.eng.eval(rho, rho * a.tri_weight)
            Lc, prim, nviol = parts[0], parts[6], int(parts[8])
            gmin = parts[9]
            tviol = float(max(0.0, -gmin)) if inst.T and math.isfinite(gmin) else 0.0
            sep_now = a.tri and it % a.tri_every == 0 and not self.stop
            need_host = sep_now or (a.ckpt_every > 0 and it % a.ckpt_every == 0)
            last = self.stop or it == a.outer or time.time() - self.t0 > a.time_limit
            if need_host or last:
                # the evaluation's PHR values are the multipliers after the update below
                pp, pt, _ = self.fetch_multipliers()
                self.sync_level_sig(pp)
                inst.lam = pt.copy() if inst.T else np.zeros(0)
            self.eng.phr_update()                       # sig <- pp, lam <- pt, VCS sig <- pv, ages
            msg = ""
            if inst.vcs and not self.stop and it < a.outer and (it % a.vcs_every == 0 or it == a.vcs_first):
                ts = time.time()
                msg += self.vcs_separation()
                tvcs += time.time() - ts
            if sep_now:
                ts = time.time()
                msg += self.tri_separation(inst.lam)
                tsep += time.time() - ts
            rec = dict(outer=it, t=round(time.time() - self.t0, 1), Laug=Lc / L, primal=prim / L, nviol=nviol,
                       tviol=tviol, ntri=int(inst.T))
            if inst.vcs:
                rec.update(vcs_nnz=[vl.nnz for vl in inst.vcs], vcs_viol=int(parts[10]))
            self.log(f"outer {it}: augmented {Lc / L:.6f}, primal {prim / L:.6f}, {nviol} pair violations, worst "
                     f"triangle {tviol:.3f}, {inst.T} triangles{msg} (sweeps {tsweep:.0f}s, triangle separation "
                     f"{tsep:.0f}s, VCS {tvcs:.0f}s)")
            hist.append(rec)
            if a.save and (last or (a.ckpt_every > 0 and it % a.ckpt_every == 0)):
                self.save(a.save.replace(".npz", "_last.npz"))
            if a.save and (it % 20 == 0 or last):
                json.dump(hist, open(a.save.replace(".npz", ".history.json"), "w"))
            if last:
                break
        # final state on the host: V and the multipliers after the last PHR update (levels and lam were fetched
        # above at the last outer; the VCS multipliers are read here)
        self.V = self.eng.get_V()
        if inst.vcs and sum(v.nnz for v in inst.vcs):
            sig_all, _ = self.eng.get_vcs_state()
            o = 0
            for vl in inst.vcs:
                vl.sig = sig_all[o:o + vl.nnz].copy()
                o += vl.nnz
        self.history = hist
        if close:
            self.eng.close()
        return hist


def _new_triangles(new, val, TI, n):
    """Candidates that are not in the pool, one per packed key (its first occurrence), in ascending key order
    (alm_vcs.py:1751-1758: drop existing, dedupe with np.unique), through liblbhost (lbh_new_triangles: the distinct
    candidates found by a sort and checked against the sorted pool keys in one merge pass; the same set and order as
    filtering first)."""
    fi = H.new_triangles(n, new, TI)
    return new[fi], val[fi]


def _take_kept(a, src, dtype):
    """out[j] = a[src[j]] where src[j] >= 0 (a kept entry), 0 for a new entry; `a` may be empty."""
    return H.take(a, src, dtype)


def _merge_ws(vl, keep, ni, nR):
    """update_ws / _merge (vcsops.py:439-464): kept old entries and new (i, R) entries, CSR sorted by (i, R), in the
    order of a stable sort of [kept, new] (liblbhost lbh_ws_merge). Returns (ptr, col, src) with src = index of the
    old entry or -1 for a new one."""
    return H.ws_merge(len(vl.lab), vl.P, vl.ptr, vl.col, keep, ni, nR)


def load_instance(D):
    """Instance directory -> Instance (alm_vcs.py:1103-1117)."""
    g = json.load(open(os.path.join(D, "graph.json")))
    n = int(g["n"])
    keys = np.load(os.path.join(D, "keys.npy")).astype(np.int64)
    wf = os.path.join(D, "weights.npy")
    W = np.load(wf).astype(np.float64) if os.path.exists(wf) else None
    u = keys // n
    v = keys % n
    d = np.bincount(np.r_[u, v], minlength=n).astype(float)
    L = 2.0 * len(keys)
    C0 = 0.0
    if os.path.exists(os.path.join(D, "vol.npy")):
        meta = json.load(open(os.path.join(D, "meta.json")))
        d = np.load(os.path.join(D, "vol.npy")).astype(float)
        L = float(int(meta["L_orig"]))
        C0 = 2.0 * meta["leaf_edges"]
        assert int(d.sum()) == int(L)
    rowptr, col, w = I.csr_from_keys(n, keys, W)
    return I.Instance(n=n, rowptr=rowptr, col=col, w=w, d=d, L=L, C0=C0)


def init_partition_V(inst, lab, k, rng, log):
    """Start V at a colouring-embedded partition matrix (alm_vcs.py:1198-1212): greedy colouring of the cluster graph
    with k colours (largest volume first, least-used free colour, else the least-used colour), one-hot rows plus
    1e-3 Gaussian noise."""
    import scipy.sparse as sp
    n = inst.n
    lab = np.unique(lab, return_inverse=True)[1]
    P = int(lab.max()) + 1
    rows = np.repeat(np.arange(n), np.diff(inst.rowptr))
    Cg = sp.csr_matrix((np.ones(inst.nnz), (lab[rows], lab[inst.col])), shape=(P, P))
    vol = np.bincount(lab, weights=inst.d, minlength=P)
    col = -np.ones(P, np.int64)
    cvol = np.zeros(k)
    for Q in np.argsort(-vol):
        nb_ = Cg.indices[Cg.indptr[Q]:Cg.indptr[Q + 1]]
        used = np.zeros(k, bool)
        used[col[nb_][col[nb_] >= 0]] = True
        free = np.flatnonzero(~used)
        col[Q] = free[np.argmin(cvol[free])] if len(free) else int(np.argmin(cvol))
        cvol[col[Q]] += vol[Q]
    V = np.zeros((n, k))
    V[np.arange(n), col[lab]] = 1.0
    V += 1e-3 * rng.standard_normal(V.shape)
    coo = Cg.tocoo()
    log(f"init partition: {P} clusters -> {k} colours, {int(((col[coo.row] == col[coo.col]) & (coo.row < coo.col)).sum())} clashing cluster "
        f"adjacencies")
    return V


def main(argv=None):
    a = parse(argv)
    if not a.save:
        raise SystemExit("--save is required")
    s = Solver(a)
    s.run()
    return 0


if __name__ == "__main__":
    sys.exit(main())
