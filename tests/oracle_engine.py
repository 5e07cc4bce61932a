"""A CPU stand-in for lagbound_amd.engine.Engine built on the FP64 oracle -- TEST INFRASTRUCTURE ONLY.

It lets the solver's host logic (working-set merges, triangle purge and dedupe, checkpoints, resume, phase switches)
run end to end on a machine without a GPU. It keeps the device's state model: multipliers live "on the device" (here,
in this object) and are replaced by the last evaluation's PHR values on phr_update(); VCS ages advance by the slack
rule (vcsops.py:300-305). Installed into lagbound_amd.alm by monkeypatching `alm.Engine` inside a test.
"""
from __future__ import annotations

import copy

import numpy as np

import oracle as O


class OracleEngine:
    def __init__(self, r=16, dtype="f32", device=0):
        self.r = int(r)
        self.dtype = dtype
        self.inst = None
        self.V = None
        self.last = None
        self.ages = None
        self.closed = False

    def close(self):
        self.closed = True              # like the device context: no calls after close

    def _open(self):
        if self.closed:
            raise RuntimeError("engine used after close()")

    def set_tuning(self, *a, **k):
        pass

    def _ws(self):
        return sum(v.nnz for v in self.inst.vcs)

    def set_instance(self, inst):
        self.inst = copy.deepcopy(inst)
        self.ages = np.zeros(self._ws(), np.uint8)
        self.last = None

    def set_vcs(self, inst):
        self.inst.vcs = copy.deepcopy(inst.vcs)
        self.ages = np.zeros(self._ws(), np.uint8)
        self.last = None

    def set_triangles(self, TI, lam):
        self.inst.TI = np.asarray(TI, np.int32).reshape(-1, 3).copy()
        self.inst.lam = np.asarray(lam, np.float64).copy()
        self.last = None

    def set_weights(self, pairlvl_weight, vc_weight=1.0):
        self.inst.pairlvl_weight = float(pairlvl_weight)
        self.inst.vc_weight = float(vc_weight)

    def set_vcs_age(self, age):
        assert len(age) == self._ws()
        self.ages = np.asarray(age, np.uint8).copy()

    def set_V(self, V):
        self._open()
        V = np.asarray(V, np.float64)
        assert V.shape == (self.inst.n, self.r)
        self.V = V.copy()

    def get_V(self):
        return self.V.copy()

    def sync(self):
        pass

    def eval(self, rho, rt=None, fetch=True):
        self._open()
        rt = rho if rt is None else rt
        parts, G, pp, pt, pv = O.eval_c(self.inst, self.V, rho, rt, threads=2)
        self.last = (rho, parts, G, pp, pt, pv)
        return parts if fetch else None

    def get_G(self):
        return self.last[2].copy()

    def get_multipliers(self):
        _, _, _, pp, pt, pv = self.last
        return pp.copy(), pt.copy(), pv.copy()

    def _slack(self):
        """Inactive-with-slack flag of each VCS entry at the current V: p == 0 and g > 0 (vcsops.py:285)."""
        out = []
        inst = self.inst
        d = inst.d
        for m, vl in enumerate(inst.vcs):
            if vl.nnz == 0:
                continue
            Y = O._W(vl.lab, d, vl.P) @ self.V
            rows = np.repeat(np.arange(inst.n), np.diff(vl.ptr))
            R = vl.col.astype(np.int64)
            s = np.einsum("ij,ij->i", self.V[rows], Y[R])
            dd = d[rows]
            g = np.where(vl.lab[rows] == R, dd * (s - dd), dd * s)
            out.append(g)
        return np.concatenate(out) if out else np.zeros(0)

    def phr_update(self):
        rho, parts, G, pp, pt, pv = self.last
        assert rho > 0
        inst = self.inst
        o = 0
        for lv in inst.levels:
            lv.sig = pp[o:o + lv.P * lv.P].reshape(lv.P, lv.P).copy()
            o += lv.P * lv.P
        if self._ws():
            g = self._slack()
            slack = (pv == 0.0) & (g > 0.0)
            self.ages = np.where(slack, np.minimum(self.ages.astype(np.int64) + 1, 255), 0).astype(np.uint8)
            o = 0
            for vl in inst.vcs:
                vl.sig = pv[o:o + vl.nnz].copy()
                o += vl.nnz
        if inst.T:
            inst.lam = pt.copy()
        self.last = None

    def get_vcs_state(self):
        sig = np.concatenate([v.sig for v in self.inst.vcs]) if self._ws() else np.zeros(0)
        return sig.copy(), self.ages.copy()

    def sweeps(self, orders, nbatch, rho, rt=None, damp=0.0, kap2=1.0):
        self.V = O.sweeps_c(self.inst, self.V, orders, nbatch, rho, rt, damp, kap2, threads=2)
        self.last = None

    def separate_triangles(self, fam=0, thresh=0.3, maxh=100, maxk=60, hub2=8, cap2=200, vtol=1e-3, cap=1000,
                           nchunk=512, nb=2000):
        return O.separate_c(self.inst, self.V, fam, thresh, maxh, maxk, hub2, cap2, vtol, cap, nchunk, nb, threads=2)

    def separate_vcs(self, m=0, nadd=1000, tol=1e-4, mode="pen"):
        vl = self.inst.vcs[m]
        return O.vcs_separate_numpy(vl, self.V, self.inst.d, self.inst.wv, nadd, tol, mode)


class OracleLagrangian:
    """C(y) X at the instance's multipliers with the C oracle (evaluation at rho = 0, G = C(y) X); the CPU stand-in
    for lagbound_amd.bound.DeviceLagrangian."""

    def __init__(self, inst, r=16):
        self.inst = inst
        self.r = r
        self.const = float(inst.C0) + (float(np.sum(inst.lam)) if inst.T else 0.0)
        self.inc = O.incidence_c(inst.n, inst.TI) if inst.T else None

    def CX(self, X):
        X = np.ascontiguousarray(X, np.float64)
        return O.eval_c(self.inst, X, 0.0, threads=2, incidence=self.inc)[1]

    def close(self):
        pass
