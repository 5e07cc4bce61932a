"""Batches of independent graphs evaluated together on one device context (lb_batch_*): the graphs are packed into
one block-diagonal graph (contiguous row ranges, vertex ids offset), each with its own L, C0, rank-1 term and pair
level; one launch of each batch kernel evaluates f and G = C(y)V for every graph.

No reference counterpart: the reference evaluates one graph per process (alm_vcs.py `augmented`, `CV`). Each graph's
numbers are those of the single-graph evaluation of the same instance.
"""
from __future__ import annotations

import numpy as np

from . import instances as I
from .engine import Engine


class Packed:
    """Block-diagonal packing of Instances that carry (at most) one pair level each and a triangle pool."""

    def __init__(self, insts):
        self.insts = insts
        self.G = len(insts)
        ns = np.array([x.n for x in insts], np.int64)
        self.row_off = np.r_[0, np.cumsum(ns)].astype(np.int64)
        self.n = int(self.row_off[-1])
        nnzs = [x.nnz for x in insts]
        self.rowptr = np.zeros(self.n + 1, np.int64)
        cols, ws, ds, labs, Ps, sigs, rqs, dsqs, TIs, lams = [], [], [], [], [], [], [], [], [], []
        weighted = any(x.w is not None for x in insts)
        e_off = 0
        for g, x in enumerate(insts):
            o = self.row_off[g]
            self.rowptr[o + 1:o + x.n + 1] = x.rowptr[1:] + e_off
            e_off += x.nnz
            cols.append(x.col.astype(np.int64) + o)
            if weighted:
                ws.append(np.ones(x.nnz) if x.w is None else x.w)
            ds.append(x.d)
            if len(x.levels) > 1:
                raise ValueError("batch mode carries one pair level per graph")
            if x.vcs:
                raise ValueError("batch mode carries no VCS levels")
            if x.levels:
                lv = x.levels[0]
                labs.append(lv.lab.astype(np.int32))
                Ps.append(lv.P)
                sigs.append(lv.sig.ravel())
                rqs.append(lv.rq)
                dsqs.append(lv.dsq)
            if x.T:
                TIs.append(x.TI.astype(np.int64) + o)
                lams.append(x.lam)
        self.col = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)
        self.w = np.concatenate(ws) if weighted else None
        self.d = np.concatenate(ds)
        self.L = np.array([x.L for x in insts])
        self.C0 = np.array([x.C0 for x in insts])
        self.has_levels = bool(labs)
        if self.has_levels:
            if len(labs) != self.G:
                raise ValueError("either every graph of the batch has a pair level or none has")
            self.lab = np.concatenate(labs)
            self.P = np.array(Ps, np.int32)
            self.sig = np.concatenate(sigs)
            self.rq = np.concatenate(rqs)
            self.dsq = np.concatenate(dsqs)
            self.pw = float(insts[0].pairlvl_weight)
        self.TI = np.concatenate(TIs).astype(np.int32) if TIs else np.zeros((0, 3), np.int32)
        self.lam = np.concatenate(lams) if lams else np.zeros(0)
        self.tri_off = np.r_[0, np.cumsum([x.T for x in insts])].astype(np.int64)

    def rows(self, g):
        return slice(int(self.row_off[g]), int(self.row_off[g + 1]))


class BatchEngine:
    """One device context holding a packed batch."""

    def __init__(self, packed: Packed, r=8, dtype="f32", device=0):
        self.p = packed
        self.eng = Engine(r=r, dtype=dtype, device=device)
        e = self.eng
        union = I.Instance(n=packed.n, rowptr=packed.rowptr, col=packed.col, w=packed.w, d=packed.d,
                           L=float(packed.L.sum()), C0=0.0)
        e.set_instance(union)
        e.set_batch(packed.row_off, packed.L, packed.C0)
        if packed.has_levels:
            e.set_batch_levels(packed.lab, packed.P, packed.sig, packed.rq, packed.dsq, packed.pw)
        if len(packed.TI):
            e.set_triangles(packed.TI, packed.lam)

    def set_V(self, V):
        self.eng.set_V(V)

    def eval(self, rho, rt=None, fetch=True):
        return self.eng.batch_eval(rho, rt, fetch)

    def get_G(self):
        return self.eng.get_G()

    def sync(self):
        self.eng.sync()

    def close(self):
        self.eng.close()
