"""Vertex order of the device copy of an instance.

The f+grad evaluation and the sweeps are bound by random 32-64 byte gathers of V rows, one per edge and two per
triangle record. In a power-law graph most gathers hit the few thousand highest-degree vertices. Numbering the
vertices by descending degree packs those rows into a small set of cache lines that stay resident in L2, where the
given (random) numbering spreads every hub row over a line it shares with three cold rows.

`VertexOrder` builds the relabelled instance: vertex v of the input becomes `new_of[v]`, and new vertex k is input
vertex `order[k]` (stable sort by descending degree). `Relabeled` wraps an engine-shaped backend (the device
context, or any object with the same methods) that holds the relabelled instance, and translates every argument and
result back to the caller's numbering, so callers never see the device order:

- V, G, sweep orders, diagonals: rows permuted (V and G by the backend's own copy loops when it has
  `set_row_order`, as the device context does);
- pair and VCS labels: permuted with their rows; VCS working-set entries (sig, age, effective multipliers) through
  an entry map;
- triangles: vertex ids mapped, pool order unchanged (so lam and pt keep their indices);
- triangle separation: run on a second, graph-only context holding the caller's CSR and V in the caller's order,
  because its candidate order, chunking and caps are defined on the given numbering (alm_vcs.py:684-797);
- VCS separation: run on the device order, rows mapped back, re-sorted by (score desc, i, R).

No reference counterpart: the reference runs on the given numbering.
"""
from __future__ import annotations

import copy

import numpy as np

from . import _host as H
from . import instances as I


class VertexOrder:
    def __init__(self, rowptr: np.ndarray):
        deg = np.diff(np.asarray(rowptr, np.int64))
        self.n = len(deg)
        dmax = int(deg.max()) if len(deg) else 0
        self.order = H.argsort_bounded(dmax - deg, dmax + 1)       # argsort(-deg, kind="stable")
        self.new_of = np.empty(self.n, np.int64)
        self.new_of[self.order] = np.arange(self.n, dtype=np.int64)

    # ---------------------------------------------------------------------------------------------- rows
    def rows_in(self, X: np.ndarray) -> np.ndarray:
        """Caller order -> device order (row k of the result is row order[k] of X)."""
        return np.ascontiguousarray(np.asarray(X)[self.order])

    def rows_out(self, Y: np.ndarray) -> np.ndarray:
        """Device order -> caller order."""
        out = np.empty_like(Y)
        out[self.order] = Y
        return out

    def ids_in(self, ids) -> np.ndarray:
        return self.new_of[np.asarray(ids, np.int64)]

    def ids_out(self, ids) -> np.ndarray:
        return self.order[np.asarray(ids, np.int64)]

    # ---------------------------------------------------------------------------------------------- structures
    def _row_gather(self, ptr: np.ndarray):
        """For a row-indexed CSR (ptr), the new ptr and, for each entry of the new layout, its entry index in the
        old layout (rows taken in `order`, entries of a row kept in their order)."""
        nptr, idx = H.row_gather(ptr, self.order)
        return nptr, idx, np.diff(nptr)

    def graph(self, inst):
        """(rowptr, col int32, w or None, d) of P A P^T with rows sorted by column. Row k of the row-permuted matrix
        B holds new_of[col] of input row order[k]; a stable grouping of B's entries by column lists, for each column
        c, the rows k ascending -- row c of B^T = P A P^T (A is symmetric)."""
        rowptr, idx, cnt = self._row_gather(inst.rowptr)
        bc = self.new_of[np.asarray(inst.col, np.int64)[idx]]
        t = H.argsort_bounded(bc, self.n)
        col = np.repeat(np.arange(self.n, dtype=np.int64), cnt)[t].astype(np.int32)
        w = None if inst.w is None else np.asarray(inst.w, np.float64)[idx][t]
        return rowptr, col, w, np.asarray(inst.d, np.float64)[self.order]

    def vcs(self, v):
        """Relabelled VCS level and the entry map (new entry position -> input entry index)."""
        ptr, idx, _ = self._row_gather(v.ptr)
        col = np.asarray(v.col, np.int32)[idx] if len(idx) else np.zeros(0, np.int32)
        sig = np.asarray(v.sig, np.float64)[idx] if len(idx) else np.zeros(0)
        nv = I.VCSLevelData(lab=np.asarray(v.lab)[self.order].astype(np.int32), P=v.P, rq=v.rq, ptr=ptr, col=col,
                            sig=sig)
        return nv, idx

    def instance(self, inst):
        """The relabelled instance (graph, pair levels, VCS levels, triangle pool) and the concatenated VCS entry
        map."""
        rowptr, col, w, d = self.graph(inst)
        J = I.Instance(n=inst.n, rowptr=rowptr, col=col, w=w, d=d, L=inst.L, C0=inst.C0,
                       pairlvl_weight=inst.pairlvl_weight, vc_weight=inst.vc_weight)
        for lv in inst.levels:
            J.levels.append(I.PairLevelData(lab=np.asarray(lv.lab)[self.order].astype(np.int32), P=lv.P, sig=lv.sig,
                                            rq=lv.rq, dsq=lv.dsq))
        maps, off = [], 0
        for v in inst.vcs:
            nv, idx = self.vcs(v)
            J.vcs.append(nv)
            maps.append(idx + off)
            off += v.nnz
        J.TI = self.triangles(inst.TI)
        J.lam = inst.lam
        vmap = np.concatenate(maps) if maps else np.zeros(0, np.int64)
        return J, vmap

    def triangles(self, TI) -> np.ndarray:
        TI = np.asarray(TI).reshape(-1, 3)
        return self.new_of[TI.astype(np.int64)].astype(np.int32) if len(TI) else np.zeros((0, 3), np.int32)


class Relabeled:
    """An engine-shaped handle whose backend holds the degree-ordered copy of the instance (order="degree") or the
    instance as given (order="given"). `make_backend()` creates a further backend of the same kind (used for the
    graph-only separation context)."""

    def __init__(self, backend, make_backend, order: str = "degree"):
        if order not in ("degree", "given"):
            raise ValueError(f"vertex order must be 'degree' or 'given', not {order!r}")
        self.b = backend
        self._make = make_backend
        self.order_mode = order
        self.vo = None
        self.vmap = None
        self.inst = None
        self._sep = None
        self._native_rows = False     # the backend permutes V / G rows in its own copies (set_row_order)

    def __getattr__(self, name):                      # everything without a translation goes to the backend
        b = self.__dict__.get("b")
        if b is None:
            raise AttributeError(name)
        return getattr(b, name)

    # ------------------------------------------------------------------------------------------------ life cycle
    def close(self):
        if self._sep is not None:
            self._sep.close()
            self._sep = None
        self.b.close()

    def _drop_sep(self):
        if self._sep is not None:
            self._sep.close()
            self._sep = None

    # ------------------------------------------------------------------------------------------------ upload
    def set_instance(self, inst):
        self._drop_sep()
        if self.order_mode == "degree" and inst.n > 1:
            self.vo = VertexOrder(inst.rowptr)
            J, self.vmap = self.vo.instance(inst)
        else:
            self.vo, self.vmap, J = None, None, inst
        self.b.set_instance(J)
        self._native_rows = False
        if self.vo is not None and hasattr(self.b, "set_row_order"):
            self.b.set_row_order(self.vo.order)
            self._native_rows = True
        self.inst = inst

    def set_levels(self, inst):
        if self.vo is None:
            return self.b.set_levels(inst)
        J = copy.copy(inst)
        J.levels = [I.PairLevelData(lab=np.asarray(lv.lab)[self.vo.order].astype(np.int32), P=lv.P, sig=lv.sig,
                                    rq=lv.rq, dsq=lv.dsq) for lv in inst.levels]
        return self.b.set_levels(J)

    def set_vcs(self, inst):
        if self.vo is None:
            return self.b.set_vcs(inst)
        J = copy.copy(inst)
        J.vcs, maps, off = [], [], 0
        for v in inst.vcs:
            nv, idx = self.vo.vcs(v)
            J.vcs.append(nv)
            maps.append(idx + off)
            off += v.nnz
        self.vmap = np.concatenate(maps) if maps else np.zeros(0, np.int64)
        return self.b.set_vcs(J)

    def _entries_in(self, x):
        return x if self.vo is None else np.asarray(x)[self.vmap]

    def _entries_out(self, x):
        if self.vo is None:
            return x
        out = np.empty_like(x)
        out[self.vmap] = x
        return out

    def set_vcs_sig(self, sigs):
        if self.vo is None:
            return self.b.set_vcs_sig(sigs)
        cat = np.concatenate([np.asarray(s, np.float64) for s in sigs]) if len(sigs) else np.zeros(0)
        return self.b.set_vcs_sig([self._entries_in(cat)])

    def set_vcs_age(self, age):
        return self.b.set_vcs_age(self._entries_in(np.asarray(age, np.uint8)))

    def get_vcs_state(self):
        sig, age = self.b.get_vcs_state()
        return self._entries_out(sig), self._entries_out(age)

    def set_triangles(self, TI, lam):
        TI = np.asarray(TI).reshape(-1, 3)
        return self.b.set_triangles(TI if self.vo is None else self.vo.triangles(TI), lam)

    def set_V(self, V):
        V = np.asarray(V, np.float64)
        assert V.shape == (self.inst.n, self.b.r), (V.shape, self.inst.n, self.b.r)
        return self.b.set_V(V if self.vo is None or self._native_rows else self.vo.rows_in(V))

    def get_V(self):
        V = self.b.get_V()
        return V if self.vo is None or self._native_rows else self.vo.rows_out(V)

    def get_G(self):
        G = self.b.get_G()
        return G if self.vo is None or self._native_rows else self.vo.rows_out(G)

    def get_multipliers(self):
        pp, pt, pv = self.b.get_multipliers()
        return pp, pt, self._entries_out(pv)

    # ------------------------------------------------------------------------------------------------ compute
    def sweeps(self, orders, nbatch, rho, rt=None, damp=0.0, kap2=1.0):
        if self.vo is not None:
            orders = self.vo.ids_in(np.atleast_2d(orders))
        return self.b.sweeps(orders, nbatch, rho, rt=rt, damp=damp, kap2=kap2)

    def _sep_backend(self):
        if self._sep is None:
            s = self._make()
            g = self.inst
            s.set_instance(I.Instance(n=g.n, rowptr=g.rowptr, col=g.col, w=g.w, d=g.d, L=g.L, C0=g.C0))
            self._sep = s
        return self._sep

    def separate_triangles(self, *args, **kw):
        if self.vo is None:
            return self.b.separate_triangles(*args, **kw)
        s = self._sep_backend()
        s.set_V(self.get_V())
        return s.separate_triangles(*args, **kw)

    def separate_vcs(self, *args, **kw):
        i, R, sc, st = self.b.separate_vcs(*args, **kw)
        if self.vo is None or len(i) == 0:
            return i, R, sc, st
        i = self.vo.ids_out(i)
        if not np.any(sc[1:] == sc[:-1]):            # no tied scores: the order (score descending) is final
            return i, R, sc, st
        o = H.tie_order(sc, i, R)                  # np.lexsort((R, i, -sc)), sorting only the tied runs
        return i[o], R[o], sc[o], st

    def _native(self, what):
        if self.vo is not None and not self._native_rows:
            raise ValueError(f"{what} needs a backend that keeps the row order itself (set_row_order)")
        return self.b

    def blk_engine(self):
        """The backend itself, for the device eigensolver blocks (their rows cross in the caller's numbering)."""
        return self._native("the device eigensolver blocks")

    def polish_begin(self, V):
        return self._native("polish_begin").polish_begin(V)

    def polish_get(self):
        return self._native("polish_get").polish_get()

    def certify_dense(self, dg):
        return self.b.certify_dense(dg if self.vo is None else self.vo.rows_in(np.asarray(dg, np.float64)))

    def set_batch(self, *args, **kw):
        if self.vo is not None:
            raise ValueError("batch mode packs the graphs as given: create the engine with order='given'")
        return self.b.set_batch(*args, **kw)
