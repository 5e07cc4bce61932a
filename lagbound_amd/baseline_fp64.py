"""The plain FP64 GPU form of one f+grad evaluation: the same mathematics as the HIP kernels with straightforward
torch operations (FP64 CSR SpMM, gathers, sparse products, index_add_ where few rows collide). It is the comparison
baseline timed beside the HIP path (`vs FP64` in the metric) and is not used by the search. Mirrors `augmented` +
`CV` (alm_vcs.py:1496-1533, 1299-1312). Sums into a few cluster rows (Y per label set, Z per VCS level) are sparse
products with the cluster-membership matrices (measured faster than index_add_ onto 64 rows, whose atomics all
collide, and than segment sums). Every sparse matrix (A, the memberships, T(pt) with its six entries per triangle,
the VCS products) is built once with 32-bit indices; a call writes only the values that depend on V."""
from __future__ import annotations

import numpy as np
import torch


def _csr(crow, col, val, shape, device):
    """FP64 CSR with 32-bit indices (64-bit only when the entries need them)."""
    it = torch.int32 if int(np.asarray(crow)[-1]) < 2 ** 31 else torch.int64
    return torch.sparse_csr_tensor(torch.as_tensor(np.asarray(crow), dtype=it), torch.as_tensor(np.asarray(col), dtype=it),
                                   torch.as_tensor(val, dtype=torch.float64), shape, device=device)


def _members(lab, P, device):
    """(P x n) CSR with a one in row Q for every member of cluster Q: M @ X sums the rows of X per cluster."""
    lab = np.asarray(lab, np.int64)
    crow = np.zeros(P + 1, np.int64)
    np.cumsum(np.bincount(lab, minlength=P), out=crow[1:])
    return _csr(crow, np.argsort(lab, kind="stable"), np.ones(len(lab)), (P, len(lab)), device)


class PlainFP64:
    def __init__(self, inst, device="cuda"):
        dt = torch.float64
        self.dev = device
        n = inst.n
        self.n = n
        data = np.ones(inst.nnz) if inst.w is None else inst.w
        self.A = _csr(inst.rowptr, inst.col, data, (n, n), device)
        self.d = torch.as_tensor(inst.d, dtype=dt, device=device)
        self.L = float(inst.L)
        self.C0 = float(inst.C0)
        self.pw = float(inst.pairlvl_weight)
        self.wv = float(inst.wv)
        self.levels = [dict(lab=torch.as_tensor(l.lab.astype(np.int64), device=device), P=l.P,
                            mem=_members(l.lab, l.P, device),
                            sig=torch.as_tensor(l.sig, dtype=dt, device=device),
                            rq=torch.as_tensor(l.rq, dtype=dt, device=device),
                            dsq=torch.as_tensor(l.dsq, dtype=dt, device=device)) for l in inst.levels]
        self.vcs = []
        for v in inst.vcs:
            rows = np.repeat(np.arange(n), np.diff(v.ptr))
            col = np.asarray(v.col, np.int64)
            zperm = np.argsort(col, kind="stable")              # entries grouped by cluster: Z = (P x n) @ V
            zcrow = np.zeros(v.P + 1, np.int64)
            np.cumsum(np.bincount(col, minlength=v.P), out=zcrow[1:])
            self.vcs.append(dict(Zm=_csr(zcrow, rows[zperm], np.zeros(len(col)), (v.P, n), device),
                                 Sm=_csr(v.ptr, v.col, np.zeros(len(col)), (n, v.P), device),
                                 lab=torch.as_tensor(v.lab.astype(np.int64), device=device), P=v.P,
                                 mem=_members(v.lab, v.P, device),
                                 zperm=torch.as_tensor(zperm, device=device),
                                 rq=torch.as_tensor(v.rq, dtype=dt, device=device),
                                 rows=torch.as_tensor(rows, device=device),
                                 col=torch.as_tensor(v.col.astype(np.int64), device=device),
                                 sig=torch.as_tensor(v.sig, dtype=dt, device=device)))
        TI = inst.TI.astype(np.int64)
        self.T = inst.T
        self.a = torch.as_tensor(TI[:, 0], device=device)
        self.b = torch.as_tensor(TI[:, 1], device=device)
        self.c = torch.as_tensor(TI[:, 2], device=device)
        self.lam = torch.as_tensor(inst.lam, dtype=dt, device=device)
        # T(pt) V as one sparse product: per triangle the six entries (a,c,+h) (a,b,-h) (b,a,-h) (b,c,-h) (c,a,+h)
        # (c,b,-h), laid out once by row (index_add_ onto the rows would make every triangle at a hub collide)
        a_, b_, c_ = TI[:, 0], TI[:, 1], TI[:, 2]
        r6 = np.concatenate([a_, a_, b_, b_, c_, c_])
        c6 = np.concatenate([c_, b_, a_, c_, a_, b_])
        tperm = np.argsort(r6, kind="stable")
        tcrow = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(r6, minlength=n), out=tcrow[1:])
        self.tperm = torch.as_tensor(tperm, device=device)
        self.Tm = _csr(tcrow, c6[tperm], np.zeros(6 * len(TI)), (n, n), device)

    def eval(self, V, rho, rt=None):
        """Returns (f as a 0-d tensor, G)."""
        rt = rho if rt is None else rt
        d = self.d
        AV = torch.sparse.mm(self.A, V)
        z = (d[:, None] * V).sum(0)                 # a column reduction (as d @ V it runs as a one-tile GEMM)
        G = AV - torch.outer(d, z) / self.L
        quad = (AV * V).sum()
        f = quad - (z @ z) / self.L + self.C0
        f0 = torch.zeros((), dtype=V.dtype, device=V.device)
        for l in self.levels:
            Y = torch.sparse.mm(l["mem"], d[:, None] * V)
            if rho > 0:
                H = Y @ Y.T - torch.diag(l["dsq"])
                rr = rho * torch.outer(l["rq"], l["rq"]) * self.pw
                pp = torch.clamp(l["sig"] - rr * H, min=0.0)
                f = f - ((pp * pp - l["sig"] * l["sig"]) / (2 * rr)).sum()
            else:
                pp = l["sig"]
            G = G + d[:, None] * (pp @ Y)[l["lab"]] - (torch.diagonal(pp)[l["lab"]] * d * d)[:, None] * V
        if self.T:
            Va, Vb, Vc = V[self.a], V[self.b], V[self.c]
            if rho > 0:
                tau = (Va * Vb).sum(1) + (Vb * Vc).sum(1) - (Va * Vc).sum(1)
                pt = torch.clamp(self.lam - rt * (1.0 - tau), min=0.0)
                f = f - ((pt * pt - self.lam * self.lam) / (2 * rt)).sum()
            else:
                pt = self.lam
            h = 0.5 * pt
            self.Tm.values().copy_(torch.cat([h, -h, -h, -h, h, -h])[self.tperm])
            G = G + torch.sparse.mm(self.Tm, V)
        for v in self.vcs:
            Y = torch.sparse.mm(v["mem"], d[:, None] * V)
            rows, col = v["rows"], v["col"]
            Vr = V[rows]
            s = (Vr * Y[col]).sum(1)
            di = d[rows]
            own = col == v["lab"][rows]
            g = torch.where(own, di * (s - di), di * s)
            if rho > 0:
                rr = rho * (self.wv * v["rq"][col] / (di * di))
                pv = torch.clamp(v["sig"] - rr * g, min=0.0)
                f = f - ((pv * pv - v["sig"] * v["sig"]) / (2 * rr)).sum()
            else:
                pv = v["sig"]
            v["Zm"].values().copy_((pv * di)[v["zperm"]])
            Z = torch.sparse.mm(v["Zm"], V)
            v["Sm"].values().copy_(pv)
            SY = torch.sparse.mm(v["Sm"], Y)
            ownv = torch.zeros_like(d).index_add_(0, rows, torch.where(own, pv, torch.zeros_like(pv)))
            G = G + 0.5 * d[:, None] * (SY + Z[v["lab"]]) - (ownv * d * d)[:, None] * V
        if rho <= 0:
            f = self.C0 + self.lam.sum() + (G * V).sum() if self.T else self.C0 + (G * V).sum() + f0
        return f, G
