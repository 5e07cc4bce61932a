"""The plain FP64 GPU form of one f+grad evaluation: the same mathematics as the HIP kernels with straightforward
torch operations (FP64 CSR SpMM, gathers, index_add_). It is the comparison baseline timed beside the HIP path
(`vs FP64` in the metric) and is not used by the search. Mirrors `augmented` + `CV` (alm_vcs.py:1496-1533,
1299-1312)."""
from __future__ import annotations

import numpy as np
import torch


class PlainFP64:
    def __init__(self, inst, device="cuda"):
        dt = torch.float64
        self.dev = device
        n = inst.n
        self.n = n
        data = np.ones(inst.nnz) if inst.w is None else inst.w
        self.A = torch.sparse_csr_tensor(torch.as_tensor(inst.rowptr), torch.as_tensor(inst.col.astype(np.int64)),
                                         torch.as_tensor(data, dtype=dt), (n, n), device=device)
        self.d = torch.as_tensor(inst.d, dtype=dt, device=device)
        self.L = float(inst.L)
        self.C0 = float(inst.C0)
        self.pw = float(inst.pairlvl_weight)
        self.wv = float(inst.wv)
        self.levels = [dict(lab=torch.as_tensor(l.lab.astype(np.int64), device=device), P=l.P,
                            sig=torch.as_tensor(l.sig, dtype=dt, device=device),
                            rq=torch.as_tensor(l.rq, dtype=dt, device=device),
                            dsq=torch.as_tensor(l.dsq, dtype=dt, device=device)) for l in inst.levels]
        self.vcs = []
        for v in inst.vcs:
            rows = np.repeat(np.arange(n), np.diff(v.ptr))
            self.vcs.append(dict(lab=torch.as_tensor(v.lab.astype(np.int64), device=device), P=v.P,
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

    def eval(self, V, rho, rt=None):
        """Returns (f as a 0-d tensor, G)."""
        rt = rho if rt is None else rt
        d = self.d
        AV = torch.sparse.mm(self.A, V)
        z = d @ V
        G = AV - torch.outer(d, z) / self.L
        quad = (AV * V).sum()
        f = quad - (z @ z) / self.L + self.C0
        f0 = torch.zeros((), dtype=V.dtype, device=V.device)
        for l in self.levels:
            Y = torch.zeros((l["P"], V.shape[1]), dtype=V.dtype, device=V.device).index_add_(0, l["lab"], d[:, None] * V)
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
            h = 0.5 * pt[:, None]
            G = G.index_add(0, self.a, h * (Vc - Vb)).index_add(0, self.b, -h * (Va + Vc)).index_add(0, self.c, h * (Va - Vb))
        for v in self.vcs:
            Y = torch.zeros((v["P"], V.shape[1]), dtype=V.dtype, device=V.device).index_add_(0, v["lab"], d[:, None] * V)
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
            Z = torch.zeros_like(Y).index_add_(0, col, (pv * di)[:, None] * Vr)
            SY = torch.zeros_like(V).index_add_(0, rows, pv[:, None] * Y[col])
            ownv = torch.zeros_like(d).index_add_(0, rows, torch.where(own, pv, torch.zeros_like(pv)))
            G = G + 0.5 * d[:, None] * (SY + Z[v["lab"]]) - (ownv * d * d)[:, None] * V
        if rho <= 0:
            f = self.C0 + self.lam.sum() + (G * V).sum() if self.T else self.C0 + (G * V).sum() + f0
        return f, G
