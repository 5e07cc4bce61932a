"""Candidate bound and its certificate at fixed multipliers y (the stages after the search).

Weak duality (est_dual_vcs.py:1-12): for multipliers y >= 0 of valid inequalities, ANY diagonal g and positive
weights s,
    max_{X psd, diag X = 1} <C(y), X> <= sum(g) + max(0, -mu) sum(s),   mu = lambda_min(Diag(g) - C(y), Diag(s)),
so L Q* <= C0 + sum(lam) + sum(g) + max(0, -mu) sum(s).

* `estimate` restates est_dual_vcs.py's production path (--polish rgd --scale vol): Riemannian gradient ascent of
  <C, W W^T> over unit rows with BB steps and Zhang-Hager nonmonotone acceptance (polish_rgd, est_dual_vcs.py:
  276-308), g_i = (C W)_i . W_i (value, 245-246), then LOBPCG on the pencil with a Jacobi preconditioner and the
  residual-hedged value theta_1 - r_1 (eig_round, 327-364). Every C(y) X product runs on the device (the liblbhip
  evaluation at rho = 0, whose half-gradient is C(y) V exactly); the scalar assembly is FP64 on the host. The result
  is an ESTIMATE: an eigensolver can miss the smallest eigenvalue.
* `certify_dense` turns (g, shift) into a proved bound for instances small enough for a dense factorisation: it
  assembles M = Diag(g + t) - C(y) in FP64, factors M = R^T R (Cholesky), and charges every rounding error --
  of the assembly (entrywise bound F) and of the factorisation (R^T R = M + E, |E| <= gamma(n+1) |R^T||R|, Higham,
  Accuracy and Stability of Numerical Algorithms, Thm 10.3) -- as one extra uniform diagonal eps >= ||E||_F + ||F||_F,
  so Diag(g + t + eps) - C(y) is PSD in exact arithmetic. The bound is summed in exact rational arithmetic and
  rounded up (exact_sum / frac_up, verify_cert.py:68-81). It is the dense analogue of the reference's verifier
  (psd_ldl / accept_price, verify_cert.py:92-104, 363-386) without the lifted tree.
"""
from __future__ import annotations

import math
import time
import warnings
from fractions import Fraction

import numpy as np

# LOBPCG runs in chunks of a few iterations and the chunk loop checks convergence itself: scipy's per-chunk
# "not reaching the requested tolerance" notices are expected (filters set once; catch_warnings is not thread-safe)
for _msg in (".*not reaching the requested tolerance.*", ".*Exited at iteration.*", ".*Exited postprocessing.*"):
    warnings.filterwarnings("ignore", message=_msg, category=UserWarning)


# ----------------------------------------------------------------------------------------------- exact arithmetic
def exact_sum(a) -> Fraction:
    """Exact rational sum of float64 values (verify_cert.exact_sum, verify_cert.py:68-77)."""
    a = np.asarray(a, np.float64).ravel()
    a = a[a != 0]
    if len(a) == 0:
        return Fraction(0)
    if not np.isfinite(a).all():
        raise ValueError("non-finite value in an exact sum")
    m, e = np.frexp(a)
    mi = (m * 2.0 ** 53).astype(np.int64)
    e = e.astype(np.int64) - 53
    tot = Fraction(0)
    for ee in np.unique(e):
        tot += Fraction(int(mi[e == ee].astype(object).sum())) * Fraction(2) ** int(ee)
    return tot


def frac_up(x: Fraction) -> float:
    """Smallest float64 >= x (verify_cert.frac_up, verify_cert.py:79-81)."""
    f = float(x)
    return f if Fraction(f) >= x else float(np.nextafter(f, np.inf))


def gamma(k: float) -> float:
    """gamma_k = k u / (1 - k u), u = 2^-53 (unit roundoff of FP64), rounded up."""
    u = 2.0 ** -53
    k = float(math.ceil(k))
    if k * u >= 0.5:
        raise ValueError("operation count too large for the gamma bound")
    return float(np.nextafter(k * u / (1 - k * u), np.inf))


# ----------------------------------------------------------------------------------------------- C(y) on the device
class DeviceLagrangian:
    """C(y) X on the device at the instance's multipliers: upload X as V, evaluate at rho = 0, read G = C(y) X."""

    def __init__(self, inst, r: int, dtype: str = "f64", device: int = 0):
        from .engine import Engine
        self.inst = inst
        self.r = int(r)
        self.eng = Engine(r=self.r, dtype=dtype, device=device)
        self.eng.set_instance(inst)
        self.const = float(inst.C0) + (float(exact_sum(inst.lam)) if inst.T else 0.0)

    def _G(self, V):
        self.eng.set_V(V)
        self.eng.eval(0.0, fetch=False)
        return self.eng.get_G()

    def CX(self, X):
        X = np.asarray(X, np.float64)
        n, b = X.shape
        out = np.empty_like(X)
        for c0 in range(0, b, self.r):
            w = min(self.r, b - c0)
            blk = np.zeros((n, self.r))
            blk[:, :w] = X[:, c0:c0 + w]
            out[:, c0:c0 + w] = self._G(blk)[:, :w]
        return out

    def close(self):
        self.eng.close()


def value(lag, W):
    """(dv, g, CW) at unit-row W: g_i = (C W)_i . W_i, dv = (C0 + sum lam + sum g) / L (est_dual_vcs.value)."""
    CW = lag.CX(W)
    g = np.einsum("ij,ij->i", CW, W)
    return (lag.const + float(g.sum())) / lag.inst.L, g, CW


# ----------------------------------------------------------------------------------------------- estimate
def polish_rgd(lag, V, iters=400, polish_time=600.0, ptol=2e-5, log=print):
    """Riemannian gradient ascent with BB steps and nonmonotone acceptance (polish_rgd, est_dual_vcs.py:276-308)."""
    L = lag.inst.L
    tp = time.time()
    f, g, CW = value(lag, V)
    Cref, Q, alpha = f, 1.0, 2.0
    nr = np.maximum(np.linalg.norm(CW, axis=1), 1e-12)
    Rr = CW - g[:, None] * V
    R = Rr / nr[:, None]
    last_chk = f
    it = 0
    while it < iters and time.time() - tp < polish_time:
        it += 1
        gr = float(np.einsum("ij,ij->", Rr, R)) / L
        for bt in range(12):
            Vn = V + alpha * R
            Vn /= np.linalg.norm(Vn, axis=1, keepdims=True)
            fn, gn, CWn = value(lag, Vn)
            if fn >= Cref + 1e-4 * alpha * gr:
                break
            alpha *= 0.5
        S_ = Vn - V
        nrn = np.maximum(np.linalg.norm(CWn, axis=1), 1e-12)
        Rrn = CWn - gn[:, None] * Vn
        Rn = Rrn / nrn[:, None]
        Y_ = Rn - R
        sy = float(np.einsum("ij,ij->", S_, Y_ * nrn[:, None]))
        ss = float(np.einsum("ij,ij->", S_, S_ * nrn[:, None]))
        yy = float(np.einsum("ij,ij->", Y_, Y_ * nrn[:, None]))
        if sy < 0:
            alpha = (ss / (-sy)) if it % 2 else (-sy / yy)
        else:
            alpha = min(2 * alpha, 50.0)
        alpha = float(np.clip(alpha, 0.05, 50.0))
        V, f, g, CW, R, Rr = Vn, fn, gn, CWn, Rn, Rrn
        Qn = 0.85 * Q + 1
        Cref = (0.85 * Q * Cref + f) / Qn
        Q = Qn
        if it % 10 == 0:
            log(f"rgd it {it} dv {f:.6f} alpha {alpha:.3g} bt {bt}")
            if f - last_chk < ptol * 2:
                break
            last_chk = f
    return V, f, g


def polish_rgd_device(lag, V, iters=400, polish_time=600.0, ptol=2e-5, log=print):
    """polish_rgd with the row work on the device (lb_polish_*: step and normalise, g and its sum, new direction and
    the Barzilai-Borwein sums on an FP64 copy of the iterate); the step rule, the backtracking and the
    Zhang-Hager reference value are the same as polish_rgd's, on the scalars read back."""
    eng = lag.eng
    L = lag.inst.L
    cst = lag.const
    tp = time.time()
    gs, grs = eng.polish_begin(V)
    f = (cst + gs) / L
    Cref, Q, alpha = f, 1.0, 2.0
    last_chk = f
    it = 0
    while it < iters and time.time() - tp < polish_time:
        it += 1
        gr = grs / L
        for bt in range(12):
            fn = (cst + eng.polish_trial(alpha)) / L
            if fn >= Cref + 1e-4 * alpha * gr:
                break
            alpha *= 0.5
        grs, sy, ss, yy = eng.polish_accept()
        if sy < 0:
            alpha = (ss / (-sy)) if it % 2 else (-sy / yy)
        else:
            alpha = min(2 * alpha, 50.0)
        alpha = float(np.clip(alpha, 0.05, 50.0))
        f = fn
        Qn = 0.85 * Q + 1
        Cref = (0.85 * Q * Cref + f) / Qn
        Q = Qn
        if it % 10 == 0:
            log(f"rgd it {it} dv {f:.6f} alpha {alpha:.3g} bt {bt}")
            if f - last_chk < ptol * 2:
                break
            last_chk = f
    V, g = eng.polish_get()
    return V, f, g


def _device_polish_ok(lag, W):
    eng = getattr(lag, "eng", None)
    return eng is not None and hasattr(eng, "polish_begin") and W.shape[1] == getattr(lag, "r", -1)


SCALES = {"ones": lambda d, L, n: np.ones(n), "vol": lambda d, L, n: np.maximum(d, 1.0),
          "sqrtvol": lambda d, L, n: np.sqrt(np.maximum(d, 1.0)) * math.sqrt(L / n)}


def pencil_lmin(lag, g, s, b, maxiter=400, chunk=20, tol=1e-9, atol=1e-5, eig_time=1200.0, X0=None, rng=None,
                log=print, tag=""):
    """LOBPCG for the smallest eigenvalues of (Diag(g) - C, Diag(s)) with a Jacobi preconditioner, in chunks until
    the residual and the change of theta_1 (in bound units, x sum(s)/L) are below atol (eig_round,
    est_dual_vcs.py:325-364). Returns (theta, resid, Xe, trace): Rayleigh quotients and Diag(s)^-1-norm residuals
    of the final block, sorted ascending."""
    import scipy.sparse.linalg as sla
    inst = lag.inst
    n = inst.n
    L = inst.L
    diagM = g + inst.d ** 2 / L                         # C_ii = -d_i^2 / L: the other parts have zero diagonal
    prec = 1.0 / (np.abs(diagM) + 1e-2 * s)

    def Mop(X):
        X = np.asarray(X, np.float64).reshape(n, -1)
        return g[:, None] * X - lag.CX(X)

    Aop = sla.LinearOperator((n, n), matvec=lambda x: Mop(x)[:, 0], matmat=Mop, dtype=float)
    Bop = sla.LinearOperator((n, n), matvec=lambda x: s * x.ravel(), matmat=lambda X: s[:, None] * X, dtype=float)
    Pop = sla.LinearOperator((n, n), matvec=lambda x: prec * x.ravel(), matmat=lambda X: prec[:, None] * X,
                             dtype=float)
    rng = np.random.default_rng(0) if rng is None else rng
    r0 = 0 if X0 is None else min(X0.shape[1], b)
    Xe = rng.standard_normal((n, b)) if r0 == 0 else np.c_[X0[:, :r0], rng.standard_normal((n, b - r0))]
    sums = float(s.sum())
    te = time.time()
    it_tot = 0
    th_prev = np.inf
    trace = []
    while True:
        ev, Xe = sla.lobpcg(Aop, Xe, B=Bop, M=Pop, largest=False, maxiter=chunk, tol=tol)
        it_tot += chunk
        Xe = Xe[:, np.argsort(ev)]
        MX = Mop(Xe)
        xs = np.einsum("ij,ij->j", Xe, s[:, None] * Xe)
        th = np.einsum("ij,ij->j", Xe, MX) / xs
        Rres = MX - th[None, :] * (s[:, None] * Xe)
        rn = np.sqrt(np.einsum("ij,ij->j", Rres, Rres / s[:, None]) / xs)
        trace.append((it_tot, time.time() - te, float(th[0]), float(rn[0])))
        log(f"  [s={tag}] lobpcg its {it_tot} theta1 {th[0]:.6e} r1 {rn[0]:.2e}  (shift-units: theta "
            f"{max(0, -th[0]) * sums / L:.6f} r {rn[0] * sums / L:.2e})")
        if rn[0] * sums / L < atol and abs(th[0] - th_prev) * sums / L < atol:
            break
        if it_tot >= maxiter or time.time() - te > eig_time:
            break
        th_prev = th[0]
    return th, rn, Xe, trace


def pencil_lmin_device(lag, g, s, b, maxiter=400, chunk=20, tol=1e-9, atol=1e-5, eig_time=1200.0, X0=None, rng=None,
                       log=print, tag=""):
    """pencil_lmin with the n-sized blocks on the device (lb_blk_*): block LOBPCG for the smallest eigenvalues of
    (Diag(g) - C, Diag(s)) with the Jacobi preconditioner of pencil_lmin. The block is as wide as the context (r
    columns: the operator costs the same for any width up to r); its Gram matrices and combinations run on the device
    and only the (3r x 3r) Rayleigh-Ritz problems on the host (the scheme of Knyazev's LOBPCG, as scipy's lobpcg
    used by est_dual_vcs.py:325-364). Convergence is tested every `chunk` iterations with pencil_lmin's rule. Returns
    (theta, resid, Xe, trace) like pencil_lmin, with r Ritz pairs (b is only the minimum width)."""
    import scipy.linalg as sl
    inst = lag.inst
    n, L, r = inst.n, inst.L, lag.r
    e = lag.eng.blk_engine()
    sums = float(s.sum())
    diagM = g + inst.d ** 2 / L
    prec = 1.0 / (np.abs(diagM) + 1e-2 * s)
    X_, AX_, W_, AW_, P_, AP_, Xn_, AXn_, Pn_, APn_, R_ = range(11)
    DG, DS, DP, DI = 0, 1, 2, 3                               # diagonals: g, s, preconditioner, 1 / s
    e.blk_alloc(11, 4)
    for k, w in ((DG, g), (DS, s), (DP, prec), (DI, 1.0 / s)):
        e.blk_set_diag(k, np.asarray(w, np.float64))
    rng = np.random.default_rng(0) if rng is None else rng
    r0 = 0 if X0 is None else min(X0.shape[1], r)
    Xh = rng.standard_normal((n, r)) if r0 == 0 else np.c_[X0[:, :r0], rng.standard_normal((n, r - r0))]
    e.blk_set(X_, Xh)
    ids = {"X": X_, "AX": AX_, "P": P_, "AP": AP_, "Xn": Xn_, "AXn": AXn_, "Pn": Pn_, "APn": APn_}

    def sym(A):
        return 0.5 * (A + A.T)

    def rr(GA, GB):
        """Rayleigh-Ritz on span(S) with Gram matrices GA = S^T M S, GB = S^T Diag(s) S: directions in which GB is
        numerically singular (converged residuals, dependent search directions) are dropped, as soft locking does."""
        lam, U = np.linalg.eigh(sym(GB))
        keep = lam > 1e-11 * lam[-1]
        T = U[:, keep] / np.sqrt(lam[keep])
        vals, Y = np.linalg.eigh(sym(T.T @ sym(GA) @ T))
        return vals[:r], (T @ Y)[:, :r]

    e.blk_op(AX_, X_, DG)
    th, C = rr(e.blk_gram([X_], [AX_]), e.blk_gram([X_], [X_], DS))
    e.blk_comb(Xn_, [X_], C)
    e.blk_comb(AXn_, [AX_], C)
    ids["X"], ids["Xn"], ids["AX"], ids["AXn"] = Xn_, X_, AXn_, AX_
    have_p = False
    te = time.time()
    th_prev = np.inf
    trace = []
    it = 0

    def residuals():
        e.blk_elem(2, R_, ids["X"], ids["AX"], DS, th)
        return np.sqrt(np.maximum(np.diag(e.blk_gram([R_], [R_], DI)), 0.0))

    while True:
        it += 1
        X, AX, P, AP = ids["X"], ids["AX"], ids["P"], ids["AP"]
        e.blk_elem(2, R_, X, AX, DS, th)                      # R = AX - Diag(s) X Theta
        rres = np.sqrt(np.maximum(np.diag(e.blk_gram([R_], [R_], DI)), 0.0))
        e.blk_elem(1, W_, R_, -1, DP)                         # W = T R
        # unit Diag(s)-norm columns; the columns of converged Ritz pairs (residual below tol) are left out (soft
        # locking, as scipy's lobpcg), so their rounding noise does not enter the search space
        dw = np.diag(e.blk_gram([W_], [W_], DS))
        sc = np.where((dw > 1e-300) & (rres > tol), 1.0 / np.sqrt(np.maximum(dw, 1e-300)), 0.0)
        e.blk_comb(R_, [W_], np.diag(sc))
        e.blk_op(AW_, R_, DG)
        W = R_
        S, AS = ([X, W, P], [AX, AW_, AP]) if have_p else ([X, W], [AX, AW_])
        th, C = rr(e.blk_gram(S, AS), e.blk_gram(S, S, DS))
        e.blk_comb(ids["Xn"], S, C)
        e.blk_comb(ids["AXn"], AS, C)
        e.blk_comb(ids["Pn"], S[1:], C[r:])
        e.blk_comb(ids["APn"], AS[1:], C[r:])
        ids["X"], ids["Xn"] = ids["Xn"], ids["X"]
        ids["AX"], ids["AXn"] = ids["AXn"], ids["AX"]
        ids["P"], ids["Pn"] = ids["Pn"], ids["P"]
        ids["AP"], ids["APn"] = ids["APn"], ids["AP"]
        have_p = True
        if it % chunk == 0 or it >= maxiter:
            rn = residuals()
            trace.append((it, time.time() - te, float(th[0]), float(rn[0])))
            log(f"  [s={tag}] device lobpcg its {it} theta1 {th[0]:.6e} r1 {rn[0]:.2e}  (shift-units: theta "
                f"{max(0, -th[0]) * sums / L:.6f} r {rn[0] * sums / L:.2e})")
            if rn[0] * sums / L < atol and abs(th[0] - th_prev) * sums / L < atol:
                break
            if it >= maxiter or time.time() - te > eig_time:
                break
            th_prev = th[0]
    rn = residuals()
    Xe = e.blk_get(ids["X"])
    o = np.argsort(th)
    return th[o], rn[o], Xe[:, o], trace


def _device_eig_ok(lag, n):
    """An FP64 device operator with the block entry points, on a graph large enough for an r-wide block."""
    eng = getattr(lag, "eng", None)
    return eng is not None and hasattr(eng, "blk_alloc") and getattr(eng, "dtype", None) == "f64" and n >= 8 * lag.r


def estimate(inst, V, lag=None, polish="rgd", rgd_iters=400, polish_time=600.0, ptol=2e-5, scale="vol",
             eig_extra=16, eig_maxiter=400, eig_tol=1e-9, eig_chunk=20, eig_atol=1e-5, eig_time=1200.0, escape=2,
             escape_min=0.002, escape_p=4, escape_ratio=0.3, seed=0, no_w_init=False, device=0, dtype="f64",
             eig_block=None, device_polish=True, eig_device=True, log=print):
    """Candidate bound at the instance's multipliers (levels' sig, lam, VCS sig): polish, then per scale s the
    pencil eigenvalue, ub = dv + max(0, -(theta_1 - r_1)) sum(s) / L, then up to `escape` rank-escape rounds
    (append the isolated bottom eigenvectors to W, re-polish, re-estimate; the best round is kept) -- est_dual_vcs.py
    :316-389. Returns the reference's result record (dv0, dv, s_<scale> records, escape_rounds, ub, ...) plus the
    best round's W, g and s (`W`, `g`, `s` keys, for certification). `lag` (anything with .inst/.const/.CX) defaults
    to the device operator. `eig_block` overrides the LOBPCG block width k + eig_extra (a narrower block costs
    less host work per iteration). With the device operator and W as wide as its rank, the polish runs its row work
    on the device (polish_rgd_device; device_polish=False keeps it on the host). With an FP64 device operator the
    pencil eigenproblem runs as a device-block LOBPCG as wide as the operator (pencil_lmin_device; eig_device=False
    keeps scipy's LOBPCG on the host). ub is an estimate, not a proof."""
    own = lag is None
    if own:
        lag = DeviceLagrangian(inst, r=V.shape[1], dtype=dtype, device=device)
    t0 = time.time()
    rng = np.random.default_rng(seed)
    n, L = inst.n, inst.L
    scales = scale.split(",")

    def do_polish(W):
        for meth in polish.split("+"):
            if meth == "rgd":
                pol = polish_rgd_device if device_polish and _device_polish_ok(lag, W) else polish_rgd
                W = pol(lag, W, iters=rgd_iters, polish_time=polish_time, ptol=ptol, log=log)[0]
            elif meth not in ("none", ""):
                raise ValueError(f"polish method {meth!r} is not supported (rgd | none)")
        return W

    def eig_round(W, dv, g):
        recs, vecs = {}, {}
        for sname in scales:
            s = SCALES[sname](inst.d, L, n)
            b = W.shape[1] + eig_extra if eig_block is None else int(eig_block)
            X0 = None
            if not no_w_init:
                Uw, Sw, _ = np.linalg.svd(W, full_matrices=False)
                X0 = Uw[:, :int((Sw > 1e-6 * Sw[0]).sum())]
            te = time.time()
            eig = pencil_lmin_device if eig_device and _device_eig_ok(lag, n) else pencil_lmin
            th, rn, Xe, trace = eig(lag, g, s, b, maxiter=eig_maxiter, chunk=eig_chunk, tol=eig_tol, atol=eig_atol,
                                    eig_time=eig_time, X0=X0, rng=rng, log=log, tag=sname)
            sums = float(s.sum())
            lmin_est = float(th[0] - rn[0])
            shift = max(0.0, -lmin_est) * sums / L
            recs[sname] = dict(theta=th[:6].tolist(), resid=rn[:6].tolist(), lmin_est=lmin_est, shift=shift,
                               ub=dv + shift, sum_s=sums, eig_time=time.time() - te, eig_trace=trace,
                               ub_ritz=dv + max(0.0, -float(th[0])) * sums / L, dv=dv, k=int(W.shape[1]))
            log(f"[s={sname}] theta {th[:4]} resid {rn[:4]} -> lmin_est {lmin_est:.3e}  shift {shift:.6f}  "
                f"UB {dv + shift:.6f}")
            vecs[sname] = (th, Xe, s)
        return recs, vecs

    try:
        W = V / np.maximum(np.linalg.norm(V, axis=1, keepdims=True), 1e-300)
        dv0 = value(lag, W)[0]
        tp = time.time()
        W = do_polish(W)
        tpol = time.time() - tp
        dv, g, _ = value(lag, W)
        log(f"POLISHED dv {dv:.6f} (start {dv0:.6f}) in {tpol:.0f}s")
        res = dict(n=n, L=L, k=int(W.shape[1]), dv0=dv0, dv=dv, polish_time=tpol)
        best, rounds = np.inf, []
        for rnd in range(escape + 1):
            recs, vecs = eig_round(W, dv, g)
            ub_r = min(r["ub"] for r in recs.values())
            rounds.append(dict(round=rnd, dv=dv, ub=ub_r, k=int(W.shape[1])))
            log(f"[round {rnd}] k {W.shape[1]} dv {dv:.6f} UB {ub_r:.6f}")
            if ub_r < best:
                best = ub_r
                res["dv"] = dv
                for sname, rec in recs.items():
                    res[f"s_{sname}"] = rec
                bs = min(recs, key=lambda q: recs[q]["ub"])
                res.update(W=W, g=g, s=vecs[bs][2], lmin=recs[bs]["lmin_est"], best_scale=bs)
            if rnd == escape:
                break
            th, Xe, _ = vecs[scales[0]]
            if recs[scales[0]]["shift"] < escape_min:
                break
            sel = [j for j in range(len(th)) if th[j] < 0 and th[j] <= escape_ratio * th[0]][:escape_p]
            if not sel:
                break
            U = Xe[:, sel]
            U = 0.5 * U / np.abs(U).max(axis=0, keepdims=True)
            W = np.c_[W, U]
            W /= np.linalg.norm(W, axis=1, keepdims=True)
            W = do_polish(W)
            dv, g, _ = value(lag, W)
            log(f"ESCAPE round {rnd + 1}: appended {len(sel)} eigenvectors, re-polished dv {dv:.6f} k {W.shape[1]}")
    finally:
        if own:
            lag.close()
    res["escape_rounds"] = rounds
    res["ub"] = best
    res["shift"] = best - res["dv"]
    res["time"] = time.time() - t0
    return res


# ----------------------------------------------------------------------------------------------- certificate
def dense_C(inst):
    """C(y) as a dense symmetric FP64 matrix from the instance's multipliers (upper triangle mirrored, so it is
    exactly symmetric), and F >= |fl(C) - C| entrywise: each entry is a sum of at most h terms of at most 3
    roundings each, so |error| <= gamma(h + 3) sum |terms|, with h the largest number of terms of any entry; the
    computed sum of magnitudes is itself within a factor (1 + gamma(h + 3)) of the exact one, hence
    F = gamma(2h + 8) fl(sum |terms|)."""
    n = inst.n
    d = inst.d
    Cm = inst.dense_A()
    absC = np.abs(Cm)
    dd = np.outer(d, d)
    dd /= inst.L
    Cm -= dd
    absC += dd
    dd = np.outer(d, d)
    h = 2
    for lv in inst.levels:
        S = lv.sig[np.ix_(lv.lab, lv.lab)]
        S *= dd
        np.fill_diagonal(S, 0.0)
        Cm += S
        absC += np.abs(S)
        h += 1
    for vl in inst.vcs:
        sig = np.zeros((n, vl.P))
        rows = np.repeat(np.arange(n), np.diff(vl.ptr))
        sig[rows, vl.col] = vl.sig
        S = sig[:, vl.lab]
        S += S.T.copy()
        S *= 0.5
        S *= dd
        np.fill_diagonal(S, 0.0)
        Cm += S
        absC += np.abs(S)
        h += 2
    if inst.T:
        a, b, c = (inst.TI[:, q].astype(np.int64) for q in range(3))
        hl = 0.5 * inst.lam
        ii = np.concatenate([a, b, b, c, a, c])
        jj = np.concatenate([b, a, c, b, c, a])
        vals = np.concatenate([-hl, -hl, -hl, -hl, hl, hl])
        key = ii * n + jj
        Cm += np.bincount(key, weights=vals, minlength=n * n).reshape(n, n)
        absC += np.bincount(key, weights=np.abs(vals), minlength=n * n).reshape(n, n)
        h += int(np.bincount(key).max())
    iu = np.triu_indices(n, 1)
    diag = np.diag(Cm).copy()
    Cs = np.zeros((n, n))
    Cs[iu] = Cm[iu]
    Cs += Cs.T.copy()
    Cs[np.diag_indices(n)] = diag
    absC *= gamma(2 * h + 8)                        # covers the rounding of absC itself as well
    return Cs, absC


def check_multipliers(inst):
    """Weak duality needs y >= 0 on every inequality (pair levels, VCS entries, triangles); the dense assembly reads
    each level's sig as the symmetric matrix it is."""
    for lv in inst.levels:
        if (lv.sig < 0).any():
            raise ValueError("negative pair-level multiplier")
        if not np.array_equal(lv.sig, lv.sig.T):
            raise ValueError("pair-level multipliers must be symmetric (h_QR = h_RQ)")
    for vl in inst.vcs:
        if (vl.sig < 0).any():
            raise ValueError("negative VCS multiplier")
    if inst.T and (inst.lam < 0).any():
        raise ValueError("negative triangle multiplier")


def certify_dense(inst, g, t=None):
    """Prove L Q* <= C0 + sum(lam) + sum(g) + sum(t) + n eps, with M = Diag(g + t) - C(y) factored by Cholesky.
    `t` (default 0) is an extra diagonal, e.g. max(0, -mu) s from the estimate plus a margin. Returns dict(passed,
    UB, eps, price) -- price = n eps / L, the rounding allowance in bound units."""
    check_multipliers(inst)
    n = inst.n
    t = np.zeros(n) if t is None else np.asarray(t, np.float64)
    Cm, F = dense_C(inst)
    M = -Cm
    dg = g + t                                          # rounded once; the exact diagonal used below is fl(g + t)
    M[np.diag_indices(n)] = dg - np.diag(Cm)
    F[np.diag_indices(n)] += gamma(1) * (np.abs(dg) + np.abs(np.diag(Cm)))
    try:
        Rm = np.linalg.cholesky(M)
    except np.linalg.LinAlgError:
        return dict(passed=False, reason="Cholesky failed: M is not positive definite at this shift")
    nF = float(np.sqrt(np.sum(F * F))) * (1 + gamma(2 * n * n + 8))
    nR = float(np.sum(Rm * Rm)) * (1 + gamma(n * n + 8))
    eps = (gamma(n + 1) * nR + nF) * 1.01
    # lambda_min(Diag(dg) - C_exact) >= -eps, hence Diag(dg + eps) - C(y) is PSD
    total = Fraction(inst.C0) + (exact_sum(inst.lam) if inst.T else 0) + exact_sum(dg) + Fraction(n) * Fraction(eps)
    UB = frac_up(total / Fraction(inst.L))
    return dict(passed=True, UB=UB, eps=eps, price=n * eps / inst.L)


def max_terms(inst):
    """Largest number of terms of any off-diagonal entry of C(y): A, the rank-1 part, one per pair level, two per
    VCS level, and the most triangles sharing one vertex pair."""
    h = 2 + len(inst.levels) + 2 * len(inst.vcs)
    if inst.T:
        n = inst.n
        a, b, c = (inst.TI[:, q].astype(np.int64) for q in range(3))
        ii = np.concatenate([np.minimum(a, b), np.minimum(b, c), np.minimum(a, c)])
        jj = np.concatenate([np.maximum(a, b), np.maximum(b, c), np.maximum(a, c)])
        h += int(np.unique(ii * n + jj, return_counts=True)[1].max())
    return h


def certify_dense_device(inst, g, t=None, device=0, eng=None):
    """certify_dense with the assembly and the Cholesky factorisation on the device (lb_certify_dense, rocSOLVER
    dpotrf in FP64); the rounding price and the exact rational total are computed here as in certify_dense."""
    from .engine import Engine
    check_multipliers(inst)
    n = inst.n
    t = np.zeros(n) if t is None else np.asarray(t, np.float64)
    dg = np.ascontiguousarray(g + t)                # the exact diagonal used below is fl(g + t)
    own = eng is None
    if own:
        eng = Engine(r=8, dtype="f64", device=device)
        eng.set_instance(inst)
    try:
        out = eng.certify_dense(dg)
    finally:
        if own:
            eng.close()
    info, sr, sab = int(out[0]), float(out[1]), float(out[2])
    if info != 0:
        return dict(passed=False, reason=f"Cholesky failed at leading minor {info}: M is not positive definite")
    gh = gamma(2 * max_terms(inst) + 8)
    diagC = inst.d ** 2 / inst.L
    Fd = gh * diagC + gamma(1) * (np.abs(dg) + diagC)
    nF = math.sqrt(gh * gh * sab * (1 + gamma(2 * n * n + 8)) + float(np.sum(Fd * Fd)) * (1 + gamma(n + 8)))
    nF *= 1 + gamma(8)
    nR = sr * (1 + gamma(n * n + 8))
    eps = (gamma(n + 1) * nR + nF) * 1.01
    total = Fraction(inst.C0) + (exact_sum(inst.lam) if inst.T else 0) + exact_sum(dg) + Fraction(n) * Fraction(eps)
    UB = frac_up(total / Fraction(inst.L))
    return dict(passed=True, UB=UB, eps=eps, price=n * eps / inst.L, device=True)


def certify_estimate(inst, est, margin=1e-7, device=None, tries=8):
    """certify_dense at the estimate's best round (g, s, lmin) with t = (factor max(0, -lmin) + margin) s, starting at
    factor 1 and doubling factor and margin while the factorisation fails (the eigensolver may have stopped above
    the smallest eigenvalue; the reference likewise starts its certificate at a multiple of the estimated shift,
    api.py:216-218). On the device when `device` is given. The result records the factor used."""
    shift = max(0.0, -est["lmin"])
    factor, mg = 1.0, margin
    res = None
    for _ in range(tries):
        t = (factor * shift + mg) * est["s"]
        if device is not None:
            res = certify_dense_device(inst, est["g"], t, device=device)
        else:
            res = certify_dense(inst, est["g"], t)
        if res["passed"]:
            res["factor"] = factor
            return res
        factor *= 2.0
        mg = max(mg * 10.0, 1e-6 * max(shift, 1e-3))
    return res


# ----------------------------------------------------------------------------------------------- checkpoints
def instance_from_checkpoint(inst, z):
    """Attach the multipliers of a solver checkpoint (alm_vcs.py:1576-1589 keys) to `inst` (graph only)."""
    from . import instances as I
    inst.levels = []
    for li in range(int(z["nlev"])):
        inst.levels.append(I.pair_level(z[f"lab{li}"], inst.d, np.asarray(z[f"sig{li}"], np.float64)))
    inst.vcs = []
    for m in range(int(z["nvcs"]) if "nvcs" in z else 0):
        inst.vcs.append(I.vcs_level(z[f"vcslab{m}"], inst.d, z[f"vcsptr{m}"], z[f"vcscol{m}"],
                                    np.asarray(z[f"vcssig{m}"], np.float64)))
    inst.TI = np.asarray(z["TI"], np.int32).reshape(-1, 3)
    inst.lam = np.asarray(z["lam"], np.float64)
    return inst
