"""Candidate bound and its certificate at fixed multipliers y (the stages after the search).

Weak duality (the reference's estimate stage, est_dual_vcs.py:1-12): for multipliers y >= 0 of valid inequalities,
ANY diagonal g and positive weights s,
    max_{X psd, diag X = 1} <C(y), X> <= sum(g) + max(0, -mu) sum(s),   mu = lambda_min(Diag(g) - C(y), Diag(s)),
so L Q* <= C0 + sum(lam) + sum(g) + max(0, -mu) sum(s).

* `estimate` follows the behaviour of est_dual_vcs.py's production path (--polish rgd --scale vol, lines 245-389):
  a preconditioned Riemannian ascent of <C, W W^T> over unit rows (`ascend`), g_i = (C W)_i . W_i, then block
  LOBPCG on the pencil (`pencil_min`) with the residual-hedged value theta_1 - r_1, and rank-escape rounds. The
  ascent's row work and the LOBPCG blocks run in the device context (lb_polish_*, lb_blk_*) or on the host, through
  one implementation of each method; the scalar assembly is FP64. The result is an ESTIMATE: an eigensolver can
  miss the smallest eigenvalue.
* `certify_dense` turns (g, shift) into a proved bound for instances small enough for a dense factorisation: it
  assembles M = Diag(g + t) - C(y) in FP64, factors M = R^T R (Cholesky), and charges every rounding error --
  of the assembly (entrywise bound F) and of the factorisation (R^T R = M + E, |E| <= gamma(n+1) |R^T||R|, Higham,
  Accuracy and Stability of Numerical Algorithms, Thm 10.3) -- as one extra uniform diagonal eps >= ||E||_F + ||F||_F,
  so Diag(g + t + eps) - C(y) is PSD in exact arithmetic. The bound is summed in exact rational arithmetic and
  rounded up (exact_sum / frac_up, as verify_cert.py:68-81). Larger instances go through the lifted sparse
  certificate of lagbound_amd.sparse_cert.
"""
from __future__ import annotations

import math
import time
from fractions import Fraction

import numpy as np


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
# The estimate has two numerical parts, each written once and run on either of two storage backends:
#   * ascent: maximise f(W) = <C(y), W W^T> over unit rows W (the oblique manifold) by preconditioned Riemannian
#     gradient steps; the rows live on the host (HostRows, any operator with .CX) or in the device context
#     (DeviceRows, lb_polish_*);
#   * pencil: the smallest eigenvalues of (Diag(g) - C(y), Diag(s)) by block LOBPCG; the blocks live on the host
#     (HostBlocks) or in the device context (DeviceBlocks, lb_blk_*), and only the small Rayleigh-Ritz problems are
#     solved on the host in either case.
# Behaviour of the reference stage (est_dual_vcs.py:245-389): the same objective, preconditioner, step rule,
# eigenproblem, convergence test and rank-escape rule; its result record keys are kept because its consumers read
# them.

ZH_ETA = 0.85          # Zhang-Hager averaging weight of the nonmonotone reference value
ARMIJO = 1e-4          # sufficient-increase constant
HALVINGS = 12          # step halvings per iteration at most
STEP_MIN, STEP_MAX = 0.05, 50.0


class HostRows:
    """Ascent iterate on the host. State per iterate: W (unit rows), CW = C(y) W, g_i = (C W)_i . W_i, Jacobi row
    weights nu_i = |(C W)_i| and the preconditioned ascent direction D = (CW - g W) / nu (the tangent gradient
    divided by the row weight)."""

    def __init__(self, lag):
        self.lag = lag
        self.cur = self.cand = None

    def _state(self, W):
        CW = self.lag.CX(W)
        g = np.einsum("ij,ij->i", CW, W)
        nu = np.maximum(np.linalg.norm(CW, axis=1), 1e-12)
        tg = CW - g[:, None] * W
        return dict(W=W, g=g, nu=nu, tg=tg, D=tg / nu[:, None])

    def begin(self, W):
        self.cur = self._state(np.array(W, np.float64))
        return float(self.cur["g"].sum()), float(np.einsum("ij,ij->", self.cur["tg"], self.cur["D"]))

    def trial(self, step):
        Wc = self.cur["W"] + step * self.cur["D"]
        Wc /= np.linalg.norm(Wc, axis=1, keepdims=True)
        self.cand = self._state(Wc)
        return float(self.cand["g"].sum())

    def accept(self):
        a, b = self.cur, self.cand
        ds, dD = b["W"] - a["W"], b["D"] - a["D"]
        nu = b["nu"][:, None]
        sums = (float(np.einsum("ij,ij->", ds, dD * nu)), float(np.einsum("ij,ij->", ds, ds * nu)),
                float(np.einsum("ij,ij->", dD, dD * nu)))
        self.cur, self.cand = b, None
        return (float(np.einsum("ij,ij->", b["tg"], b["D"])),) + sums

    def result(self):
        return self.cur["W"], self.cur["g"]


class DeviceRows:
    """Ascent iterate in the device context (lb_polish_*): the same state as HostRows, on an FP64 copy of W; the
    calls return the same scalars."""

    def __init__(self, eng):
        self.eng = eng

    def begin(self, W):
        return self.eng.polish_begin(W)

    def trial(self, step):
        return self.eng.polish_trial(step)

    def accept(self):
        return self.eng.polish_accept()

    def result(self):
        return self.eng.polish_get()


def rows_backend(lag, W, on_device=True):
    """DeviceRows when the operator is a device context as wide as W, else HostRows."""
    eng = getattr(lag, "eng", None)
    if on_device and eng is not None and hasattr(eng, "polish_begin") and W.shape[1] == getattr(lag, "r", -1):
        return DeviceRows(eng)
    return HostRows(lag)


def bb_step(k, sy, ss, yy, prev):
    """Barzilai-Borwein step (Barzilai & Borwein, IMA J. Numer. Anal. 8, 1988) for an ascent, in the row-weighted
    metric: with curvature sy < 0 along the last step, odd iterations take the long step ss / |sy| and even ones the
    short step |sy| / yy (alternating BB); otherwise the step doubles. Clipped to [STEP_MIN, STEP_MAX]."""
    if sy < 0.0:
        step = ss / -sy if k % 2 else -sy / yy
    else:
        step = min(2.0 * prev, STEP_MAX)
    return float(min(max(step, STEP_MIN), STEP_MAX))


def ascend(lag, W, iters=400, time_limit=600.0, ptol=2e-5, on_device=True, log=print):
    """Maximise f(W) = (const + <C(y), W W^T>) / L over unit rows: W+ = normalise(W + t D) with the Jacobi-
    preconditioned direction D of the backend, t from bb_step, accepted by the nonmonotone Armijo test of Zhang &
    Hager (SIAM J. Optim. 14, 2004): f(W+) >= C_k + ARMIJO t <grad, D> / L with the reference value
    C_{k+1} = (ZH_ETA Q_k C_k + f_{k+1}) / Q_{k+1}, Q_{k+1} = ZH_ETA Q_k + 1; up to HALVINGS halvings of t. Stops
    after `iters` iterations, after `time_limit` seconds, or when 10 iterations gain less than 2 ptol.
    Returns (W, f, g)."""
    rows = rows_backend(lag, W, on_device)
    L, const = lag.inst.L, lag.const
    gsum, slope = rows.begin(W)
    f = (const + gsum) / L
    ref, qk, step = f, 1.0, 2.0
    t0 = time.time()
    f_mark = f
    for k in range(1, iters + 1):
        if time.time() - t0 >= time_limit:
            break
        for h in range(HALVINGS):
            f_try = (const + rows.trial(step)) / L
            if f_try >= ref + ARMIJO * step * slope / L:
                break
            step *= 0.5
        slope, sy, ss, yy = rows.accept()
        f = f_try
        step = bb_step(k, sy, ss, yy, step)
        q_next = ZH_ETA * qk + 1.0
        ref = (ZH_ETA * qk * ref + f) / q_next
        qk = q_next
        if k % 10 == 0:
            log(f"ascent {k}: value {f:.6f}, step {step:.3g}, halvings {h}")
            if f - f_mark < 2.0 * ptol:
                break
            f_mark = f
    W, g = rows.result()
    return W, f, g


SCALES = {"ones": lambda d, L, n: np.ones(n), "vol": lambda d, L, n: np.maximum(d, 1.0),
          "sqrtvol": lambda d, L, n: np.sqrt(np.maximum(d, 1.0)) * math.sqrt(L / n)}


class HostBlocks:
    """LOBPCG blocks as host arrays (n x b). The operator is M X = Diag(g) X - C(y) X through lag.CX; `diags` maps
    a name to an n-vector used as a diagonal weight."""

    def __init__(self, lag, g, diags, width):
        self.lag, self.g, self.diags, self.b = lag, g, diags, int(width)
        self.buf = {}

    def put(self, name, X):
        self.buf[name] = np.array(X, np.float64)

    def get(self, name):
        return self.buf[name]

    def apply(self, dst, src):
        X = self.buf[src]
        self.buf[dst] = self.g[:, None] * X - self.lag.CX(X)

    def gram(self, A, B, diag=None):
        SA = np.hstack([self.buf[a] for a in A])
        SB = np.hstack([self.buf[b] for b in B])
        if diag is not None:
            SB = self.diags[diag][:, None] * SB
        return SA.T @ SB

    def comb(self, dst, srcs, C):
        self.buf[dst] = np.hstack([self.buf[s] for s in srcs]) @ C

    def resid(self, dst, X, MX, theta, diag):
        self.buf[dst] = self.buf[MX] - self.diags[diag][:, None] * self.buf[X] * theta[None, :]

    def scale(self, dst, src, diag):
        self.buf[dst] = self.diags[diag][:, None] * self.buf[src]


class DeviceBlocks:
    """LOBPCG blocks in the device context (lb_blk_*). A block of width b is m = ceil(b / r) device blocks of the
    context's r columns; the library takes at most 6 device blocks per side of a Gram matrix or combination, so wider
    products are assembled from several calls."""
    NAMES = ("X", "MX", "W", "MW", "P", "MP", "X1", "MX1", "P1", "MP1", "R")
    SIDE = 6

    def __init__(self, eng, r, g, diags, width):
        self.e, self.r = eng, int(r)
        self.m = -(-int(width) // self.r)
        self.b = self.m * self.r
        if len(self.NAMES) * self.m > 64:
            raise ValueError(f"LOBPCG block of {width} columns needs more than 64 device blocks of {r}")
        self.dix = {"g": 0}
        eng.blk_alloc(len(self.NAMES) * self.m, len(diags) + 1)
        eng.blk_set_diag(0, np.asarray(g, np.float64))
        for k, (name, w) in enumerate(diags.items()):
            self.dix[name] = k + 1
            eng.blk_set_diag(k + 1, np.asarray(w, np.float64))
        self.ids = {nm: list(range(k * self.m, (k + 1) * self.m)) for k, nm in enumerate(self.NAMES)}

    def put(self, name, X):
        for j, bid in enumerate(self.ids[name]):
            self.e.blk_set(bid, np.ascontiguousarray(X[:, j * self.r:(j + 1) * self.r]))

    def get(self, name):
        return np.hstack([self.e.blk_get(bid) for bid in self.ids[name]])

    def apply(self, dst, src):
        for a, b in zip(self.ids[dst], self.ids[src]):
            self.e.blk_op(a, b, self.dix["g"])

    def gram(self, A, B, diag=None):
        ia = [bid for a in A for bid in self.ids[a]]
        ib = [bid for b in B for bid in self.ids[b]]
        r, S = self.r, self.SIDE
        out = np.empty((len(ia) * r, len(ib) * r))
        dg = -1 if diag is None else self.dix[diag]
        for p in range(0, len(ia), S):
            for q in range(0, len(ib), S):
                ca, cb = ia[p:p + S], ib[q:q + S]
                out[p * r:(p + len(ca)) * r, q * r:(q + len(cb)) * r] = self.e.blk_gram(ca, cb, dg)
        return out

    def comb(self, dst, srcs, C):
        isrc = [bid for s in srcs for bid in self.ids[s]]
        r, S = self.r, self.SIDE
        for j, bid in enumerate(self.ids[dst]):
            cols = C[:, j * r:(j + 1) * r]
            for p in range(0, len(isrc), S):
                chunk = isrc[p:p + S]
                self.e.blk_comb(bid, chunk, np.ascontiguousarray(cols[p * r:(p + len(chunk)) * r]), accumulate=p > 0)

    def resid(self, dst, X, MX, theta, diag):
        r = self.r
        for j, (d_, x_, m_) in enumerate(zip(self.ids[dst], self.ids[X], self.ids[MX])):
            self.e.blk_elem(2, d_, x_, m_, self.dix[diag], theta[j * r:(j + 1) * r])

    def scale(self, dst, src, diag):
        for d_, s_ in zip(self.ids[dst], self.ids[src]):
            self.e.blk_elem(1, d_, s_, -1, self.dix[diag])


def _ritz(GA, GB, k):
    """Rayleigh-Ritz on a subspace S from its Gram matrices GA = S^T M S and GB = S^T Diag(s) S: B-orthonormalise
    through the eigendecomposition of GB, leaving out the directions in which GB is numerically singular (dependent
    search directions, converged residuals), and return the k smallest Ritz values with their coefficients."""
    lam, U = np.linalg.eigh(0.5 * (GB + GB.T))
    keep = lam > 1e-11 * lam[-1]
    T = U[:, keep] / np.sqrt(lam[keep])
    H = T.T @ (0.5 * (GA + GA.T)) @ T
    vals, Y = np.linalg.eigh(0.5 * (H + H.T))
    k = min(k, len(vals))
    return vals[:k], (T @ Y)[:, :k]


def lobpcg_pencil(space, X0, s, L, maxiter=400, chunk=20, tol=1e-9, atol=1e-5, time_limit=1200.0, log=print, tag=""):
    """Block LOBPCG (Knyazev, SIAM J. Sci. Comput. 23(2), 2001) for the smallest eigenvalues of the pencil
    (M, Diag(s)) with the preconditioner "prec" of `space` (a HostBlocks or DeviceBlocks holding the diagonals "s",
    "is" = 1/s and "prec"). Each iteration: residuals R = M X - Diag(s) X Theta; W = prec R, columns scaled to unit
    Diag(s)-norm and those of pairs with residual below tol left out (soft locking); Rayleigh-Ritz on [X, W, P].
    Every `chunk` iterations the first pair is tested: stop when its Diag(s)^-1-norm residual and the change of its
    value, both times sum(s) / L (bound units), are below atol. Returns (theta, resid, X, trace) in ascending order."""
    b = space.b
    sums = float(np.sum(s))
    space.put("X", X0)
    space.apply("MX", "X")
    theta, C = _ritz(space.gram(["X"], ["MX"]), space.gram(["X"], ["X"], "s"), b)
    names = {nm: nm for nm in ("X", "MX", "X1", "MX1", "P", "MP", "P1", "MP1")}

    def rotate(pairs):
        # the loop writes the next X / MX / P / MP into the spare buffers, then swaps the names
        for a, c in pairs:
            names[a], names[c] = names[c], names[a]

    space.comb(names["X1"], [names["X"]], C)
    space.comb(names["MX1"], [names["MX"]], C)
    rotate((("X", "X1"), ("MX", "MX1")))

    def residual_norms():
        space.resid("R", names["X"], names["MX"], theta, "s")
        return np.sqrt(np.maximum(np.diag(space.gram(["R"], ["R"], "is")), 0.0))

    trace, t0, prev, it, searched = [], time.time(), np.inf, 0, False
    while True:
        it += 1
        rn = residual_norms()
        space.scale("W", "R", "prec")
        dw = np.diag(space.gram(["W"], ["W"], "s"))
        live = (dw > 1e-300) & (rn > tol)
        space.comb("R", ["W"], np.diag(np.where(live, 1.0 / np.sqrt(np.maximum(dw, 1e-300)), 0.0)))
        space.apply("MW", "R")
        S = [names["X"], "R"] + ([names["P"]] if searched else [])
        MS = [names["MX"], "MW"] + ([names["MP"]] if searched else [])
        theta, C = _ritz(space.gram(S, MS), space.gram(S, S, "s"), b)
        space.comb(names["X1"], S, C)
        space.comb(names["MX1"], MS, C)
        space.comb(names["P1"], S[1:], C[b:])
        space.comb(names["MP1"], MS[1:], C[b:])
        rotate((("X", "X1"), ("MX", "MX1"), ("P", "P1"), ("MP", "MP1")))
        searched = True
        if it % chunk == 0 or it >= maxiter:
            rn = residual_norms()
            trace.append((it, time.time() - t0, float(theta[0]), float(rn[0])))
            log(f"  pencil [{tag}] iteration {it}: theta_1 {theta[0]:.6e}, residual {rn[0]:.2e} (bound units: "
                f"shift {max(0.0, -theta[0]) * sums / L:.6f}, residual {rn[0] * sums / L:.2e})")
            if rn[0] * sums / L < atol and abs(theta[0] - prev) * sums / L < atol:
                break
            if it >= maxiter or time.time() - t0 > time_limit:
                break
            prev = theta[0]
    rn = residual_norms()
    X = space.get(names["X"])
    o = np.argsort(theta)
    return theta[o], rn[o], X[:, o], trace


def dense_pencil(lag, g, s, k):
    """The k smallest eigenpairs of (Diag(g) - C(y), Diag(s)) by a dense generalized eigensolver, with C(y) built
    column by column through the operator (for graphs too small for a block iteration). Residuals are those of the
    computed pairs, as lobpcg_pencil reports them."""
    import scipy.linalg as sl
    n = lag.inst.n
    C = lag.CX(np.eye(n))
    M = np.diag(g) - 0.5 * (C + C.T)
    vals, X = sl.eigh(M, np.diag(s))
    k = min(k, n)
    vals, X = vals[:k], X[:, :k]
    R = M @ X - s[:, None] * X * vals[None, :]
    xs = np.einsum("ij,ij->j", X, s[:, None] * X)
    rn = np.sqrt(np.einsum("ij,ij->j", R, R / s[:, None]) / xs)
    return vals, rn, X, []


def pencil_min(lag, g, s, b, X0=None, rng=None, maxiter=400, chunk=20, tol=1e-9, atol=1e-5, time_limit=1200.0,
               on_device=True, log=print, tag=""):
    """Smallest eigenvalues of (Diag(g) - C(y), Diag(s)) with a block of b columns: the starting block holds the
    given directions X0 (e.g. an orthonormal basis of the ascent's W) and random columns up to b. On the device
    (DeviceBlocks) when the operator is an FP64 device context on a graph of at least 8 r rows, else on the host;
    dense for graphs of at most 4 b rows. The preconditioner is Jacobi: 1 / (|M_ii| + 0.01 s_i), with
    M_ii = g_i + d_i^2 / L (C has -d_i^2 / L on its diagonal). Returns (theta, resid, X, trace)."""
    inst = lag.inst
    n, L = inst.n, inst.L
    rng = np.random.default_rng(0) if rng is None else rng
    if n <= 4 * b:
        return dense_pencil(lag, g, s, b)
    prec = 1.0 / (np.abs(g + inst.d ** 2 / L) + 1e-2 * s)
    diags = {"s": s, "is": 1.0 / s, "prec": prec}
    eng = getattr(lag, "eng", None)
    if on_device and eng is not None and hasattr(eng, "blk_alloc") and getattr(eng, "dtype", None) == "f64" \
            and n >= 8 * lag.r:
        space = DeviceBlocks(eng.blk_engine(), lag.r, g, diags, b)
    else:
        space = HostBlocks(lag, g, diags, b)
    width = space.b
    k0 = 0 if X0 is None else min(X0.shape[1], width)
    cols = [X0[:, :k0]] if k0 else []
    cols.append(rng.standard_normal((n, width - k0)))
    return lobpcg_pencil(space, np.hstack(cols), s, L, maxiter=maxiter, chunk=chunk, tol=tol, atol=atol,
                         time_limit=time_limit, log=log, tag=tag)


def escape_directions(theta, X, ratio=0.3, count=4):
    """Rank escape (Journee, Bach, Absil & Sepulchre, SIAM J. Optim. 20(5), 2010: a stationary point of the rank-k
    problem that is not optimal has a direction of negative curvature, and adding a column along it escapes): the
    eigenvectors of at most `count` Ritz values that are negative and at most ratio theta_1, each scaled to max-abs
    entry 1/2, as new columns for W. None when no value qualifies."""
    pick = [j for j in range(len(theta)) if theta[j] < 0 and theta[j] <= ratio * theta[0]][:count]
    if not pick:
        return None
    U = X[:, pick]
    return 0.5 * U / np.abs(U).max(axis=0, keepdims=True)


def _fp64_operator(lag, device):
    """The operator for the FP64 assembly: `lag` itself unless it is an FP32 / BF16 device context, in which case an
    FP64 context of the same instance (returned as (operator, owned))."""
    eng = getattr(lag, "eng", None)
    if eng is None or getattr(eng, "dtype", "f64") == "f64":
        return lag, False
    return DeviceLagrangian(lag.inst, r=lag.r, dtype="f64", device=device), True


def estimate(inst, V, lag=None, polish="rgd", rgd_iters=400, polish_time=600.0, ptol=2e-5, scale="vol",
             eig_extra=16, eig_maxiter=400, eig_tol=1e-9, eig_chunk=20, eig_atol=1e-5, eig_time=1200.0, escape=2,
             escape_min=0.002, escape_p=4, escape_ratio=0.3, seed=0, no_w_init=False, device=0, dtype="f64",
             eig_block=None, device_polish=True, eig_device=True, log=print):
    """Candidate bound at the instance's multipliers. Weak duality: for any g and s > 0,
    L Q* <= C0 + sum(lam) + sum(g) + max(0, -lambda_min(Diag(g) - C, Diag(s))) sum(s).
    Steps: ascend from V (the search's embedding) so that g_i = (C W)_i . W_i is a good diagonal; per scale s, the
    pencil's smallest eigenvalue mu, hedged by its residual (theta_1 - r_1), gives ub = dv + max(0, -mu) sum(s) / L;
    then up to `escape` rank-escape rounds (new columns from escape_directions, ascent again, new estimate), keeping
    the best. The ascent runs on `lag`'s precision (dtype: the search's), the values g, dv and the pencil on an FP64
    operator (the FP64 assembly). The pencil block is eig_block columns, or W's width + eig_extra.
    Returns the record of the reference stage (dv0, dv, s_<scale>, escape_rounds, ub, ...) plus the best round's W, g,
    s and lmin for the certificate. ub is an estimate: the eigensolver may stop above the smallest eigenvalue."""
    own = lag is None
    if own:
        lag = DeviceLagrangian(inst, r=V.shape[1], dtype=dtype, device=device)
    asm, own_asm = _fp64_operator(lag, device)
    t_start = time.time()
    rng = np.random.default_rng(seed)
    n, L = inst.n, inst.L
    scales = scale.split(",")
    methods = [m for m in polish.split("+") if m not in ("none", "")]
    for m in methods:
        if m != "rgd":
            raise ValueError(f"polish method {m!r} is not supported (rgd | none)")

    def improve(W):
        for _ in methods:
            W = ascend(lag, W, iters=rgd_iters, time_limit=polish_time, ptol=ptol, on_device=device_polish, log=log)[0]
        return W

    def pencil_round(W, dv, g):
        """One estimate per scale at the diagonal g: (records by scale, (theta, X, s) by scale)."""
        recs, eig = {}, {}
        for name in scales:
            s = SCALES[name](inst.d, L, n)
            b = W.shape[1] + eig_extra if eig_block is None else int(eig_block)
            X0 = None
            if not no_w_init:
                U, sv, _ = np.linalg.svd(W, full_matrices=False)
                X0 = U[:, :int(np.count_nonzero(sv > 1e-6 * sv[0]))]
            t = time.time()
            theta, rn, X, trace = pencil_min(asm, g, s, b, X0=X0, rng=rng, maxiter=eig_maxiter, chunk=eig_chunk,
                                             tol=eig_tol, atol=eig_atol, time_limit=eig_time, on_device=eig_device,
                                             log=log, tag=name)
            sums = float(s.sum())
            mu = float(theta[0] - rn[0])
            shift = max(0.0, -mu) * sums / L
            recs[name] = dict(theta=theta[:6].tolist(), resid=rn[:6].tolist(), lmin_est=mu, shift=shift,
                              ub=dv + shift, sum_s=sums, eig_time=time.time() - t, eig_trace=trace,
                              ub_ritz=dv + max(0.0, -float(theta[0])) * sums / L, dv=dv, k=int(W.shape[1]),
                              block=int(X.shape[1]))
            log(f"scale {name}: lambda_min ~ {mu:.3e} (theta_1 {theta[0]:.3e}, residual {rn[0]:.1e}); "
                f"shift {shift:.6f}; candidate {dv + shift:.6f}")
            eig[name] = (theta, X, s)
        return recs, eig

    try:
        W = V / np.maximum(np.linalg.norm(V, axis=1, keepdims=True), 1e-300)
        dv0 = value(asm, W)[0]
        t = time.time()
        W = improve(W)
        t_pol = time.time() - t
        dv, g, _ = value(asm, W)
        log(f"ascent: value {dv0:.6f} -> {dv:.6f} in {t_pol:.1f}s")
        res = dict(n=n, L=L, k=int(W.shape[1]), dv0=dv0, dv=dv, polish_time=t_pol)
        best, history = np.inf, []
        for rnd in range(escape + 1):
            recs, eig = pencil_round(W, dv, g)
            ub = min(r_["ub"] for r_ in recs.values())
            history.append(dict(round=rnd, dv=dv, ub=ub, k=int(W.shape[1])))
            if ub < best:
                best = ub
                top = min(recs, key=lambda q: recs[q]["ub"])
                res.update({f"s_{q}": rec for q, rec in recs.items()})
                res.update(dv=dv, W=W, g=g, s=eig[top][2], lmin=recs[top]["lmin_est"], best_scale=top)
            first = scales[0]
            if rnd == escape or recs[first]["shift"] < escape_min:
                break
            U = escape_directions(eig[first][0], eig[first][1], escape_ratio, escape_p)
            if U is None:
                break
            W = np.hstack([W, U])
            W /= np.linalg.norm(W, axis=1, keepdims=True)
            W = improve(W)
            dv, g, _ = value(asm, W)
            log(f"rank escape {rnd + 1}: {U.shape[1]} columns added (k = {W.shape[1]}), value {dv:.6f}")
    finally:
        if own_asm:
            asm.close()
        if own:
            lag.close()
    res.update(escape_rounds=history, ub=best, shift=best - res["dv"], time=time.time() - t_start)
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
