"""Continuous-batching bound service on one GPU: requests (graphs) arrive over time, are admitted as worker slots
and device memory free up, and each runs the whole bound pipeline with its V and multipliers resident on the device:
instance build -> seeded label-propagation levels -> device ALM search (pair level, VCS level, triangle and VCS
separation, PHR updates; lagbound_amd.alm.Solver in process) -> candidate bound (device C(y)X products, FP64
assembly; lagbound_amd.bound.estimate) -> FP64 certificate: the dense one (assembly and Cholesky on the device) up to
3000 vertices, the lifted sparse one (lagbound_amd.sparse_cert: host LDL^T, dense core on the device) above.

Concurrency: one worker per slot, each with its own device contexts (own HIP stream), so the work of requests that
are in flight at the same time overlaps on the GPU. By default a worker is a process of its own (spawned and warmed
up before the stream starts; one request at a time, so the device library is called from one thread per process and
the host parts of concurrent requests -- LOBPCG, merges, label propagation -- do not share one interpreter lock).
`workers="thread"` keeps the slots as threads of this process; it is refused while the solver runs on the device
library (one run of device thread slots ended in a host segmentation fault whose cause is not established) and
serves only with a host engine in its place (the CPU tests' oracle stand-in). Admission
is by projected HBM footprint, in this process: a request waits until the sum of the footprints in flight plus its
own fits the budget. Every computation of a request depends only on its own inputs and seeds, so its result is the
same alone and in the mix (up to the summation order of floating-point atomics inside the device reductions).

No reference counterpart: the reference computes one bound per process (api.upper_bound, src/lagbound/api.py).
"""
from __future__ import annotations

import math
import multiprocessing as mp
import os
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import alm
from . import bound as B
from . import instances as I


@dataclass
class Request:
    rid: int
    n: int
    keys: np.ndarray                 # sorted unique u*n+v, u < v
    arrival: float = 0.0             # seconds after the stream starts
    seed: int = 0


@dataclass
class Params:
    k: int = 16
    dtype: str = "f32"
    options: tuple = ()              # kernel switches of the search's device library ("NAME=VALUE", lb_set_option)
    outer: int = 24
    nbatch: int = 0                  # 0: the reference's rule (nbatch_rule)
    tri_every: int = 4
    tri_add: int = 0                 # 0: 2 n
    vcs_every: int = 3
    P: int = 64                      # label-propagation start clusters
    rgd_iters: int = 120
    eig_maxiter: int = 60
    eig_block: int = 8               # LOBPCG block width (the smallest eigenvalue only is needed)
    lpa_iters: int = 3
    escape: int = 0
    cert_max_n: int = 10 ** 9        # certify requests up to this many vertices (0: none)
    cert_dense_max_n: int = 3000     # dense FP64 certificate up to this size, the lifted sparse one above
    cert_max_core: int = 170_000     # largest dense core of the sparse certificate (8 m^2 bytes: 215 GiB)
    cert_margin: float = 1e-7
    cert_on_device: bool = True
    blas_threads: int = 2            # host BLAS threads per slot while serving (slots x threads ~ the host cores)
    hw_queues: int = 1               # process slots: GPU_MAX_HW_QUEUES of each worker (0: the runtime's default)


@dataclass
class Result:
    rid: int
    n: int
    m: int
    arrival: float
    start: float
    finish: float
    lb: float                        # modularity of the level partition (a feasible point)
    dv: float
    ub: float
    certified: bool
    UB: float | None
    device_s: float
    ub_search: float | None = None   # the candidate bound in the search context's precision (same W and shift)
    du_fp32_fp64: float | None = None  # |ub_search - ub|: FP32 against FP64 assembly of one result
    cert_note: str | None = None     # why a result is not certified
    stages: dict = field(default_factory=dict)   # seconds per stage (search, estimate, certificate) and core rows
    log: list = field(default_factory=list)

    @property
    def latency(self):
        return self.finish - self.arrival


def nbatch_rule(n):
    """Solver batches as the reference's upper_bound sets them (api.py:247-250): 512 from 32,768 rows, else
    min(n, max(8, n / 64))."""
    return 512 if n >= 512 * 64 else min(n, max(8, n // 64))


def footprint_bytes(n, m, k, p: Params):
    """Projected device memory of one request: the solver context (V, G and sweep snapshots in FP32, CSR, the
    triangle pool at its cap, VCS working set and separation scratch) plus the FP64 estimate context with the
    device eigensolver's eleven n x k blocks and four diagonals."""
    tri_cap = (p.tri_add or 2 * n) * max(1, p.outer // max(1, p.tri_every))
    solver = 6 * n * k * 4 + 2 * m * 12 + tri_cap * (16 * 3 + 16) + 8 * n * 24 + 64 * n
    est = 3 * n * k * 8 + 2 * m * 12 + tri_cap * (24 * 3 + 16) + 8 * n * 24 + 11 * n * k * 8 + 4 * n * 8
    cert = 0
    if 0 < n <= p.cert_max_n:
        # dense: M and its magnitudes (FP64); sparse: the dense core's Schur complement, at most about n / 2 rows on
        # power-law graphs (measured 0.41-0.45 n of the vertices)
        cert = 16 * n * n if n <= p.cert_dense_max_n else 8 * (n // 2) ** 2
    return int(1.25 * (solver + est)) + cert + (32 << 20)


def modularity(inst, lab):
    u = np.repeat(np.arange(inst.n), np.diff(inst.rowptr))
    same = lab[u] == lab[inst.col]
    w = np.ones(inst.nnz) if inst.w is None else inst.w
    vol = np.bincount(lab, weights=inst.d)
    return float((inst.C0 + w[same].sum() - (vol ** 2).sum() / inst.L) / inst.L)


def search_request(req: Request, p: Params, device=0, log=None):
    """Instance, label-propagation levels and the device search of one request: returns (inst with the final
    multipliers, labels, the finished alm.Solver, whose context still holds V and the multipliers)."""
    log = log or (lambda *a: None)
    inst = I.build_instance(req.n, req.keys)
    lab = I.label_propagation(inst.rowptr, inst.col, inst.n, req.seed, P=p.P, iters=p.lpa_iters)
    nb = p.nbatch or nbatch_rule(inst.n)
    argv = ["--tag", f"req{req.rid}", "--k", str(p.k), "--nbatch", str(nb), "--inner", "1", "--est_every", "0",
            "--mom", "0", "--outer", str(p.outer), "--time_limit", "1e9", "--seed", str(req.seed),
            "--tri_add", str(p.tri_add or 2 * inst.n), "--tri1_add", "0", "--tri_every", str(p.tri_every),
            "--tri_max", "16000000", "--vcs_every", str(p.vcs_every), "--vcs_add", str(4 * inst.n),
            "--vcs_max", str(16 * inst.n), "--pw_rule", "default", "--dtype", p.dtype, "--device", str(device)]
    for o in p.options:
        argv += ["--option", o]
    a = alm.parse(argv)
    s = alm.Solver(a, inst=inst, labels=[lab], vcs_labels=[lab], init_labels=lab, log=log, signals=False)
    s.run(close=False)
    return inst, lab, s


def search_precision_bound(s, inst, est):
    """The candidate bound recomputed with the search context's own evaluation (its storage precision: FP32 for the
    FP32 search) at the estimate's W and shift: (C0 + sum(lam) + sum_i (C W)_i . W_i) / L + shift. Against the
    FP64-assembled est["ub"] it gives |U_fp32 - U_fp64| of one result. None when W is wider than the context (after
    rank-escape rounds)."""
    W = est["W"]
    if W.shape[1] != s.eng.r:
        return None
    s.eng.set_V(W)
    parts = s.eng.eval(0.0)
    return float(parts[0]) / inst.L + est["shift"]


def solve_request(req: Request, p: Params, device=0):
    """The bound pipeline of one request; returns (Result fields without timing, seconds spent inside the device
    library's calls)."""
    logs = []
    log = lambda *s: logs.append(" ".join(str(x) for x in s))
    t0 = time.perf_counter()
    inst, lab, s = search_request(req, p, device, log)
    t1 = time.perf_counter()
    try:
        lag = B.DeviceLagrangian(inst, r=p.k, dtype="f64", device=device)
        try:
            est = B.estimate(inst, s.V, lag=lag, rgd_iters=p.rgd_iters, eig_maxiter=p.eig_maxiter,
                             eig_block=p.eig_block, escape=p.escape, log=log)
        finally:
            lag.close()
        ub_s = search_precision_bound(s, inst, est)
    finally:
        s.eng.close()
    t_dev = s.eng.lib_s + lag.eng.lib_s
    t2 = time.perf_counter()
    cert = None
    dev = device if p.cert_on_device else None
    if inst.n <= p.cert_dense_max_n and inst.n <= p.cert_max_n:
        cert = B.certify_estimate(inst, est, margin=p.cert_margin, device=dev)
    elif inst.n <= p.cert_max_n:
        from . import sparse_cert as SC
        cert = SC.certify_estimate_sparse(inst, est, margin=p.cert_margin, device=dev, max_core=p.cert_max_core,
                                          log=log)
    why = None if (cert and cert["passed"]) else (cert.get("reason") if cert else f"n = {inst.n} above cert_max_n")
    stages = dict(search_s=round(t1 - t0, 3), estimate_s=round(t2 - t1, 3), cert_s=round(time.perf_counter() - t2, 3),
                  cert_core=int(cert.get("info", {}).get("core", 0)) if cert else 0)
    return dict(n=inst.n, m=inst.nnz // 2, stages=stages, lb=modularity(inst, lab.astype(np.int64)), dv=est["dv"], ub=est["ub"],
                ub_search=ub_s, du_fp32_fp64=(abs(ub_s - est["ub"]) if ub_s is not None else None),
                certified=bool(cert and cert["passed"]), UB=(cert["UB"] if cert and cert["passed"] else None),
                cert_note=why, log=logs), t_dev


def prepare_device(p: Params, device=0):
    """One tiny device certificate, so rocBLAS / rocSOLVER are opened (dlopen, kernel registration, first handle)
    before any request runs and never while other slots launch kernels."""
    if not (p.cert_on_device and p.cert_max_n > 0):
        return
    keys = np.array([0 * 3 + 1, 0 * 3 + 2, 1 * 3 + 2], np.int64)          # a triangle
    tiny = I.build_instance(3, keys)
    B.certify_dense_device(tiny, np.zeros(3), np.ones(3), device=device)


def _worker_process(conn, p: Params, device: int):
    """A slot's process: limit host BLAS threads, open the device and the solver libraries, report ready, then solve
    one request per message until None arrives. Replies ("ok", (out, device seconds)) or ("error", text)."""
    try:
        if p.hw_queues > 0:                 # read by the HIP runtime when it initialises, which has not happened yet
            os.environ["GPU_MAX_HW_QUEUES"] = str(int(p.hw_queues))
        from threadpoolctl import threadpool_limits
        threadpool_limits(limits=p.blas_threads)
        prepare_device(p, device)
        conn.send(("ready", None))
    except Exception as exc:
        conn.send(("error", repr(exc)))
        return
    while True:
        req = conn.recv()
        if req is None:
            return
        try:
            conn.send(("ok", solve_request(req, p, device)))
        except Exception as exc:
            conn.send(("error", repr(exc)))


class _ProcessSlots:
    """`count` spawned worker processes, started and warmed up together; `call(k, req)` runs a request on slot k."""

    def __init__(self, count, p: Params, device: int, start_timeout=600.0):
        ctx = mp.get_context("spawn")
        self.conns, self.procs = [], []
        for _ in range(count):
            a, b = ctx.Pipe()
            pr = ctx.Process(target=_worker_process, args=(b, p, device), daemon=True)
            pr.start()
            b.close()
            self.conns.append(a)
            self.procs.append(pr)
        try:
            for k, c in enumerate(self.conns):
                if not c.poll(start_timeout):
                    raise RuntimeError(f"serving slot {k} did not start within {start_timeout:.0f} s")
                st, msg = c.recv()
                if st != "ready":
                    raise RuntimeError(f"serving slot {k} failed to start: {msg}")
        except BaseException:
            self.close()
            raise

    def call(self, k, req):
        self.conns[k].send(req)
        st, payload = self.conns[k].recv()
        if st != "ok":
            raise RuntimeError(payload)
        return payload

    def close(self):
        for c in self.conns:
            try:
                c.send(None)
            except (OSError, BrokenPipeError):
                pass
        for pr in self.procs:
            pr.join(timeout=30)
            if pr.is_alive():
                pr.terminate()
                pr.join(timeout=10)
        for c in self.conns:
            c.close()
        self.conns, self.procs = [], []


def _device_solver():
    """True when the solver's engine is the device library's (not a host stand-in installed by a test)."""
    from . import engine
    return alm.Engine is engine.Engine


class Server:
    """Continuous batching: `slots` workers (processes by default, or threads), HBM-footprint admission against
    `budget_bytes`."""

    def __init__(self, p: Params, slots=8, budget_bytes=None, device=0, progress=None, workers="process"):
        if workers not in ("process", "thread"):
            raise ValueError(f"workers must be 'process' or 'thread', not {workers!r}")
        if workers == "thread" and _device_solver():
            raise RuntimeError("thread slots are not available with the device library (one such run ended in a host "
                               "segmentation fault of unknown cause); use workers='process'")
        self.progress = progress                 # callable(Result) after each completed request
        self.p = p
        self.slots = slots
        self.workers = workers
        self.device = device
        self.budget = int(budget_bytes) if budget_bytes is not None else 192 << 30  # of the 288 GB of HBM3E
        self.used = 0
        self.cv = threading.Condition()
        self.results = {}
        self.errors = {}

    def _admit(self, fp):
        with self.cv:
            while self.used + fp > self.budget and self.used > 0:
                self.cv.wait()
            self.used += fp

    def _release(self, fp):
        with self.cv:
            self.used -= fp
            self.cv.notify_all()

    def run(self, requests, realtime=True):
        """Serve `requests` (sorted by arrival). Returns {rid: Result}. Process slots are started and warmed up
        before the stream's clock starts and stopped when it ends."""
        self.start_s = 0.0
        if self.workers == "thread":
            from threadpoolctl import threadpool_limits
            with threadpool_limits(limits=self.p.blas_threads):
                prepare_device(self.p, self.device)
                return self._run(requests, realtime, lambda k, req: solve_request(req, self.p, self.device))
        t = time.perf_counter()
        slots = _ProcessSlots(self.slots, self.p, self.device)
        self.start_s = time.perf_counter() - t
        try:
            return self._run(requests, realtime, slots.call)
        finally:
            slots.close()

    def _run(self, requests, realtime, solve):
        q = queue.Queue()
        t0 = time.perf_counter()

        def worker(k):
            while True:
                req = q.get()
                if req is None:
                    return
                fp = footprint_bytes(req.n, len(req.keys), self.p.k, self.p)
                self._admit(fp)
                start = time.perf_counter() - t0
                try:
                    out, tdev = solve(k, req)
                    res = Result(rid=req.rid, arrival=req.arrival, start=start, finish=time.perf_counter() - t0,
                                 device_s=tdev, **out)
                    self.results[req.rid] = res
                    if self.progress:
                        self.progress(res)
                except Exception as exc:                      # reported per request, never swallowed
                    self.errors[req.rid] = repr(exc)
                finally:
                    self._release(fp)

        ths = [threading.Thread(target=worker, args=(k,), daemon=True) for k in range(self.slots)]
        for t in ths:
            t.start()
        for req in requests:
            if realtime:
                dt = req.arrival - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
            q.put(req)
        for _ in ths:
            q.put(None)
        for t in ths:
            t.join()
        if self.errors:
            raise RuntimeError(f"{len(self.errors)} request(s) failed: {self.errors}")
        return self.results


def poisson_stream(count, rate, n_lo, n_hi, kbar_lo, kbar_hi, seed):
    """Poisson arrivals at `rate` req/s of Chung-Lu graphs (gamma 2.1) with n log-uniform in [n_lo, n_hi] and mean
    degree uniform in [kbar_lo, kbar_hi]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    reqs = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        kb = float(rng.uniform(kbar_lo, kbar_hi))
        n2, keys = I.chung_lu(n, 2.1, kb, seed * 100003 + i)
        reqs.append(Request(rid=i, n=n2, keys=keys, arrival=float(t[i]), seed=i))
    return reqs


def summarize(results, wall):
    lat = np.array([r.latency for r in results.values()])
    svc = np.array([r.finish - r.start for r in results.values()])
    wait = np.array([r.start - r.arrival for r in results.values()])
    cert = [r for r in results.values() if r.certified]
    # certified bracket lb <= Q* <= UB: lb is the modularity of the level partition, UB the certified bound
    gaps = np.array([r.UB - r.lb for r in cert]) if cert else np.zeros(1) * np.nan
    return {"requests": len(results), "wall_s": round(wall, 3), "solves_per_s": round(len(results) / wall, 4),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
            "service_p50_s": round(float(np.percentile(svc, 50)), 4),
            "queue_wait_p50_s": round(float(np.percentile(wait, 50)), 4),
            "device_call_share": round(float(sum(r.device_s for r in results.values()) / max(svc.sum(), 1e-9)), 4),
            "certified_gap_UB_minus_lb": {"p10": round(float(np.percentile(gaps, 10)), 5),
                                          "p50": round(float(np.percentile(gaps, 50)), 5),
                                          "p90": round(float(np.percentile(gaps, 90)), 5)},
            "certified": len(cert),
            "du_fp32_fp64_max": (max(du) if (du := [r.du_fp32_fp64 for r in results.values()
                                                      if r.du_fp32_fp64 is not None]) else None),
            "uncertified": {str(r.rid): r.cert_note for r in results.values() if not r.certified},
            # where a request's service time goes: search, candidate bound, FP64 certificate (p50 / p99 over the
            # requests), and the latency a request would have without its certificate stage
            "stage_s": {k: {"p50": round(float(np.percentile(v, 50)), 4), "p99": round(float(np.percentile(v, 99)), 4)}
                        for k in ("search_s", "estimate_s", "cert_s")
                        if len(v := np.array([r.stages.get(k, 0.0) for r in results.values()]))},
            "latency_before_cert_p99_s": round(float(np.percentile(
                lat - np.array([r.stages.get("cert_s", 0.0) for r in results.values()]), 99)), 4)}


def main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(prog="lagbound_amd.serve", description="Serve a Poisson stream of bound requests "
                                 "on one GPU and print the latency summary as one JSON line.")
    ap.add_argument("--count", type=int, default=24)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--n_lo", type=int, default=1000)
    ap.add_argument("--n_hi", type=int, default=200000)
    ap.add_argument("--kbar_lo", type=float, default=6.0)
    ap.add_argument("--kbar_hi", type=float, default=30.0)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--workers", choices=["process"], default="process",
                    help="one process per slot (thread slots serve only with a host engine)")
    ap.add_argument("--blas_threads", type=int, default=None, help="host BLAS threads per slot (default: Params)")
    ap.add_argument("--hw_queues", type=int, default=None, help="process slots: hardware queues per worker")
    ap.add_argument("--budget_gb", type=float, default=192.0)
    ap.add_argument("--outer", type=int, default=24)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="a kernel switch of the search's device library (lb_set_option), e.g. sweep_graph=1")
    ap.add_argument("--inline", action="store_true",
                    help="solve the requests one after another in this process (no slots, no arrival times): one "
                         "JSON line per request; the form to run under a profiler")
    a = ap.parse_args(argv)
    t = time.perf_counter()
    reqs = poisson_stream(a.count, a.rate, a.n_lo, a.n_hi, a.kbar_lo, a.kbar_hi, a.seed)
    gen = time.perf_counter() - t
    p = Params(outer=a.outer, options=tuple(a.option))
    if a.blas_threads:
        p.blas_threads = a.blas_threads
    if a.hw_queues is not None:
        p.hw_queues = a.hw_queues

    if a.inline:
        prepare_device(p)
        for req in reqs:
            t0 = time.perf_counter()
            out, t_dev = solve_request(req, p)
            out.pop("log", None)
            out.update(rid=req.rid, service_s=round(time.perf_counter() - t0, 3), device_lib_s=round(t_dev, 3))
            print(json.dumps(out), flush=True)
        return 0

    def progress(r):
        print(f"[serve] rid {r.rid} n {r.n} m {r.m} latency {r.latency:.3f}s service {r.finish - r.start:.3f}s "
              f"lb {r.lb:.5f} ub {r.ub:.5f} certified {r.certified} stages {r.stages}", file=sys.stderr, flush=True)

    srv = Server(p, slots=a.slots, budget_bytes=a.budget_gb * 2 ** 30, progress=progress, workers=a.workers)
    res = srv.run(reqs)
    wall = max(r.finish for r in res.values())
    out = summarize(res, wall)
    out.update(stream=dict(count=a.count, rate=a.rate, n=[a.n_lo, a.n_hi], kbar=[a.kbar_lo, a.kbar_hi], seed=a.seed,
                           slots=a.slots, workers=a.workers, blas_threads=p.blas_threads, budget_gb=a.budget_gb,
                           outer=a.outer, hw_queues=p.hw_queues, generate_s=round(gen, 2),
                           slots_start_s=round(srv.start_s, 2)))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
