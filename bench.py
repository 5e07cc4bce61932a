#!/usr/bin/env python3
"""Benchmark: Lagrangian f+grad evaluations per second on the metric instance (one MI355X).

Step = one f+grad evaluation (f as an FP64 scalar and G = C(y) V, n x r) at fixed multipliers, rho = 0.02,
rt = rho * tri_weight (tri_weight 1), on the resident instance: Chung-Lu, n ~ 1M, gamma 2.1, mean degree ~13,
r = 16, T ~ 3.4M wedge triangles, one pair level and one VCS level (P = 64, 4 working-set entries per row), FP32 V.
Inputs are resident in HBM before the timed region; the timed region is K back-to-back evaluations bracketed by
device synchronisation.

Also reported on the same instance: the HIP path with BF16-stored V and with FP64 V, the plain FP64 GPU form
(torch FP64 CSR SpMM, gathers and sparse products), and the FP64 C/OpenMP oracle on the host cores (cpu_baseline). The
metric's second half, batched bound latency, comes from a Poisson stream of whole bound requests (`serving`:
p50 / p99 latency, solves/s, every result certified by the FP64 certificate), from the batched evaluation of 512
uneven graphs in one launch per kernel (`batch`), and from the batched search of 512 such graphs with every graph's
FP64 certificate (`batch_search`).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu] [--no-extra] [--no-serve] [--no-batch]
       [--no-batch-search]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Lagrangian f+grad evals/s (n=1M,r=16) vs FP64; batched bound p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def alg_bytes(inst, r, s_v, s_g, s_c):
    """Compulsory bytes of one evaluation: every input read once, every output written once.
    V + G + col + rowptr (+ weights) + d + labels + triangles (TI + lam) + VCS (ptr + col + sig)."""
    n, nnz, T = inst.n, inst.nnz, inst.T
    b = n * r * s_v + n * r * s_g + nnz * 4 + (n + 1) * 8 + n * s_c
    if inst.w is not None:
        b += nnz * s_c
    b += (len(inst.levels) + len(inst.vcs)) * n * 4
    b += T * (12 + s_c)
    for v in inst.vcs:
        b += (n + 1) * 8 + v.nnz * (4 + s_c)
    return b


def time_engine(eng, rho, steps, warmup):
    import numpy as np  # noqa: F401
    for _ in range(warmup):
        eng.eval(rho, fetch=False)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.eval(rho, fetch=False)
    eng.sync()
    return (time.perf_counter() - t0) / steps


def large_graph(tuned, rho, steps):
    """The HBM-scale evaluation: Chung-Lu n = 2.4M, gamma 2.2, mean degree 50 (~120M CSR entries, int64 row
    pointers), 8M wedge triangles, pair + VCS levels over 64 LPA clusters; BF16-stored V (FP32 accumulation) and
    FP32 V, with the given launch shapes (None: the library's defaults). Reports evals/s and compulsory GB/s against
    the HBM peak."""
    from lagbound_amd import instances as I
    from lagbound_amd import tune as TU
    from lagbound_amd.engine import Engine
    t0 = time.time()
    log("building the n=2.4M graph (gamma 2.2, mean degree 50, 8M triangles)")
    inst, V = I.metric_instance(2_400_000, 2.2, 50.0, T=8_000_000)
    out = {"n": inst.n, "nnz": inst.nnz, "T": inst.T, "pair_P": inst.levels[0].P, "vcs_ws": inst.vcs[0].nnz,
           "r": V.shape[1], "rho": rho, "built_s": round(time.time() - t0, 1)}
    log(f"n={inst.n} nnz={inst.nnz} T={inst.T} built in {out['built_s']}s")
    for dt, sv in (("bf16", 2), ("f32", 4)):
        e = Engine(r=V.shape[1], dtype=dt)
        if tuned:
            TU.apply(e, tuned)
        e.set_instance(inst)
        e.set_V(V)
        e.eval(rho)
        sec = time_engine(e, rho, max(20, steps // 4), 5)
        ms, _ = e.time_eval(rho, iters=20, per_kernel=False)
        b = alg_bytes(inst, V.shape[1], sv, 4, 4)
        gbs = b / (ms * 1e-3) / 1e9
        out["launch_shapes"] = tuned or "library defaults"
        out[dt] = {"evals_per_s": round(1.0 / sec, 2), "ms_per_eval": round(sec * 1e3, 4),
                   "ms_per_eval_events": round(ms, 4), "alg_bytes": b, "achieved_GBs": round(gbs, 1),
                   "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"n=2.4M {dt}: {sec * 1e3:.3f} ms/eval, {gbs:.0f} GB/s compulsory")
        e.close()
    return out


def summarize_pmc(fetch_dir, write_dir, hitmiss_dir=None, workload="", tuning="", commands=()):
    """Per-evaluation HBM traffic of the f+grad chain from rocprofv3 counter runs of this script with --calibrate:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) summed over the chain's kernels and averaged over the evaluations
    (one lb_k_finalize per evaluation), each scaled by the factor that makes the copy probe of the same run
    (lbp_copy_kernel, 256 MiB read + 256 MiB written per launch) read / write its known bytes; optional TCC hit /
    miss counts per kernel. Returns the dict stored as profiles/eval_metric_f32_traffic.json."""
    import csv
    import glob
    import collections

    def load(d, name):
        f = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
        if not f:
            raise FileNotFoundError(f"no counter_collection.csv under {d}")
        per = collections.defaultdict(lambda: collections.defaultdict(float))
        for row in csv.DictReader(open(f[0])):
            if not row["Counter_Name"].startswith(name):
                continue
            per[row["Kernel_Name"]][row["Dispatch_Id"]] += float(row["Counter_Value"])
        return per

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    evals = max(len(v) for k, v in fetch.items() if k.startswith("lb_k_finalize"))
    import numpy as np
    probe = lambda per: np.mean([x for k, v in per.items() if k.startswith("lbp_copy_kernel") for x in v.values()])  # noqa: E731
    kib = (256 << 20) / 1024.0
    ff, wf = kib / probe(fetch), kib / probe(write)
    hm = {}
    if hitmiss_dir:
        hit, miss = load(hitmiss_dir, "TCC_HIT"), load(hitmiss_dir, "TCC_MISS")
        hm = {k: (sum(hit[k].values()) / evals, sum(miss.get(k, {}).values()) / evals) for k in hit}
    per_kernel, tf, tw = {}, 0.0, 0.0
    for k in fetch:
        # the chain's kernels only: each runs at least once per evaluation
        if not (k.startswith("lb_k_") or k.startswith("void lb_k_")) or len(fetch[k]) < evals:
            continue
        name = k.replace("void ", "").split("(")[0]
        fk, wk = sum(fetch[k].values()) / evals, sum(write.get(k, {}).values()) / evals
        e = {"fetch_KiB_raw": round(fk, 1), "write_KiB_raw": round(wk, 1), "fetch_bytes": int(fk * 1024 * ff),
             "write_bytes": int(wk * 1024 * wf)}
        if k in hm:
            h_, m_ = hm[k]
            e.update(tcc_hit=int(h_), tcc_miss=int(m_), l2_hit_rate=round(h_ / max(h_ + m_, 1.0), 3))
        per_kernel[name] = e
        tf += e["fetch_bytes"]
        tw += e["write_bytes"]
    return {"workload": workload, "tuning": tuning, "bytes_per_eval": int(tf + tw), "fetch_bytes_per_eval": int(tf),
            "write_bytes_per_eval": int(tw), "calibration": {
                "probe": "lbp_copy_kernel, 268435456 bytes read + 268435456 bytes written per launch",
                "fetch_KiB_per_launch": round(float(probe(fetch)), 1), "write_KiB_per_launch": round(float(probe(write)), 1),
                "fetch_factor": round(ff, 4), "write_factor": round(wf, 4)},
            "evals_profiled": evals, "per_kernel": per_kernel, "commands": list(commands),
            "note": "FETCH_SIZE / WRITE_SIZE summed over the kernels of one evaluation, scaled by the factors that make "
                    "the copy probe read and write its known byte counts."}


def gather_probe(n):
    """The evaluation's row gathers on the metric instance with nothing around them (liblbprobe lbp_gather): the
    neighbour rows of every edge in CSR order and the two other rows of every triangle record in incidence order,
    under the degree numbering the engine uses; FP32 rows (64 B, S = 4 lanes) and BF16-sized rows (32 B, S = 2), from
    global memory only or with the highest-degree rows staged in LDS (as many as a workgroup's 160 KB hold)."""
    import ctypes
    import numpy as np
    from lagbound_amd import instances as I
    from lagbound_amd.vertex_order import VertexOrder
    inst, _ = I.metric_instance(n=n)
    vo = VertexOrder(inst.rowptr)
    g, _ = vo.instance(inst)
    TI = np.asarray(g.TI, np.int64)
    own = np.concatenate([TI[:, 0], TI[:, 1], TI[:, 2]])
    u = np.concatenate([TI[:, 1], TI[:, 0], TI[:, 0]])
    w = np.concatenate([TI[:, 2], TI[:, 2], TI[:, 1]])
    o = np.argsort(own, kind="stable")
    rec = np.stack([u[o], w[o]], axis=1).ravel()
    idx = np.ascontiguousarray(np.concatenate([np.asarray(g.col, np.int64), rec]), np.int32)
    lp = ctypes.CDLL(os.path.join(ROOT, "lagbound_amd", "liblbprobe.so"))
    lp.lbp_gather.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    cnt = np.bincount(idx, minlength=g.n)
    res = {"gathers": int(len(idx)), "n": int(g.n), "runs": []}
    for S in (4, 2):
        H = (160 * 1024) // (S * 16) - 16
        for h, grid in ((0, 1024), (0, 256), (H, 256)):
            for eb in (4, 8):
                ms = ctypes.c_double()
                rc = lp.lbp_gather(idx.ctypes.data, len(idx), g.n, S, eb, grid, h, 5, ctypes.byref(ms))
                r_ = {"row_bytes": 16 * S, "lds_rows": h, "lds_share": round(float(cnt[:h].sum() / len(idx)), 3),
                      "eb": eb, "workgroups": grid, "rc": rc, "us": round(ms.value * 1e3, 1),
                      "rows_per_s": round(len(idx) / (ms.value * 1e-3), -6) if rc == 0 else None}
                res["runs"].append(r_)
                log(f"gather probe: {r_}")
    # reference points: uniform indices over rows that fit L2 (2 MB), the Infinity Cache (64 MB) and neither (256 MB)
    rng = np.random.default_rng(0)
    for nrows in (32768, 1 << 20, 1 << 22):
        uni = np.ascontiguousarray(rng.integers(0, nrows, len(idx)), np.int32)
        ms = ctypes.c_double()
        rc = lp.lbp_gather(uni.ctypes.data, len(uni), nrows, 4, 4, 1024, 0, 5, ctypes.byref(ms))
        r_ = {"uniform_rows": nrows, "row_bytes": 64, "eb": 4, "workgroups": 1024, "rc": rc, "us": round(ms.value * 1e3, 1),
              "rows_per_s": round(len(uni) / (ms.value * 1e-3), -6) if rc == 0 else None}
        res["runs"].append(r_)
        log(f"gather probe: {r_}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--rho", type=float, default=0.02)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-CPU oracle baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the BF16 / FP64 / plain-FP64 comparison runs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--tuning", default="", help="heavy_edges,heavy_tris,seg_len,row_grid")
    ap.add_argument("--option", action="append", default=[], help="name=value kernel switch (lb_set_option)")
    ap.add_argument("--no-autotune", action="store_true", help="use the library's default launch shapes")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream (batched bound latency)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-graph batched evaluation")
    ap.add_argument("--no-batch-search", action="store_true",
                    help="skip the 512-graph batched search with per-graph certificates")
    ap.add_argument("--serve-count", type=int, default=16)
    ap.add_argument("--no-large", action="store_true", help="skip the n=2.4M BF16 evaluation")
    ap.add_argument("--only-large", action="store_true",
                    help="run only the n=2.4M evaluation (with --tuning if given) and print its JSON object")
    ap.add_argument("--summarize-pmc", nargs="+", metavar="DIR", default=None,
                    help="no GPU work: write the traffic summary of rocprofv3 counter runs (FETCH_SIZE dir, WRITE_SIZE "
                         "dir, optional TCC_HIT/TCC_MISS dir) to --pmc-out")
    ap.add_argument("--pmc-out", default=os.path.join(ROOT, "profiles", "eval_metric_f32_traffic.json"))
    ap.add_argument("--pmc-note", action="append", default=[], help="a command line recorded in the summary")
    ap.add_argument("--gather-probe", action="store_true",
                    help="only the row-gather probe on the metric instance (measurement of the gather floor and of "
                         "LDS staging of the hub rows); prints its JSON object")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain copy kernel of known size (for scaling hardware byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.summarize_pmc:
        dirs = a.summarize_pmc + [None] * (3 - len(a.summarize_pmc))
        res = summarize_pmc(dirs[0], dirs[1], dirs[2], tuning=a.tuning, commands=a.pmc_note,
                            workload="f+grad eval, Chung-Lu gamma=2.1 mean-degree~13 (metric instance M), r=16, FP32 V, "
                                     "rho=0.02, vertices numbered by descending degree on the device")
        with open(a.pmc_out, "w") as f:
            json.dump(res, f, indent=1)
        log(f"traffic {res['bytes_per_eval'] / 1e9:.3f} GB per evaluation over {res['evals_profiled']} evaluations "
            f"-> {a.pmc_out}")
        return
    if a.gather_probe:
        print(json.dumps({"workload": "row gathers of the metric instance", **gather_probe(a.n)}), flush=True)
        return
    if a.only_large:
        tl = None
        if a.tuning:
            h, ht, sl, rg = [int(x) for x in a.tuning.split(",")]
            tl = dict(heavy_edges=h, heavy_tris=ht, seg_len=sl, row_grid=rg, row_order=1)
        print(json.dumps({"workload": "f+grad eval, Chung-Lu n=2.4M gamma=2.2 mean-degree~50",
                          **large_graph(tl, a.rho, a.steps)}), flush=True)
        return
    import numpy as np
    from lagbound_amd import instances as I
    from lagbound_amd.engine import Engine

    t0 = time.time()
    inst, V = I.metric_instance(n=a.n)
    r = V.shape[1]
    log(f"instance n={inst.n} nnz={inst.nnz} T={inst.T} ws={inst.vcs[0].nnz} built in {time.time() - t0:.1f}s")
    rho = a.rho

    eng = Engine(r=r, dtype="f32")
    tune = [int(x) for x in a.tuning.split(",")] if a.tuning else None
    tuned = None
    if tune:
        eng.set_tuning(*tune)
    elif not a.no_autotune:
        # launch shapes picked by the autotuner on this instance (hub threshold, segment length, row grid)
        from lagbound_amd import tune as TU
        t1 = time.time()
        tuned, _ = TU.autotune(inst, V, rho, r=r, grid={"heavy": [64, 128, 256, 512], "seg_len": [256, 512, 1024],
                                                        "row_grid": [2048, 4096]}, iters=10)
        TU.apply(eng, tuned)
        log(f"autotuned {tuned} in {time.time() - t1:.1f}s")
    for opt in a.option:
        k_, v_ = opt.split("=")
        eng.set_option(k_, int(v_))
    eng.set_instance(inst)
    eng.set_V(V)
    eng.eval(rho)
    sec = time_engine(eng, rho, a.steps, a.warmup)
    value = 1.0 / sec
    log(f"f32: {sec * 1e3:.4f} ms/eval, {value:.1f} evals/s")
    ms_ev, kms = eng.time_eval(rho, iters=max(20, a.steps // 4), per_kernel=True)
    log("kernel ms:", json.dumps({k: round(v, 5) for k, v in kms.items()}))
    B = alg_bytes(inst, r, 4, 4, 4)
    ach = B / (ms_ev * 1e-3) / 1e9
    dom = max(kms, key=kms.get)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "eval_metric_f32_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_eval")
        except Exception:
            traffic = None
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(sec * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "f+grad eval, Chung-Lu gamma=2.1 mean-degree~13 (metric instance M)", "n": inst.n,
                   "nnz": inst.nnz, "r": r, "T": inst.T, "pair_P": inst.levels[0].P, "vcs_P": inst.vcs[0].P,
                   "vcs_ws": inst.vcs[0].nnz, "rho": rho, "V": "fp32, fp64 global sums",
                   "launch_shapes": tuned if tuned else ({"tuning": a.tuning} if a.tuning else "library defaults")},
        "roofline": {"bound": "hbm", "kernel": "one f+grad evaluation (lb_k_colsum .. lb_k_finalize chain)",
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": "profiles/eval_metric_f32_traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE "
                                       "per evaluation, scaled by the copy probe's factors (not measured in this run)",
                     "alg_bytes": B, "ms_per_eval_events": round(ms_ev, 5),
                     "dominant_kernel": dom, "kernel_ms": {k: round(v, 5) for k, v in kms.items()},
                     "kernel_ms_note": "per-kernel HIP events with the chain on one stream; the timed evaluation runs "
                                       "lb_k_segs on a second stream beside lb_k_colsum .. lb_k_reduce(Z)"},
    }
    if not a.no_extra:
        extra = {}
        for dt, sv, sc in (("bf16", 2, 4), ("f64", 8, 8)):
            e2 = Engine(r=r, dtype=dt)
            if tuned:                                   # the FP32 run's launch shapes, so the variants compare
                TU.apply(e2, tuned)
            elif tune:
                e2.set_tuning(*tune)
            e2.set_instance(inst)
            e2.set_V(V)
            s2 = time_engine(e2, rho, max(20, a.steps // 2), 5)
            ms2, _ = e2.time_eval(rho, iters=20, per_kernel=False)
            b2 = alg_bytes(inst, r, sv, 4 if dt == "bf16" else 8, sc)
            extra[dt] = {"evals_per_s": round(1.0 / s2, 2), "ms_per_eval": round(s2 * 1e3, 5),
                         "hbm_frac": round(b2 / (ms2 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            log(f"{dt}: {s2 * 1e3:.4f} ms/eval")
            e2.close()
        # the plain FP64 GPU form runs in a process of its own (torch brings its own HIP runtime)
        import subprocess
        try:
            res = subprocess.run([sys.executable, "-m", "lagbound_amd._fp64_baseline_main", "--n", str(a.n),
                                  "--rho", str(rho), "--steps", "10"], capture_output=True, text=True, timeout=600,
                                 cwd=ROOT)
            line = [l for l in res.stdout.splitlines() if l.startswith("{")]
            if res.returncode != 0 or not line:
                raise RuntimeError(res.stderr[-300:])
            pf = json.loads(line[-1])
            extra["plain_fp64_gpu"] = pf
            out["speedup_vs_plain_fp64_gpu"] = round(value * pf["ms_per_eval"] * 1e-3, 2)
            # beside it, the like-for-like ratios: the FP64 HIP path against the plain FP64 form, and the FP32
            # headline against the FP64 HIP path
            if "f64" in extra:
                out["speedup_f64_hip_vs_plain_fp64_gpu"] = round(pf["ms_per_eval"] / extra["f64"]["ms_per_eval"], 2)
                out["speedup_f32_vs_f64_hip"] = round(value * extra["f64"]["ms_per_eval"] * 1e-3, 3)
            log(f"plain fp64 gpu: {pf['ms_per_eval']:.3f} ms/eval")
        except Exception as exc:  # the comparison leg is reported, never silently replaced
            extra["plain_fp64_gpu"] = {"error": repr(exc)[:300]}
        # exact reduction (device) on a 300k-vertex power-law graph with planted twins and leaves; the reference's
        # reduction is a serial host loop, timed here as the single-threaded C oracle on the same graph
        from lagbound_amd import reduce as RD
        nr, kr = I.chung_lu(300000, 2.1, 13, 11)
        nr, kr = I.add_twins_and_leaves(nr, kr, 5000, 20000, 12)
        RD.reduce_graph(nr, kr)                                        # warm-up (module load, allocations)
        tr = []
        for _ in range(5):
            t1 = time.perf_counter()
            resr = RD.reduce_graph(nr, kr)
            tr.append(time.perf_counter() - t1)
        red = {"n": nr, "m": len(kr), "groups": resr["N"], "ms_gpu_call": round(1e3 * float(np.median(tr)), 3),
               "note": "host keys in, host results out (PCIe and context set-up included)"}
        if not a.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            t1 = time.perf_counter()
            ro = O.reduce_c(nr, kr)
            red["ms_cpu_oracle_1thread"] = round(1e3 * (time.perf_counter() - t1), 3)
            red["bit_exact_vs_oracle"] = all(np.array_equal(np.asarray(ro[k]), np.asarray(resr[k])) for k in ro)
        extra["reduction_300k"] = red
        log(f"reduction n={nr} m={len(kr)}: {red['ms_gpu_call']:.2f} ms (device call)")
        if not a.no_large:
            extra["chung_lu_2400k"] = large_graph(None, rho, a.steps)     # the library's default launch shapes
        out["variants"] = extra
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cores = os.cpu_count() or 1
        inc = O.incidence_c(inst.n, inst.TI)
        O.eval_c(inst, V, rho, threads=cores, incidence=inc)   # warm-up (page-in)
        ts = []
        t_start = time.perf_counter()
        while len(ts) < 3 or (time.perf_counter() - t_start < a.cpu_seconds and len(ts) < 50):
            t1 = time.perf_counter()
            O.eval_c(inst, V, rho, threads=cores, incidence=inc)
            ts.append(time.perf_counter() - t1)
        med = float(np.median(ts))
        out["cpu_baseline"] = {"value": round(1.0 / med, 4), "unit": "evals/s", "cores": cores, "kind": "port",
                               "sample": f"{len(ts)} full FP64 evaluations of the same instance (C/OpenMP oracle, "
                                         f"reference loop structure), median {med:.3f}s"}
        out["speedup_vs_cpu"] = round(value * med, 1)
        log(f"cpu oracle ({cores} threads): {med:.3f} s/eval")
        t1s = []
        for _ in range(2):                                  # the 1-thread figure (two evaluations)
            t1 = time.perf_counter()
            O.eval_c(inst, V, rho, threads=1, incidence=inc)
            t1s.append(time.perf_counter() - t1)
        out["cpu_baseline"]["single_thread"] = {"value": round(1.0 / min(t1s), 4), "unit": "evals/s", "cores": 1,
                                                "sample": "2 full FP64 evaluations, best"}
        log(f"cpu oracle (1 thread): {min(t1s):.3f} s/eval")
    # batched bound latency: a Poisson stream of whole bound requests (lagbound_amd.serve) and the batched
    # evaluation of 512 uneven graphs (lagbound_amd.batch), each in a process of its own
    import subprocess

    def _module_json(args, timeout):
        res = subprocess.run([sys.executable, "-m"] + args, capture_output=True, text=True, timeout=timeout, cwd=ROOT)
        line = [l for l in res.stdout.splitlines() if l.startswith("{")]
        if res.returncode != 0 or not line:
            return {"error": (res.stderr or "")[-400:]}
        return json.loads(line[-1])

    if not a.no_serve:
        out["serving"] = _module_json(["lagbound_amd.serve", "--count", str(a.serve_count), "--rate", "200"], 900)
        sv = out["serving"]
        if "latency_p50_s" in sv:
            log(f"serving: p50 {sv['latency_p50_s']} s, p99 {sv['latency_p99_s']} s, {sv['solves_per_s']} solves/s")
    if not a.no_batch:
        out["batch"] = _module_json(["lagbound_amd.batch", "--count", "512"], 900)
        bt = out["batch"]
        if "ms_per_batch_eval" in bt:
            log(f"batch: {bt['ms_per_batch_eval']} ms per evaluation of {bt['graphs']} graphs "
                f"({bt['graph_evals_per_s']} graph evaluations/s)")
    if not a.no_batch_search:
        # the batched search over 512 graphs of the same distribution (captured iterations of evaluation +
        # per-graph step, PHR updates, per-graph triangle budgets), then every graph's FP64 certificate
        out["batch_search"] = _module_json(["lagbound_amd.batch", "--count", "512", "--search", "1"], 1200)
        bs = out["batch_search"]
        if "search_s" in bs:
            log(f"batched search: {bs['search_s']} s ({bs['ms_per_captured_iteration']} ms per captured iteration of "
                f"{bs['graphs']} graphs), certificates {bs['finish_s']} s, certified {bs['certified']}")
    if a.calibrate:
        import ctypes
        lp = ctypes.CDLL(os.path.join(ROOT, "lagbound_amd", "liblbprobe.so"))
        lp.lbp_copy.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        cms = ctypes.c_double()
        nbytes = 256 << 20
        rc = lp.lbp_copy(nbytes, 5, ctypes.byref(cms))
        out["copy_probe"] = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": round(cms.value, 5),
                             "GBps": round(2 * nbytes / (cms.value * 1e-3) / 1e9, 1), "rc": rc}
        log(f"copy probe: {2 * nbytes / (cms.value * 1e-3) / 1e9:.1f} GB/s (read+write)")
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
