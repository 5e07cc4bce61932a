"""Times the plain FP64 GPU form (baseline_fp64.PlainFP64) on the metric instance in a process of its own (torch
bundles its own HIP runtime, which must not share a process with liblbhip's). Prints one JSON line."""
import argparse
import json
import sys
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--rho", type=float, default=0.02)
    ap.add_argument("--steps", type=int, default=20)
    a = ap.parse_args()
    import torch
    from lagbound_amd import instances as I
    from lagbound_amd.baseline_fp64 import PlainFP64
    inst, V = I.metric_instance(n=a.n)
    pl = PlainFP64(inst, device="cuda")
    Vt = torch.as_tensor(V, dtype=torch.float64, device="cuda")
    for _ in range(3):
        f, G = pl.eval(Vt, a.rho)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        f, G = pl.eval(Vt, a.rho)
    torch.cuda.synchronize()
    s = (time.perf_counter() - t0) / a.steps
    print(json.dumps({"evals_per_s": round(1.0 / s, 3), "ms_per_eval": round(s * 1e3, 4), "f": float(f),
                      "form": "torch FP64: CSR SpMM, gathers, sparse products for the cluster sums and the triangle scatter",
                      "n": inst.n}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
