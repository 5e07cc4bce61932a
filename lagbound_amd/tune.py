"""Autotuner of the single-graph evaluation's launch shapes: hub thresholds (rows with more edges / triangle records
than the threshold are split into wavefront segments), segment length, row-kernel grid and the row order switch.
Each candidate is uploaded to a fresh context, warmed up and timed with device events (lb_time_eval); the fastest
is returned and can be applied to any engine with `apply`. No reference counterpart (the reference's parallel
shapes are numba prange and fixed chunk sizes)."""
from __future__ import annotations

import itertools

from .engine import Engine
from .vertex_order import VertexOrder

DEFAULT_GRID = {
    "heavy": [64, 128, 256, 512],
    "seg_len": [256, 512, 1024],
    "row_grid": [2048, 4096],
    "row_order": [1],
}


def candidates(grid=None):
    g = dict(DEFAULT_GRID, **(grid or {}))
    for h, s, rg, ro in itertools.product(g["heavy"], g["seg_len"], g["row_grid"], g["row_order"]):
        if s >= h:                                  # a segment at least as long as the hub threshold
            yield dict(heavy_edges=h, heavy_tris=h, seg_len=s, row_grid=rg, row_order=ro)


def apply(eng: Engine, cfg):
    """Set a configuration on an engine (takes effect at its next set_instance)."""
    eng.set_tuning(cfg["heavy_edges"], cfg["heavy_tris"], cfg["seg_len"], cfg["row_grid"])
    eng.set_option("row_order", cfg["row_order"])


def autotune(inst, V, rho=0.02, r=None, dtype="f32", iters=20, grid=None, device=0, log=None, order="degree"):
    """Time every candidate on (inst, V) in the vertex order the engines use (relabelled once here); returns
    (best config, [(config, ms per evaluation)])."""
    r = r or V.shape[1]
    if order == "degree":
        vo = VertexOrder(inst.rowptr)
        inst, _ = vo.instance(inst)
        V = vo.rows_in(V)
    results = []
    for cfg in candidates(grid):
        e = Engine(r=r, dtype=dtype, device=device, order="given")
        try:
            apply(e, cfg)
            e.set_instance(inst)
            e.set_V(V)
            e.eval(rho)
            ms, _ = e.time_eval(rho, iters=iters, per_kernel=False)
        finally:
            e.close()
        results.append((cfg, ms))
        if log:
            log(f"tune {cfg}: {ms:.4f} ms")
    best = min(results, key=lambda x: x[1])[0]
    return best, results
