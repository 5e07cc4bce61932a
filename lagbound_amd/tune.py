"""Autotuner of the evaluation's launch shapes. Single graph: hub thresholds (rows with more edges / triangle records
than the threshold are split into wavefront segments), segment length, row-kernel grid and the row order switch;
each candidate is uploaded to a fresh context, warmed up and timed with device events (lb_time_eval). Batch of
graphs (autotune_batch): the light-row threshold of the batch row kernels (rows up to it run one slot each, longer
rows one wavefront each) and the grid caps of both kernels; each candidate is set on one batch context and timed as
captured replays of the batch evaluation. The fastest is returned and applied with `apply` / `apply_batch`. No
reference counterpart (the reference's parallel shapes are numba prange and fixed chunk sizes)."""
from __future__ import annotations

import itertools

from .engine import Engine
from .vertex_order import VertexOrder

DEFAULT_GRID = {
    "heavy": [64, 128, 256, 512],
    "seg_len": [256, 512, 1024],
    "row_grid": [2048, 4096],
    "row_order": [0, 1],
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


BATCH_GRID = {
    "batch_light": [32, 64, 128, 256, 512],
    "batch_light_grid": [2048, 8192, 32768],
    "batch_heavy_grid": [1024, 8192],
}


def batch_candidates(grid=None):
    g = dict(BATCH_GRID, **(grid or {}))
    for lt, lg, hg in itertools.product(g["batch_light"], g["batch_light_grid"], g["batch_heavy_grid"]):
        yield dict(batch_light=lt, batch_light_grid=lg, batch_heavy_grid=hg)


def apply_batch(eng: Engine, cfg):
    """Set a batch configuration on an engine holding a batch (takes effect at its next evaluation)."""
    for k, v in cfg.items():
        eng.set_option(k, int(v))


def autotune_batch(be, rho=0.02, reps=20, grid=None, log=None):
    """Time every batch candidate on a BatchEngine `be` (its current V and multipliers) as `reps` captured replays of
    the batch evaluation (host clock around a synchronized run); leaves the fastest applied. Returns (best config,
    [(config, ms per evaluation)])."""
    import time
    results = []
    for cfg in batch_candidates(grid):
        apply_batch(be.eng, cfg)
        be.eng.capture_eval(rho)
        be.eng.replay_eval(3, fetch=False)
        be.sync()
        t = time.perf_counter()
        be.eng.replay_eval(reps, fetch=False)
        be.sync()
        ms = 1e3 * (time.perf_counter() - t) / reps
        results.append((cfg, ms))
        if log:
            log(f"tune batch {cfg}: {ms:.4f} ms")
    best = min(results, key=lambda x: x[1])[0]
    apply_batch(be.eng, best)
    return best, results
