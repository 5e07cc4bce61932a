"""Drop-in for lagbound's reduction, solver and estimate stage calls (pipeline.Workspace.reduce, pipeline.py:96-103;
Workspace.solve, pipeline.py:105-136; Workspace.estimate, pipeline.py:156-170).

`solve(ws, ...)` takes the reference's Workspace (anything with `.data`, `.root`, `.runs` and `.run`-style
subprocess conventions) and runs the device solver `python -m lagbound_amd.alm` with exactly the production flag
set of the reference (main phase or proximal phase), then returns (last checkpoint path, pair weight) parsed from
the `-> pairlvl_weight X` log line -- the same return value as Workspace.solve. `estimate(ws, ...)` runs
`python -m lagbound_amd.estimate` with the reference's estimate arguments and returns its result JSON.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys

TAG = "g"


class PipelineError(RuntimeError):
    """The solver stage failed; the message ends with the tail of its log."""


def solver_args(name, pairs, vc, lb, time_limit, outer, nbatch, runs_rel="runs", init=None, resume=None,
                pair_weight=None, proximal=False, max_clusters=6000, ckpt_every=None, tag=TAG + "__rx"):
    """The argument list of the production launch (pipeline.py:116-128)."""
    args = ["--tag", tag, "--pair_labels"] + list(pairs) + [
        "--max_clusters", max_clusters, "--k", 32, "--nbatch", nbatch, "--inner", 1, "--est_every", 0,
        "--tri_max", 16000000, "--mom", 0, "--lb", lb, "--shift_p", 0, "--time_limit", time_limit, "--seed", 0,
        "--vcs_labels", vc, "--vcs_max", 150000000, "--vcs_add", 10000000, "--outer", outer,
        "--save", os.path.join(runs_rel, name + ".npz")]
    if proximal:
        args += ["--resume", resume, "--tri_add", 0, "--tri1_add", 0, "--tri_every", 1000000,
                 "--pairlvl_weight", pair_weight, "--phase2_at_it", 1, "--phase2_w", pair_weight,
                 "--phase2_inner", 4, "--phase2_tri_every", 1000000, "--phase2_est_every", 0,
                 "--ckpt_every", ckpt_every or 20]
    else:
        args += ["--tri_add", 600000, "--tri1_add", 300000, "--tri_every", 9, "--init_partition", init,
                 "--pw_rule", "default", "--ckpt_every", ckpt_every or 50]
    return [str(x) for x in args]


def _run_module(ws, module, args, log_name):
    """Run `python -m module args` in the workspace root with DATA_DIR set (Workspace.run, pipeline.py:58-75);
    raise PipelineError with the log's tail on failure. Returns the log path."""
    env = dict(os.environ, DATA_DIR=ws.data + os.sep)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    log_path = os.path.join(ws.runs, log_name)
    with open(log_path, "w") as fh:
        proc = subprocess.run([sys.executable, "-m", module] + [str(x) for x in args], cwd=ws.root, env=env,
                              stdout=fh, stderr=subprocess.STDOUT)
    if proc.returncode != 0:
        with open(log_path) as fh:
            tail = "".join(fh.readlines()[-15:])
        raise PipelineError(f"{module} failed (exit {proc.returncode}); log {log_path}:\n{tail}")
    return log_path


def reduce(ws, label_files, tag=TAG):
    """Workspace.reduce with the device reduction (pipeline.py:96-99): `python -m lagbound_amd.reduce <tag> labels`
    in the workspace root; returns the reduced tag."""
    _run_module(ws, "lagbound_amd.reduce", [tag] + list(label_files), "reduce.log")
    return tag + "__rx"


def estimate(ws, name, ckpt, lb, init_v=None, save_w=None, polish_time=900, eig_time=900, extra_args=(),
             tag=TAG + "__rx"):
    """Workspace.estimate with the device operator (pipeline.py:156-170): the reference's argument list, run as
    `python -m lagbound_amd.estimate`; returns the result JSON as a dict with its path under "json"."""
    out = os.path.join(ws.runs, name + ".json")
    args = ["--tag", tag, "--ckpt", ckpt, "--lb", lb, "--polish", "rgd", "--polish_time", polish_time,
            "--scale", "vol", "--eig_time", eig_time, "--out", out]
    if init_v:
        args += ["--initV", init_v]
    if save_w:
        args += ["--saveW", save_w]
    _run_module(ws, "lagbound_amd.estimate", args + list(extra_args), name + ".log")
    import json
    with open(out) as fh:
        res = json.load(fh)
    res["json"] = out
    return res


def solve(ws, name, pairs, vc, lb, time_limit, outer, nbatch, init=None, resume=None, pair_weight=None,
          proximal=False, max_clusters=6000, ckpt_every=None, extra_args=(), k=None):
    """Workspace.solve with the device solver. `extra_args` (e.g. ["--dtype", "bf16"]) are appended; `k` overrides
    the rank (the device supports 8, 16, 32)."""
    args = solver_args(name, pairs, vc, lb, time_limit, outer, nbatch, init=init, resume=resume,
                       pair_weight=pair_weight, proximal=proximal, max_clusters=max_clusters, ckpt_every=ckpt_every)
    if k is not None:
        args[args.index("--k") + 1] = str(k)
    args += [str(x) for x in extra_args]
    log_path = _run_module(ws, "lagbound_amd.alm", args, name + ".log")
    with open(log_path) as fh:
        weight = parse_weight(fh.read(), pair_weight)
    return os.path.join(ws.runs, name + "_last.npz"), weight


def parse_weight(log_text, default=None):
    """The last `-> pairlvl_weight X` value of a solver log (pipeline.py:131-135), or `default`."""
    weight = default
    for hit in re.finditer(r"-> pairlvl_weight ([0-9.eE+-]+)", log_text):
        weight = float(hit.group(1))
    return weight


class Workspace:
    """Minimal working directory with the reference's layout (data/<tag>/..., runs/) for running the solver stage
    on an instance without the rest of the pipeline."""

    def __init__(self, root):
        self.root = os.path.abspath(root)
        self.data = os.path.join(self.root, "data")
        self.runs = os.path.join(self.root, "runs")
        os.makedirs(self.runs, exist_ok=True)

    def solve(self, *args, **kw):
        return solve(self, *args, **kw)

    def estimate(self, *args, **kw):
        return estimate(self, *args, **kw)

    def reduce(self, label_files):
        return reduce(self, label_files)

    def reduced_labels(self, label_file, tag=TAG):
        """Path of a label file restricted to the reduced instance (pipeline.py:101-103)."""
        return os.path.join(self.data, tag + "__rx", os.path.basename(label_file))
