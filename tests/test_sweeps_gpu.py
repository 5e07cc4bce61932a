"""Parity of lb_sweeps (batched row steps on the device) with the FP64 oracle (same orders, same batches, batch
snapshot semantics). Tolerances: FP64 V 1e-10 max abs; FP32 V 2e-4 after the sweeps (FP32 row arithmetic, errors
compound over steps); BF16-stored V 5e-2 (every stored row is rounded to 8 bits of mantissa while the oracle keeps
FP64), plus unit rows to storage precision."""
import numpy as np
import pytest

import oracle as O
from conftest import small_instance, chung_lu_instance

pytestmark = pytest.mark.gpu


def _run(inst, V, dtype, r, orders, nbatch, rho, damp, tuning=None, options=None):
    from lagbound_amd.engine import Engine
    eng = Engine(r=r, dtype=dtype)
    if tuning:
        eng.set_tuning(*tuning)
    for k, v in (options or {}).items():
        eng.set_option(k, v)
    eng.set_instance(inst)
    eng.set_V(V)
    V0 = eng.get_V()
    print(f"[sweeps] dtype={dtype} n={inst.n} nbatch={nbatch} rho={rho} nsw={len(orders)}", flush=True)
    eng.sweeps(orders, nbatch, rho, damp=damp)
    W = eng.get_V()
    eng.close()
    return V0, W


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-10), ("f32", 2e-4)])
@pytest.mark.parametrize("rho,nbatch", [(0.0, 4), (0.3, 8), (0.3, 1)])
def test_small_sweeps(dtype, tol, rho, nbatch):
    inst, V = small_instance(n=60, seed=3, k=16, sig_scale=0.05)
    rng = np.random.default_rng(4)
    orders = np.stack([rng.permutation(inst.n) for _ in range(2)])
    V0, W = _run(inst, V, dtype, 16, orders, nbatch, rho, 0.1)
    Wo = O.sweeps_c(inst, V0, orders, nbatch, rho, damp=0.1)
    assert np.abs(W - Wo).max() <= tol
    assert np.abs(np.linalg.norm(W, axis=1) - 1).max() <= (1e-12 if dtype == "f64" else 1e-5)


@pytest.mark.parametrize("step_wave", [1, 0])
@pytest.mark.parametrize("dtype,tol", [("f64", 1e-9), ("f32", 5e-4)])
def test_chung_lu_sweeps_with_hubs(dtype, tol, step_wave):
    """Hub rows (segment path) inside batches; weighted/unweighted edges; three sweeps; with the row step as one
    wavefront per batch row (the default) and as one slot per row."""
    inst, V = chung_lu_instance(n=20000, seed=5, k=16)
    rng = np.random.default_rng(6)
    orders = np.stack([rng.permutation(inst.n) for _ in range(3)])
    V0, W = _run(inst, V, dtype, 16, orders, 16, 0.02, 0.0, options={"step_wave": step_wave})
    Wo = O.sweeps_c(inst, V0, orders, 16, 0.02, damp=0.0)
    assert np.abs(W - Wo).max() <= tol


@pytest.mark.parametrize("r", [8, 32])
@pytest.mark.parametrize("step_wave", [1, 0])
def test_sweeps_other_ranks(r, step_wave):
    """Ranks 8 and 32 (2 and 8 lanes per row slot, so 32 and 8 slots per wavefront in the wavefront-per-row step)."""
    inst, V = small_instance(n=70, seed=12, k=r, ntri=60, sig_scale=0.05)
    rng = np.random.default_rng(13)
    orders = np.stack([rng.permutation(inst.n) for _ in range(2)])
    V0, W = _run(inst, V, "f64", r, orders, 6, 0.3, 0.1, options={"step_wave": step_wave})
    Wo = O.sweeps_c(inst, V0, orders, 6, 0.3, damp=0.1)
    assert np.abs(W - Wo).max() <= 1e-10


def test_segments_forced_small():
    """Tiny hub thresholds for the evaluation's tables and for the sweeps' own (rows above 2 edges / records split
    into 3-item segments), so the segment paths of both run on a small weighted instance."""
    inst, V = small_instance(n=50, seed=6, k=8, weighted=True, ntri=80)
    rng = np.random.default_rng(7)
    orders = np.stack([rng.permutation(inst.n) for _ in range(2)])
    V0, W = _run(inst, V, "f64", 8, orders, 5, 0.5, 0.2, tuning=(2, 2, 3), options={"sweep_heavy": 2, "sweep_seg": 3})
    Wo = O.sweeps_c(inst, V0, orders, 5, 0.5, damp=0.2)
    assert np.abs(W - Wo).max() <= 1e-10


def test_bf16_sweeps():
    inst, V = chung_lu_instance(n=5000, seed=8, k=16)
    rng = np.random.default_rng(9)
    orders = rng.permutation(inst.n)[None, :]
    V0, W = _run(inst, V, "bf16", 16, orders, 8, 0.02, 0.0)
    Wo = O.sweeps_c(inst, V0, orders, 8, 0.02, damp=0.0)
    assert np.abs(W - Wo).max() <= 5e-2
    assert np.abs(np.linalg.norm(W, axis=1) - 1).max() <= 2e-2


def test_sweeps_then_eval_consistent():
    """After sweeps the resident V feeds the evaluation unchanged (no stale aggregates)."""
    from lagbound_amd.engine import Engine
    inst, V = chung_lu_instance(n=8000, seed=10, k=16)
    eng = Engine(r=16, dtype="f64")
    eng.set_instance(inst)
    eng.set_V(V)
    rng = np.random.default_rng(11)
    eng.sweeps(rng.permutation(inst.n)[None, :], 8, 0.02)
    parts = eng.eval(0.02)
    po, Go, _, _, _ = O.eval_c(inst, eng.get_V(), 0.02)
    assert abs(parts[0] - po[0]) <= 1e-10 * max(1.0, abs(po[0]))
    eng.close()


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_captured_sweeps_equal_eager(dtype):
    """A sweep's per-batch launches run as a captured graph (option sweep_graph 1) or eagerly (the default): the same
    kernels with the same arguments, so V after the sweeps agrees to the run-to-run rounding of the FP64 LDS adds in
    the aggregate increments (1e-12 absolute; FP32 V is the same bit for bit in practice). Three calls per engine: the
    second replays the graph captured by the first (new orders, so new hub-segment lists through the device copy of
    their bounds), the third changes rho and must capture again; the third call is also checked against the oracle
    from the same start."""
    from lagbound_amd.engine import Engine
    inst, V = chung_lu_instance(n=6000, seed=5, k=16, P=16)
    rng = np.random.default_rng(9)
    calls = [(np.stack([rng.permutation(inst.n) for _ in range(2)]), 0.3), (np.stack([rng.permutation(inst.n)]), 0.3),
             (np.stack([rng.permutation(inst.n)]), 0.5)]
    out = {}
    for graph in (1, 0):
        eng = Engine(r=16, dtype=dtype)
        eng.set_option("sweep_graph", graph)
        eng.set_instance(inst)
        eng.set_V(V)
        V0 = eng.get_V()
        Ws = []
        for orders, rho in calls:
            eng.sweeps(orders, 16, rho, damp=0.1)
            Ws.append(eng.get_V())
        out[graph] = (V0, Ws)
        eng.close()
    for a, b in zip(out[1][1], out[0][1]):
        assert np.abs(a - b).max() <= (1e-12 if dtype == "f64" else 1e-6)
    Ws = out[1][1]
    Wo = O.sweeps_c(inst, Ws[-2], calls[-1][0], 16, calls[-1][1], damp=0.1)
    tol = 1e-10 if dtype == "f64" else 2e-4
    assert np.abs(Ws[-1] - Wo).max() <= tol


@pytest.mark.skip(reason="the fused sweep form (option sweep_fused 1) is switched off: one run of the 16-request "
                         "serving stream with it ended in a GPU memory fault inside a sweep, cause not established")
@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_fused_sweeps_equal_five_kernel_form(dtype):
    """Option sweep_fused 1 (the default: two launches per batch, the rows adding their increments to the cluster sums
    themselves, hub rows finished by whichever of their parts arrives last) against sweep_fused 0 (five launches: pair
    rows, hub segments, row steps, aggregate increments, VCS post step). Same row arithmetic; only the FP64 adds into
    the sums run in another order, so V agrees to 1e-12 (FP64 V) or to a unit of the storage precision where such a
    difference moves a rounding (FP32 2e-6, BF16 3e-2). Small hub thresholds (rows above 8 edges or records split
    into 5-item segments, so most rows are hubs with several segments of both kinds), three sweeps, pair and VCS
    levels; the fused form is also captured as a graph and replayed (second call), and its last call is checked
    against the oracle."""
    from lagbound_amd.engine import Engine
    inst, V = chung_lu_instance(n=5000, seed=21, k=16, P=16)
    rng = np.random.default_rng(22)
    calls = [np.stack([rng.permutation(inst.n) for _ in range(2)]), np.stack([rng.permutation(inst.n)])]
    out = {}
    for fused, graph in ((1, 0), (0, 0), (1, 1)):
        eng = Engine(r=16, dtype=dtype)
        for k, v in {"sweep_fused": fused, "sweep_graph": graph, "sweep_heavy": 8, "sweep_seg": 5}.items():
            eng.set_option(k, v)
        eng.set_instance(inst)
        eng.set_V(V)
        Ws = [eng.get_V()]
        print(f"[sweeps] fused={fused} graph={graph} dtype={dtype} n={inst.n}", flush=True)
        for orders in calls:
            eng.sweeps(orders, 12, 0.05, damp=0.1)
            Ws.append(eng.get_V())
        out[(fused, graph)] = Ws
        eng.close()
    tol = {"f64": 1e-12, "f32": 2e-6, "bf16": 3e-2}[dtype]
    for key in ((1, 1), (0, 0)):
        for a, b in zip(out[(1, 0)][1:], out[key][1:]):
            assert np.abs(a - b).max() <= tol, key
    Ws = out[(1, 0)]
    Wo = O.sweeps_c(inst, Ws[1], calls[1], 12, 0.05, damp=0.1)
    otol = {"f64": 1e-10, "f32": 2e-4, "bf16": 5e-2}[dtype]
    assert np.abs(Ws[2] - Wo).max() <= otol


@pytest.mark.parametrize("nbatch", [12, 64])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_sweep_launch_forms_agree(dtype, nbatch):
    """The sweep batch in two launches (the default: row steps adding their increments to a per-batch delta buffer,
    with a snapshot of Y; then the deltas applied to the aggregates, the VCS post step, the next batch's pair rows and
    hub segments, the last two reading the snapshot plus the deltas and, for a row just stepped, its new value from the
    staging rows -- above 128 rows per batch; below, and with option sweep_delta 0, each workgroup sums the batch's
    increments in its own LDS instead), in three or four (option sweep_two
    0: pair rows with hub segments, row steps, then the aggregate increments with the VCS post step in one launch for
    batches of at most 256 rows) and in five (pair rows, hub segments, row steps, aggregate increments, VCS post step
    apart). Same per-wavefront arithmetic, so V agrees to the run-to-run rounding of the FP64 aggregate adds (1e-12
    on FP64 V, 1e-6 on FP32 V). Small hub thresholds (most rows are hubs with several segments of both kinds), two
    calls of sweeps at 12 batches (417 rows) and 64 (78 rows); the default form is also checked against the oracle."""
    from lagbound_amd.engine import Engine
    inst, V = chung_lu_instance(n=5000, seed=31, k=16, P=16)
    rng = np.random.default_rng(32)
    calls = [np.stack([rng.permutation(inst.n) for _ in range(2)]), np.stack([rng.permutation(inst.n)])]
    forms = {"two": {}, "two_lds": {"sweep_delta": 0}, "three": {"sweep_two": 0},
             "five": {"sweep_two": 0, "sweep_pair_segs": 0, "sweep_post_merge": 0}}
    out = {}
    for name, opts in forms.items():
        eng = Engine(r=16, dtype=dtype)
        for k, v in dict(opts, sweep_heavy=8, sweep_seg=5).items():
            eng.set_option(k, v)
        eng.set_instance(inst)
        eng.set_V(V)
        Ws = [eng.get_V()]
        print(f"[sweeps] form={name} dtype={dtype} n={inst.n} nbatch={nbatch}", flush=True)
        for orders in calls:
            eng.sweeps(orders, nbatch, 0.05, damp=0.1)
            Ws.append(eng.get_V())
        out[name] = Ws
        eng.close()
    for name in ("two", "two_lds", "three"):
        for a, b in zip(out[name][1:], out["five"][1:]):
            assert np.abs(a - b).max() <= (1e-12 if dtype == "f64" else 1e-6), name
    Wo = O.sweeps_c(inst, out["two"][1], calls[1], nbatch, 0.05, damp=0.1)
    assert np.abs(out["two"][2] - Wo).max() <= (1e-10 if dtype == "f64" else 2e-4)
