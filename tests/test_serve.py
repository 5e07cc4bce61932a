"""Serving loop host logic on the CPU (device engine replaced by the FP64 oracle, thread slots): admission by
footprint, every request answered, the same results alone and in the mix, latency summary fields; process slots
start, report a request that fails in a worker, and stop."""
import numpy as np
import pytest

from lagbound_amd import alm, serve
from lagbound_amd import bound as B
from oracle_engine import OracleEngine, OracleLagrangian


@pytest.fixture
def cpu_engine(monkeypatch):
    monkeypatch.setattr(alm, "Engine", OracleEngine)
    monkeypatch.setattr(B, "DeviceLagrangian", lambda inst, r, dtype="f64", device=0: _Lag(inst, r))


class _Lag(OracleLagrangian):
    def __init__(self, inst, r):
        super().__init__(inst, r)
        self.eng = OracleEngine(r)
        self.eng.lib_s = 0.0


def _params():
    # requests above 250 vertices take the lifted sparse certificate (host core), the others the dense one
    return serve.Params(k=8, outer=6, tri_every=3, vcs_every=2, rgd_iters=20, eig_maxiter=40, cert_max_n=400,
                        cert_dense_max_n=250, cert_on_device=False)


def test_footprint_and_rule():
    p = serve.Params()
    assert serve.footprint_bytes(10 ** 6, 13 * 10 ** 6, 16, p) > serve.footprint_bytes(10 ** 4, 10 ** 5, 16, p)
    assert serve.nbatch_rule(1000) == 15 and serve.nbatch_rule(40000) == 512 and serve.nbatch_rule(100) == 8


def test_alone_equals_mixed_and_admission(cpu_engine):
    OracleEngine.lib_s = 0.0
    reqs = serve.poisson_stream(5, 200.0, 150, 400, 6, 12, seed=3)
    p = _params()
    mixed = serve.Server(p, slots=3, workers="thread").run(reqs, realtime=False)
    # a budget below any single footprint serialises the requests, but every one is still served
    tight = serve.Server(p, slots=3, budget_bytes=1, workers="thread").run(reqs, realtime=False)
    assert sorted(mixed) == sorted(tight) == [r.rid for r in reqs]
    for r in reqs:
        alone = serve.Server(p, slots=1, workers="thread").run([r], realtime=False)[r.rid]
        for res in (mixed[r.rid], tight[r.rid]):
            assert res.ub == pytest.approx(alone.ub, rel=1e-12)
            assert res.dv == pytest.approx(alone.dv, rel=1e-12)
            assert res.certified == alone.certified
        assert alone.certified and alone.UB >= alone.lb
    summ = serve.summarize(mixed, wall=1.0)
    for key in ("latency_p50_s", "latency_p99_s", "solves_per_s", "certified_gap_UB_minus_lb", "certified"):
        assert key in summ


def test_process_slots_report_worker_errors():
    """Process slots on a host without a device: both workers start (no certificate warm-up), each request fails
    inside its worker (no device context can be created) and the failure reaches the caller."""
    reqs = serve.poisson_stream(2, 200.0, 100, 200, 6, 8, seed=4)
    p = serve.Params(k=8, outer=2, cert_max_n=0)
    with pytest.raises(RuntimeError, match="2 request"):
        serve.Server(p, slots=2).run(reqs, realtime=False)
    with pytest.raises(ValueError):
        serve.Server(p, workers="fiber")


def test_thread_slots_refused_with_the_device_library():
    """Thread slots serve only with a host engine: with the device library as the solver's engine they are refused
    (their one device run ended in a host segmentation fault of unknown cause)."""
    with pytest.raises(RuntimeError, match="thread slots"):
        serve.Server(serve.Params(), workers="thread")
