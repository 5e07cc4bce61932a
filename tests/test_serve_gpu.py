"""Serving on the device: overlapping requests give the same bounds as the same requests served alone (tolerance
1e-6 relative: the device reductions add FP64 partials with atomics in places, so the last bits may differ between
runs), every small request is certified, and the latency summary is complete. The process-slot server (the default)
is checked against each request solved alone in this process; the thread-slot variant is skipped, see its reason."""
import pytest

from lagbound_amd import serve

pytestmark = pytest.mark.gpu


@pytest.mark.skip(reason="one host segmentation fault inside lb_eval_fgrad while three workers were in device calls "
                         "(one in its first lb_certify_dense); cause not established, so serve.Server refuses thread "
                         "slots with the device library (the alone-versus-mixed check of thread slots runs on the CPU "
                         "in test_serve.py)")
def test_alone_equals_mixed_on_device():
    reqs = serve.poisson_stream(6, 200.0, 800, 3000, 6, 20, seed=5)
    p = serve.Params(outer=12, tri_every=4, vcs_every=3, rgd_iters=60, eig_maxiter=100, cert_max_n=3000)
    print(f"[serve] mixed: {len(reqs)} requests, n {[r.n for r in reqs]}", flush=True)
    mixed = serve.Server(p, slots=3, workers="thread").run(reqs, realtime=True)
    summ = serve.summarize(mixed, wall=max(r.finish for r in mixed.values()))
    print(f"[serve] {summ}", flush=True)
    for r in reqs:
        alone = serve.Server(p, slots=1, workers="thread").run([r], realtime=False)[r.rid]
        m = mixed[r.rid]
        print(f"[serve] rid {r.rid} n {r.n}: ub mixed {m.ub:.9f} alone {alone.ub:.9f} UB {alone.UB}", flush=True)
        assert m.ub == pytest.approx(alone.ub, rel=1e-6)
        assert m.dv == pytest.approx(alone.dv, rel=1e-6)
        assert alone.certified and m.certified
        assert alone.UB >= alone.lb and m.UB >= m.lb


def test_process_slots_alone_equals_mixed_on_device():
    reqs = serve.poisson_stream(6, 200.0, 800, 3000, 6, 20, seed=5)
    p = serve.Params(outer=12, tri_every=4, vcs_every=3, rgd_iters=60, eig_maxiter=100, cert_max_n=3000)
    print(f"[serve] process slots, mixed: {len(reqs)} requests, n {[r.n for r in reqs]}", flush=True)
    srv = serve.Server(p, slots=3)
    mixed = srv.run(reqs, realtime=True)
    summ = serve.summarize(mixed, wall=max(r.finish for r in mixed.values()))
    print(f"[serve] slots started in {srv.start_s:.2f} s; {summ}", flush=True)
    for key in ("latency_p50_s", "latency_p99_s", "solves_per_s", "service_p50_s", "certified"):
        assert key in summ
    serve.prepare_device(p)
    for r in reqs:
        out, _ = serve.solve_request(r, p)
        m = mixed[r.rid]
        print(f"[serve] rid {r.rid} n {r.n}: ub mixed {m.ub:.9f} alone {out['ub']:.9f} UB {out['UB']}", flush=True)
        assert m.ub == pytest.approx(out["ub"], rel=1e-6)
        assert m.dv == pytest.approx(out["dv"], rel=1e-6)
        assert out["certified"] and m.certified
        assert out["UB"] >= out["lb"] and m.UB >= m.lb
