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


@pytest.mark.timeout(1100)
def test_c5_stream_certified_and_alone_equals_mixed():
    """The serving workload at its own size: 16 requests arriving as a Poisson stream at 200 req/s, Chung-Lu graphs
    with n log-uniform in [1,000, 200,000] and mean degree uniform in [6, 30], 8 process slots. Every result is
    certified (dense FP64 certificate up to 3,000 vertices, the lifted sparse certificate above) or states why not,
    and states |U_fp32 - U_fp64| (the FP32 search context's own evaluation of the bound against the FP64 assembly);
    the smallest, the median and the largest request solved alone in this process give the same bound as in the
    mix (1e-6 relative: FP64 atomics in a few device reductions change the summation order)."""
    reqs = serve.poisson_stream(16, 200.0, 1000, 200000, 6, 30, seed=1)
    p = serve.Params()
    print(f"[c5] {len(reqs)} requests, n {sorted(r.n for r in reqs)}", flush=True)
    srv = serve.Server(p, slots=8, progress=lambda r: print(
        f"[c5] rid {r.rid} n {r.n}: latency {r.latency:.2f} s ub {r.ub:.6f} UB {r.UB} certified {r.certified} "
        f"|dU| {r.du_fp32_fp64} note {r.cert_note}", flush=True))
    mixed = srv.run(reqs, realtime=True)
    summ = serve.summarize(mixed, wall=max(r.finish for r in mixed.values()))
    print(f"[c5] {summ}", flush=True)
    assert len(mixed) == len(reqs)
    for r in mixed.values():
        assert r.certified or r.cert_note, r.rid
        if r.certified:
            assert r.UB >= r.lb and r.UB >= r.ub - 1e-9
        assert r.du_fp32_fp64 is not None and r.du_fp32_fp64 <= 1e-4 * abs(r.ub)
    assert summ["certified"] == len(reqs)
    serve.prepare_device(p)
    by_n = sorted(reqs, key=lambda r: r.n)
    for r in (by_n[0], by_n[len(by_n) // 2], by_n[-1]):
        out, _ = serve.solve_request(r, p)
        m = mixed[r.rid]
        print(f"[c5] alone rid {r.rid} n {r.n}: ub {out['ub']:.9f} (mixed {m.ub:.9f}) UB {out['UB']} (mixed {m.UB})",
              flush=True)
        assert m.ub == pytest.approx(out["ub"], rel=1e-6)
        assert out["certified"] == m.certified
