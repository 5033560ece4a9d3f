"""Serving: continuous batching across requests of mixed sizes.

Each request's candidates must be identical alone and in the mix, and equal
to the batched search of the same input; admission respects the HBM budget;
a failing request does not disturb the others."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def same(a, b):
    key = lambda c: (int(c["dm_idx"]), float(c["acc"]), int(c["nh"]), float(c["freq"]), float(c["snr"]),
                     int(c["nassoc"]))
    return sorted(map(key, a)) == sorted(map(key, b))


def make_requests():
    import peasoup_amd.workloads as W
    specs = [
        W.Workload("a", 2, 1 << 17, 64e-6, 11, 0.0, -50, 50, injections=[W.Injection(1, 5.3e-3, 10.0, 18.0)]),
        W.Workload("b", 3, 1 << 16, 64e-6, 12, 3.0, -100, 100, injections=[W.Injection(0, 3.1e-3, -30.0, 20.0)]),
        W.Workload("c", 1, 1 << 19, 64e-6, 13, 5.0, -50, 50, injections=[W.Injection(0, 11e-3, 5.0, 25.0)]),
        W.Workload("d", 2, 1 << 17, 64e-6, 14, 0.0, -200, 200),
    ]
    return [(w, W.generate_host(w)) for w in specs]


@pytest.fixture(scope="module")
def reqs(psx):
    return make_requests()


def test_alone_equals_in_the_mix(psx, reqs):
    import peasoup_amd.workloads as W
    from peasoup_amd.serve import Server
    alone = []
    for w, series in reqs:
        with Server(64 << 30, batch=8) as srv:
            rid = srv.submit(w.search_args(), series, w.dm_list())
            st = srv.wait(rid)
            assert st["state"] == "done", st
            alone.append(srv.result(rid))
        direct, _ = W.batch_search(w.search_args(), series, w.dm_list(), None, W.make_batch_opts(8))
        assert same(alone[-1], direct), w.name
    with Server(64 << 30, batch=8) as srv:
        ids = [srv.submit(w.search_args(), s, w.dm_list()) for w, s in reqs]
        # a second copy of request a, submitted while the others run
        ids.append(srv.submit(reqs[0][0].search_args(), reqs[0][1], reqs[0][0].dm_list()))
        sts = [srv.wait(i) for i in ids]
        assert all(s["state"] == "done" for s in sts), sts
        mixed = [srv.result(i) for i in ids]
        st = srv.stats()
        assert st["completed"] == len(ids) and st["failed"] == 0
        assert st["engines"] >= 3
    for k in range(len(reqs)):
        assert same(mixed[k], alone[k]), reqs[k][0].name
    assert same(mixed[-1], alone[0])


def test_budget_admission_and_rejection(psx, reqs):
    from peasoup_amd.serve import Server
    w, series = reqs[0]
    one = 2 * w.size * 4  # resident series of one 2-DM request
    # engine buffers are counted too: allow one engine plus one request only
    import ctypes as C
    import peasoup_amd as P
    budget = 3 * one + 600 * (1 << 20)
    with Server(budget, batch=8, max_active=0) as srv:
        ids = [srv.submit(w.search_args(), series, w.dm_list()) for _ in range(3)]
        for i in ids:
            assert srv.wait(i)["state"] == "done"
        assert srv.stats()["used_bytes"] <= budget
    with Server(1 << 20, batch=8) as srv:
        with pytest.raises(P.PsxError) as e:
            srv.submit(w.search_args(), series, w.dm_list())
        assert e.value.code == 6  # PSX_ERR_BUDGET


def test_failing_request_is_isolated(psx, reqs):
    import peasoup_amd.workloads as W
    from peasoup_amd.serve import Server
    w, series = reqs[0]
    bad = W.make_search_args(w.size, w.tsamp, acc_start=-5e9, acc_end=5e9)  # |a tsamp / 2c| n >= 1
    with Server(64 << 30, batch=8) as srv:
        good1 = srv.submit(w.search_args(), series, w.dm_list())
        badid = srv.submit(bad, series, w.dm_list(), accs=np.array([5e9], np.float32))
        good2 = srv.submit(w.search_args(), series, w.dm_list())
        assert srv.wait(good1)["state"] == "done"
        assert srv.wait(good2)["state"] == "done"
        assert srv.wait(badid)["state"] == "failed"
        assert same(srv.result(good1), srv.result(good2))


def test_request_with_zap_list(psx, reqs):
    """A request's zap list is copied at submit and applied in its prologues;
    requests with and without one share an engine."""
    import peasoup_amd.workloads as W
    from peasoup_amd.serve import Server
    w, series = reqs[0]
    zl = [(50.0, 0.1), (100.0, 0.15), (1.0 / 5.3e-3, 0.2)]
    za = W.make_search_args(w.size, w.tsamp, acc_start=w.acc_lo, acc_end=w.acc_hi, birdies=zl)
    direct, _ = W.batch_search(za, series, w.dm_list(), None, W.make_batch_opts(8))
    plain, _ = W.batch_search(w.search_args(), series, w.dm_list(), None, W.make_batch_opts(8))
    with Server(64 << 30, batch=8) as srv:
        i1 = srv.submit(W.make_search_args(w.size, w.tsamp, acc_start=w.acc_lo, acc_end=w.acc_hi, birdies=zl),
                        series, w.dm_list())
        i2 = srv.submit(w.search_args(), series, w.dm_list())
        s1, s2 = srv.wait(i1), srv.wait(i2)
        assert s1["state"] == "done" and s2["state"] == "done", (s1, s2)
        got1, got2 = srv.result(i1), srv.result(i2)
    assert same(got1, direct) and same(got2, plain)
    assert not same(direct, plain)


def test_tuned_schedule_gives_same_results(psx, reqs):
    """A per-length schedule's batch size changes only how the trials are
    launched, never the candidates; its FFT decomposition also changes the
    FFT's rounding, so the candidates then agree to float rounding (a bin
    within rounding of the threshold may fall on either side)."""
    from candmatch import compare
    from peasoup_amd.serve import Server
    w, series = reqs[0]  # 2^17 samples

    def run(sched):
        with Server(64 << 30, batch=8, schedules=sched) as srv:
            rid = srv.submit(w.search_args(), series, w.dm_list())
            assert srv.wait(rid)["state"] == "done"
            return srv.result(rid, forest=True)

    ref = run(None)
    got = run({w.size: {"batch": 3}})
    assert same(got, ref) and len(got) == len(ref)
    got = run({w.size: {"batch": 3, "decomp": [(64, 16), (1024, 8)]}})
    assert len(ref) > 0 and len(got) > 0
    assert not compare(got, ref, w.min_snr, check_assoc=False)
    # the injected 5.3 ms pulsar of DM trial 1 is among them
    assert np.any((got["dm_idx"] == 1) & (got["snr"] > 10) &
                  (np.abs(got["freq"] * 5.3e-3 - np.round(got["freq"] * 5.3e-3)) < 0.01))


def test_serving_sizes_alone_and_in_the_mix(psx):
    """Serving-configuration sizes: beams of 16 DM trials at 2^20, 2^22 and
    2^24 samples (64 us, red noise, a pulsar in each), a few accelerations
    each, under the tuned per-length schedules.  Each request's candidates
    are the same alone and in the mix, and equal the batched search of the
    same series with the same schedule."""
    import peasoup_amd.workloads as W
    from peasoup_amd import tuning
    from peasoup_amd.serve import Server
    table = tuning.table(psx.device_info(0)["name"])
    specs = []
    for k, n in enumerate((1 << 20, 1 << 22, 1 << 24)):
        w = W.Workload("beam%d" % n, 16, n, 64e-6, 0x5EA5D0 + k, 5.0, -50.0, 50.0)
        accs_all = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
        j = int(0.62 * (accs_all.size - 1))
        w.injections = [W.Injection(3 + 5 * k, 0.0047 * (k + 1), float(accs_all[j]), 20.0, 0.3)]
        accs = accs_all[sorted({0, j, accs_all.size - 1})]
        specs.append((w, W.generate_device(w), accs))
    psx.synchronize()
    direct = []
    for w, d, accs in specs:
        t = table.get(w.size, {})
        opts = W.make_batch_opts(t.get("batch") or 16, decomp=t.get("decomp"))
        direct.append(W.batch_search(w.search_args(), d, w.dm_list(), accs, opts, nsamps=w.size)[0])
        print("beam %d: %d candidates" % (w.size, len(direct[-1])), flush=True)
    alone = []
    for w, d, accs in specs:
        with Server(200e9, batch=16, schedules=table) as srv:
            rid = srv.submit(w.search_args(), d, w.dm_list(), accs=accs, nsamps=w.size)
            assert srv.wait(rid)["state"] == "done"
            alone.append(srv.result(rid))
    with Server(200e9, batch=16, schedules=table) as srv:
        # the largest first, so the others arrive while it runs
        ids = {}
        for k in (2, 0, 1):
            w, d, accs = specs[k]
            ids[k] = srv.submit(w.search_args(), d, w.dm_list(), accs=accs, nsamps=w.size)
        sts = {k: srv.wait(i) for k, i in ids.items()}
        assert all(s["state"] == "done" for s in sts.values()), sts
        mixed = {k: srv.result(i) for k, i in ids.items()}
        assert srv.stats()["engines"] == 3
    for k, (w, _, _) in enumerate(specs):
        assert len(direct[k]) > 0
        assert same(alone[k], direct[k]), w.name
        assert same(mixed[k], alone[k]), w.name
        inj = w.injections[0]
        c = direct[k][direct[k]["dm_idx"] == inj.trial]
        r = c["freq"] * inj.period
        assert np.any((c["snr"] > 9) & (np.abs(r - np.round(r)) < 0.01)), (w.name, c[:5])


def test_mixed_graph_dispatch_same_results(psx, reqs):
    """One graph launch per mixed-length batch (each length's batch graph a
    child of it) returns every request's candidates exactly as the
    per-length streams do, and as the request alone."""
    from peasoup_amd.serve import DISPATCH_MIXED_GRAPH, DISPATCH_STREAMS, Server
    results = {}
    for mode in (DISPATCH_STREAMS, DISPATCH_MIXED_GRAPH):
        with Server(64 << 30, batch=8, dispatch=mode) as srv:
            ids = [srv.submit(w.search_args(), s, w.dm_list()) for w, s in reqs]
            sts = [srv.wait(i) for i in ids]
            assert all(st["state"] == "done" for st in sts), (mode, sts)
            results[mode] = [srv.result(i, forest=True) for i in ids]
            assert srv.stats()["engines"] >= 3
    for k, (w, _) in enumerate(reqs):
        a, b = results[DISPATCH_STREAMS][k], results[DISPATCH_MIXED_GRAPH][k]
        assert same(a, b) and len(a) == len(b), w.name
    assert sum(len(r) for r in results[DISPATCH_MIXED_GRAPH]) > 0
    import peasoup_amd as P
    with pytest.raises(P.PsxError):
        Server(64 << 30, batch=8, use_graph=False, dispatch=DISPATCH_MIXED_GRAPH)
