"""The bound-filtered harmonic pass returns exactly the candidates of the
exact-gather pass (and of the plain Worker): it sums bfloat16 upper
bounds of bin pairs first and sums the spectrum exactly only where a bound
crosses the threshold.  Cases: every fold count, thresholds low enough that
most bins cross (bounds and exact values both), trial counts that do and do
not take the XCD-grouped trial order (multiples of 8), both spectrum
layouts, and the headline length."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def same(a, b):
    key = lambda c: (int(c["dm_idx"]), float(c["acc"]), int(c["nh"]), float(c["freq"]), float(c["snr"]),
                     int(c["nassoc"]))
    return sorted(map(key, a)) == sorted(map(key, b))


def both_forms(W, w, series, accs, batch, **kw):
    out = []
    for form in (W.HARMONIC_BOUND, W.HARMONIC_EXACT):
        got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs,
                                W.make_batch_opts(batch, harmonic_pass=form, **kw), forest=True)
        out.append(got)
    return out


@pytest.mark.parametrize("nharm", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("batch", [8, 5])
def test_bound_equals_exact_every_fold_count(psx, nharm, batch):
    import peasoup_amd.workloads as W
    w = W.Workload("b", 2, 1 << 16, 64e-6, 0x5EB000 + nharm, 3.0, -30.0, 30.0, nharmonics=nharm, min_snr=5.0,
                   injections=[W.Injection(1, 3.3e-3, 5.0, 15.0, 0.1)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    bound, exact = both_forms(W, w, series, accs, batch)
    print("nharm %d batch %d: %d candidates" % (nharm, batch, len(exact)), flush=True)
    assert len(exact) > 0 and set(exact["nh"]) <= set(range(nharm + 1))
    assert same(bound, exact) and len(bound) == len(exact)


@pytest.mark.parametrize("thr", [1.0, 2.5, 4.0])
def test_bound_equals_exact_low_thresholds(psx, thr):
    """Thresholds at which most bins' bounds cross and many exact values do
    too: the exact recomputation runs in every wave."""
    import peasoup_amd.workloads as W
    w = W.Workload("lo", 2, 1 << 17, 64e-6, 0x5EB0A0, 5.0, -20.0, 20.0, min_snr=thr,
                   injections=[W.Injection(0, 0.0123, 5.0, 20.0, 0.3)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    bound, exact = both_forms(W, w, series, accs, 16)
    print("thr %.1f: %d candidates" % (thr, len(exact)), flush=True)
    assert len(exact) > 10
    assert same(bound, exact) and len(bound) == len(exact)
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs, forest=True)
    assert same(bound, ref)


@pytest.mark.parametrize("thr", [1.5, 3.0, 4.5])
def test_bound_rows_layout_low_thresholds(psx, thr):
    """The row-layout form writes no spectrum: the bins whose bound crosses
    are recomputed from the FFT output for the exact sums.  At these
    thresholds most or many bins take that path, and every candidate's S/N
    must equal the exact-gather form's (which reads the spectrum written by
    the same pass) bit for bit."""
    import peasoup_amd.workloads as W
    n = 1 << 18
    w = W.Workload("rlo", 2, n, 64e-6, 0x5EB0E0, 5.0, -20.0, 20.0, min_snr=thr,
                   injections=[W.Injection(1, 0.0123, 5.0, 20.0, 0.3)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)[:8]
    bound, exact = both_forms(W, w, series, accs, 8, decomp=[(128, 16), (1024, 4)], spectrum_layout=W.SPECTRUM_ROWS)
    print("rows thr %.1f: %d candidates" % (thr, len(exact)), flush=True)
    assert len(exact) > 10
    assert same(bound, exact) and len(bound) == len(exact)


@pytest.mark.parametrize("layout", ["natural", "rows"])
def test_bound_equals_exact_both_layouts(psx, layout):
    import peasoup_amd.workloads as W
    n = 1 << 22
    w = W.Workload("lay", 1, n, 64e-6, 0x5EB0B0, 5.0, -50.0, 50.0, min_snr=6.0,
                   injections=[W.Injection(0, 0.0071, 20.0, 25.0, 0.2)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    accs = accs[np.argsort(np.abs(accs - 20.0))[:8]]
    lay = W.SPECTRUM_NATURAL if layout == "natural" else W.SPECTRUM_ROWS
    bound, exact = both_forms(W, w, series, accs, 8, spectrum_layout=lay)
    print("%s: %d candidates" % (layout, len(exact)), flush=True)
    assert len(exact) > 0
    assert same(bound, exact) and len(bound) == len(exact)


def test_bound_headline_length_matches_worker(psx):
    """2^23 at 76 us with red noise and a pulsar, 16 trials per launch (the
    XCD-grouped order), against the plain Worker."""
    import peasoup_amd.workloads as W
    w = W.Workload("h23", 2, 1 << 23, 76e-6, 0x5EB0C0, 5.0, -100, 100,
                   injections=[W.Injection(1, 0.01234, 35.0, 25.0, 0.3)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    accs = accs[np.argsort(np.abs(accs - 35.0))[:16]]
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs, forest=True)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(16), forest=True)
    print("2^23: %d candidates" % len(ref), flush=True)
    assert same(got, ref) and len(got) == len(ref)
    assert np.any((got["dm_idx"] == 1) & (got["snr"] > 10))


@pytest.mark.parametrize("n,decomp,layout", [
    (1 << 17, None, "auto"),
    (1 << 20, [(1024, 8), (512, 16)], "natural"),
    (1 << 21, [(1024, 8), (1024, 4)], "rows"),
    (1 << 22, [(1024, 8), (2048, 4)], "rows"),
    (3 << 21, [(2048, 8), (1536, 4)], "rows"),
    (3 << 21, [(2048, 8), (1536, 4)], "natural"),
    (1 << 23, [(1024, 16), (4096, 4)], "rows"),
    (1 << 23, [(2048, 8), (2048, 4)], "rows"),
    (1 << 20, [(1024, 8), (512, 16)], "rows"),
    (1 << 21, [(32, 16), (128, 16), (256, 16)], "rows"),
    (3 << 19, [(48, 16), (128, 16), (128, 16)], "rows"),
])
def test_bound_spectrum_audit(psx, n, decomp, layout):
    """Every bound entry the spectrum stage leaves is at least max(P[2u],
    P[2u+1]) of the exact spectra of the same FFT output (row-layout tiles,
    their edge rows and bin n/2, and the pooling pass after the natural
    layout), rarely one bfloat16 step above the tight value, and the
    fundamental's appended candidates are exactly the bins above threshold
    in its window."""
    import ctypes as C
    import peasoup_amd.workloads as W
    lay = {"auto": W.SPECTRUM_AUTO, "rows": W.SPECTRUM_ROWS, "natural": W.SPECTRUM_NATURAL}[layout]
    w = W.Workload("aud", 1, n, 64e-6, 0x5EB0D0, 5.0, -50.0, 50.0, min_snr=3.0,
                   injections=[W.Injection(0, 0.0071, 20.0, 25.0, 0.2)])
    d = W.generate_device(w)
    accs = np.ascontiguousarray(W.accel_list(n, w.tsamp, -50.0, 50.0)[:8], np.float32)
    psx.register_signatures({"psx_batch_bound_audit": (C.c_int, [C.POINTER(W.SearchArgs), C.POINTER(W.BatchOpts),
                                                                 C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32,
                                                                 C.c_void_p])})
    res = np.zeros(7, np.uint64)
    opts = W.make_batch_opts(8, decomp=decomp, spectrum_layout=lay)
    psx.check(psx.lib().psx_batch_bound_audit(C.byref(w.search_args()), C.byref(opts), d.ptr, n,
                                              accs.ctypes.data, accs.size, res.ctypes.data))
    entries, invalid, fundamentals, appended, crossing, window, loose = (int(x) for x in res)
    print("n %d %s %s: %d entries, %d invalid, %d loose, %d fundamentals, %d appended, %d of %d bins cross"
          % (n, decomp, layout, entries, invalid, loose, fundamentals, appended, crossing, window), flush=True)
    assert window > 0 and crossing <= window
    # the bounds of approximate arithmetic rarely land a step above the tight one
    assert loose <= entries // 100
    assert entries == accs.size * ((n // 2 + 2) // 2)
    assert invalid == 0
    assert fundamentals > 0 and appended == fundamentals
