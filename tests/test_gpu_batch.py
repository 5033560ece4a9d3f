"""The batched, fused, graph-launched search returns the plain Worker's
candidates exactly: every thresholded value is computed with the same
arithmetic, and host clustering/distilling is unchanged.  Results must not
depend on how trials are grouped into batches (alone vs in the mix)."""
import numpy as np
import pytest

from candmatch import compare

pytestmark = pytest.mark.gpu


def same(a, b):
    key = lambda c: (int(c["dm_idx"]), float(c["acc"]), int(c["nh"]), float(c["freq"]), float(c["snr"]),
                     int(c["nassoc"]))
    return sorted(map(key, a)) == sorted(map(key, b))


@pytest.fixture(scope="module")
def cfg1(psx):
    import peasoup_amd.workloads as W
    w = W.config("cpu8x2^17")
    return w, W.generate_host(w), W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)


# batch 0: the library's default for the length (64 trials per launch here)
@pytest.mark.parametrize("batch,graph", [(16, True), (3, True), (1, False), (5, False), (0, True)])
def test_batched_equals_worker(psx, cfg1, batch, graph):
    import peasoup_amd.workloads as W
    w, series, accs = cfg1
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(batch, use_graph=graph))
    assert same(got, ref)


def test_kept_engine_gives_the_same_candidates(psx, cfg1):
    """keep_engine: a call leaves its engine behind, the next call with the
    same arguments and options reuses it (buffers, plans, launch graphs),
    calls with other options or another length replace it, and
    psx_batch_release frees it; every call returns the Worker's candidates."""
    import peasoup_amd.workloads as W
    w, series, accs = cfg1
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs)
    w2 = W.Workload("kept2", 3, 1 << 16, 64e-6, 0x5EA5E2, 5.0, -40.0, 40.0,
                    injections=[W.Injection(1, 3.3e-3, 10.0, 15.0, 0.2)])
    s2 = W.generate_host(w2)
    a2 = W.accel_list(w2.size, w2.tsamp, w2.acc_lo, w2.acc_hi)
    ref2, _ = W.worker_search(w2.search_args(), s2, w2.dm_list(), a2)
    try:
        for batch in (16, 16, 5, 5, 16):
            got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs,
                                    W.make_batch_opts(batch, keep_engine=True))
            assert same(got, ref), batch
        got2, _ = W.batch_search(w2.search_args(), s2, w2.dm_list(), a2, W.make_batch_opts(16, keep_engine=True))
        assert same(got2, ref2)
        got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(16, keep_engine=True))
        assert same(got, ref)
    finally:
        W.release_batch_engine()
    W.release_batch_engine()
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(16))
    assert same(got, ref)


def test_batched_vs_oracle(psx, orc, cfg1):
    import peasoup_amd.workloads as W
    w, series, accs = cfg1
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs)
    ref = []
    for d in range(w.ndm):
        p = orc.make_params(w.size, np.float32(w.tsamp), dm=float(d), dm_idx=d, nharmonics=w.nharmonics)
        ref.append(orc.search_dm(series[d], p, accs, nthreads=4)[0])
    assert not compare(got, np.concatenate(ref), w.min_snr)


def test_final_candidates_vs_oracle(psx, orc):
    """Series -> searches -> main()'s DM/harmonic distilling, scoring and limit:
    the batched forests equal the plain Worker's, the library's finalize equals
    the oracle's on them bit for bit, and the final list matches the oracle's
    end-to-end computation.  One pulsar is seen at three neighbouring DM
    trials (DM distilling) and a second at twice its frequency (harmonic
    distilling); the threshold lets noise candidates through."""
    import peasoup_amd.postprocess as pp
    import peasoup_amd.workloads as W
    w = W.Workload("final", 6, 1 << 17, 64e-6, 0x5EA5F1, 0.0, -20.0, 20.0, min_snr=6.0,
                   injections=[W.Injection(1, 4.1e-3, 0.0, 12.0, 0.1), W.Injection(2, 4.1e-3, 0.0, 18.0, 0.1),
                               W.Injection(3, 4.1e-3, 0.0, 12.0, 0.1), W.Injection(4, 2.05e-3, 0.0, 14.0, 0.1)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, forest=True)
    plain, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs, forest=True)
    assert same(got, plain) and np.array_equal(got["nassoc"], plain["nassoc"])
    ref = []
    for d in range(w.ndm):
        p = orc.make_params(w.size, np.float32(w.tsamp), dm=float(d), dm_idx=d, nharmonics=w.nharmonics,
                            min_snr=w.min_snr)
        ref.append(orc.search_dm(series[d], p, accs, nthreads=4, forest=True)[0])
    ref = np.concatenate(ref)
    args = pp.make_final_args(freq_tol=1e-4, max_harm=16.0, tsamp=w.tsamp, cfreq=1400.0, foff=0.390625,
                              nchans=1024, limit=1000)
    fin, sc = pp.finalize(got, args)
    ofin, osc = orc.finalize(got.view(orc.CAND_DTYPE), 1e-4, 16.0, 1400.0, 0.390625, args.bw, 1000)
    assert np.array_equal(fin.view(np.uint8), ofin.view(np.uint8))
    assert np.array_equal(sc["is_adjacent"], osc["is_adjacent"])
    assert np.array_equal(sc["ddm_count_ratio"], osc["ddm_count_ratio"])
    rfin, _ = orc.finalize(ref, 1e-4, 16.0, 1400.0, 0.390625, args.bw, 1000)
    tops, rtops = W.top_level(fin), orc.top_level(rfin)
    assert len(tops) > 0 and len(tops) < len(W.top_level(got))
    assert not compare(W.top_level(got), orc.top_level(ref), w.min_snr)
    # association counts can move with a near-threshold candidate; the final
    # top-level candidates and their S/N may not
    assert not compare(tops, rtops, w.min_snr, check_assoc=False)


def test_zapped_search(psx, orc):
    """A zap list in the prologue (pipeline_multi.cu:222-226): the plain and
    batched paths agree, both match the oracle, and a pulsar sitting on a
    birdie is no longer found at its fundamental."""
    import peasoup_amd.workloads as W
    w = W.Workload("zap", 3, 1 << 17, 64e-6, 0x5EA52A, 0.0, -10.0, 10.0,
                   injections=[W.Injection(1, 0.02, 0.0, 20.0, 0.1)])  # 50 Hz
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    zl = [(50.0, 0.1), (100.0, 0.15), (150.0, 0.15), (200.0, 0.15), (250.0, 0.15)]
    args = W.make_search_args(w.size, w.tsamp, acc_start=w.acc_lo, acc_end=w.acc_hi, birdies=zl)
    plain, _ = W.worker_search(args, series, w.dm_list(), accs)
    got, _ = W.batch_search(args, series, w.dm_list(), accs)
    assert same(got, plain)
    ref = []
    for d in range(w.ndm):
        p = orc.make_params(w.size, np.float32(w.tsamp), dm=float(d), dm_idx=d, nharmonics=w.nharmonics, birdies=zl)
        ref.append(orc.search_dm(series[d], p, accs, nthreads=4)[0])
    assert not compare(got, np.concatenate(ref), w.min_snr)
    unzapped, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs)
    near50 = lambda c: np.any((np.abs(c["freq"] - 50.0) < 0.1) & (c["nh"] == 0) & (c["dm_idx"] == 1))
    assert near50(unzapped) and not near50(got)


@pytest.mark.parametrize("nharm", [1, 2, 3, 5])
def test_batched_equals_worker_every_fold_count(psx, nharm):
    """Each fold count has its own fused harmonic kernel."""
    import peasoup_amd.workloads as W
    w = W.Workload("h", 2, 1 << 16, 64e-6, 0x5EA5A0 + nharm, 0.0, -30.0, 30.0, nharmonics=nharm, min_snr=6.0,
                   injections=[W.Injection(1, 3.3e-3, 5.0, 15.0, 0.1)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs, forest=True)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(4), forest=True)
    assert same(got, ref) and len(got) == len(ref)
    assert len(got) > 0 and set(got["nh"]) <= set(range(nharm + 1))


def test_batched_device_input_and_plan(psx, cfg1):
    import peasoup_amd.workloads as W
    w, series, accs = cfg1
    d = psx.DeviceArray.from_numpy(series)
    a, _ = W.batch_search(w.search_args(), d, w.dm_list(), accs, nsamps=w.size)
    b, _ = W.batch_search(w.search_args(), series, w.dm_list(), None)
    assert same(a, b)


def test_batched_2pow23_matches_worker(psx):
    """Headline length 2^23 at 76 us with red noise and an injected pulsar."""
    import peasoup_amd.workloads as W
    w = W.Workload("t", 2, 1 << 23, 76e-6, 99, 5.0, -100, 100,
                   injections=[W.Injection(1, 0.01234, 35.0, 25.0, 0.3)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)[::16]
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs)
    got, secs = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(4))
    assert same(got, ref)
    assert np.any((got["dm_idx"] == 1) & (got["snr"] > 10))


@pytest.mark.parametrize("n", [3 * (1 << 21), 1 << 22])
def test_batched_uneven_sizes_match_worker(psx, n):
    import peasoup_amd.workloads as W
    w = W.Workload("t", 1, n, 76e-6, 7, 5.0, -50, 50, injections=[W.Injection(0, 0.0071, -20.0, 20.0)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)[::6]
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(3))
    assert same(got, ref)


# Two-pass plans whose last pass has narrow column tiles write the spectrum
# as contiguous rows (R = first-pass length rows of L) and the spectrum pass
# reads that transposed layout: tile interiors, the halo rows, the self-paired
# rows 0 and R/2, partial row tiles (R/2 - 1 not a multiple of 64) and
# radix-3 lengths.  For the same FFT decomposition the candidates must equal
# those of the natural layout exactly (a different decomposition, as the
# plain Worker's default can be, rounds the FFT differently).
def _plan_desc(psx, W, w, series0, accs, opts):
    import ctypes as C
    stage_ms = np.zeros(8)
    ns = C.c_int(8)
    desc = C.create_string_buffer(128)
    d = psx.DeviceArray.from_numpy(np.ascontiguousarray(series0, np.float32))
    psx.register_signatures({"psx_batch_profile": (C.c_int, [C.POINTER(W.SearchArgs), C.POINTER(W.BatchOpts),
                                                             C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32,
                                                             C.c_int, C.c_void_p, C.POINTER(C.c_int), C.c_char_p,
                                                             C.c_size_t])})
    acc32 = np.ascontiguousarray(accs, np.float32)
    psx.check(psx.lib().psx_batch_profile(C.byref(w.search_args()), C.byref(opts), d.ptr, w.size,
                                          acc32.ctypes.data, acc32.size, 1, stage_ms.ctypes.data, C.byref(ns),
                                          desc, 128))
    return desc.value.decode()


@pytest.mark.parametrize("n,decomp", [
    (1 << 17, [(32, 16), (2048, 8)]),         # R = 32: one partial row tile
    (1 << 20, [(256, 16), (2048, 4)]),        # R = 256: two row tiles
    (3 << 19, [(384, 16), (2048, 8)]),        # R = 384: R/2 - 1 = 191, last tile partial
    (3 << 19, [(512, 16), (1536, 8)]),        # radix-3 row length
    (1 << 23, [(1024, 16), (4096, 4)]),       # the 2^23 workload's plan
    (1 << 20, [(1024, 8), (512, 16)]),        # 16-column last pass (the tuned 2^20 plan)
    (1 << 21, [(32, 16), (128, 16), (256, 16)]),   # three passes: R = 4096 rows of 256
    (3 << 19, [(48, 16), (128, 16), (128, 16)]),   # three passes, radix-3 first: R = 6144
])
def test_row_layout_equals_natural_layout(psx, n, decomp):
    import peasoup_amd.workloads as W
    w = W.Workload("rows", 2, n, 64e-6, 0x5EA5C0 + n % 97, 0.0, -50.0, 50.0,
                   injections=[W.Injection(1, 0.0123, 20.0, 15.0, 0.2)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    accs = accs[:: max(1, accs.size // 12)]
    rows = W.make_batch_opts(4, decomp=decomp, spectrum_layout=W.SPECTRUM_ROWS)
    nat = W.make_batch_opts(4, decomp=decomp, spectrum_layout=W.SPECTRUM_NATURAL)
    assert _plan_desc(psx, W, w, series[0], accs, rows).endswith(" rows")
    assert not _plan_desc(psx, W, w, series[0], accs, nat).endswith(" rows")
    ref, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, nat, forest=True)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, rows, forest=True)
    assert len(ref) > 0
    assert same(got, ref)


def test_row_layout_default_plan_matches_worker(psx):
    """The default 2^21 plan takes the row layout automatically and equals
    the plain Worker (same decomposition) bit for bit."""
    import peasoup_amd.workloads as W
    w = W.Workload("rows", 2, 1 << 21, 64e-6, 0x5EA5C7, 0.0, -50.0, 50.0,
                   injections=[W.Injection(0, 0.0077, 18.0, 12.0, 0.2)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    accs = accs[:: max(1, accs.size // 10)]
    opts = W.make_batch_opts(4)
    assert _plan_desc(psx, W, w, series[0], accs, opts).endswith(" rows")
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs, forest=True)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, opts, forest=True)
    assert len(ref) > 0
    assert same(got, ref)


def test_row_layout_unavailable_is_an_error(psx):
    import peasoup_amd.workloads as W
    w = W.Workload("rows", 1, 1 << 17, 64e-6, 0x5EA5C8, 0.0, -5.0, 5.0)
    series = W.generate_host(w)
    # no contiguous-row kernel exists for this plan's last pass (64 points,
    # 16 columns), so it has no row-layout form
    opts = W.make_batch_opts(2, decomp=[(16 * 2, 16), (32, 16), (64, 16)], spectrum_layout=W.SPECTRUM_ROWS)
    with pytest.raises(psx.PsxError, match="row-layout"):
        W.batch_search(w.search_args(), series, w.dm_list(), None, opts)


def test_profile_times_three_pass_plans_on_valid_spectra(psx):
    """psx_batch_profile launches each stage several times back to back; the
    middle pass of a three-pass plan works in place, so it is launched once
    (a second launch would transform its own output and every later stage
    would be timed on garbage: the harmonic pass then finds bins above
    threshold everywhere and runs many times slower).  Its spectrum and
    harmonic stages take about as long as a two-pass plan's."""
    import ctypes as C
    import peasoup_amd.workloads as W
    n = 1 << 19
    w = W.Workload("prof3", 1, n, 64e-6, 0x5EA5D9, 5.0, -50.0, 50.0)
    d = W.generate_device(w)
    accs = np.ascontiguousarray(W.accel_list(n, w.tsamp, -50.0, 50.0)[:16], np.float32)
    psx.register_signatures({"psx_batch_profile": (C.c_int, [C.POINTER(W.SearchArgs), C.POINTER(W.BatchOpts),
                                                             C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32,
                                                             C.c_int, C.c_void_p, C.POINTER(C.c_int), C.c_char_p,
                                                             C.c_size_t])})
    args = w.search_args()

    def stages(decomp):
        ms = np.zeros(8)
        ns = C.c_int(8)
        desc = C.create_string_buffer(128)
        opts = W.make_batch_opts(16, decomp=decomp)
        psx.check(psx.lib().psx_batch_profile(C.byref(args), C.byref(opts), d.ptr, n, accs.ctypes.data, accs.size,
                                              3, ms.ctypes.data, C.byref(ns), desc, 128))
        return ms[:ns.value]

    two = stages([(512, 16), (512, 16)])
    three = stages([(64, 16), (64, 16), (64, 16)])
    assert len(two) == 4 and len(three) == 5
    assert three[-1] < 4 * two[-1] + 0.05, (two, three)   # harmonic sums + threshold
    assert three[-2] < 4 * two[-2] + 0.05, (two, three)   # spectrum pass


def test_many_bins_above_threshold_match_worker(psx):
    """A threshold low enough that most bins of every fold cross it: the
    harmonic pass's append path (taken by every wave) and the host's
    clustering of tens of thousands of bins per spectrum give the plain
    Worker's candidates exactly."""
    import peasoup_amd.workloads as W
    w = W.Workload("low", 2, 1 << 17, 64e-6, 0x5EA5DA, 5.0, -20.0, 20.0, min_snr=1.0,
                   injections=[W.Injection(0, 0.0123, 5.0, 20.0, 0.3)])
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs, forest=True)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(16), forest=True)
    assert len(ref) > 100
    assert same(got, ref)


def test_async_prologue_baseline_loop_matches_worker(psx):
    """The batched search runs the DM prologues ahead on their own stream
    with a fixed number of device iterations of remove_baseline's loop
    (kernels.cu:463-479); a DM whose loop has not stopped by then is redone
    with the host-driven loop.  Series needing 1, 2, 3 and 102 iterations
    (the loop's cap: a quarter of the samples at 1e10) give the plain
    Worker's candidates exactly."""
    import peasoup_amd.workloads as W
    n = 1 << 17
    w = W.Workload("bl", 4, n, 64e-6, 0x5EA5D0, 0.0, -20.0, 20.0,
                   injections=[W.Injection(d, 0.0111, 25.0, 15.0, 0.2) for d in range(4)])
    series = W.generate_host(w).astype(np.float32)
    series[0] -= np.float32(series[0].mean())                    # ~1 iteration
    series[1] += np.float32(5.0)                                 # 2 iterations
    series[2] = np.float32(2.0 ** 22) + np.float32(0.5) * (np.arange(n) % 2).astype(np.float32)  # 3
    series[3][: n // 4] = np.float32(1e10)                       # loop cap: host fallback
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs, forest=True)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(8), forest=True)
    assert len(ref) > 0
    assert same(got, ref)


def test_batched_overflow_is_an_error(psx):
    """A threshold below the noise floor floods the candidate buffer: the
    search must fail loudly rather than drop candidates."""
    import peasoup_amd.workloads as W
    w = W.Workload("t", 1, 1 << 16, 64e-6, 5, 0.0, 0, 0)
    series = W.generate_host(w)
    args = W.make_search_args(w.size, w.tsamp, acc_start=0, acc_end=0, min_snr=-10.0)
    with pytest.raises(psx.PsxError):
        W.batch_search(args, series, w.dm_list(), None, W.make_batch_opts(1, cand_capacity=10000))


@pytest.mark.parametrize("n", [1 << 25, 1 << 26])
def test_batched_long_three_pass_matches_worker(psx, n):
    """Lengths 2^25 and 2^26 (three FFT passes, fused first pass, row-layout
    hand-over; tile-major resampling at 2^26) against the plain Worker."""
    import peasoup_amd.workloads as W
    w = W.Workload("t", 1, n, 64e-6, 41, 5.0, -50, 50, injections=[W.Injection(0, 0.0023, 12.0, 25.0, 0.7)])
    series = W.generate_host(w)
    full = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    k = int(np.argmin(np.abs(full - 12.0)))  # the trials around the injected acceleration
    accs = full[k - 2:k + 2]
    ref, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, W.make_batch_opts(2))
    assert same(got, ref)
    assert np.any(got["snr"] > 10)


@pytest.mark.parametrize("n,plan_desc", [(1 << 25, "64/C128x64/C128x4096/C4"), (1 << 26, "128/C128x64/C128x4096/C4")])
def test_fft_long_default_plans_vs_exact(psx, orc, n, plan_desc):
    """The default plans beyond 2^24 samples (two short strided passes and
    a 4,096-point contiguous last pass) against an exact double DFT."""
    x = np.random.default_rng(n.bit_length()).standard_normal(n).astype(np.float32)
    plan = psx.FFTPlan(n, psx.PSX_FFT_R2C)
    assert plan_desc in plan.describe(), plan.describe()
    d_x, d_X = psx.DeviceArray.from_numpy(x), psx.DeviceArray(n // 2 + 1, np.complex64)
    plan.execute(d_x, d_X)
    got = d_X.numpy()
    ref = orc.fft_r2c_f64(x.astype(np.float64))
    err = float(np.max(np.abs(got - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2)))
    assert err < 2e-6, (plan.describe(), err)
