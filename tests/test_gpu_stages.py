"""Per-stage parity of the HIP kernels against the CPU oracle.

Each test runs one C-ABI entry point on seeded inputs and compares with the
oracle's restatement of the same reference routine.  Integer/index work and
the elementwise stages whose arithmetic is restated operation by operation
are compared bit for bit; the FFT is compared with an exact double DFT.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(20261202)


def dev(P, a):
    return P.DeviceArray.from_numpy(np.ascontiguousarray(a))


def out(P, shape, dtype=np.float32):
    return P.DeviceArray(shape, dtype)


# ---------------------------------------------------------------- resample
@pytest.mark.parametrize("n", [1 << 17, 4194304, 1 << 23])
@pytest.mark.parametrize("acc,tsamp", [(0.0, 64e-6), (-50.0, 64e-6), (125.5, 64e-6), (99.3, 76e-6),
                                       (-200.0, 76e-6)])
def test_resample_ii_bit_exact(psx, orc, n, acc, tsamp):
    x = RNG.standard_normal(n).astype(np.float32)
    d_in, d_out = dev(psx, x), out(psx, n)
    psx.check(psx.lib().psx_resample(d_in.ptr, d_out.ptr, n, acc, tsamp, None))
    got = d_out.numpy()
    ref = orc.resample2(x, acc, tsamp)
    assert np.array_equal(got, ref)
    if acc == 0.0:
        assert np.array_equal(got, x)


def test_resample_index_pattern(psx, orc):
    """src/resampling_test.cpp pattern: x[i] = i % 451, N = 4194304, a = 125.5."""
    n = 4194304
    x = (np.arange(n) % 451).astype(np.float32)
    d_in, d_out = dev(psx, x), out(psx, n)
    psx.check(psx.lib().psx_resample(d_in.ptr, d_out.ptr, n, 125.5, 0.000064, None))
    got = d_out.numpy()
    assert np.array_equal(got, orc.resample2(x, 125.5, 0.000064))
    psx.check(psx.lib().psx_resample_v1(d_in.ptr, d_out.ptr, n, 125.5, 0.000064, None))
    got1 = d_out.numpy()
    assert np.array_equal(got1, orc.resample1(x, 125.5, 0.000064))


def test_resample_rejects_bad_args(psx):
    n = 1024
    d_in, d_out = out(psx, n), out(psx, n)
    rc = psx.lib().psx_resample(d_in.ptr, d_in.ptr, n, 1.0, 64e-6, None)
    assert rc == 1
    rc = psx.lib().psx_resample(d_in.ptr, d_out.ptr, n, 1e12, 1.0, None)
    assert rc == 1
    assert b"must be" in psx.lib().psx_last_error_string()


# --------------------------------------------------------------------- FFT
FFT_SIZES = [16, 32, 64, 128, 1024, 8192, 1 << 13 | 0, 1 << 14, 1 << 17, 1 << 20, 1 << 22, 1 << 23,
             48, 96, 384, 3072, 6144, 3 * (1 << 13), 3 * (1 << 17), 3 * (1 << 21)]


def fft_err(got, ref):
    ref = ref.astype(np.complex128)
    scale = np.sqrt(np.mean(np.abs(ref) ** 2))
    return float(np.max(np.abs(got.astype(np.complex128) - ref)) / scale)


@pytest.mark.parametrize("n", sorted(set(FFT_SIZES)))
def test_fft_r2c_vs_exact(psx, orc, n):
    x = RNG.standard_normal(n).astype(np.float32)
    plan = psx.FFTPlan(n, psx.PSX_FFT_R2C)
    d_x, d_X = dev(psx, x), out(psx, n // 2 + 1, np.complex64)
    plan.execute(d_x, d_X)
    got = d_X.numpy()
    ref = orc.fft_r2c_f64(x.astype(np.float64))
    err = fft_err(got, ref)
    assert err < 2e-6, (plan.describe(), err)


def _tuned_decomps():
    """Every FFT decomposition the shipped schedule table names."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "peasoup_amd", "tuned.json")
    with open(path) as f:
        table = json.load(f)
    out = set()
    for per_dev in table.values():
        for n, e in per_dev.items():
            if e.get("decomp"):
                out.add((int(n), tuple(tuple(x) for x in e["decomp"])))
    return sorted(out)


# named decompositions beyond the defaults: strided radix-3 columns whose last
# radix is at most 8 (the factored four-step twiddle), radix-3 rows, three
# passes with a radix-3 first pass, and the 2^23 default and a three-pass
# alternative named explicitly
EXTRA_DECOMPS = [
    (3 << 20, ((384, 16), (4096, 4))),
    (3 << 20, ((1536, 8), (1024, 4))),
    (3 << 20, ((1024, 16), (1536, 4))),
    (3 << 21, ((96, 16), (128, 16), (256, 16))),
    (1 << 23, ((1024, 16), (4096, 4))),
    (1 << 23, ((128, 16), (128, 16), (256, 16))),
]


@pytest.mark.parametrize("n,decomp", _tuned_decomps() + EXTRA_DECOMPS)
def test_fft_r2c_decomposition_vs_exact(psx, orc, n, decomp):
    """Each schedule's FFT decomposition, built as a plain plan, is within
    the same 2e-6 of the exact double DFT as the default plans."""
    assert len(_tuned_decomps()) >= 3
    x = RNG.standard_normal(n).astype(np.float32)
    plan = psx.FFTPlan(n, psx.PSX_FFT_R2C, decomp=list(decomp))
    desc = plan.describe()
    assert "passes=" + "x".join("%d/C%d" % (L, c) for L, c in decomp) + " " in desc, desc
    d_x, d_X = dev(psx, x), out(psx, n // 2 + 1, np.complex64)
    plan.execute(d_x, d_X)
    err = fft_err(d_X.numpy(), orc.fft_r2c_f64(x.astype(np.float64)))
    assert err < 2e-6, (desc, err)


@pytest.mark.parametrize("n", [16, 1024, 6144, 1 << 17, 3 * (1 << 17), 1 << 23])
def test_fft_c2r_vs_exact(psx, orc, n):
    x = RNG.standard_normal(n).astype(np.float32)
    X = orc.fft_r2c(x)
    plan = psx.FFTPlan(n, psx.PSX_FFT_C2R)
    d_X, d_y = dev(psx, X), out(psx, n)
    plan.execute(d_X, d_y)
    got = d_y.numpy()
    ref = orc.fft_c2r(X, n)
    err = np.max(np.abs(got - ref)) / np.sqrt(np.mean(ref.astype(np.float64) ** 2))
    assert err < 2e-6, (plan.describe(), err)


def test_fft_batched(psx, orc):
    n, b = 1 << 14, 5
    x = RNG.standard_normal((b, n)).astype(np.float32)
    plan = psx.FFTPlan(n, psx.PSX_FFT_R2C, batch=b)
    d_x, d_X = dev(psx, x), out(psx, (b, n // 2 + 1), np.complex64)
    plan.execute(d_x, d_X)
    got = d_X.numpy()
    for i in range(b):
        assert fft_err(got[i], orc.fft_r2c_f64(x[i].astype(np.float64))) < 2e-6


def test_fft_impulse_and_tone(psx):
    n = 1 << 16
    x = np.zeros(n, np.float32)
    x[0] = 1.0
    plan = psx.FFTPlan(n, psx.PSX_FFT_R2C)
    d_x, d_X = dev(psx, x), out(psx, n // 2 + 1, np.complex64)
    plan.execute(d_x, d_X)
    assert np.allclose(d_X.numpy(), 1.0, atol=1e-6)
    k0 = 1234
    x = np.cos(2 * np.pi * k0 * np.arange(n) / n).astype(np.float32)
    d_x.upload(x)
    plan.execute(d_x, d_X)
    X = d_X.numpy()
    assert abs(X[k0].real - n / 2) < 1e-3 * n
    X[k0] = 0
    assert np.max(np.abs(X)) < 1e-3 * n


def test_fft_rejects_bad_sizes(psx):
    p = C.c_void_p()
    assert psx.lib().psx_fft_plan_create(1000, 0, 1, C.byref(p)) == 1
    assert psx.lib().psx_fft_plan_create(5 * 1024, 0, 1, C.byref(p)) == 1


# --------------------------------------------------------- spectrum forming
def rand_spectrum(nb, zeros=True):
    X = (RNG.standard_normal(nb) + 1j * RNG.standard_normal(nb)).astype(np.complex64) * 100
    if zeros:
        X[:7] = 0
        X[nb // 3] = 0
    return X


@pytest.mark.parametrize("nb", [1, 2, 65537, (1 << 22) + 1])
def test_form_amplitude_interbin(psx, orc, nb):
    X = rand_spectrum(nb, zeros=nb > 8)
    d_X, d_A = dev(psx, X), out(psx, nb)
    psx.check(psx.lib().psx_form_amplitude(d_X.ptr, d_A.ptr, nb, None))
    assert np.array_equal(d_A.numpy(), orc.form_amplitude(X))
    psx.check(psx.lib().psx_form_interbin(d_X.ptr, d_A.ptr, nb, None))
    assert np.array_equal(d_A.numpy(), orc.form_interbin(X))


# ----------------------------------------------------------- dereddening
@pytest.mark.parametrize("count", [1, 2, 3, 4, 5, 9, 10, 65537, (1 << 22) + 1])
def test_median_scrunch5(psx, orc, count):
    x = RNG.standard_normal(count).astype(np.float32)
    if count > 20:
        x[10:15] = 1.5  # ties
    d_x, d_o = dev(psx, x), out(psx, max(count // 5, 1))
    psx.check(psx.lib().psx_median_scrunch5(d_x.ptr, count, d_o.ptr, None))
    ref = orc.median_scrunch5(x)
    assert np.array_equal(d_o.numpy()[:ref.size], ref)


# the last two stretch to 2^25 + 1 outputs (2^26 samples), where the float
# index x = i*step is inexact for i >= 2^24 (src/kernels.cu:1019-1020)
@pytest.mark.parametrize("inc,outc", [(2, 10), (13107, 65537), (838860, (1 << 22) + 1), (33554, (1 << 22) + 1),
                                      (6710886, (1 << 25) + 1), (268435, (1 << 25) + 1)])
def test_linear_stretch(psx, orc, inc, outc):
    x = RNG.standard_normal(inc).astype(np.float32)
    d_x, d_o = dev(psx, x), out(psx, outc)
    psx.check(psx.lib().psx_linear_stretch(d_x.ptr, inc, d_o.ptr, outc, None))
    assert np.array_equal(d_o.numpy(), orc.linear_stretch(x, outc))


@pytest.mark.parametrize("n,tsamp", [(1 << 17, 64e-6), (1 << 23, 76e-6), (3 * (1 << 21), 76e-6), (1 << 26, 64e-6)])
def test_deredden_median(psx, orc, n, tsamp):
    nb = n // 2 + 1
    A = np.abs(RNG.standard_normal(nb)).astype(np.float32) * 10
    tobs = np.float32(np.float32(n) * np.float32(tsamp))
    bw = float(np.float32(1.0 / tobs))
    d_A, d_m = dev(psx, A), out(psx, nb)
    d_w = out(psx, psx.lib().psx_deredden_work_floats(nb))
    psx.check(psx.lib().psx_deredden_median(d_A.ptr, nb, bw, 0.05, 0.5, d_m.ptr, d_w.ptr, None))
    assert np.array_equal(d_m.numpy(), orc.calculate_median(A, bw))


def test_divide_c_by_f(psx, orc):
    nb = 100003
    X = rand_spectrum(nb)
    f = (np.abs(RNG.standard_normal(nb)) + 0.1).astype(np.float32)
    d_X, d_f = dev(psx, X), dev(psx, f)
    psx.check(psx.lib().psx_divide_c_by_f(d_X.ptr, d_f.ptr, nb, None))
    assert np.array_equal(d_X.numpy(), orc.divide_c_by_f(X, f))


DEFAULT_ZAPLIST = [(50.0, 0.1), (100.0, 0.15), (150.0, 0.15), (200.0, 0.15), (250.0, 0.15)]  # misc/default_zaplist.txt


@pytest.mark.parametrize("nb,bw", [((1 << 16) + 1, np.float32(1.0 / ((1 << 17) * np.float32(64e-6)))),
                                   ((1 << 22) + 1, np.float32(1.0 / ((1 << 23) * np.float32(76e-6)))),
                                   (2001, np.float32(0.25))])
def test_zap_birdies(psx, orc, nb, bw):
    X = rand_spectrum(nb)
    top = float(bw) * (nb - 1)
    birdies = DEFAULT_ZAPLIST + [(0.01, 0.5), (top - 0.3 * float(bw), 2 * float(bw)), (top + 5.0, 1.0),
                                 (float(bw) * 1000, 0.0), (float(bw) * 1000.5, 0.0)]
    b = np.asarray(birdies, np.float32)
    d_X, d_f, d_w = dev(psx, X), dev(psx, np.ascontiguousarray(b[:, 0])), dev(psx, np.ascontiguousarray(b[:, 1]))
    psx.check(psx.lib().psx_zap_birdies(d_X.ptr, d_f.ptr, d_w.ptr, len(b), bw, nb, None))
    ref = orc.zap_birdies(X, bw, birdies)
    assert np.array_equal(d_X.numpy(), ref)
    assert np.count_nonzero(ref != X) > 0


# ------------------------------------------------------------------ stats
@pytest.mark.parametrize("n,first", [(100, 0), ((1 << 22) + 1, 0), (1000003, 17)])
def test_mean_rms(psx, orc, n, first):
    x = (RNG.standard_normal(n) * 3 + 1.25).astype(np.float32)
    d_x = dev(psx, x)
    d_w = out(psx, psx.lib().psx_reduce_work_bytes(), np.uint8)
    m, r = C.c_float(), C.c_float()
    psx.check(psx.lib().psx_mean_rms(d_x.ptr, n, first, C.byref(m), C.byref(r), d_w.ptr, None))
    om, orr, _ = orc.stats(x, first)
    assert abs(m.value - om) <= 2 * np.spacing(np.float32(abs(om)))
    assert abs(r.value - orr) <= 2 * np.spacing(np.float32(orr))


def test_normalise(psx, orc):
    n = (1 << 20) + 3
    p = RNG.standard_normal(n).astype(np.float32) * 7
    d_p = dev(psx, p)
    psx.check(psx.lib().psx_normalise(d_p.ptr, np.float32(1.5), np.float32(3.25), n, None))
    assert np.array_equal(d_p.numpy(), orc.normalise(p, 1.5, 3.25))


@pytest.mark.parametrize("offset", [0.0, 0.3, 1000.0])
def test_remove_baseline(psx, orc, offset):
    n = 1 << 20
    x = (RNG.standard_normal(n) + offset).astype(np.float32)
    d_x = dev(psx, x)
    d_w = out(psx, psx.lib().psx_reduce_work_bytes(), np.uint8)
    psx.check(psx.lib().psx_remove_baseline(d_x.ptr, n, d_w.ptr, None))
    got = d_x.numpy()
    ref = orc.remove_baseline(x)
    assert np.max(np.abs(got - ref)) <= 4 * np.spacing(np.float32(max(offset, 1.0)))


def test_fill_and_convert(psx):
    n = 10007
    d = out(psx, n)
    psx.check(psx.lib().psx_fill(d.ptr, 0, n, 2.5, None))
    psx.check(psx.lib().psx_fill(d.ptr, 100, 200, 0.0, None))
    h = d.numpy()
    assert np.all(h[:100] == 2.5) and np.all(h[100:200] == 0) and np.all(h[200:] == 2.5)
    u = RNG.integers(0, 256, n).astype(np.uint8)
    d_u = dev(psx, u)
    psx.check(psx.lib().psx_convert_to_f32(d_u.ptr, psx.PSX_U8, d.ptr, n, None))
    assert np.array_equal(d.numpy(), u.astype(np.float32))


# --------------------------------------------------------- harmonic sums
@pytest.mark.parametrize("nfolds", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("nb", [33, (1 << 16) + 1, (1 << 22) + 1])
def test_harmonic_sum(psx, orc, nfolds, nb):
    P = RNG.standard_normal(nb).astype(np.float32)
    d_P = dev(psx, P)
    folds = [out(psx, nb) for _ in range(nfolds)]
    arr = (C.c_void_p * nfolds)(*[f.ptr for f in folds])
    psx.check(psx.lib().psx_harmonic_sum(d_P.ptr, arr, nfolds, nb, None))
    ref = orc.harmonic_sum(P, nfolds)
    for k in range(nfolds):
        assert np.array_equal(folds[k].numpy(), ref[k]), k


def test_harmonic_sum_comb(psx):
    """src/harmonic_sum_test.cpp:21-27 comb (x[i]=1 if i%32==0): closed form."""
    nb = 1 << 20
    P = np.zeros(nb, np.float32)
    P[::32] = 1
    d_P = dev(psx, P)
    folds = [out(psx, nb) for _ in range(4)]
    arr = (C.c_void_p * 4)(*[f.ptr for f in folds])
    psx.check(psx.lib().psx_harmonic_sum(d_P.ptr, arr, 4, nb, None))
    i = np.arange(nb)
    val = P.astype(np.float64).copy()
    scale = [2 ** -0.5, 0.5, 8 ** -0.5, 0.25]
    fr = [[1], [3, 1], [1, 3, 5, 7], list(range(1, 16, 2))]
    for k in range(4):
        m = k + 1
        for j in fr[k]:
            val += P[(i * j + (1 << (m - 1))) >> m]
        assert np.allclose(folds[k].numpy(), val * scale[k], rtol=1e-6, atol=1e-6)


def test_harmonic_sum_rejects(psx):
    d = out(psx, 16)
    arr = (C.c_void_p * 6)(*([d.ptr] * 6))
    assert psx.lib().psx_harmonic_sum(d.ptr, arr, 6, 16, None) == 1
    assert psx.lib().psx_harmonic_sum(d.ptr, arr, 0, 16, None) == 1


# ----------------------------------------------------------- peak finding
def run_peaks(psx, p, start, end, thr, cap):
    n = p.size
    d_p = dev(psx, p)
    d_i, d_s = out(psx, max(cap, 1), np.int32), out(psx, max(cap, 1))
    d_c = out(psx, 1, np.uint32)
    d_w = out(psx, psx.lib().psx_find_peaks_work_bytes(max(end - start, 1)), np.uint8)
    psx.check(psx.lib().psx_find_peaks(d_p.ptr, start, end, thr, d_i.ptr, d_s.ptr, cap, d_c.ptr, d_w.ptr, None))
    c = int(d_c.numpy()[0])
    k = min(c, cap)
    return c, d_i.numpy()[:k], d_s.numpy()[:k]


@pytest.mark.parametrize("n,start,end,thr", [(1000, 0, 1000, 2.0), ((1 << 22) + 1, 13, (1 << 22) + 1, 4.0),
                                             ((1 << 22) + 1, 100000, 3000000, 3.0), (5000, 6000, 5000, 0.0),
                                             (70000, 0, 70000, 10.0)])
def test_find_peaks(psx, orc, n, start, end, thr):
    p = RNG.standard_normal(n).astype(np.float32)
    if n > 1000:
        p[n // 2: n // 2 + 40] = 9.0  # a dense cluster
    c, idx, snr = run_peaks(psx, p, start, end, thr, n)
    ri, rs = orc.threshold(p, min(end, n), start, thr)
    assert c == ri.size
    assert np.array_equal(idx, ri) and np.array_equal(snr, rs)


def test_find_peaks_overflow_reports_count(psx, orc):
    n = 50000
    p = np.full(n, 10.0, np.float32)
    c, idx, snr = run_peaks(psx, p, 0, n, 1.0, 1000)
    assert c == n
    assert np.array_equal(idx, np.arange(1000))


# ------------------------------------------------------------ synthetic
def test_synth_deterministic_and_white(psx):
    n = 1 << 20
    d = out(psx, n)
    w = out(psx, psx.lib().psx_synth_work_bytes(n), np.uint8)
    psx.check(psx.lib().psx_synth_series(d.ptr, n, 64e-6, 12345, 3, 0.0, None, 0, w.ptr, None))
    a = d.numpy()
    psx.check(psx.lib().psx_synth_series(d.ptr, n, 64e-6, 12345, 3, 0.0, None, 0, w.ptr, None))
    assert np.array_equal(a, d.numpy())
    assert abs(a.mean()) < 1e-6 and abs(a.std() - 1) < 0.01
    psx.check(psx.lib().psx_synth_series(d.ptr, n, 64e-6, 12345, 4, 5.0, None, 0, w.ptr, None))
    b = d.numpy()
    assert not np.array_equal(a, b) and abs(b.mean()) < 1e-4
