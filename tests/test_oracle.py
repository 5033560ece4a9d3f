"""Pins for the CPU oracle (runs without a GPU).

Each oracle function is checked against something that does not restate it:
the reference's own golden output, numpy's FFT, brute force on small inputs,
closed forms, or the reference's own distiller compiled into oracle/_ref.
"""
import ctypes as C
import itertools
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
RNG = np.random.default_rng(7)
f32 = np.float32


# ------------------------------------------------------ acceleration plan
def overview_params():
    with open(os.path.join(GOLD, "overview_search_params.json")) as f:
        return json.load(f)


def test_accel_plan_matches_overview_golden():
    """example_output/overview.xml:91-10850 - 10,758 trials at DM = cdm = 54.5."""
    p = overview_params()
    gold = np.load(os.path.join(GOLD, "overview_accel_list.npy"))
    # pipeline_multi.cu:412-421: tsamp, cfreq*1e6, foff*1e6 from float header fields
    fch1, foff = f32(p["fch1"]), f32(p["foff"])
    cfreq = f32(fch1 + foff * f32(p["nchans"]) / f32(2)) if foff < 0 else f32(fch1 - foff * p["nchans"] / 2)
    acc = O.accel_list(p["acc_start"], p["acc_end"], f32(p["acc_tol"]), f32(p["acc_pulse_width"] * 1e-6),
                       p["size"], f32(p["tsamp"]), float(f32(float(cfreq) * 1e6)), float(f32(float(foff) * 1e6)),
                       p["dm"], p["cdm"])
    assert acc.size == gold.size == p["count"]
    assert np.array_equal(acc, gold)


def test_accel_plan_closed_form_and_cases():
    # step = 2 * w * 24 c / T^2 * sqrt(tol^2 - 1), w = sqrt(pw^2 + tsamp^2) when dm == cdm
    n, ts, pw, tol = 1 << 23, f32(76e-6), f32(64e-6), f32(1.1)
    step = O.accel_step(-100, 100, tol, pw, n, ts, 1.4e9, 1e6, 0.0, 0.0)
    T = float(f32(n) * ts)
    w = np.sqrt(float(pw) ** 2 + float(ts) ** 2)
    closed = 2 * w * 24 * 299792458.0 / T ** 2 * np.sqrt(float(tol) ** 2 - 1)
    assert abs(step - closed) / closed < 1e-6
    acc = O.accel_list(-100, 100, tol, pw, n, ts, 1.4e9, 1e6, 0.0, 0.0)
    assert acc.size == 127 and acc[0] == -100 and acc[-1] == 100 and 0.0 in acc
    assert np.all(np.diff(acc) > 0)
    # equal bounds -> single trial; range not spanning zero -> no zero
    assert list(O.accel_list(5, 5, tol, pw, n, ts, 1.4e9, 1e6)) == [5.0]
    a = O.accel_list(10, 20, tol, pw, n, ts, 1.4e9, 1e6)
    assert a[0] == 10 and a[-1] == 20 and 0.0 not in a
    # dispersive smearing widens the step away from cdm
    far = O.accel_step(-100, 100, tol, pw, n, ts, 1.4e9, 1e6, 100.0, 0.0)
    assert far > step


# --------------------------------------------------------------- resample
def test_resample_identity_and_equivalence():
    n = 1 << 16
    x = RNG.standard_normal(n).astype(f32)
    assert np.array_equal(O.resample2(x, 0.0, 64e-6), x)
    assert np.array_equal(O.resample1(x, 0.0, 64e-6), x)
    # src/resampling_test.cpp: resample and resampleII agree on x[i] = i % 451
    n = 4194304
    x = (np.arange(n) % 451).astype(f32)
    a = O.resample1(x, 125.5, 0.000064)
    b = O.resample2(x, 125.5, 0.000064)
    assert np.count_nonzero(np.abs(a - b) > 1e-4) <= 16


def test_resample_index_closed_form():
    """i + i*af*(i-n) rounded to nearest; fma vs unfused differ only at ties."""
    n, acc, ts = 1 << 20, 77.7, f32(64e-6)
    af = float(f32(f32(acc) * ts)) / (2 * 299792458.0)
    idx = np.arange(0, n, 997, dtype=np.int64)
    exact = idx + idx * af * (idx - n)
    got = np.array([O.resample2_index(i, n, acc, ts) for i in idx])
    assert np.all(np.abs(got - exact) <= 0.5 + 1e-9)
    nofma = np.array([O.resample2_index(i, n, acc, ts, fma=False) for i in idx])
    assert np.count_nonzero(got != nofma) <= 2


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 4, 6, 8, 12, 16, 24, 48, 96, 1024, 3072, 1 << 15, 3 * (1 << 14)])
def test_fft_vs_numpy(n):
    x = RNG.standard_normal(n)
    got = O.fft_r2c_f64(x)
    ref = np.fft.rfft(x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.sqrt(n) * max(1.0, np.max(np.abs(ref)))


def test_fft_brute_force_and_parseval():
    n = 96
    x = RNG.standard_normal(n)
    k = np.arange(n // 2 + 1)[:, None]
    dft = (x[None, :] * np.exp(-2j * np.pi * k * np.arange(n)[None, :] / n)).sum(1)
    assert np.allclose(O.fft_r2c_f64(x), dft, atol=1e-10)
    n = 1 << 12
    x = RNG.standard_normal(n)
    X = O.fft_r2c_f64(x)
    e = np.abs(X[0]) ** 2 + np.abs(X[-1]) ** 2 + 2 * np.sum(np.abs(X[1:-1]) ** 2)
    assert abs(e / n - np.sum(x ** 2)) < 1e-9 * np.sum(x ** 2)


def test_fft_complex_and_roundtrip():
    m = 3 * 256
    z = RNG.standard_normal(m) + 1j * RNG.standard_normal(m)
    assert np.allclose(O.fft_c2c_f64(z, -1), np.fft.fft(z), atol=1e-10)
    assert np.allclose(O.fft_c2c_f64(z, +1), np.fft.ifft(z) * m, atol=1e-10)
    n = 1 << 14
    x = RNG.standard_normal(n).astype(f32)
    y = O.fft_c2r(O.fft_r2c(x), n)
    assert np.max(np.abs(y / n - x)) < 1e-5


def test_fft_unsupported_size():
    with pytest.raises(ValueError):
        O.fft_r2c(np.zeros(10, f32))


# ------------------------------------------------------- spectrum forming
def test_form_closed_forms():
    X = np.array([0, 3 + 4j, -5 + 12j, 0, 1e-10 + 0j], np.complex64)
    assert np.array_equal(O.form_amplitude(X), np.array([0, 5, 13, 0, 1e-10], f32))
    A = O.form_interbin(X)
    # bin 1: max(|X1|^2, |X1-X0|^2/2) = 25; bin 3: max(0, |X3 - X2|^2/2) = 169/2
    assert A[1] == 5 and A[2] == 13 and A[3] == f32(np.sqrt(f32(84.5)))
    X = (RNG.standard_normal(1000) + 1j * RNG.standard_normal(1000)).astype(np.complex64)
    ref = np.sqrt(np.maximum(np.abs(X.astype(np.complex128)) ** 2,
                             0.5 * np.abs(np.diff(np.concatenate([[0], X.astype(np.complex128)]))) ** 2))
    assert np.allclose(O.form_interbin(X), ref, rtol=3e-7)


# ------------------------------------------------------------ dereddening
def median5_network(a, b, c, d, e):
    """The seven compare-exchanges the device kernel uses (kernels_dev.hpp)."""
    def cs(x, y):
        return min(x, y), max(x, y)
    a, b = cs(a, b)
    d, e = cs(d, e)
    a, d = cs(a, d)
    b, e = cs(b, e)
    b, c = cs(b, c)
    c, d = cs(c, d)
    b, c = cs(b, c)
    return c


def test_median5_network_exhaustive():
    for vals in [(1, 2, 3, 4, 5), (1, 1, 2, 3, 3), (2, 2, 2, 1, 3), (0, 0, 0, 0, 1)]:
        for perm in itertools.permutations(vals):
            assert median5_network(*perm) == sorted(perm)[2]


@pytest.mark.parametrize("count", [1, 2, 3, 4, 5, 7, 10, 12345])
def test_median_scrunch5_brute(count):
    x = RNG.standard_normal(count).astype(f32)
    got = O.median_scrunch5(x)
    if count == 1:
        ref = x[:1]
    elif count == 2:
        ref = [f32(0.5) * (x[0] + x[1])]
    elif count in (3, 4):
        s = np.sort(x)
        ref = [s[1]] if count == 3 else [f32(0.5) * (s[1] + s[2])]
    else:
        ref = np.median(x[: 5 * (count // 5)].reshape(-1, 5), axis=1).astype(f32)
    assert np.array_equal(got, np.asarray(ref, f32))


def test_linear_stretch_closed_form():
    x = np.array([0, 10, 20, 40], f32)
    y = O.linear_stretch(x, 7)
    assert np.allclose(y, [0, 5, 10, 15, 20, 30, 40])
    x = RNG.standard_normal(1000).astype(f32)
    y = O.linear_stretch(x, 100000)
    pos = np.arange(100000) * ((1000 - 1) / (100000 - 1))
    assert np.allclose(y, np.interp(pos, np.arange(1000), x), atol=2e-4)


def test_calculate_median_zones():
    n, ts = 1 << 17, f32(64e-6)
    nb = n // 2 + 1
    bw = float(f32(1.0 / (f32(n) * ts)))
    A = np.abs(RNG.standard_normal(nb)).astype(f32)
    med = O.calculate_median(A, bw)
    pos5, pos25 = int(0.05 / bw), int(0.5 / bw)
    s5 = O.linear_stretch(O.median_scrunch5(A), nb)
    s25 = O.linear_stretch(O.median_scrunch5(O.median_scrunch5(A)), nb)
    s125 = O.linear_stretch(O.median_scrunch5(O.median_scrunch5(O.median_scrunch5(A))), nb)
    assert np.array_equal(med[:pos5], s5[:pos5])
    assert np.array_equal(med[pos5:pos25], s25[pos5:pos25])
    assert np.array_equal(med[pos25:], s125[pos25:])
    X = (RNG.standard_normal(nb) + 1j * RNG.standard_normal(nb)).astype(np.complex64)
    D = O.divide_c_by_f(X, med + 1)
    assert np.all(D[:5] == 0) and np.allclose(D[5:], X[5:] / (med[5:] + 1), rtol=1e-6)


# ------------------------------------------------------------------ stats
def test_stats_and_normalise():
    x = (RNG.standard_normal(100001) * 2 + 0.5).astype(f32)
    m, r, s = O.stats(x)
    assert abs(m - x.astype(np.float64).mean()) < 1e-6
    assert abs(r - np.sqrt(np.mean(x.astype(np.float64) ** 2))) < 1e-6
    assert abs(s - x.astype(np.float64).std()) < 1e-5
    y = O.normalise(x, m, s)
    assert np.array_equal(y, ((x - f32(m)) / f32(s)).astype(f32))
    z = O.remove_baseline(x + f32(100))
    assert abs(z.astype(np.float64).mean()) < 1e-4


# --------------------------------------------------------- harmonic sums
def harmonic_brute(P, nfolds):
    nb = P.size
    i = np.arange(nb)
    order = [[1], [3, 1], [1, 3, 5, 7], list(range(1, 16, 2)), list(range(1, 32, 2))]
    scale = [1 / np.sqrt(2), 0.5, 1 / np.sqrt(8), 0.25, 1 / np.sqrt(32)]
    val = P.copy()
    out = []
    for k in range(nfolds):
        for j in order[k]:
            idx = np.floor(i * (j / 2.0 ** (k + 1)) + 0.5).astype(np.int64)  # (int)(idx*j/2^m + 0.5)
            val = (val + P[idx]).astype(f32)
        out.append((val.astype(np.float64) * scale[k]).astype(f32))
    return np.array(out)


@pytest.mark.parametrize("nfolds", [1, 2, 3, 4, 5])
def test_harmonic_sum_brute(nfolds):
    P = RNG.standard_normal(20001).astype(f32)
    assert np.array_equal(O.harmonic_sum(P, nfolds), harmonic_brute(P, nfolds))


# ------------------------------------------------------------ peak finder
def test_threshold_and_clustering():
    p = np.zeros(200, f32)
    p[[10, 12, 40, 41, 75, 130, 150]] = [8, 9, 7.5, 8.5, 20, 8, 9]
    idx, snr = O.threshold(p, 200, 0, 7.0)
    assert list(idx) == [10, 12, 40, 41, 75, 130, 150]
    # clusters continue while i - (index of the last new maximum) < 30:
    #  10, 12 (new max), 40 (28 from 12), 41 (29 from 12, not a new max) | 75 (63 from 12)
    #  | 130, 150 (new max, 20 from 130)
    bw = 1.0
    pi, ps, pf = O.find_candidates(p, bw, 0, 7.0, 0.0, 1e9, 30)
    assert list(pi) == [12, 75, 150] and list(ps) == [9, 20, 9]
    nyq = f32(bw * 200)
    assert pf[0] == f32(12 * f32(1.0 / 200 * float(nyq)))


def test_find_candidates_freq_window():
    nb = 1001
    p = np.full(nb, 10.0, f32)
    bw = 0.5
    # start = (int)((float)(2*(nb-1)) * (fmin/nyquist) * 2^nh); end = min(nb, (int)(fmax/bw * 2^nh))
    for nh in (0, 1, 3):
        pi, ps, pf = O.find_candidates(p, bw, nh, 7.0, 2.0, 100.0, 30)
        start = int(float(f32(f32(2 * (nb - 1)) * (f32(2.0) / f32(bw * nb)))) * 2.0 ** nh)
        end = min(nb, int(100.0 / bw * 2 ** nh))
        expect = np.arange(start, end)
        # one cluster per 30 bins: only the first bin of each flat run is kept
        assert pi[0] == expect[0]
        assert pi[-1] <= expect[-1]


# ------------------------------------------------------------ distillers
REF_SO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref.so")


def _ref():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = C.CDLL(REF_SO)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.ref_harmonic_distill.argtypes = [fp, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, fp, ip]
    L.ref_accel_distill.argtypes = [fp, C.c_int, C.c_float, C.c_float, C.c_int, fp, ip]
    L.ref_dm_distill.argtypes = [fp, C.c_int, C.c_float, C.c_int, fp, ip]
    return L


def random_cands(n, rng):
    base = rng.uniform(1, 50, size=max(1, n // 6))
    rows = []
    for _ in range(n):
        f = base[rng.integers(base.size)] * rng.choice([1, 2, 3, 0.5, 0.25, 4, 1.5]) * (1 + rng.normal(0, 5e-5))
        rows.append((rng.choice([10.0, 10.5]), rng.integers(0, 3), rng.uniform(-50, 50), rng.integers(0, 5),
                     rng.uniform(7, 30), f))
    return rows


def rows_to_orc(rows):
    a = np.zeros(len(rows), O.CAND_DTYPE)
    for i, r in enumerate(rows):
        a[i] = (r[0], r[1], r[2], r[3], r[4], r[5], 0, 0)
    return a


def ref_rows(rows):
    return np.array(rows, np.float32).reshape(-1, 6)


@pytest.mark.parametrize("seed", range(6))
def test_distillers_match_reference_build(seed):
    L = _ref()
    rng = np.random.default_rng(seed)
    rows = random_cands(300, rng)
    R = ref_rows(rows)
    oc = rows_to_orc(rows)
    outr = np.zeros_like(R)
    na = np.zeros(len(rows), np.int32)
    for keep, frac in [(0, 1), (1, 0), (1, 1)]:
        n = L.ref_harmonic_distill(R, len(rows), 1e-4, 16.0, keep, frac, outr, na)
        got = O.harmonic_distill(oc, 1e-4, 16.0, keep, frac)
        assert got.size == n
        assert np.array_equal(got["freq"], outr[:n, 5]) and np.array_equal(got["snr"], outr[:n, 4])
        assert np.array_equal(got["nassoc"], na[:n])
    n = L.ref_accel_distill(R, len(rows), 637.5, 1e-4, 1, outr, na)
    got = O.accel_distill(oc, 637.5, 1e-4, True)
    assert got.size == n and np.array_equal(got["freq"], outr[:n, 5]) and np.array_equal(got["nassoc"], na[:n])
    n = L.ref_dm_distill(R, len(rows), 1e-4, 1, outr, na)
    got = O.dm_distill(oc, 1e-4, True)
    assert got.size == n and np.array_equal(got["freq"], outr[:n, 5]) and np.array_equal(got["nassoc"], na[:n])


def test_distillers_closed_cases():
    rows = [(0, 0, 0, 0, 20, 10.0), (0, 0, 0, 1, 15, 20.0), (0, 0, 0, 0, 12, 10.002), (0, 0, 0, 0, 11, 10.0005),
            (0, 0, 0, 0, 10, 13.0)]
    out = O.harmonic_distill(rows_to_orc(rows), 1e-4, 16, False, True)
    # 20 Hz (nh=1) is 2x 10 Hz; 10.0005 is within 1e-4 of 10 Hz; 10.002 and 13 are unrelated
    assert sorted(out["freq"].tolist()) == sorted([10.0, float(np.float32(10.002)), 13.0])


# ----------------------------------------------------- oracle end to end
def test_oracle_trial_finds_injected_tone():
    n, ts = 1 << 15, f32(64e-6)
    t = np.arange(n) * float(ts)
    x = RNG.standard_normal(n) + 0.08 * np.cos(2 * np.pi * 37.0 * t)
    p = O.make_params(n, ts, min_snr=6.0)
    tim, mean, rms, std = O.prologue(x.astype(f32), p)
    cands = O.trial(tim, p, 0.0, mean, std)
    assert cands.size > 0
    best = cands[np.argmax(cands["snr"])]
    assert abs(best["freq"] - 37.0) < 2 * (1 / (n * float(ts)))


# ------------------------------------------- float FFT of the CPU baseline
@pytest.mark.parametrize("n", [16, 1024, 8192, 6144, 3 << 13, 1 << 17, 3 << 17, 1 << 20, 1 << 23])
def test_fast_fft_against_exact(n):
    """The CPU baseline's float four-step FFT agrees with the exact double
    DFT to single-precision rounding (2e-6 of the RMS bin, the bound the GPU
    FFT is held to), and its C2R inverts it (times n)."""
    x = RNG.standard_normal(n).astype(np.float32)
    X = O.fft_r2c_fast(x)
    ref = O.fft_r2c_f64(x.astype(np.float64))
    scale = np.sqrt(np.mean(np.abs(ref) ** 2))
    assert np.max(np.abs(X.astype(np.complex128) - ref)) / scale < 2e-6
    y = O.fft_c2r_fast(X, n)
    assert np.max(np.abs(y / n - x)) / np.sqrt(np.mean(x.astype(np.float64) ** 2)) < 4e-6
    # impulse: a flat spectrum; the closed form of a tone
    d = np.zeros(n, np.float32)
    d[0] = 1.0
    assert np.allclose(O.fft_r2c_fast(d), 1.0, atol=1e-6)
    k = n // 8
    tone = np.cos(2 * np.pi * k * np.arange(n) / n).astype(np.float32)
    T = O.fft_r2c_fast(tone)
    assert abs(T[k] - n / 2) < 1e-3 * n and np.max(np.abs(np.delete(T, k))) < 1e-4 * n


def test_fast_fft_search_matches_exact():
    """A search with the float FFT finds the exact-FFT search's candidates
    (2^17 samples, an injected 5.3 ms pulsar, red noise)."""
    import sys
    sys.path.insert(0, HERE)
    from candmatch import compare
    n = 1 << 17
    t = np.arange(n) * 64e-6
    rng = np.random.default_rng(11)
    x = rng.standard_normal(n) + 0.5 * np.exp(-0.5 * ((t % 5.3e-3 - 2e-3) / 1.2e-4) ** 2)
    x = (x + 0.02 * np.cumsum(rng.standard_normal(n))).astype(np.float32)
    accs = np.array([-20.0, 0.0, 20.0], np.float32)
    exact, _ = O.search_dm(x, O.make_params(n, f32(64e-6)), accs, nthreads=2)
    fast, _ = O.search_dm(x, O.make_params(n, f32(64e-6), fast_fft=True), accs, nthreads=2)
    assert len(exact) > 0
    assert not compare(fast, exact, 7.0)
