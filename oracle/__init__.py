"""CPU oracle for the acceleration-search core (test infrastructure only).

ctypes bindings over ``oracle/liboracle.so`` (built from ``oracle.cpp`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package; the product library
never calls it.  Each function restates a reference routine; the citations
are in ``oracle.h``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class Cand(C.Structure):
    _fields_ = [("dm", C.c_float), ("dm_idx", C.c_int32), ("acc", C.c_float),
                ("nh", C.c_int32), ("snr", C.c_float), ("freq", C.c_float),
                ("nassoc", C.c_int32), ("_pad", C.c_int32)]


CAND_DTYPE = np.dtype([("dm", "<f4"), ("dm_idx", "<i4"), ("acc", "<f4"), ("nh", "<i4"),
                       ("snr", "<f4"), ("freq", "<f4"), ("nassoc", "<i4"), ("_pad", "<i4")])


class SearchParams(C.Structure):
    _fields_ = [("size", C.c_uint64), ("tsamp", C.c_float), ("dm", C.c_float),
                ("dm_idx", C.c_int32), ("min_snr", C.c_float), ("min_freq", C.c_float),
                ("max_freq", C.c_float), ("nharmonics", C.c_int32), ("freq_tol", C.c_float),
                ("max_harm", C.c_float), ("min_gap", C.c_int32),
                ("boundary_5_freq", C.c_float), ("boundary_25_freq", C.c_float), ("nbirdies", C.c_int32),
                ("birdie_freq", C.c_void_p), ("birdie_width", C.c_void_p), ("fast_fft", C.c_int32)]


def make_params(size, tsamp, dm=0.0, dm_idx=0, min_snr=7.0, min_freq=0.1, max_freq=1100.0,
                nharmonics=4, freq_tol=1e-4, max_harm=16.0, min_gap=30,
                boundary_5_freq=0.05, boundary_25_freq=0.5, birdies=None, fast_fft=False):
    """fast_fft: the float four-step FFT (CPU baseline timing) instead of the
    exact double DFT the checks use."""
    p = SearchParams(size, tsamp, dm, dm_idx, min_snr, min_freq, max_freq, nharmonics,
                     freq_tol, max_harm, min_gap, boundary_5_freq, boundary_25_freq, 0, None, None,
                     1 if fast_fft else 0)
    if birdies is not None and len(birdies):
        b = np.asarray(birdies, np.float32).reshape(-1, 2)
        p._birdies = (np.ascontiguousarray(b[:, 0]), np.ascontiguousarray(b[:, 1]))
        p.nbirdies = len(b)
        p.birdie_freq = p._birdies[0].ctypes.data
        p.birdie_width = p._birdies[1].ctypes.data
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(_LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        candp = np.ctypeslib.ndpointer(CAND_DTYPE, flags="C_CONTIGUOUS")
        u64, f32, i32 = C.c_uint64, C.c_float, C.c_int
        sig = {
            "orc_accel_list": (C.c_int, [f32, f32, f32, f32, u64, f32, f32, f32, f32, f32, f32p, i32]),
            "orc_accel_step": (C.c_float, [f32, f32, f32, f32, u64, f32, f32, f32, f32, f32]),
            "orc_resample2_index": (u64, [u64, u64, f32, f32]),
            "orc_resample2_index_nofma": (u64, [u64, u64, f32, f32]),
            "orc_resample2": (None, [f32p, f32p, u64, f32, f32]),
            "orc_resample1": (None, [f32p, f32p, u64, f32, f32]),
            "orc_fft_r2c": (C.c_int, [f32p, f32p, u64]),
            "orc_fft_r2c_f64": (C.c_int, [f64p, f64p, u64]),
            "orc_fft_c2r": (C.c_int, [f32p, f32p, u64]),
            "orc_fft_c2c_f64": (C.c_int, [f64p, u64, i32]),
            "orc_fft_r2c_fast": (C.c_int, [f32p, f32p, u64]),
            "orc_fft_c2r_fast": (C.c_int, [f32p, f32p, u64]),
            "orc_form_amplitude": (None, [f32p, f32p, u64]),
            "orc_form_interbin": (None, [f32p, f32p, u64]),
            "orc_median_scrunch5": (u64, [f32p, u64, f32p]),
            "orc_linear_stretch": (None, [f32p, u64, f32p, u64]),
            "orc_calculate_median": (None, [f32p, u64, C.c_double, f32, f32, f32p]),
            "orc_divide_c_by_f": (None, [f32p, f32p, u64]),
            "orc_zap_birdies": (None, [f32p, u64, f32, f32p, f32p, i32]),
            "orc_stats": (None, [f32p, u64, u64, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float)]),
            "orc_normalise": (None, [f32p, f32, f32, u64]),
            "orc_remove_baseline": (None, [f32p, u64]),
            "orc_harmonic_sum": (None, [f32p, f32p, u64, C.c_uint]),
            "orc_threshold": (C.c_int, [f32p, i32, i32, f32, i32p, f32p, i32]),
            "orc_find_candidates": (C.c_int, [f32p, C.c_uint32, C.c_double, C.c_uint32, f32, f32, f32,
                                              i32, i32p, f32p, f32p, i32]),
            "orc_harmonic_distill": (C.c_int, [candp, i32, f32, f32, i32, i32, candp]),
            "orc_accel_distill": (C.c_int, [candp, i32, f32, f32, i32, candp]),
            "orc_dm_distill": (C.c_int, [candp, i32, f32, i32, candp]),
            "orc_prologue": (C.c_int, [f32p, u64, C.POINTER(SearchParams), f32p,
                                       C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
            "orc_trial": (C.c_int, [f32p, C.POINTER(SearchParams), f32, f32, f32, candp, i32,
                                    C.c_void_p, C.c_void_p, C.c_void_p, i32, C.POINTER(C.c_int)]),
            "orc_search_dm": (C.c_int, [f32p, u64, C.POINTER(SearchParams), f32p, i32, candp, i32, i32,
                                        C.c_void_p]),
            "orc_search_accels": (C.c_int, [f32p, C.POINTER(SearchParams), f32p, i32, f32, f32, candp,
                                            i32, i32]),
            "orc_score": (C.c_int, [candp, i32, f32, f32, f32, f32p, i32p]),
            "orc_finalize": (C.c_int, [candp, i32, f32, f32, f32, f32, f32, i32, candp, i32, f32p, i32p,
                                       C.POINTER(C.c_int)]),
            "orc_dm_list": (C.c_int, [f32, f32, f32, f32, f32, f32, i32, f32, f32p, i32]),
            "orc_delay_table": (None, [i32, f32, f32, f32, f32p]),
            "orc_dedisperse": (C.c_long, [np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS"), C.c_long, i32,
                                          i32, f32p, i32, C.c_void_p, f32, f32, f32, f32p]),
            "orc_hardware_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- wrappers
def accel_list(acc_lo, acc_hi, tol, pulse_width, nsamps, tsamp, cfreq, ch_bw, dm=0.0, cdm=0.0):
    n = lib().orc_accel_list(acc_lo, acc_hi, tol, pulse_width, nsamps, tsamp, cfreq, ch_bw, dm, cdm,
                             np.zeros(1, np.float32), 0)
    out = np.zeros(max(n, 1), np.float32)
    lib().orc_accel_list(acc_lo, acc_hi, tol, pulse_width, nsamps, tsamp, cfreq, ch_bw, dm, cdm, out, n)
    return out[:n]


def accel_step(acc_lo, acc_hi, tol, pulse_width, nsamps, tsamp, cfreq, ch_bw, dm=0.0, cdm=0.0):
    return lib().orc_accel_step(acc_lo, acc_hi, tol, pulse_width, nsamps, tsamp, cfreq, ch_bw, dm, cdm)


def resample2(x, acc, tsamp):
    x = _f32(x)
    out = np.empty_like(x)
    lib().orc_resample2(x, out, x.size, acc, tsamp)
    return out


def resample1(x, acc, tsamp):
    x = _f32(x)
    out = np.empty_like(x)
    lib().orc_resample1(x, out, x.size, acc, tsamp)
    return out


def resample2_index(i, n, acc, tsamp, fma=True):
    f = lib().orc_resample2_index if fma else lib().orc_resample2_index_nofma
    return f(int(i), int(n), acc, tsamp)


def fft_r2c(x):
    """Unnormalised forward real DFT (double internally), complex64 output of n/2+1 bins."""
    x = _f32(x)
    out = np.empty(2 * (x.size // 2 + 1), np.float32)
    if lib().orc_fft_r2c(x, out, x.size):
        raise ValueError("unsupported FFT size %d" % x.size)
    return out.view(np.complex64)


def fft_r2c_fast(x):
    """The CPU baseline's float four-step R2C (unnormalised), n/2+1 bins."""
    x = _f32(x)
    out = np.empty(2 * (x.size // 2 + 1), np.float32)
    if lib().orc_fft_r2c_fast(x, out, x.size):
        raise ValueError("unsupported FFT size %d" % x.size)
    return out.view(np.complex64)


def fft_c2r_fast(X, n):
    X = np.ascontiguousarray(X, dtype=np.complex64).view(np.float32)
    out = np.empty(n, np.float32)
    if lib().orc_fft_c2r_fast(X, out, n):
        raise ValueError("unsupported FFT size %d" % n)
    return out


def fft_r2c_f64(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(2 * (x.size // 2 + 1), np.float64)
    if lib().orc_fft_r2c_f64(x, out, x.size):
        raise ValueError("unsupported FFT size %d" % x.size)
    return out.view(np.complex128)


def fft_c2r(X, n):
    X = np.ascontiguousarray(X, dtype=np.complex64).view(np.float32)
    out = np.empty(n, np.float32)
    if lib().orc_fft_c2r(X, out, n):
        raise ValueError("unsupported FFT size %d" % n)
    return out


def fft_c2c_f64(z, sign=-1):
    buf = np.ascontiguousarray(z, dtype=np.complex128).copy().view(np.float64)
    if lib().orc_fft_c2c_f64(buf, buf.size // 2, sign):
        raise ValueError("unsupported FFT size")
    return buf.view(np.complex128)


def form_amplitude(X):
    X = np.ascontiguousarray(X, dtype=np.complex64)
    out = np.empty(X.size, np.float32)
    lib().orc_form_amplitude(X.view(np.float32), out, X.size)
    return out


def form_interbin(X):
    X = np.ascontiguousarray(X, dtype=np.complex64)
    out = np.empty(X.size, np.float32)
    lib().orc_form_interbin(X.view(np.float32), out, X.size)
    return out


def median_scrunch5(x):
    x = _f32(x)
    out = np.zeros(max(x.size // 5, 1), np.float32)
    n = lib().orc_median_scrunch5(x, x.size, out)
    return out[:n]


def linear_stretch(x, out_count):
    x = _f32(x)
    out = np.empty(out_count, np.float32)
    lib().orc_linear_stretch(x, x.size, out, out_count)
    return out


def calculate_median(A, bin_width, b5=0.05, b25=0.5):
    A = _f32(A)
    out = np.zeros_like(A)
    lib().orc_calculate_median(A, A.size, bin_width, b5, b25, out)
    return out


def divide_c_by_f(X, f):
    X = np.ascontiguousarray(X, dtype=np.complex64).copy()
    lib().orc_divide_c_by_f(X.view(np.float32), _f32(f), X.size)
    return X


def stats(p, first=0):
    p = _f32(p)
    m, r, s = C.c_float(), C.c_float(), C.c_float()
    lib().orc_stats(p, p.size, first, C.byref(m), C.byref(r), C.byref(s))
    return m.value, r.value, s.value


def normalise(p, mean, sigma):
    p = _f32(p).copy()
    lib().orc_normalise(p, mean, sigma, p.size)
    return p


def remove_baseline(x):
    x = _f32(x).copy()
    lib().orc_remove_baseline(x, x.size)
    return x


def harmonic_sum(P, nfolds):
    P = _f32(P)
    out = np.zeros((nfolds, P.size), np.float32)
    lib().orc_harmonic_sum(P, out, P.size, nfolds)
    return out


def threshold(p, n, start, thr):
    p = _f32(p)
    cap = p.size
    idx = np.zeros(cap, np.int32)
    snr = np.zeros(cap, np.float32)
    c = lib().orc_threshold(p, n, start, thr, idx, snr, cap)
    return idx[:c], snr[:c]


def find_candidates(p, bin_width, nh, thr, fmin, fmax, min_gap=30):
    p = _f32(p)
    cap = p.size + 1
    idx = np.zeros(cap, np.int32)
    snr = np.zeros(cap, np.float32)
    frq = np.zeros(cap, np.float32)
    n = lib().orc_find_candidates(p, p.size, bin_width, nh, thr, fmin, fmax, min_gap, idx, snr, frq, cap)
    if n < 0:
        raise RuntimeError("peak finder overflow (%d)" % n)
    return idx[:n], snr[:n], frq[:n]


def _cands(rows):
    a = np.zeros(len(rows), CAND_DTYPE)
    for i, r in enumerate(rows):
        a[i] = tuple(r) + (0,) * (8 - len(r))
    return a


def harmonic_distill(cands, tol, max_harm, keep_related=False, fractional=True):
    c = np.ascontiguousarray(cands, dtype=CAND_DTYPE)
    out = np.zeros(max(c.size, 1), CAND_DTYPE)
    n = lib().orc_harmonic_distill(c, c.size, tol, max_harm, int(keep_related), int(fractional), out)
    return out[:n]


def accel_distill(cands, tobs, tol, keep_related=True):
    c = np.ascontiguousarray(cands, dtype=CAND_DTYPE)
    out = np.zeros(max(c.size, 1), CAND_DTYPE)
    n = lib().orc_accel_distill(c, c.size, tobs, tol, int(keep_related), out)
    return out[:n]


def dm_distill(cands, tol, keep_related=True):
    c = np.ascontiguousarray(cands, dtype=CAND_DTYPE)
    out = np.zeros(max(c.size, 1), CAND_DTYPE)
    n = lib().orc_dm_distill(c, c.size, tol, int(keep_related), out)
    return out[:n]


SCORE_DTYPE = np.dtype([("ddm_count_ratio", "<f4"), ("ddm_snr_ratio", "<f4"), ("is_adjacent", "<i4"),
                        ("is_physical", "<i4")])


def _scores(ratios, flags, n):
    out = np.zeros(n, SCORE_DTYPE)
    out["ddm_count_ratio"] = ratios[0:2 * n:2]
    out["ddm_snr_ratio"] = ratios[1:2 * n:2]
    out["is_adjacent"] = flags[0:2 * n:2]
    out["is_physical"] = flags[1:2 * n:2]
    return out


def score(forest, cfreq, foff, bw):
    """CandidateScorer over the top-level candidates of a forest."""
    c = np.ascontiguousarray(forest, dtype=CAND_DTYPE)
    ratios = np.zeros(2 * max(c.size, 1), np.float32)
    flags = np.zeros(2 * max(c.size, 1), np.int32)
    n = lib().orc_score(c, c.size, cfreq, foff, bw, ratios, flags)
    if n < 0:
        raise ValueError("malformed candidate forest")
    return _scores(ratios, flags, n)


def finalize(forest, freq_tol, max_harm, cfreq, foff, bw, limit, cap=None):
    """main()'s post-processing: (final forest, scores of its top-level candidates)."""
    c = np.ascontiguousarray(forest, dtype=CAND_DTYPE)
    cap = cap or max(4 * c.size, 1024)
    out = np.zeros(cap, CAND_DTYPE)
    ratios = np.zeros(2 * max(c.size, 1), np.float32)
    flags = np.zeros(2 * max(c.size, 1), np.int32)
    nfin = C.c_int()
    n = lib().orc_finalize(c, c.size, freq_tol, max_harm, cfreq, foff, bw, limit, out, cap, ratios, flags,
                           C.byref(nfin))
    if n < 0:
        raise ValueError("malformed candidate forest")
    if n > cap:
        raise RuntimeError("orc_finalize: %d records exceed cap %d" % (n, cap))
    return out[:n], _scores(ratios, flags, nfin.value)


def prologue(series, params):
    series = _f32(series)
    tim = np.empty(params.size, np.float32)
    m, r, s = C.c_float(), C.c_float(), C.c_float()
    rc = lib().orc_prologue(series, series.size, C.byref(params), tim, C.byref(m), C.byref(r), C.byref(s))
    if rc:
        raise RuntimeError("orc_prologue failed (%d)" % rc)
    return tim, m.value, r.value, s.value


def trial(tim, params, acc, mean, std, raw=False, cap=1 << 20):
    tim = _f32(tim)
    out = np.zeros(cap, CAND_DTYPE)
    rn = np.zeros(cap, np.int32)
    ri = np.zeros(cap, np.int32)
    rs = np.zeros(cap, np.float32)
    rc = C.c_int(0)
    n = lib().orc_trial(tim, C.byref(params), acc, mean, std, out, cap, rn.ctypes.data, ri.ctypes.data,
                        rs.ctypes.data, cap, C.byref(rc))
    if n < 0:
        raise RuntimeError("orc_trial failed (%d)" % n)
    if raw:
        k = min(rc.value, cap)
        return out[:n], (rn[:k], ri[:k], rs[:k])
    return out[:n]


def zap_birdies(X, bin_width, birdies):
    """X: complex64 spectrum (modified copy returned); birdies: (n, 2) freq, half width."""
    X = np.ascontiguousarray(X, np.complex64).copy()
    b = np.asarray(birdies, np.float32).reshape(-1, 2)
    lib().orc_zap_birdies(X.view(np.float32), X.size, bin_width, np.ascontiguousarray(b[:, 0]),
                          np.ascontiguousarray(b[:, 1]), len(b))
    return X


def top_level(forest):
    """Top-level records of a candidate forest (nassoc = subtree records after each)."""
    na = np.asarray(forest["nassoc"], np.int64)
    idx, i = [], 0
    while i < len(forest):
        idx.append(i)
        i += 1 + int(na[i])
    return forest[np.asarray(idx, np.int64)]


def search_accels(tim, params, accs, mean, std, nthreads=1, cap=1 << 20, forest=False):
    tim = _f32(tim)
    accs = _f32(accs)
    out = np.zeros(cap, CAND_DTYPE)
    n = lib().orc_search_accels(tim, C.byref(params), accs, accs.size, mean, std, out, cap, nthreads)
    if n < 0:
        raise RuntimeError("orc_search_accels failed (%d)" % n)
    if n > cap:
        raise RuntimeError("orc_search_accels: %d records exceed cap %d" % (n, cap))
    return out[:n] if forest else top_level(out[:n])


def search_dm(series, params, accs, nthreads=1, cap=1 << 20, forest=False):
    series = _f32(series)
    accs = _f32(accs)
    out = np.zeros(cap, CAND_DTYPE)
    t = np.zeros(3, np.float64)
    n = lib().orc_search_dm(series, series.size, C.byref(params), accs, accs.size, out, cap, nthreads,
                            t.ctypes.data)
    if n < 0:
        raise RuntimeError("orc_search_dm failed (%d)" % n)
    if n > cap:
        raise RuntimeError("orc_search_dm: %d records exceed cap %d" % (n, cap))
    return (out[:n] if forest else top_level(out[:n])), t


def hardware_threads():
    return lib().orc_hardware_threads()


def dm_list(dm_start, dm_end, tsamp, pulse_width_us, f0, df, nchans, tol):
    """libdedisp's DM list (dedisp_generate_dm_list), float32."""
    n = lib().orc_dm_list(dm_start, dm_end, tsamp, pulse_width_us, f0, df, nchans, tol, np.zeros(1, np.float32), 0)
    if n < 0:
        raise RuntimeError("DM list does not converge")
    out = np.zeros(n, np.float32)
    lib().orc_dm_list(dm_start, dm_end, tsamp, pulse_width_us, f0, df, nchans, tol, out, n)
    return out


def delay_table(nchans, tsamp, f0, df):
    t = np.zeros(nchans, np.float32)
    lib().orc_delay_table(nchans, tsamp, f0, df, t)
    return t


def dedisperse(packed, nsamps, nchans, nbits, dms, tsamp, f0, df, killmask=None):
    """Direct dedispersion of packed samples: [ndm, nout] float32."""
    packed = np.ascontiguousarray(packed, np.uint8)
    dms = _f32(dms)
    km = None if killmask is None else np.ascontiguousarray(killmask, np.int32)
    table = delay_table(nchans, tsamp, f0, df)
    nout = nsamps - int(np.rint(dms[-1] * table[-1]))
    out = np.zeros((dms.size, max(nout, 1)), np.float32)
    n = lib().orc_dedisperse(packed, nsamps, nchans, nbits, dms, dms.size, None if km is None else km.ctypes.data,
                             tsamp, f0, df, out)
    if n < 0:
        raise ValueError("series shorter than the dispersion delay")
    return out[:, :n]
