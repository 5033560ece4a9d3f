"""Synthetic workloads of the acceleration-search benchmark configurations.

Series are generated on the device by ``psx_synth_series`` (counter-based
Philox RNG, so every trial is reproducible from (seed, trial index)).  The
injection parameters of each configuration are drawn on the host from a
numpy generator seeded with the configuration seed.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

import peasoup_amd as P

C_LIGHT = 299792458.0


@dataclass
class Injection:
    trial: int
    period: float
    accel: float
    snr: float
    phase0: float = 0.0
    duty: float = 0.05


@dataclass
class Workload:
    name: str
    ndm: int
    size: int
    tsamp: float
    seed: int
    red_amp: float = 0.0
    acc_lo: float = -50.0
    acc_hi: float = 50.0
    nharmonics: int = 4
    min_snr: float = 7.0
    injections: List[Injection] = field(default_factory=list)
    nsamps: Optional[int] = None  # real samples per series (< size pads with zeros)

    @property
    def n_real(self):
        return self.nsamps if self.nsamps else self.size

    def search_args(self):
        return make_search_args(self.size, self.tsamp, acc_start=self.acc_lo, acc_end=self.acc_hi,
                                nharmonics=self.nharmonics, min_snr=self.min_snr)

    def dm_list(self):
        return np.arange(self.ndm, dtype=np.float32)


class SearchArgs(C.Structure):
    _fields_ = [("size", C.c_uint64), ("tsamp", C.c_float), ("min_snr", C.c_float), ("min_freq", C.c_float),
                ("max_freq", C.c_float), ("nharmonics", C.c_uint32), ("freq_tol", C.c_float),
                ("max_harm", C.c_float), ("acc_start", C.c_float), ("acc_end", C.c_float),
                ("acc_tol", C.c_float), ("acc_pulse_width", C.c_float), ("cdm", C.c_float),
                ("cfreq", C.c_float), ("ch_bw", C.c_float), ("min_gap", C.c_int32), ("nbirdies", C.c_uint32),
                ("birdie_freq", C.c_void_p), ("birdie_width", C.c_void_p)]


CAND_DTYPE = np.dtype([("dm", "<f4"), ("dm_idx", "<i4"), ("acc", "<f4"), ("nh", "<i4"), ("snr", "<f4"),
                       ("freq", "<f4"), ("nassoc", "<i4"), ("reserved", "<i4")])


def make_search_args(size, tsamp, acc_start=-50.0, acc_end=50.0, acc_tol=1.1, acc_pulse_width=64.0,
                     nharmonics=4, min_snr=7.0, min_freq=0.1, max_freq=1100.0, freq_tol=1e-4, max_harm=16.0,
                     cdm=0.0, cfreq=1.4e9, ch_bw=1e6, min_gap=30, birdies=None):
    """birdies: optional (n, 2) frequencies and half-widths in Hz (a zap list)."""
    a = SearchArgs(size, tsamp, min_snr, min_freq, max_freq, nharmonics, freq_tol, max_harm, acc_start,
                   acc_end, acc_tol, acc_pulse_width, cdm, cfreq, ch_bw, min_gap, 0, None, None)
    if birdies is not None and len(birdies):
        b = np.asarray(birdies, np.float32).reshape(-1, 2)
        a._birdies = (np.ascontiguousarray(b[:, 0]), np.ascontiguousarray(b[:, 1]))  # kept alive with a
        a.nbirdies = len(b)
        a.birdie_freq = a._birdies[0].ctypes.data
        a.birdie_width = a._birdies[1].ctypes.data
    return a


def read_zaplist(path):
    """Zap list lines "<frequency> <half width>" (Hz); blank lines skipped
    (include/transforms/birdiezapper.hpp:35-57)."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if len(parts) < 2:
                raise ValueError("zap list line without a width: %r" % line)
            rows.append((float(parts[0]), float(parts[1])))
    return np.array(rows, np.float32).reshape(-1, 2)


def accel_list(size, tsamp, acc_lo, acc_hi, tol=1.1, pulse_width_us=64.0):
    """AccelerationPlan list with dm = cdm = 0 (include/utils/utils.hpp:198-232),
    restated in numpy float32 arithmetic."""
    f32 = np.float32
    ts, pw, tol = f32(tsamp), f32(pulse_width_us * 1e-6), f32(tol)
    if acc_hi == acc_lo:
        return np.array([acc_hi], f32)
    broad = f32(np.sqrt(f32(f32(0) + pw * pw + ts * ts)))
    factor = f32(np.sqrt(np.float64(tol * tol) - 1.0))
    tobs = f32(f32(size) * ts)
    step = f32(2.0 * np.float64(broad) * 24.0 * C_LIGHT / np.float64(tobs * tobs) * np.float64(factor))
    out = [f32(0.0)] if (acc_hi >= 0 and acc_lo <= 0) else []
    acc = f32(acc_lo)
    while acc <= f32(acc_hi):
        out.append(acc)
        acc = f32(acc + step)
    out.append(f32(acc_hi))
    return np.unique(np.array(out, f32))


def config(name):
    """The benchmark configurations (plain CPU case, plain single-GPU scale, ...)."""
    if name == "cpu8x2^17":
        return Workload("cpu8x2^17", 8, 1 << 17, 64e-6, 0x5EA501, 0.0, -50.0, 50.0,
                        injections=[Injection(4, 5.3e-3, 10.0, 15.0, 0.13)])
    if name == "gpu64x2^23":
        w = Workload("gpu64x2^23", 64, 1 << 23, 76e-6, 0x5EA502, 5.0, -100.0, 100.0)
        rng = np.random.default_rng(w.seed)
        trials = rng.choice(w.ndm, 6, replace=False)
        for t in trials:
            w.injections.append(Injection(int(t), float(rng.uniform(2e-3, 200e-3)), float(rng.uniform(-80, 80)),
                                          float(rng.uniform(8, 30)), float(rng.uniform())))
        return w
    if name == "long32x2^26":
        w = Workload("long32x2^26", 32, 1 << 26, 64e-6, 0x5EA504, 5.0, -50.0, 50.0)
        rng = np.random.default_rng(w.seed)
        for t in rng.choice(w.ndm, 4, replace=False):
            w.injections.append(Injection(int(t), float(rng.uniform(1.5e-3, 30e-3)), float(rng.uniform(-50, 50)),
                                          float(rng.uniform(8, 30)), float(rng.uniform())))
        return w
    raise KeyError(name)


class SeriesGenerator:
    """Generates workload series into device buffers."""

    def __init__(self, n_max):
        self.work = P.DeviceArray(P.lib().psx_synth_work_bytes(n_max), np.uint8)

    def generate(self, d_out, n, tsamp, seed, trial, red_amp, injections, stream=None):
        inj = [i for i in injections if i.trial == trial]
        arr = (P.Pulsar * max(1, len(inj)))()
        for k, i in enumerate(inj):
            arr[k] = P.Pulsar(i.period, i.accel, i.snr, i.phase0, i.duty)
        ptr = d_out.ptr if isinstance(d_out, P.DeviceArray) else d_out
        P.check(P.lib().psx_synth_series(ptr, n, tsamp, seed, trial, red_amp, arr if inj else None, len(inj),
                                         self.work.ptr, stream))


def generate_host(w: Workload, trials=None):
    """All (or selected) series of a workload as a host array [ntrials, n_real]."""
    trials = list(range(w.ndm)) if trials is None else list(trials)
    gen = SeriesGenerator(w.n_real)
    d = P.DeviceArray(w.n_real)
    out = np.empty((len(trials), w.n_real), np.float32)
    for k, t in enumerate(trials):
        gen.generate(d, w.n_real, w.tsamp, w.seed, t, w.red_amp, w.injections)
        out[k] = d.numpy()
    return out


class BatchOpts(C.Structure):
    _fields_ = [("batch", C.c_uint32), ("cand_capacity", C.c_uint32), ("use_graph", C.c_int32),
                ("fft_npass", C.c_int32), ("fft_len", C.c_int32 * 3), ("fft_cols", C.c_int32 * 3),
                ("spectrum_layout", C.c_int32), ("keep_engine", C.c_int32), ("harmonic_pass", C.c_int32)]


SPECTRUM_AUTO, SPECTRUM_NATURAL, SPECTRUM_ROWS = 0, 1, 2
HARMONIC_BOUND, HARMONIC_EXACT = 0, 1


def make_batch_opts(batch=16, cand_capacity=0, use_graph=True, decomp=None, spectrum_layout=SPECTRUM_AUTO,
                    keep_engine=False, harmonic_pass=HARMONIC_BOUND):
    o = BatchOpts(batch, cand_capacity, 1 if use_graph else 0, 0)
    o.spectrum_layout = spectrum_layout
    o.keep_engine = 1 if keep_engine else 0
    o.harmonic_pass = harmonic_pass
    if decomp:
        o.fft_npass = len(decomp)
        for q, (L, Cc) in enumerate(decomp):
            o.fft_len[q] = L
            o.fft_cols[q] = Cc
    return o


P.register_signatures({
    "psx_batch_release": (C.c_int, []),
    "psx_batch_search": (C.c_int, [C.POINTER(SearchArgs), C.POINTER(BatchOpts), C.c_void_p, C.c_int, C.c_uint64,
                                   C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                   C.POINTER(C.c_uint32), C.c_void_p]),
    "psx_worker_search": (C.c_int, [C.POINTER(SearchArgs), C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32,
                                    C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                    C.c_void_p]),
})


def top_level(forest):
    """Top-level records of a candidate forest (each record is followed by the
    nassoc records of its associated candidates, in pre-order)."""
    na = np.asarray(forest["nassoc"], np.int64)
    idx = []
    i, n = 0, len(forest)
    while i < n:
        idx.append(i)
        i += 1 + int(na[i])
    return forest[np.asarray(idx, np.int64)]


def worker_search(args, series, dm_list, accs=None, cap=1 << 20, forest=False):
    """Plain Worker sequence (one kernel per reference call) on host series.
    Returns the top-level candidates, or the whole forest with forest=True."""
    series = np.ascontiguousarray(series, np.float32)
    dm_list = np.ascontiguousarray(dm_list, np.float32)
    out = np.zeros(cap, CAND_DTYPE)
    cnt = C.c_uint32()
    secs = np.zeros(3, np.float64)
    a = None if accs is None else np.ascontiguousarray(accs, np.float32)
    P.check(P.lib().psx_worker_search(C.byref(args), series.ctypes.data, series.shape[-1], dm_list.ctypes.data,
                                      dm_list.size, None if a is None else a.ctypes.data,
                                      0 if a is None else a.size, out.ctypes.data, cap, C.byref(cnt),
                                      secs.ctypes.data))
    out = out[:cnt.value]
    return (out if forest else top_level(out)), secs


def release_batch_engine():
    """Frees the engine that batch_search calls with keep_engine left behind."""
    P.check(P.lib().psx_batch_release())


def batch_search(args, series, dm_list, accs=None, opts=None, nsamps=None, cap=1 << 20, forest=False):
    """Batched fused search.  series: host array [ndm, nsamps] or a DeviceArray
    holding ndm x nsamps samples (then nsamps is required).  Returns the
    top-level candidates, or the whole forest with forest=True."""
    dm_list = np.ascontiguousarray(dm_list, np.float32)
    on_dev = isinstance(series, P.DeviceArray)
    if on_dev:
        ptr, ns = series.ptr, nsamps
    else:
        series = np.ascontiguousarray(series, np.float32)
        ptr, ns = series.ctypes.data, series.shape[-1]
    out = np.zeros(cap, CAND_DTYPE)
    cnt = C.c_uint32()
    secs = np.zeros(3, np.float64)
    a = None if accs is None else np.ascontiguousarray(accs, np.float32)
    o = opts if opts is not None else make_batch_opts()
    P.check(P.lib().psx_batch_search(C.byref(args), C.byref(o), ptr, 1 if on_dev else 0, ns, dm_list.ctypes.data,
                                     dm_list.size, None if a is None else a.ctypes.data, 0 if a is None else a.size,
                                     out.ctypes.data, cap, C.byref(cnt), secs.ctypes.data))
    out = out[:cnt.value]
    return (out if forest else top_level(out)), secs


def generate_device(w: Workload, trials=None):
    """Series of a workload generated straight into one device buffer."""
    trials = list(range(w.ndm)) if trials is None else list(trials)
    gen = SeriesGenerator(w.n_real)
    d = P.DeviceArray((len(trials), w.n_real))
    for k, t in enumerate(trials):
        gen.generate(d.offset(k * w.n_real), w.n_real, w.tsamp, w.seed, t, w.red_amp, w.injections)
    return d
