#!/usr/bin/env python3
"""DM x acceleration trials/s of the acceleration-search core on one MI355X.

Workload (one step): 64 dedispersed DM trials of 2^23 float32 samples at
tsamp = 76 us - N(0,1) noise, 1/f^2 red noise (5x white amplitude at 0.1 Hz)
and 6 injected accelerated pulsars - searched over the acceleration plan
-100..+100 m/s^2 (tol 1.1, 64 us pulse width: 127 trials per DM, 8128 trials
per step) with 4 harmonic folds and S/N threshold 7.  Series are generated
on the device before timing, so inputs are resident in HBM.  A step is the
whole search: the 64 per-DM prologues plus all 8128 trials through the
batched, fused, graph-launched path, with host clustering and distilling.

Also reported on the same inputs: the plain one-call-per-kernel Worker form
on the GPU (one DM, all its trials), the CPU oracle on the host cores (DM 0,
all its trials, whose candidates are compared with the GPU step's), the
long-segment, uneven-length and serving workloads, per-launch times of the fused kernels from
HIP events on the batch stream, and the roofline of the dominant kernel.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
METRIC = "DM×accel trials/sec at 2^23 samples vs plain GPU; p50/p99 segment latency"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def stage_bytes(n, npass, bound=True, fused=True):
    """Algorithmic HBM bytes per trial of each fused launch: FFT pass 0 reads
    the n-sample series through the resampler and writes n/2 complex; every
    other FFT pass reads and writes n/2 complex; the spectrum pass reads n/2
    complex and writes n/2+1 floats.  Bound-filtered harmonic form: the
    row-layout spectrum pass writes one 16-bit bound per two bins instead of
    the floats (the harmonic pass recomputes the rare bins it sums exactly
    from the FFT output); after the natural layout the floats are written and
    a pooling pass reads them again to write the bounds.  The harmonic pass
    reads the bounds once (its run copies hit the caches; the exact sums of
    the rare crossing bins and the candidates are negligible).  Exact-gather
    form: the harmonic pass reads the n/2+1 floats once."""
    m, nb = n // 2, n // 2 + 1
    qb = 2 * ((nb + 1) // 2)
    b = [4 * n + 8 * m] + [16 * m] * (npass - 1)
    if bound:
        b += [8 * m + qb if fused else 8 * m + 4 * nb + 4 * nb + qb, qb]
    else:
        b += [8 * m + 4 * nb, 4 * nb]
    return b


def profile_stages(P, W, args_c, opts, d_series, nsamps, accs, reps=10):
    P.register_signatures({"psx_batch_profile": (C.c_int, [C.POINTER(W.SearchArgs), C.POINTER(W.BatchOpts),
                                                           C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32, C.c_int,
                                                           C.c_void_p, C.POINTER(C.c_int), C.c_char_p,
                                                           C.c_size_t])})
    accs = np.ascontiguousarray(accs, np.float32)
    ms = np.zeros(8, np.float64)
    ns = C.c_int(8)
    desc = C.create_string_buffer(128)
    P.check(P.lib().psx_batch_profile(C.byref(args_c), C.byref(opts), d_series, nsamps, accs.ctypes.data,
                                      accs.size, reps, ms.ctypes.data, C.byref(ns), desc, 128))
    return ms[:ns.value].tolist(), desc.value.decode()


def series_of(d_series, w, dm):
    """One DM trial's series copied from the device buffer of the workload."""
    import peasoup_amd as P
    out = np.empty(w.n_real, np.float32)
    P.check(P.lib().psx_memcpy_d2h(out.ctypes.data, d_series.offset(dm * w.n_real), out.nbytes, None))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads for the CPU baseline: every CPU this process may run on, up to
    the 16-CPU share a one-GPU process is given on the benchmark host."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline(O, W, w, series0, accs, threads):
    """Host-core baseline: DM 0's prologue and all of its acceleration
    trials, one trial per thread, through the reference sequence restated in
    C++ with a float four-step FFT (the FP32 transform cuFFT performs in the
    reference; FFTW3f is not installed).  Its FFT plan (twiddle tables) is
    built before the clock starts, as an FFTW plan would be.  Returns the
    baseline record and DM 0's accel-distilled candidates."""
    p = O.make_params(w.size, np.float32(w.tsamp), nharmonics=w.nharmonics, min_snr=w.min_snr, fast_fft=True)
    O.fft_r2c_fast(np.zeros(w.size, np.float32))
    t0 = time.perf_counter()
    tim, mean, rms, std = O.prologue(series0, p)
    t_pro = time.perf_counter() - t0
    t0 = time.perf_counter()
    cands = O.search_accels(tim, p, accs, mean, std, nthreads=threads)
    t_acc = time.perf_counter() - t0
    rec = {"value": round(len(accs) / (t_pro + t_acc), 3), "unit": "trials/s", "cores": threads,
           "host_cpus": os.cpu_count(), "cpu_model": cpu_model(), "kind": "port",
           "sample": "DM 0 of the workload: prologue (%.2f s, one thread) + all %d acceleration trials "
                     "(%.2f s, %d threads), wall time of both; float four-step CPU FFT (FFTW3f absent)" %
                     (t_pro, len(accs), t_acc, threads)}
    return rec, cands


def serve_stream(P, W, nreq=64, rate=4.0, seed=0x5EA505, budget=200e9, batch=16, dms=(16, 256), schedules=None,
                 log=log):
    """Bounded Poisson stream of segment-search requests through the server.

    Each request is one beam: DM count uniform in dms ({16..256}), fft size
    uniform in {2^20..2^24}, tsamp 64 us, N(0,1) + red noise, accel range
    +-50 or +-200 m/s^2, admitted against an HBM budget; latency = completion
    - arrival on the library clock, queueing included.  The beams' series are
    windows of a resident pool (dms[1] series per fft size, two of them with
    an injected pulsar) at a random offset, so a beam carries 0-2 pulsars and
    the pool (32 GB) is generated once before the stream starts; the server
    copies each admitted beam into its own resident buffers as it would a
    dedispersed beam."""
    from peasoup_amd.serve import Server
    rng = np.random.default_rng(seed)
    sizes = [1 << k for k in range(20, 25)]
    pools = {}
    t0 = time.perf_counter()
    for k, n in enumerate(sizes):
        w = W.Workload("pool%d" % n, dms[1], n, 64e-6, seed + 1 + k, 5.0)
        for t in rng.choice(dms[1], 2, replace=False):
            w.injections.append(W.Injection(int(t), float(rng.uniform(1.5e-3, 0.2)), float(rng.uniform(-40, 40)),
                                            float(rng.uniform(8, 30)), float(rng.random())))
        pools[n] = W.generate_device(w)
    P.synchronize()
    log("serving: generated the %.1f GB series pool in %.1f s" %
        (sum(d.nbytes for d in pools.values()) / 1e9, time.perf_counter() - t0))
    reqs = []
    t = 0.0
    for k in range(nreq):
        t += rng.exponential(1.0 / rate)
        n = int(sizes[rng.integers(0, len(sizes))])
        ndm = int(rng.integers(dms[0], dms[1] + 1))
        acc = 50.0 if rng.random() < 0.5 else 200.0
        off = int(rng.integers(0, dms[1] - ndm + 1))
        w = W.Workload("req%d" % k, ndm, n, 64e-6, 0, 5.0, -acc, acc)
        reqs.append((t, w, pools[n].view(off * n, (ndm, n))))
    ntrials = sum(w.ndm * W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi).size for _, w, _ in reqs)
    log("serving: %d requests, %d trials" % (nreq, ntrials))
    with Server(budget, batch=batch, schedules=schedules) as srv:
        ids = []
        t0 = time.perf_counter()
        for ta, w, d in reqs:
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(w.search_args(), d, w.dm_list(), nsamps=w.size))
        sts = []
        for k, i in enumerate(ids):
            sts.append(srv.wait(i, timeout=1200))
            if (k + 1) % 8 == 0 and k + 1 < len(ids):
                log("serving: %d of %d requests done at %.1f s" % (k + 1, len(ids), time.perf_counter() - t0))
        wall = time.perf_counter() - t0
        for i in ids:
            srv.result(i)
        stats = srv.stats()
    bad = [s for s in sts if s["state"] != "done"]
    if bad:
        raise RuntimeError("serving: %d requests failed: %s" % (len(bad), bad[:2]))
    lat = np.array([s["latency"] for s in sts])
    first = min(s["t_submit"] for s in sts)
    last = max(s["t_done"] for s in sts)
    res = {"requests": nreq, "arrival_rate_per_s": rate, "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 4), "max_latency_s": round(float(lat.max()), 4),
           "trials": int(ntrials), "trials_per_s": round(ntrials / (last - first), 1),
           "wall_s": round(wall, 2), "engines": stats["engines"], "hbm_budget_bytes": int(budget),
           "stream": "Poisson %.1f req/s; DMs U{%d..%d}; fft size U{2^20..2^24}; tsamp 64 us; accel +-50 or "
                     "+-200 m/s^2; red noise + 0-2 pulsars" % (rate, dms[0], dms[1])}
    log("serving", res)
    return res


def uneven_batched(P, W, ndm=512, seed=0x5EA503, batch=16, schedules=None, log=log, reps=3):
    """Segmented searches of uneven length: ndm DM trials, an equal mix of fft
    sizes 2^22, 3*2^21 and 2^23 at 76 us (each segment 90% of its fft size,
    so the series are zero padded after baseline removal), red noise and
    injections, accel -50..+50 m/s^2.  The three size classes are submitted
    together to the server, in both dispatch forms on the same inputs,
    alternating reps times: per-size engines launching on their own streams,
    and one graph launch per mixed-size batch (each size's batch graph a
    child of it).  The faster form's best time is the line's value."""
    from peasoup_amd.serve import DISPATCH_MIXED_GRAPH, DISPATCH_STREAMS, Server
    sizes = [1 << 22, 3 << 21, 1 << 23]
    per = [ndm // 3 + (1 if k < ndm % 3 else 0) for k in range(3)]
    reqs = []
    for k, (n, nd) in enumerate(zip(sizes, per)):
        w = W.Workload("seg%d" % n, nd, n, 76e-6, seed + k, 5.0, -50.0, 50.0, nsamps=int(n * 0.9))
        rng = np.random.default_rng(seed + k)
        for t in rng.choice(nd, 2, replace=False):
            w.injections.append(W.Injection(int(t), float(rng.uniform(2e-3, 0.2)), float(rng.uniform(-40, 40)),
                                            float(rng.uniform(8, 30)), float(rng.random())))
        reqs.append((w, W.generate_device(w)))
    P.synchronize()
    ntrials = sum(w.ndm * W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi).size for w, _ in reqs)
    forms = {"per-size streams": DISPATCH_STREAMS, "mixed-size graph": DISPATCH_MIXED_GRAPH}
    best, ncand = {}, {}
    for _ in range(reps):
        for name, mode in forms.items():
            with Server(200e9, batch=batch, schedules=schedules, dispatch=mode) as srv:
                # warm-up: the whole workload once (engines, batch graphs and,
                # for the mixed form, the mixed-batch graphs exist afterwards)
                for i in [srv.submit(w.search_args(), d, w.dm_list(), nsamps=w.n_real) for w, d in reqs]:
                    srv.wait(i, timeout=600)
                    srv.result(i)
                t0 = time.perf_counter()
                ids = [srv.submit(w.search_args(), d, w.dm_list(), nsamps=w.n_real) for w, d in reqs]
                sts = [srv.wait(i, timeout=600) for i in ids]
                dt = time.perf_counter() - t0
                if any(s["state"] != "done" for s in sts):
                    raise RuntimeError("uneven batch failed: %s" % sts)
                ncand[name] = sum(len(srv.result(i)) for i in ids)
            best[name] = min(best.get(name, dt), dt)
    if len(set(ncand.values())) != 1:
        raise RuntimeError("uneven batch: dispatch forms disagree on the candidate count: %s" % ncand)
    form = min(best, key=best.get)
    dt = best[form]
    res = {"trials": int(ntrials), "dm_trials": ndm, "seconds": round(dt, 3), "trials_per_s": round(ntrials / dt, 1),
           "candidates": ncand[form], "form": form,
           "forms_trials_per_s": {k: round(ntrials / v, 1) for k, v in best.items()},
           "workload": "%d DM trials, equal mix of fft sizes 2^22, 3*2^21, 2^23 (segments "
                       "90%% of the fft size, zero padded), 76 us, accel -50..+50" % ndm}
    log("uneven", res)
    return res


def long_segment(P, W, ndm=2, batch=16, log=log):
    """Long-segment configuration (32 DM trials x 2^26 samples at 64 us, red
    noise, 4 binary-like pulsars, accel -50..+50 m/s^2: 3,093 trials per DM),
    sampled on its first ndm DM trials (the injected trials among them keep
    their pulsars); one warm-up batch first, whose engine (buffers, plans,
    graphs) the timed call keeps, as the headline steps do."""
    w = W.config("long32x2^26")
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    w.ndm = ndm
    w.injections = [i for i in w.injections if i.trial < ndm]
    d = W.generate_device(w)
    P.synchronize()
    opts = W.make_batch_opts(batch, keep_engine=True)
    W.batch_search(w.search_args(), d, w.dm_list()[:1], accs[:batch], opts, nsamps=w.size)
    t0 = time.perf_counter()
    c, s = W.batch_search(w.search_args(), d, w.dm_list(), accs, opts, nsamps=w.size)
    dt = time.perf_counter() - t0
    P.lib().psx_batch_release()
    # per-launch stage times of one full batch (HIP events), after the timing
    ms, desc = profile_stages(P, W, w.search_args(), W.make_batch_opts(batch), d.ptr, w.size, accs[:batch], reps=2)
    n = ndm * accs.size
    res = {"trials": n, "seconds": round(dt, 3), "trials_per_s": round(n / dt, 1), "candidates": int(c.size),
           "fft_plan": desc, "stage_ms_per_launch": [round(x, 3) for x in ms],
           "sample": "%d of the 32 DM trials x 2^26 samples (67,108,864), %d accelerations each, batch %d, "
                     "engine kept from the warm-up batch" % (ndm, accs.size, batch)}
    log("long segment", res)
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ndm", type=int, default=64)
    ap.add_argument("--batch", type=int, default=0, help="trials per fused batch (0 = tuned/default)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--decomp", default="",
                    help="FFT passes as len/cols,len/cols[,...] (default: the tuned or library default plan)")
    ap.add_argument("--harmonic", choices=("bound", "exact"), default="bound",
                    help="harmonic pass: bound-filtered (default) or exact gathers of every addend")
    ap.add_argument("--no-keep", action="store_true",
                    help="build and free the search engine in every step instead of keeping it between steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--no-long", action="store_true", help="skip the 2^26 long-segment sample")
    ap.add_argument("--serve-dms", default="16,256", help="DM count range of serving requests (lo,hi)")
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for rocprofv3: a few fused batches of the workload plus a 1 GiB "
                         "copy kernel (calibrates FETCH_SIZE/WRITE_SIZE), no timing line")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (DM trials would shard across "
              "GPUs with no collective on the hot path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import peasoup_amd as P
    import peasoup_amd.workloads as W
    from peasoup_amd import tuning

    if not P.gpu_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    dev = P.device_info(0)
    log("device", dev)
    w = W.config("gpu64x2^23")
    w.ndm = a.ndm
    w.injections = [i for i in w.injections if i.trial < w.ndm]
    args_c = w.search_args()
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    ntrials = w.ndm * accs.size
    tuned = tuning.lookup(w.size, dev["name"])
    batch = a.batch or tuned.get("batch", 16)
    bound = a.harmonic == "bound"
    decomp = tuned.get("decomp")
    if a.decomp:
        decomp = [tuple(int(x) for x in p.split("/")) for p in a.decomp.split(",")]
    opts = W.make_batch_opts(batch, use_graph=not a.no_graph, decomp=decomp, keep_engine=not a.no_keep,
                             harmonic_pass=W.HARMONIC_BOUND if bound else W.HARMONIC_EXACT)
    log("workload %s: %d DMs x %d accels = %d trials, batch %d, decomp %s" %
        (w.name, w.ndm, accs.size, ntrials, batch, decomp))

    if a.profile_run:
        w.ndm = 1
        d_series = W.generate_device(w)
        # copy kernel of known size: 2^28 floats read and written (1 GiB each way)
        src = P.DeviceArray(1 << 28)
        dst = P.DeviceArray(1 << 28)
        for _ in range(3):
            P.check(P.lib().psx_convert_to_f32(src.ptr, P.PSX_F32, dst.ptr, 1 << 28, None))
        P.synchronize()
        stage_ms, desc = profile_stages(P, W, args_c, opts, d_series.ptr, w.size, accs[:batch], reps=5)
        log("profile run: plan %s, stage ms %s" % (desc, stage_ms))
        return

    t0 = time.perf_counter()
    d_series = W.generate_device(w)
    P.synchronize()
    log("generated %.1f GB of series in %.1f s" % (d_series.nbytes / 1e9, time.perf_counter() - t0))

    def step():
        return W.batch_search(args_c, d_series, w.dm_list(), accs, opts, nsamps=w.size)

    first = None
    for i in range(a.warmup):
        c, s = step()
        first = c
        log("warmup %d: %d candidates, prologue %.3f s, trials %.3f s" % (i, c.size, s[0], s[1]))
    P.synchronize()
    t0 = time.perf_counter()
    secs = []
    for i in range(a.steps):
        c, s = step()
        secs.append(s.tolist())
        if first is not None and not np.array_equal(c, first):
            log("WARNING: step %d candidates differ from the warm-up's" % i)
        first = c
    P.synchronize()
    dt = time.perf_counter() - t0
    value = a.steps * ntrials / dt
    log("timed %d steps: %.3f s -> %.1f trials/s" % (a.steps, dt, value))
    W.release_batch_engine()

    # per-launch profile of one batch (HIP events on the batch stream)
    stage_ms, desc = profile_stages(P, W, args_c, opts, d_series.ptr, w.size, accs[:batch], reps=10)
    nb_trials = min(batch, accs.size)
    npass = len(stage_ms) - 2
    sbytes = [b * nb_trials for b in stage_bytes(w.size, npass, bound, fused=desc.endswith(" rows"))]
    names = (["fft_pass0_resample"] + ["fft_pass%d" % q for q in range(1, npass)] +
             (["spectrum", "harm_bound"] if bound else ["spec_post", "harm_peaks"]))
    stages = {nm: {"ms": round(ms, 4), "GBps": round(bb / (ms * 1e-3) / 1e9, 1), "bytes": bb}
              for nm, ms, bb in zip(names, stage_ms, sbytes)}
    dom = int(np.argmax(stage_ms))
    achieved = sbytes[dom] / (stage_ms[dom] * 1e-3) / 1e9
    traffic = tuning.traffic(w.size, names[dom], nb_trials)
    roofline = {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": sbytes[dom], "launch_trials": nb_trials, "fft_plan": desc}
    log("stages", stages)
    bound_stats = None
    if bound:
        # how many bins the bound-filtered pass sums exactly on this workload
        # (DM 0, the first 8 accelerations), and a host check of its bounds
        P.register_signatures({"psx_batch_bound_audit": (C.c_int, [C.POINTER(W.SearchArgs), C.POINTER(W.BatchOpts),
                                                                    C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32,
                                                                    C.c_void_p])})
        res = np.zeros(7, np.uint64)
        acc8 = np.ascontiguousarray(accs[:8], np.float32)
        P.check(P.lib().psx_batch_bound_audit(C.byref(args_c), C.byref(opts), d_series.ptr, w.size,
                                              acc8.ctypes.data, acc8.size, res.ctypes.data))
        bound_stats = {"bins": int(res[5]), "bins_summed_exactly": int(res[4]),
                       "fraction_summed_exactly": float(res[4]) / max(1, int(res[5])),
                       "bound_entries_checked": int(res[0]), "bound_entries_invalid": int(res[1]),
                       "bound_entries_loose": int(res[6]),
                       "sample": "DM 0, %d accelerations" % acc8.size}
        log("bound audit", bound_stats)

    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "trials/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dt / a.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "%d DM trials x 2^23 float32 samples, tsamp 76 us, N(0,1) + red noise + %d pulsars; "
                               "accel plan -100..+100 m/s^2 tol 1.1 (127 trials/DM); 4 harmonic folds; S/N 7"
                               % (w.ndm, len(w.injections)),
                   "fft_size": w.size, "ndm": w.ndm, "trials_per_step": ntrials, "batch": batch,
                   "fft_plan": desc, "graphs": not a.no_graph, "engine_kept_between_steps": not a.no_keep,
                   "harmonic_pass": a.harmonic,
                   "device": dev["name"]},
        "roofline": roofline, "stages": stages, "bound_filter": bound_stats,
        "algorithmic_GBps": round(value * (12 * w.size + 68 * (w.size // 2 + 1)) / 1e9, 1),
        "candidates": int(first.size) if first is not None else None,
    }

    series0 = None
    if not a.no_plain:
        series0 = series_of(d_series, w, 0)
        c, s = W.worker_search(args_c, series0[None, :], np.zeros(1, np.float32), accs)
        plain = accs.size / s[1]
        result["plain_gpu"] = {"trials_per_s": round(plain, 2), "prologue_s": round(s[0], 4),
                               "sample": "DM 0, all %d trials, one library call per reference call" % accs.size}
        result["speedup_vs_plain"] = round(value / plain, 2)
        log("plain GPU: %.1f trials/s" % plain)
    if not a.no_cpu:
        import oracle as O
        if series0 is None:
            series0 = series_of(d_series, w, 0)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from candmatch import compare
        cb, ref0 = cpu_baseline(O, W, w, series0, accs, cpu_threads())
        result["cpu_baseline"] = cb
        result["speedup_vs_cpu"] = round(value / cb["value"], 1)
        log("cpu baseline", cb)
        # the timed step's candidates of DM 0 and of the first DM with an
        # injected pulsar against the oracle's (all 127 trials of each, the
        # exact double-precision FFT), and the baseline's own DM-0 list
        dms = [0] + sorted({i.trial for i in w.injections if i.trial != 0})[:1]
        refs = {}
        for dm in dms:
            p = O.make_params(w.size, np.float32(w.tsamp), dm=float(dm), dm_idx=dm, nharmonics=w.nharmonics,
                              min_snr=w.min_snr)
            refs[dm] = O.search_dm(series_of(d_series, w, dm), p, accs,
                                   nthreads=cpu_threads())[0]
        problems, keys = [], set()
        for dm in dms:
            got_d = first[first["dm_idx"] == dm]
            problems += compare(got_d, refs[dm], w.min_snr)
            keys |= {(dm, float(c["acc"]), int(c["nh"]), float(c["freq"])) for c in got_d}
            keys |= {(dm, float(c["acc"]), int(c["nh"]), float(c["freq"])) for c in refs[dm]}
        base_problems = compare(first[first["dm_idx"] == 0], ref0, w.min_snr)
        result["parity"] = {"compared": len(keys), "mismatches": len(problems),
                            "baseline_mismatches": len(base_problems),
                            "reference": "oracle, DM trials %s, all %d trials each" % (dms, accs.size),
                            "rule": "keys (accel, harmonic, bin) exact, S/N within 1e-3, one-sided only "
                                    "within 1e-3 of the threshold"}
        log("parity", result["parity"], problems[:5])
    if not a.no_uneven or not a.no_serve or not a.no_long:
        del d_series
    if not a.no_long:
        result["long_segment"] = long_segment(P, W, batch=tuning.lookup(1 << 26, dev["name"]).get("batch", 16))
    if not a.no_uneven:
        result["uneven_batched"] = uneven_batched(P, W, batch=batch, schedules=tuning.table(dev["name"]))
    if not a.no_serve:
        lo, hi = (int(x) for x in a.serve_dms.split(","))
        result["serving"] = serve_stream(P, W, nreq=a.serve_requests, batch=batch, dms=(lo, hi),
                                         schedules=tuning.table(dev["name"]))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
