"""Autotuner and tuned-schedule table for the batched search.

For a series length n the tuner times candidate schedules - FFT
decompositions (pass lengths and column tiles) and trials per batch - with
psx_batch_profile (HIP events around every fused launch) and keeps the
fastest per trial.  Results are stored per (device name, n) in
peasoup_amd/tuned.json (shipped, measured on MI355X) and in the user cache
(~/.cache/peasoup_amd/tuned.json) for new shapes.  Measured HBM traffic of
profiled kernels (rocprofv3 FETCH_SIZE/WRITE_SIZE) is kept in
profiles/traffic.json and reported by bench.py.
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHIPPED = os.path.join(HERE, "tuned.json")
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "peasoup_amd", "tuned.json")
TRAFFIC = os.path.join(os.path.dirname(HERE), "profiles", "traffic.json")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(n, device):
    """Tuned schedule {"batch": B, "decomp": [(L, C), ...]} or {}."""
    for table in (_load(CACHE), _load(SHIPPED)):
        e = table.get(device, {}).get(str(n))
        if e:
            return {"batch": e["batch"], "decomp": [tuple(x) for x in e["decomp"]] if e.get("decomp") else None}
    return {}


def table(device):
    """Every tuned size for the device: {n: {"batch", "decomp"}} (user cache first)."""
    out = {}
    for t in (_load(SHIPPED), _load(CACHE)):
        for n, e in t.get(device, {}).items():
            out[int(n)] = {"batch": e["batch"], "decomp": [tuple(x) for x in e["decomp"]] if e.get("decomp") else None}
    return out


def store(n, device, entry, path=CACHE):
    table = _load(path)
    table.setdefault(device, {})[str(n)] = entry
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def traffic(n, kernel, trials):
    """Measured HBM bytes per launch of `kernel` at length n (or None)."""
    e = _load(TRAFFIC).get(str(n), {}).get(kernel)
    if not e:
        return None
    return int(e["bytes_per_trial"] * trials)


POW2_LENS = [32, 64, 128, 256, 512, 1024, 2048, 4096]
TILES = {32: [16], 64: [16, 8], 128: [16, 8], 256: [16, 8, 4], 512: [16, 8, 4], 1024: [16, 8, 4, 2],
         2048: [8, 4, 2], 4096: [4, 2],
         48: [16], 96: [16, 8], 192: [16, 8], 384: [16, 8, 4], 768: [16, 8, 4], 1536: [8, 4], 3072: [4]}
LENS = sorted(TILES)


def candidates(n, limit=12):
    """Two-pass decompositions (M2 strided, M1 contiguous) of n/2 (2^k or
    3*2^k), then three-pass ones: short 16-column passes, and two short
    strided passes before a long contiguous one."""
    m = n // 2
    out = []
    for m1 in LENS:
        m2 = m // m1
        if m1 * m2 != m or m2 not in TILES:
            continue
        for c0 in TILES[m2]:
            for c1 in TILES[m1]:
                out.append([(m2, c0), (m1, c1)])
    out.sort(key=lambda d: (max(d[0][0], d[1][0]), -d[0][1] - d[1][1]))
    out = out[:limit]
    # three passes of short transforms with 16-column tiles: 128-byte
    # segments everywhere, for one more pass of traffic
    short = [48, 64, 96, 128, 192, 256, 384, 512]
    for q in short:
        for m2 in short:
            m1 = m // (q * m2)
            if q * m2 * m1 == m and m1 in short:
                out.append([(q, 16), (m2, 16), (m1, 16)])
    # two short strided passes and a long contiguous last pass (the shape of
    # the default plan beyond 2^24 samples)
    for m1 in (4096, 3072, 2048, 1536, 1024):
        for q in short:
            for m2 in short:
                if q * m2 * m1 == m and q >= m2:
                    for c1 in TILES[m1][:2]:
                        out.append([(q, 16), (m2, 16), (m1, c1)])
    return out


def tune(n, tsamp, d_series, accs, batches=(16, 32, 64), reps=5, limit=12, log=print):
    """Times candidate schedules on one prologued series; returns the best."""
    import peasoup_amd as P
    import peasoup_amd.workloads as W
    P.register_signatures({"psx_batch_profile": (C.c_int, [C.POINTER(W.SearchArgs), C.POINTER(W.BatchOpts),
                                                           C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32, C.c_int,
                                                           C.c_void_p, C.POINTER(C.c_int), C.c_char_p,
                                                           C.c_size_t])})
    args = W.make_search_args(n, tsamp)
    # every batch full, as in a search of many DM trials (short acceleration
    # lists of short series would otherwise time part-empty batches)
    accs = np.ascontiguousarray(np.resize(np.asarray(accs, np.float32), max(batches)))
    best = None
    for dec in [None] + candidates(n, limit):
        for b in batches:
            opts = W.make_batch_opts(b, decomp=dec)
            ms = np.zeros(8)
            ns = C.c_int(8)
            desc = C.create_string_buffer(128)
            rc = P.lib().psx_batch_profile(C.byref(args), C.byref(opts), d_series, n, accs.ctypes.data,
                                           b, reps, ms.ctypes.data, C.byref(ns), desc, 128)
            if rc:
                log("skip", dec, b, P.lib().psx_last_error_string().decode())
                continue
            per_trial = ms[:ns.value].sum() / b
            log("n=%d batch=%d plan=%s: %.4f ms/trial %s" % (n, b, desc.value.decode(), per_trial,
                                                           np.round(ms[:ns.value], 4).tolist()))
            if best is None or per_trial < best[0]:
                best = (per_trial, b, dec, desc.value.decode())
    return {"batch": best[1], "decomp": best[2], "ms_per_trial": best[0], "plan": best[3]}


def main(argv=None):
    """Tune the given lengths on device 0 and write a table (JSON).

    python -m peasoup_amd.tuning --sizes 8388608,4194304 --out tuned.json
    """
    import argparse
    import sys
    import peasoup_amd as P
    import peasoup_amd.workloads as W
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="8388608")
    ap.add_argument("--tsamp", type=float, default=76e-6)
    ap.add_argument("--batches", default="16,32,64")
    ap.add_argument("--limit", type=int, default=12)
    ap.add_argument("--out", default=CACHE)
    a = ap.parse_args(argv)
    dev = P.device_info(0)["name"]
    for n in [int(x) for x in a.sizes.split(",")]:
        w = W.Workload("tune", 1, n, a.tsamp, 0x5EA5AA, 5.0, -100, 100)
        d = W.generate_device(w)
        accs = W.accel_list(n, a.tsamp, -100, 100)
        best = tune(n, a.tsamp, d.ptr, accs, batches=[int(b) for b in a.batches.split(",")], limit=a.limit,
                    log=lambda *m: print(*m, file=sys.stderr, flush=True))
        entry = {"batch": best["batch"], "decomp": best["decomp"], "ms_per_trial": round(best["ms_per_trial"], 5),
                 "plan": best["plan"]}
        store(n, dev, entry, a.out)
        print(json.dumps({"n": n, **entry}), flush=True)


if __name__ == "__main__":
    main()
