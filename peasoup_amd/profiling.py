"""Reduce rocprofv3 CSV output to the summaries kept under profiles/.

Two inputs:

* a kernel trace with statistics (``rocprofv3 --kernel-trace --stats
  --output-format csv``) -> per-kernel calls / average / total duration;
* two counter runs of ``bench.py --profile-run`` (``--pmc FETCH_SIZE``, then
  ``--pmc WRITE_SIZE``) -> HBM bytes per dispatch of each fused kernel,
  calibrated by the run's copy kernel (2^28 floats read and 2^28 written, so
  1 GiB each way): factor = counter reading of the copy / 1 GiB, and a
  kernel's traffic = its reading / factor.  The per-trial traffic of the
  batched stages is written to profiles/traffic.json, which bench.py reports
  as roofline.traffic;
* one counter run of ``bench.py --profile-run`` with the SQ instruction
  counters (``--pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_VMEM_RD
  GRBM_GUI_ACTIVE``) -> instructions per wave of each fused kernel and the
  fraction of the CUs' issue cycles its VALU and LDS instructions take (a
  CU issues at most one wave64 VALU instruction per cycle over its four
  SIMDs; GRBM_GUI_ACTIVE counts the busy cycles of every XCD).

* counter runs of per-unit busy/stall cycle counters (``TA_TA_BUSY_sum``,
  ``TD_TD_BUSY_sum``, ``TCP_PENDING_STALL_CYCLES_sum``, ...) -> each as a share
  of the kernel's cycles on every CU.

Usage:
  python -m peasoup_amd.profiling stats <dir>            # kernel stats table
  python -m peasoup_amd.profiling traffic <fetch_dir> <write_dir> --n 8388608 --trials 16 [--out traffic.json]
  python -m peasoup_amd.profiling issue <dir> [--cus 256 --xcds 8]
  python -m peasoup_amd.profiling busy <dir> [<dir> ...] [--cus 256 --xcds 8]
"""
import argparse
import csv
import glob
import json
import os
import re
import sys

COPY_BYTES = float(1 << 30)
COPY_KERNEL = "k_convert"


def _rows(path_glob):
    rows = []
    for p in sorted(glob.glob(path_glob, recursive=True)):
        with open(p, newline="") as f:
            rows.extend(csv.DictReader(f))
    return rows


def _col(row, *names):
    for n in names:
        if n in row:
            return row[n]
    raise KeyError("none of %s in %s" % (names, list(row)))


def short_name(kernel):
    """Kernel name without its argument list, keeping template arguments."""
    k = kernel.strip().replace("(anonymous namespace)::", "").replace("psx::", "")
    if k.endswith(")"):
        depth = 0
        for pos in range(len(k) - 1, -1, -1):
            depth += {")": 1, "(": -1}.get(k[pos], 0)
            if depth == 0:
                k = k[:pos]
                break
    return k.replace("void ", "", 1) if k.startswith("void ") else k


def kernel_stats(d):
    """[(name, calls, avg_us, total_ms, percent)] from a --stats run."""
    out = []
    for r in _rows(os.path.join(d, "**", "*kernel_stats.csv")):
        out.append((short_name(_col(r, "Name", "KernelName", "Kernel_Name")), int(_col(r, "Calls")),
                    float(_col(r, "AverageNs")) / 1e3, float(_col(r, "TotalDurationNs")) / 1e6,
                    float(_col(r, "Percentage"))))
    out.sort(key=lambda x: -x[3])
    return out


def launch_stats(d):
    """[(name, grid, calls, avg_us, total_ms)] per kernel and grid size from a
    --kernel-trace run: separates the batched launches of a kernel from its
    other launches (the prologue's plain FFTs use the same kernels)."""
    acc = {}
    for r in _rows(os.path.join(d, "**", "*kernel_trace.csv")):
        name = short_name(_col(r, "Kernel_Name"))
        grid = int(_col(r, "Grid_Size_X")) * int(_col(r, "Grid_Size_Y")) * int(_col(r, "Grid_Size_Z"))
        dur = int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))
        a = acc.setdefault((name, grid), [0, 0])
        a[0] += 1
        a[1] += dur
    out = [(k[0], k[1], n, t / n / 1e3, t / 1e6) for k, (n, t) in acc.items()]
    out.sort(key=lambda x: -x[4])
    return out


def stage_of(name):
    """Batched-stage label of a fused kernel name (or None)."""
    if "k_fft_pass<" in name:
        return "fft_pass0_resample" if re.search(r",\s*true>\s*$", name) else "fft_pass"
    # the spectrum pass: natural layout (k_spec_post) or the transposed row
    # layout (k_spec_rows, whose last row of workgroups also computes the
    # layout's edge bins)
    if "k_spec_post" in name or "k_spec_rows" in name:
        return "spec_post"
    if "k_harm_peaks" in name:
        return "harm_peaks"
    if "k_pool_peaks" in name:
        return "pool_peaks"
    if "k_harm_bound" in name:
        return "harm_bound"
    return None


def counter_by_kernel(d, counter):
    """{kernel: [(grid, value_bytes, dispatch_id), ...]} (counter in KiB per dispatch)."""
    res = {}
    for r in _rows(os.path.join(d, "**", "*counter_collection.csv")):
        if _col(r, "Counter_Name") != counter:
            continue
        name = short_name(_col(r, "Kernel_Name"))
        grid = int(_col(r, "Grid_Size")) if "Grid_Size" in r else 0
        res.setdefault(name, []).append((grid, float(_col(r, "Counter_Value")) * 1024.0,
                                         int(_col(r, "Dispatch_Id"))))
    return res


def traffic(fetch_dir, write_dir, trials):
    """Per-stage HBM bytes per trial (fetch + write, copy-calibrated) of the
    batched launches (the largest grid of each kernel)."""
    out = {"calibration": {}}
    per = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        by = counter_by_kernel(d, counter)
        copy = [v for k, vals in by.items() if COPY_KERNEL in k for (_, v, _) in vals]
        if not copy:
            raise RuntimeError("no %s dispatch of the copy kernel in %s" % (counter, d))
        factor = (sum(copy) / len(copy)) / COPY_BYTES
        out["calibration"][counter] = {"copy_reading_bytes": sum(copy) / len(copy), "copy_bytes": COPY_BYTES,
                                       "factor": factor}
        # batched launches start with the resampling first pass; the
        # prologue's plain FFTs run before it
        first = min((i for k, vals in by.items() if stage_of(k) == "fft_pass0_resample" for _, _, i in vals),
                    default=0)
        fft_rank = []
        for k, vals in by.items():
            st = stage_of(k)
            vals = [x for x in vals if x[2] >= first]
            if st is None or not vals:
                continue
            gmax = max(g for g, _, _ in vals)
            big = [(v, i) for g, v, i in vals if g == gmax]
            avg = sum(v for v, _ in big) / len(big) / factor
            if st == "fft_pass":
                fft_rank.append((min(i for _, i in big), k, avg))
                continue
            per.setdefault(st, {"kernel": k})[counter] = avg
        for q, (_, k, avg) in enumerate(sorted(fft_rank), start=1):
            per.setdefault("fft_pass%d" % q, {"kernel": k})[counter] = avg
    for st, e in per.items():
        e["bytes_per_launch"] = e.get("FETCH_SIZE", 0.0) + e.get("WRITE_SIZE", 0.0)
        e["bytes_per_trial"] = e["bytes_per_launch"] / trials
    out["stages"] = per
    return out


def _dispatches(d):
    """{dispatch_id: (kernel, grid, {counter: value})} of a counter run."""
    res = {}
    for r in _rows(os.path.join(d, "**", "*counter_collection.csv")):
        i = int(_col(r, "Dispatch_Id"))
        e = res.setdefault(i, (short_name(_col(r, "Kernel_Name")), int(_col(r, "Grid_Size")), {}))
        c = _col(r, "Counter_Name")
        e[2][c] = e[2].get(c, 0.0) + float(_col(r, "Counter_Value"))
    return res


def issue(d, cus=256, xcds=8):
    """Per batched stage: instructions per wave and the share of the CUs'
    issue cycles taken by VALU and LDS instructions (averaged over the
    stage's batched launches)."""
    disp = _dispatches(d)
    first = min((i for i, (k, _, _) in disp.items() if stage_of(k) == "fft_pass0_resample"), default=0)
    groups = {}
    for i, (k, g, c) in disp.items():
        if i >= first and stage_of(k) is not None:
            groups.setdefault(k, []).append((g, i, c))
    per, fft_rank = {}, []
    for k, vals in groups.items():
        gmax = max(g for g, _, _ in vals)
        big = [(i, c) for g, i, c in vals if g == gmax]
        avg = {n: sum(c[n] for _, c in big) / len(big) for n in big[0][1]}
        waves = avg["SQ_WAVES"]
        cycles = avg["GRBM_GUI_ACTIVE"] / xcds
        e = {"kernel": k, "launches": len(big), "waves": waves, "cycles": cycles,
             "valu_per_wave": avg["SQ_INSTS_VALU"] / waves, "lds_per_wave": avg["SQ_INSTS_LDS"] / waves,
             "vmem_rd_per_wave": avg["SQ_INSTS_VMEM_RD"] / waves,
             "valu_issue": avg["SQ_INSTS_VALU"] / (cus * cycles), "lds_issue": avg["SQ_INSTS_LDS"] / (cus * cycles)}
        st = stage_of(k)
        if st == "fft_pass":
            fft_rank.append((min(i for i, _ in big), e))
        else:
            per[st] = e
    for q, (_, e) in enumerate(sorted(fft_rank, key=lambda x: x[0]), start=1):
        per["fft_pass%d" % q] = e
    return per


def busy(dirs, cus=256, xcds=8):
    """Per batched stage: each busy/stall counter of the runs in dirs
    (per-unit cycle counts summed over the device, e.g. TA_TA_BUSY_sum,
    TD_TD_BUSY_sum) as a share of the kernel's cycles on every CU
    (GRBM_GUI_ACTIVE / xcds per run), and a *_LATENCY counter over its
    request counter as cycles per request, averaged over the stage's batched
    launches."""
    per, fft_rank = {}, {}
    for d in dirs:
        disp = _dispatches(d)
        first = min((i for i, (k, _, _) in disp.items() if stage_of(k) == "fft_pass0_resample"), default=0)
        groups = {}
        for i, (k, g, c) in disp.items():
            if i >= first and stage_of(k) is not None:
                groups.setdefault(k, []).append((g, i, c))
        ranks = []
        for k, vals in groups.items():
            gmax = max(g for g, _, _ in vals)
            big = [(i, c) for g, i, c in vals if g == gmax]
            avg = {n: sum(c[n] for _, c in big) / len(big) for n in big[0][1]}
            cycles = avg.pop("GRBM_GUI_ACTIVE") / xcds
            shares = {}
            for n, v in avg.items():
                base = n.replace("_sum", "")
                if base.endswith("_LATENCY"):
                    req = avg.get(n.replace("_LATENCY", ""))
                    if req:
                        shares[base + "_PER_REQ"] = v / req  # cycles per request
                elif base.endswith(("BUSY", "CYCLES", "STALL")):
                    shares[base] = v / (cus * cycles)
            st = stage_of(k)
            if st == "fft_pass":
                ranks.append((min(i for i, _ in big), k, shares))
            else:
                per.setdefault(st, {"kernel": k}).update(shares)
        for q, (_, k, shares) in enumerate(sorted(ranks, key=lambda x: x[0]), start=1):
            per.setdefault("fft_pass%d" % q, {"kernel": k}).update(shares)
    return per


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s1 = sub.add_parser("stats")
    s1.add_argument("dir")
    s1.add_argument("--top", type=int, default=25)
    s2 = sub.add_parser("traffic")
    s2.add_argument("fetch_dir")
    s2.add_argument("write_dir")
    s2.add_argument("--n", type=int, required=True)
    s2.add_argument("--trials", type=int, required=True)
    s2.add_argument("--out")
    s3 = sub.add_parser("issue")
    s3.add_argument("dir")
    s3.add_argument("--cus", type=int, default=256)
    s3.add_argument("--xcds", type=int, default=8)
    s4 = sub.add_parser("busy")
    s4.add_argument("dirs", nargs="+")
    s4.add_argument("--cus", type=int, default=256)
    s4.add_argument("--xcds", type=int, default=8)
    a = ap.parse_args(argv)
    if a.cmd == "busy":
        per = busy(a.dirs, a.cus, a.xcds)
        names = sorted({n for e in per.values() for n in e if n != "kernel"})
        print("%-20s %-36s" % ("stage", "kernel") + "".join(" %30s" % n for n in names))
        for st, e in sorted(per.items()):
            print("%-20s %-36s" % (st, e["kernel"][:36]) +
                  "".join(" %30s" % (("%.0f" if n.endswith("_PER_REQ") else "%.3f") % e[n] if n in e else "-")
                    for n in names))
        return
    if a.cmd == "issue":
        print("%-20s %-36s %9s %10s %9s %8s %9s %10s %9s" % ("stage", "kernel", "waves", "cycles", "VALU/wave",
                                                           "LDS/wave", "VMEM/wave", "VALU issue", "LDS issue"))
        for st, e in sorted(issue(a.dir, a.cus, a.xcds).items()):
            print("%-20s %-36s %9d %10.0f %9.0f %8.1f %9.1f %10.2f %9.2f" % (
                st, e["kernel"][:36], e["waves"], e["cycles"], e["valu_per_wave"], e["lds_per_wave"],
                e["vmem_rd_per_wave"], e["valu_issue"], e["lds_issue"]))
        return
    if a.cmd == "stats":
        print("per kernel (rocprofv3 --stats):")
        print("%-60s %8s %12s %12s %7s" % ("kernel", "calls", "avg_us", "total_ms", "pct"))
        for name, calls, avg, tot, pct in kernel_stats(a.dir)[:a.top]:
            print("%-60s %8d %12.2f %12.3f %7.2f" % (name[:60], calls, avg, tot, pct))
        ls = launch_stats(a.dir)
        if ls:
            print("\nper kernel and grid size (kernel trace):")
            print("%-60s %10s %8s %12s %12s" % ("kernel", "grid", "calls", "avg_us", "total_ms"))
            for name, grid, calls, avg, tot in ls[:a.top]:
                print("%-60s %10d %8d %12.2f %12.3f" % (name[:60], grid, calls, avg, tot))
        return
    t = traffic(a.fetch_dir, a.write_dir, a.trials)
    print(json.dumps(t, indent=1))
    if a.out:
        table = {}
        if os.path.exists(a.out):
            with open(a.out) as f:
                table = json.load(f)
        table[str(a.n)] = {st: {"bytes_per_trial": round(e["bytes_per_trial"]), "kernel": e["kernel"],
                                "fetch_bytes_per_launch": round(e.get("FETCH_SIZE", 0.0)),
                                "write_bytes_per_launch": round(e.get("WRITE_SIZE", 0.0))}
                           for st, e in t["stages"].items()}
        table[str(a.n)]["calibration"] = t["calibration"]
        with open(a.out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
