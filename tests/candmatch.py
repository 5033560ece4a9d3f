"""Candidate-list comparison with an explicit near-threshold exclusion band.

Candidates are keyed by (dm_idx, acc, nh, freq); the frequency is an exact
function of the peak bin, so matching keys means identical (bin, harmonic,
accel).  S/N must agree within SNR_RTOL.  A candidate present on one side only
is tolerated only when its S/N lies within BAND of the threshold (a float
rounding difference can move a bin across the threshold there).
"""
import numpy as np

SNR_RTOL = 1e-3
BAND = 1e-3


def _key(c):
    return (int(c["dm_idx"]), float(c["acc"]), int(c["nh"]), float(c["freq"]))


def compare(got, ref, thr, check_assoc=True):
    g = {_key(c): c for c in got}
    r = {_key(c): c for c in ref}
    problems = []
    for k in sorted(set(g) | set(r)):
        if k in g and k in r:
            a, b = g[k], r[k]
            if abs(float(a["snr"]) - float(b["snr"])) > SNR_RTOL * abs(float(b["snr"])):
                problems.append(("snr", k, float(a["snr"]), float(b["snr"])))
            elif check_assoc and int(a["nassoc"]) != int(b["nassoc"]):
                problems.append(("nassoc", k, int(a["nassoc"]), int(b["nassoc"])))
        else:
            c = g.get(k, r.get(k))
            if abs(float(c["snr"]) - thr) > BAND * thr:
                problems.append(("only_" + ("got" if k in g else "ref"), k, float(c["snr"])))
    return problems
