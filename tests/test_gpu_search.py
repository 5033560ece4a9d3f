"""End-to-end parity: the Worker-equivalent on the GPU vs the CPU oracle.

Workload: 8 DM trials x 2^17 samples at 64 us, N(0,1) noise, trial 4 with an
injected 5.3 ms pulsar (5% duty, +10 m/s^2, S/N 15); accelerations from the
acceleration plan over -50..+50 m/s^2 (tol 1.1, 64 us); 4 harmonic folds;
threshold 7.  Candidate sets must match exactly (key = dm, accel, harmonic,
bin) with S/N within 1e-3, except inside the near-threshold band.
"""
import numpy as np
import pytest

from candmatch import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1(psx):
    import peasoup_amd.workloads as W
    w = W.config("cpu8x2^17")
    series = W.generate_host(w)
    accs = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    return w, series, accs


def oracle_dm(orc, w, series, accs, d):
    p = orc.make_params(w.size, np.float32(w.tsamp), dm=float(d), dm_idx=d, min_snr=w.min_snr,
                        nharmonics=w.nharmonics)
    cands, _ = orc.search_dm(series[d], p, accs, nthreads=4)
    return cands


def test_worker_matches_oracle_cfg1(psx, orc, cfg1):
    import peasoup_amd.workloads as W
    w, series, accs = cfg1
    assert list(accs) == [-50.0, 0.0, 50.0]
    got, secs = W.worker_search(w.search_args(), series, w.dm_list(), accs)
    ref = np.concatenate([oracle_dm(orc, w, series, accs, d) for d in range(w.ndm)])
    problems = compare(got, ref, w.min_snr)
    assert not problems, problems[:20]
    # the injected pulsar is found in trial 4 near 1/5.3 ms (or a harmonic of it)
    f0 = 1 / 5.3e-3
    hits = got[(got["dm_idx"] == 4) & (got["snr"] > 10)]
    assert hits.size > 0
    ratios = hits["freq"] / f0
    assert np.any(np.abs(ratios - np.round(ratios)) < 0.01)


def test_worker_uses_acceleration_plan(psx, cfg1):
    import peasoup_amd.workloads as W
    w, series, accs = cfg1
    got_plan, _ = W.worker_search(w.search_args(), series[:2], w.dm_list()[:2])
    got_list, _ = W.worker_search(w.search_args(), series[:2], w.dm_list()[:2], accs)
    assert np.array_equal(got_plan, got_list)


def test_worker_padding_and_segment(psx, orc):
    """Series shorter than the FFT size are baseline-subtracted then zero padded
    (pipeline_multi.cu:174-187)."""
    import peasoup_amd.workloads as W
    n, size, ts = 100000, 1 << 17, np.float32(64e-6)
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(n) + 0.5 + 0.1 * np.sin(np.arange(n) * 0.37)).astype(np.float32)
    args = W.make_search_args(size, float(ts), min_snr=6.0)
    accs = W.accel_list(size, float(ts), -50, 50)
    got, _ = W.worker_search(args, x[None, :], np.zeros(1, np.float32), accs)
    p = orc.make_params(size, ts, min_snr=6.0)
    ref, _ = orc.search_dm(x, p, accs, nthreads=2)
    assert not compare(got, ref, 6.0)
