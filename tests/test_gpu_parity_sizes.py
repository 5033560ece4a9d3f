"""End-to-end parity with the oracle at the benchmark lengths.

The reference sequence (src/pipeline_multi.cu:155-295: baseline, R2C,
amplitude, running-median dereddening, interbin, stats, C2R, then per
acceleration resample, R2C, interbin, normalise, harmonic sums, peak finding,
harmonic distilling, and the per-DM acceleration distilling) is run by the
CPU oracle on the same series, and the candidate lists of the batched path
(with the schedule the benchmark uses for that length) and of the plain
Worker path must match it: keys (DM, accel, harmonic, bin) exact, S/N within
1e-3, a candidate present on one side only allowed only within 1e-3 of the
threshold (tests/candmatch.py).

Lengths: 2^23 at 76 us with red noise (the 64 x 2^23 workload's shape),
3*2^21 and 2^22 with 90% real samples (zero-padded segments of the uneven
workload), 2^26 at 64 us (the long-segment workload), 2^20 and 2^24 at
64 us (the ends of the serving workload's range), and 2^22 with five
harmonic folds at the example run's settings.  Each test takes a
subset of the acceleration plan that includes the injected pulsar's
acceleration, so the pulsar and its harmonics are found and the noise
candidates around the threshold are compared as well.
"""
import numpy as np
import pytest

from candmatch import compare

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 8


def _accel_subset(accs, n_every, inj_acc):
    """Every n_every-th acceleration plus the one nearest the injection."""
    idx = set(range(0, accs.size, n_every)) | {accs.size - 1, int(np.argmin(np.abs(accs - inj_acc)))}
    return accs[sorted(idx)]


def _oracle(orc, w, series, accs):
    ref = []
    for d in range(w.ndm):
        p = orc.make_params(w.size, np.float32(w.tsamp), dm=float(d), dm_idx=d, min_snr=w.min_snr,
                            nharmonics=w.nharmonics)
        ref.append(orc.search_dm(series[d], p, accs, nthreads=ORACLE_THREADS)[0])
    return np.concatenate(ref)


def _tuned_opts(psx, W, n, batch=None):
    from peasoup_amd import tuning
    t = tuning.lookup(n, psx.device_info(0)["name"])
    return W.make_batch_opts(batch or t.get("batch", 16), decomp=t.get("decomp")), t


def _check(psx, orc, w, accs, inj_freq, batch=None):
    import peasoup_amd.workloads as W
    series = W.generate_host(w)
    print("parity %s: n=%d nsamps=%d ndm=%d naccs=%d" % (w.name, w.size, w.n_real, w.ndm, accs.size), flush=True)
    ref = _oracle(orc, w, series, accs)
    opts, tuned = _tuned_opts(psx, W, w.size, batch)
    got, _ = W.batch_search(w.search_args(), series, w.dm_list(), accs, opts)
    problems = compare(got, ref, w.min_snr)
    assert not problems, ("batched, plan %s" % tuned, problems[:10])
    plain, _ = W.worker_search(w.search_args(), series, w.dm_list(), accs)
    problems = compare(plain, ref, w.min_snr)
    assert not problems, ("plain", problems[:10])
    # the injected pulsar is among the matched candidates (fundamental or a harmonic)
    hits = ref[(ref["snr"] > 10)]
    ratios = hits["freq"] / inj_freq
    assert np.any(np.abs(ratios - np.round(ratios)) < 1e-3 * np.maximum(np.round(ratios), 1)), hits[:10]
    return got, ref


def test_parity_2pow23_red_noise(psx, orc):
    """64 x 2^23 workload shape: 76 us, red noise (5x white at 0.1 Hz), one
    pulsar; 2 DM trials x 10 accelerations over -100..+100 m/s^2."""
    import peasoup_amd.workloads as W
    w = W.Workload("c2-2dm", 2, 1 << 23, 76e-6, 0x5EA5C2, 5.0, -100.0, 100.0)
    accs_all = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    assert accs_all.size == 127
    inj_acc = float(accs_all[83])
    w.injections = [W.Injection(1, 0.0173, inj_acc, 22.0, 0.4)]
    accs = _accel_subset(accs_all, 16, inj_acc)
    got, ref = _check(psx, orc, w, accs, 1 / 0.0173)
    assert len(ref) > 5


@pytest.mark.parametrize("n", [3 * (1 << 21), 1 << 22])
def test_parity_uneven_segments(psx, orc, n):
    """Uneven-length workload: 90% real samples, zero padded after baseline
    removal (pipeline_multi.cu:174-187), 76 us, accel -50..+50 m/s^2."""
    import peasoup_amd.workloads as W
    w = W.Workload("seg%d" % n, 2, n, 76e-6, 0x5EA5C3 + n % 97, 5.0, -50.0, 50.0, nsamps=int(n * 0.9))
    accs_all = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    inj_acc = float(accs_all[accs_all.size // 3])
    w.injections = [W.Injection(0, 0.0061, inj_acc, 20.0, 0.7)]
    accs = _accel_subset(accs_all, 5, inj_acc)
    _check(psx, orc, w, accs, 1 / 0.0061)


@pytest.mark.parametrize("n", [1 << 25, 1 << 26])
def test_parity_long_segments(psx, orc, n):
    """Long segments: 2^25 and 2^26 samples at 64 us with red noise, one
    binary pulsar; 1 DM trial x 3 accelerations around it (the default
    three-pass plans: two short strided passes and a 4,096-point contiguous
    one; tile-major resampling at 2^26)."""
    import peasoup_amd.workloads as W
    w = W.Workload("long", 1, n, 64e-6, 0x5EA5C4, 5.0, -50.0, 50.0)
    accs_all = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    k = int(np.argmin(np.abs(accs_all - 21.0)))
    w.injections = [W.Injection(0, 0.00427, float(accs_all[k]), 25.0, 0.1)]
    accs = accs_all[k - 1:k + 2]
    _check(psx, orc, w, accs, 1 / 0.00427)


@pytest.mark.parametrize("n,acc_range", [(1 << 20, 50.0), (1 << 24, 200.0)])
def test_parity_serving_lengths(psx, orc, n, acc_range):
    """Serving-workload lengths at the ends of its range: 2^20 and 2^24
    samples at 64 us with red noise, accel +-50 or +-200 m/s^2."""
    import peasoup_amd.workloads as W
    w = W.Workload("serve%d" % n, 2, n, 64e-6, 0x5EA5C5 + n % 89, 5.0, -acc_range, acc_range)
    accs_all = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi)
    inj_acc = float(accs_all[(2 * accs_all.size) // 3])
    w.injections = [W.Injection(1, 0.0093, inj_acc, 20.0, 0.6)]
    accs = _accel_subset(accs_all, max(1, accs_all.size // 6), inj_acc)
    _check(psx, orc, w, accs, 1 / 0.0093)


def test_parity_five_folds_example_run_settings(psx, orc):
    """The example run's search settings (5 harmonic folds, S/N 6, acc_tol
    1.11, example_output/overview.xml:40-62) at 2^22 samples."""
    import peasoup_amd.workloads as W
    w = W.Workload("five", 2, 1 << 22, 64e-6, 0x5EA5C6, 5.0, -150.0, 150.0, nharmonics=5, min_snr=6.0)
    accs_all = W.accel_list(w.size, w.tsamp, w.acc_lo, w.acc_hi, tol=1.11)
    inj_acc = float(accs_all[accs_all.size // 4])
    w.injections = [W.Injection(0, 0.0031, inj_acc, 18.0, 0.2)]
    accs = _accel_subset(accs_all, max(1, accs_all.size // 5), inj_acc)
    got, ref = _check(psx, orc, w, accs, 1 / 0.0031)
    print("harmonic numbers of the matched candidates:", np.bincount(ref["nh"], minlength=6), flush=True)
    assert set(got["nh"]) <= set(range(6))
