"""Dedispersion on the GPU, and the tutorial filterbank end to end.

psx_dedisperse equals the oracle's direct dedispersion bit for bit (integer
channel sums, one float division).  End to end, tutorial.fil (the
reference's example data: a 250 ms pulsar) is read, dedispersed over the DM
list the reference's example candidates.peasoup was made with, and searched
with the batched path at the accelerations recorded in that file; the
detections it lists must be found again with their S/N.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
TUTORIAL = os.path.join(GOLD, "tutorial.fil")
f32 = np.float32


def _pack(samples, nbits):
    ns, nc = samples.shape
    per = 8 // nbits
    s = samples.reshape(ns, nc // per, per).astype(np.uint32)
    b = np.zeros((ns, nc // per), np.uint32)
    for k in range(per):
        b |= s[:, :, k] << (k * nbits)
    return b.astype(np.uint8).reshape(-1)


@pytest.mark.parametrize("nbits,nchans,nsamps,ndm", [(8, 64, 5000, 7), (2, 64, 20000, 12), (4, 96, 7001, 5),
                                                     (1, 256, 9000, 3)])
def test_dedisperse_vs_oracle(psx, orc, nbits, nchans, nsamps, ndm):
    import peasoup_amd.sigproc as sp
    import ctypes as C
    rng = np.random.default_rng(nbits * 1000 + nchans)
    v = rng.integers(0, 1 << nbits, size=(nsamps, nchans))
    packed = _pack(v, nbits)
    dt, f0, df = f32(0.000256), f32(1500.0), f32(-300.0 / nchans)
    table = orc.delay_table(nchans, dt, f0, df)
    dms = np.sort(rng.uniform(0, 0.8 * nsamps / table[-1], ndm)).astype(np.float32)
    km = np.ones(nchans, np.int32)
    km[rng.choice(nchans, 3, replace=False)] = 0
    ref = orc.dedisperse(packed, nsamps, nchans, nbits, dms, dt, f0, df, killmask=km)
    nout = ref.shape[1]
    d_in = psx.DeviceArray.from_numpy(packed)
    d_out = psx.DeviceArray((ndm, nout + 3), np.float32)
    work = psx.DeviceArray(int(psx.lib().psx_dedisp_work_bytes(nsamps, nchans, ndm)), np.uint8)
    print("dedisperse nbits=%d nchans=%d nsamps=%d ndm=%d nout=%d" % (nbits, nchans, nsamps, ndm, nout), flush=True)
    psx.check(psx.lib().psx_dedisperse(d_in.ptr, nsamps, nchans, nbits, dms.ctypes.data, ndm, km.ctypes.data, dt,
                                       f0, df, d_out.ptr, nout + 3, work.ptr, None))
    got = d_out.numpy()[:, :nout]
    assert np.array_equal(got, ref)


def test_dedisperse_rejects_unsorted_list(psx):
    import peasoup_amd.sigproc as sp
    fil = sp.Filterbank(TUTORIAL)
    dd = sp.Dedisperser(fil)
    dd.set_dm_list([50.0, 10.0])
    with pytest.raises(psx.PsxError):
        dd.dedisperse(0, 20000)


def test_tutorial_end_to_end(psx, orc):
    import peasoup_amd.postprocess as pp
    import peasoup_amd.sigproc as sp
    import peasoup_amd.workloads as W
    fil = sp.Filterbank(TUTORIAL)
    dd = sp.Dedisperser(fil)
    dms = dd.generate_dm_list(0.0, 260.0, 64.0, 1.1)
    trials, nout = dd.dedisperse()
    packed, n = fil.load()
    ref = orc.dedisperse(packed, n, fil.nchans, fil.nbits, dms, fil.get_tsamp(), fil.get_fch1(), fil.get_foff())
    assert np.array_equal(trials.numpy(), ref)

    size = 1 << 17  # Utils::prev_power_of_two(187520)
    args = W.make_search_args(size, float(fil.get_tsamp()), min_snr=9.0, nharmonics=4,
                              cfreq=float(fil.get_cfreq()) * 1e6, ch_bw=float(fil.get_foff()) * 1e6)
    accs = np.array([-5.0, 0.0, 5.0], np.float32)
    forest, _ = W.batch_search(args, trials, dms, accs, nsamps=nout, forest=True)
    plain, _ = W.worker_search(args, trials.numpy(), dms, accs, forest=True)
    assert len(forest) == len(plain)

    # every detection the example file lists is found again, S/N within 1%
    mine = {}
    for r in forest:
        mine.setdefault((int(r["dm_idx"]), float(r["acc"]), int(r["nh"])), []).append((float(r["freq"]),
                                                                                     float(r["snr"])))
    recs = pp.read_candidates(os.path.join(GOLD, "example_candidates.peasoup"))
    missing, off = [], []
    for _, _, hits in recs:
        for h in hits:
            cands = mine.get((int(h["dm_idx"]), float(h["acc"]), int(h["nh"])), [])
            best = min(cands, key=lambda fs: abs(fs[0] - float(h["freq"])), default=None)
            if best is None or abs(best[0] - float(h["freq"])) > 1e-5 * float(h["freq"]):
                missing.append((int(h["dm_idx"]), float(h["acc"]), int(h["nh"]), float(h["freq"]), float(h["snr"])))
            elif abs(best[1] - float(h["snr"])) > 0.01 * float(h["snr"]):
                off.append((int(h["dm_idx"]), float(h["nh"]), float(h["freq"]), float(h["snr"]), best[1]))
    ntotal = sum(len(h) for _, _, h in recs)
    print("tutorial: %d detections in the example file, %d missing, %d with S/N off" %
          (ntotal, len(missing), len(off)), missing[:5], off[:5], flush=True)
    assert not missing and not off
