"""SIGPROC reading and dedispersion (CPU).

Pins: the tutorial filterbank's header; libdedisp's DM list against the 59
DM trials recorded in the reference's example candidates.peasoup (searched
from that file with dm_tol 1.1 and a 64 us pulse width); the delay table's
closed form; the oracle's direct dedispersion against dispersed pulses that
must land in a single sample at their DM and against hand-packed sub-byte
samples.  The library's host-side entry points are compared with the oracle.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
TUTORIAL = os.path.join(GOLD, "tutorial.fil")
f32 = np.float32


@pytest.fixture(scope="module")
def sp(psx_host):
    import peasoup_amd.sigproc as sp
    return sp


def test_tutorial_header(sp):
    fil = sp.Filterbank(TUTORIAL)
    h = fil.header.as_dict()
    assert (h["nchans"], h["nbits"], h["nifs"], h["data_type"]) == (64, 2, 1, 1)
    assert h["tsamp"] == 0.00032 and h["fch1"] == 1510.0 and h["foff"] == -1.09 and h["tstart"] == 50000.0
    assert h["source_name"] == "P: 250.000000000000 ms, DM: 30.000"
    assert h["size"] == 244
    # no nsamples keyword: from the file size, 3,000,320 data bytes / 16 bytes per sample
    assert h["nsamples"] == 187520
    # Filterbank::get_cfreq with foff < 0: fch1 + foff * nchans / 2 in float
    assert fil.get_cfreq() == f32(f32(1510.0) + f32(-1.09) * f32(64) / f32(2))
    raw, n = fil.load(100, 3)
    with open(TUTORIAL, "rb") as f:
        f.seek(244 + 100 * 16)
        assert raw.tobytes() == f.read(48) and n == 3
    tail, n = fil.load(187518, 10)   # past the end: whole samples only
    assert n == 2 and tail.size == 32


def _fixture_dms():
    from peasoup_amd.postprocess import read_candidates
    dms = {}
    for _, _, hits in read_candidates(os.path.join(GOLD, "example_candidates.peasoup")):
        for r in hits:
            dms[int(r["dm_idx"])] = f32(r["dm"])
    return dms


def test_dm_list_golden(orc, sp):
    dms = _fixture_dms()
    assert len(dms) == 59 and max(dms) == 58
    args = (0.0, 260.0, f32(0.00032), 64.0, f32(1510.0), f32(-1.09), 64, 1.1)
    ref = orc.dm_list(*args)
    got = sp.generate_dm_list(*args)
    assert np.array_equal(ref, got)
    assert ref[0] == 0.0 and ref[-2] < 260.0 <= ref[-1] and np.all(np.diff(ref) > 0) and ref.size > 58
    for i, dm in dms.items():
        assert ref[i] == dm, (i, ref[i], dm)


@pytest.mark.parametrize("args", [(2.0, 100.0, 64e-6, 40.0, 1400.0, -0.390625, 1024, 1.25),
                                  (0.0, 2000.0, 54.613e-6, 64.0, 1712.0, -0.208984375, 4096, 1.11)])
def test_dm_list_oracle_vs_library(orc, sp, args):
    a = (args[0], args[1], f32(args[2]), args[3], f32(args[4]), f32(args[5]), args[6], args[7])
    assert np.array_equal(orc.dm_list(*a), sp.generate_dm_list(*a))


def test_delay_table(orc, sp):
    nchans, dt, f0, df = 64, f32(0.00032), f32(1510.0), f32(-1.09)
    t = orc.delay_table(nchans, dt, f0, df)
    assert np.array_equal(t, sp.delay_table(nchans, dt, f0, df))
    f = 1510.0 + np.arange(nchans) * -1.09
    exact = 4.15e3 / 0.00032 * (1.0 / f ** 2 - 1.0 / 1510.0 ** 2)
    assert t[0] == 0.0 and np.allclose(t, exact, rtol=2e-5, atol=1e-6)


def _pack(samples, nbits):
    """[nsamps, nchans] values -> packed bytes, lowest bits first."""
    ns, nc = samples.shape
    per = 8 // nbits
    s = samples.reshape(ns, nc // per, per).astype(np.uint32)
    b = np.zeros((ns, nc // per), np.uint32)
    for k in range(per):
        b |= s[:, :, k] << (k * nbits)
    return b.astype(np.uint8).reshape(-1)


def test_oracle_dedisperse_pulse_known_answer(orc):
    nchans, nsamps, dt, f0, df = 32, 600, f32(0.001), f32(1400.0), f32(-4.0)
    table = orc.delay_table(nchans, dt, f0, df)
    dms = np.array([0.0, 25.0, 50.0, 75.0], np.float32)
    t0 = 37
    x = np.zeros((nsamps, nchans), np.uint8)
    for c in range(nchans):
        x[t0 + int(np.rint(f32(50.0) * table[c])), c] = 255
    out = orc.dedisperse(_pack(x, 8), nsamps, nchans, 8, dms, dt, f0, df)
    assert out.shape == (4, nsamps - int(np.rint(dms[-1] * table[-1])))
    assert out[2, t0] == 1.0 and np.count_nonzero(out[2]) == 1     # all channels align at DM 50
    assert out[0].max() < 0.1 and out[3].max() < 0.2                # smeared elsewhere
    assert np.isclose(out[0].sum(), 1.0) and np.isclose(out[3].sum(), 1.0)  # flux is conserved
    # a masked channel drops out of the sum
    km = np.ones(nchans, np.int32)
    km[5] = 0
    out2 = orc.dedisperse(_pack(x, 8), nsamps, nchans, 8, dms, dt, f0, df, killmask=km)
    assert out2[2, t0] == f32(31 * 255) / f32(255 * 32)


@pytest.mark.parametrize("nbits", [1, 2, 4])
def test_oracle_dedisperse_sub_byte(orc, nbits):
    """Zero DM: the output is the channel sum of the unpacked values."""
    rng = np.random.default_rng(nbits)
    nchans, nsamps = 16, 40
    v = rng.integers(0, 1 << nbits, size=(nsamps, nchans))
    out = orc.dedisperse(_pack(v, nbits), nsamps, nchans, nbits, [0.0], f32(0.001), f32(1400.0), f32(-1.0))
    want = v.sum(axis=1).astype(np.float32) / f32(((1 << nbits) - 1) * nchans)
    assert np.array_equal(out[0], want)
