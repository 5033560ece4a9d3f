"""Program-level post-processing (DM + harmonic distilling, scoring, --limit)
and the candidates.peasoup format (CPU only: these are host computations).

Pins:
  * oracle restatement == the reference's own scorer.hpp/distiller.hpp built
    into oracle/_ref (random forests, ties included);
  * the library's psx_finalize == the oracle, bit for bit;
  * hand-computed scorer cases and a hand-worked finalize example;
  * the reference's example output: its 1,250 final candidates are all
    "physical" at foff = -8.5 MHz, and their byte offsets in
    candidates.peasoup follow from their association counts;
  * the reference's example candidates.peasoup (FOLD records) reads and
    re-writes byte for byte.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_SO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref.so")

FIELDS = ["dm", "dm_idx", "acc", "nh", "snr", "freq", "nassoc"]


@pytest.fixture(scope="module")
def pp():
    import peasoup_amd.postprocess as pp
    return pp


def _ref():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = C.CDLL(REF_SO)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    f = C.c_float
    L.ref_score.argtypes = [fp, ip, C.c_int, f, f, f, f, fp, ip]
    L.ref_finalize.argtypes = [fp, ip, C.c_int, f, f, f, f, f, f, C.c_int, fp, ip, C.c_int, fp, ip,
                               C.POINTER(C.c_int)]
    return L


def _rows(forest):
    r = np.zeros((len(forest), 6), np.float32)
    for k, name in enumerate(["dm", "dm_idx", "acc", "nh", "snr", "freq"]):
        r[:, k] = forest[name]
    return np.ascontiguousarray(r), np.ascontiguousarray(forest["nassoc"], np.int32)


def _from_rows(rows, nassoc, dtype):
    out = np.zeros(len(rows), dtype)
    for k, name in enumerate(["dm", "dm_idx", "acc", "nh", "snr", "freq"]):
        out[name] = rows[:, k]
    out["nassoc"] = nassoc
    return out


def _tree(dm, idx, acc, nh, snr, freq, kids=()):
    return ((dm, idx, acc, nh, snr, freq), list(kids))


def _flatten(trees, dtype):
    recs = []

    def walk(t):
        pos = len(recs)
        recs.append(list(t[0]) + [0])
        for k in t[1]:
            walk(k)
        recs[pos][6] = len(recs) - pos - 1

    for t in trees:
        walk(t)
    out = np.zeros(len(recs), dtype)
    for k, name in enumerate(FIELDS):
        out[name] = [r[k] for r in recs]
    return out


def random_forest(seed, dtype, ndm=24, npsr=3, nnoise=80, tol=1e-4):
    """Per-DM accel-distilled candidates as the searches return them: pulsars
    (and harmonics) seen at neighbouring DM trials, noise, accel associates,
    some equal S/N values."""
    rng = np.random.default_rng(seed)
    dms = np.cumsum(rng.uniform(0.3, 2.0, ndm)).astype(np.float32)
    trees = []
    for _ in range(npsr):
        f0 = float(rng.uniform(0.5, 200.0))
        d0 = int(rng.integers(0, ndm))
        for d in range(max(0, d0 - 4), min(ndm, d0 + 5)):
            for h in (1.0, 2.0, 3.0, 0.5, 4.0):
                freq = np.float32(f0 * h * (1 + rng.normal(0, 0.3 * tol)))
                snr = np.float32(60.0 / (1 + abs(d - d0)) / h + rng.normal(0, 0.5))
                kids = []
                for _k in range(int(rng.integers(0, 4))):
                    kid_kids = []
                    if rng.random() < 0.2:
                        kid_kids = [_tree(dms[d], d, float(rng.choice([-5., 5.])), int(rng.integers(0, 5)),
                                          np.float32(snr * 0.7), np.float32(freq * (1 + rng.normal(0, tol))))]
                    # mostly the same DM trial (accel associates); some from other trials, as
                    # after DM distilling
                    dk = int(np.clip(d + rng.choice([0, 0, 0, 1, -1, 2, -3]), 0, ndm - 1))
                    kids.append(_tree(dms[dk], dk, float(rng.choice([-10., -5., 5., 10.])), int(rng.integers(0, 5)),
                                      np.float32(snr * rng.uniform(0.6, 0.99)),
                                      np.float32(freq * (1 + rng.normal(0, tol))), kid_kids))
                trees.append(_tree(dms[d], d, 0.0, int(rng.integers(0, 5)), snr, freq, kids))
    for _ in range(nnoise):
        d = int(rng.integers(0, ndm))
        snr = np.float32(rng.choice([7.5, 8.0, 9.25]) if rng.random() < 0.3 else rng.uniform(6, 12))
        trees.append(_tree(dms[d], d, float(rng.uniform(-50, 50)), int(rng.integers(0, 5)), snr,
                           np.float32(rng.uniform(0.1, 1000.0))))
    # DM order, as the searches return them
    trees.sort(key=lambda t: t[0][1])
    return _flatten(trees, dtype)


SCORER_CASES = [
    # cfreq, foff, bw (MHz): the example observation (foff < 0), and positive foff
    (807.43359375, -8.5, 544.0),
    (1400.0, 0.390625, 400.0),
    (350.0, 0.1, 100.0),
]


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("case", SCORER_CASES)
def test_oracle_score_matches_reference_build(orc, seed, case):
    L = _ref()
    forest = random_forest(seed, orc.CAND_DTYPE)
    rows, na = _rows(forest)
    ntop = len(orc.top_level(forest))
    ratios = np.zeros(2 * ntop, np.float32)
    flags = np.zeros(2 * ntop, np.int32)
    cfreq, foff, bw = case
    n = L.ref_score(rows, na, len(forest), 64e-6, cfreq, foff, bw, ratios, flags)
    assert n == ntop
    got = orc.score(forest, cfreq, foff, bw)
    assert np.array_equal(got["ddm_count_ratio"].view(np.uint32), ratios[0::2].view(np.uint32))
    assert np.array_equal(got["ddm_snr_ratio"].view(np.uint32), ratios[1::2].view(np.uint32))
    assert np.array_equal(got["is_adjacent"], flags[0::2])
    assert np.array_equal(got["is_physical"], flags[1::2])
    assert 0 < got["is_adjacent"].sum() < ntop  # both outcomes occur


def _ref_finalize(L, forest, tol, max_harm, cfreq, foff, bw, limit, dtype):
    rows, na = _rows(forest)
    cap = 4 * len(forest) + 16
    orows = np.zeros((cap, 6), np.float32)
    ona = np.zeros(cap, np.int32)
    ratios = np.zeros(2 * len(forest) + 2, np.float32)
    flags = np.zeros(2 * len(forest) + 2, np.int32)
    nfin = C.c_int()
    n = L.ref_finalize(rows, na, len(forest), tol, max_harm, 64e-6, cfreq, foff, bw, limit, orows, ona, cap,
                       ratios, flags, C.byref(nfin))
    assert 0 <= n <= cap
    return _from_rows(orows[:n], ona[:n], dtype), ratios, flags, nfin.value


def _same_forest(a, b):
    assert len(a) == len(b)
    for name in FIELDS:
        x, y = np.asarray(a[name]), np.asarray(b[name])
        if x.dtype == np.float32:
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
        else:
            assert np.array_equal(x, y), name


@pytest.mark.parametrize("seed,limit", [(11, 1000), (12, 15), (13, 0), (14, 1000)])
def test_oracle_finalize_matches_reference_build(orc, seed, limit):
    L = _ref()
    cfreq, foff, bw = SCORER_CASES[seed % 3]
    forest = random_forest(seed, orc.CAND_DTYPE)
    ref, ratios, flags, nfin = _ref_finalize(L, forest, 1e-4, 16.0, cfreq, foff, bw, limit, orc.CAND_DTYPE)
    got, sc = orc.finalize(forest, 1e-4, 16.0, cfreq, foff, bw, limit)
    _same_forest(got, ref)
    assert len(sc) == nfin and nfin <= limit
    assert np.array_equal(sc["ddm_count_ratio"].view(np.uint32), ratios[0:2 * nfin:2].view(np.uint32))
    assert np.array_equal(sc["ddm_snr_ratio"].view(np.uint32), ratios[1:2 * nfin:2].view(np.uint32))
    assert np.array_equal(sc["is_adjacent"], flags[0:2 * nfin:2])
    assert np.array_equal(sc["is_physical"], flags[1:2 * nfin:2])
    if limit >= 1000:
        # the distillers merged something and every input record survives somewhere
        assert nfin < len(orc.top_level(forest))
        assert len(got) >= len(forest)


@pytest.mark.parametrize("seed,limit", [(21, 1000), (22, 7), (23, 1000)])
def test_product_finalize_matches_oracle(psx_host, orc, pp, seed, limit):
    cfreq, foff, bw = SCORER_CASES[seed % 3]
    forest = random_forest(seed, orc.CAND_DTYPE)
    ref, rsc = orc.finalize(forest, 1e-4, 16.0, cfreq, foff, bw, limit)
    args = pp.FinalArgs(1e-4, 16.0, 64e-6, cfreq, foff, bw, limit, 0)
    got, sc = pp.finalize(forest.view(pp.CAND_DTYPE) if forest.dtype != pp.CAND_DTYPE else forest, args)
    _same_forest(got, ref)
    assert len(sc) == len(rsc)
    for name in ("ddm_count_ratio", "ddm_snr_ratio", "is_adjacent", "is_physical"):
        assert np.array_equal(np.asarray(sc[name]).view(np.uint32), np.asarray(rsc[name]).view(np.uint32)), name
    tops = orc.top_level(ref)
    assert np.array_equal(sc["nassoc"], tops["nassoc"])


def test_finalize_worked_example(psx_host, pp):
    """Five candidates worked through by hand (freq_tol 1e-4, max_harm 16)."""
    A = _tree(10.0, 5, 0.0, 1, 20.0, 10.0)
    B = _tree(11.0, 6, 5.0, 1, 15.0, 10.0005)   # same period one DM trial up: DM-distilled into A
    Cc = _tree(10.0, 5, -5.0, 2, 12.0, 20.0)    # second harmonic of A: harmonic-distilled into A
    D = _tree(30.0, 9, 0.0, 0, 8.0, 33.3)       # unrelated
    E = _tree(30.0, 9, 0.0, 0, 8.0, 1000.0)     # unrelated, same S/N as D
    forest = _flatten([A, Cc, B, D, E], pp.CAND_DTYPE)
    cfreq, foff, bw = 1400.0, 0.390625, 400.0
    out, sc = pp.finalize(forest, pp.FinalArgs(1e-4, 16.0, 64e-6, cfreq, foff, bw, 1000, 0))
    assert list(out["freq"][:3]) == [np.float32(10.0), np.float32(10.0005), np.float32(20.0)]
    assert list(out["nassoc"][:3]) == [2, 0, 0]
    assert len(sc) == 3 and sorted(out["freq"][3:]) == [np.float32(33.3), np.float32(1000.0)]
    # A: associates at dm_idx 6 (adjacent) and 5
    assert sc["is_adjacent"][0] == 1 and sc["ndirect"][0] == 2
    # band smearing per unit DM: 4150 (1200^-2 - 1600^-2) s; window 1/(10 Hz * that) = 79 DM units,
    # so both associates (|ddm| = 1 and 0) are inside
    assert sc["ddm_count_ratio"][0] == 1.0 and sc["ddm_snr_ratio"][0] == 1.0
    # physical: 8300 * 0.390625 / 1400^3 s per DM = 1.18 us; 10 DM -> 11.8 us << 100 ms
    assert sc["is_physical"][0] == 1
    # unassociated candidates count as adjacent (unique) with ratios 1
    assert np.all(sc["is_adjacent"][1:] == 1) and np.all(sc["ddm_count_ratio"][1:] == 1.0)
    out1, sc1 = pp.finalize(forest, pp.FinalArgs(1e-4, 16.0, 64e-6, cfreq, foff, bw, 1, 0))
    assert len(sc1) == 1 and len(out1) == 3


def test_scorer_hand_cases(orc):
    cfreq, foff, bw = 1400.0, 0.390625, 400.0
    chan = 8300.0 * foff / cfreq ** 3          # s per unit DM
    band = 4150.0 * (1 / 1200.0 ** 2 - 1 / 1600.0 ** 2)
    dm = 1000.0
    f_phys, f_unphys = 0.9 / (dm * chan), 1.1 / (dm * chan)
    width = 1.0 / (5.0 * band)                 # at 5 Hz: ~159 DM units
    trees = [
        _tree(dm, 100, 0, 0, 10.0, f_phys),
        _tree(dm, 100, 0, 0, 10.0, f_unphys),
        # associates at idx 100 only: "unique" -> adjacent
        _tree(50.0, 100, 0, 0, 10.0, 5.0, [_tree(50.0, 100, 5, 0, 9.0, 5.0)]),
        # associates two trials away only -> not adjacent
        _tree(50.0, 100, 0, 0, 10.0, 5.0, [_tree(60.0, 102, 5, 0, 9.0, 5.0)]),
        # one trial below -> adjacent, even after a far one
        _tree(50.0, 100, 0, 0, 10.0, 5.0, [_tree(60.0, 102, 5, 0, 9.0, 5.0), _tree(49.0, 99, 5, 0, 9.0, 5.0)]),
        # DM window: one associate inside (|ddm| = 0.9 width), one outside (1.1 width)
        _tree(50.0, 100, 0, 0, 10.0, 5.0, [_tree(50.0 + 0.9 * width, 7, 5, 0, 6.0, 5.0),
                                          _tree(50.0 - 1.1 * width, 3, 5, 0, 4.0, 5.0)]),
    ]
    f = _flatten(trees, orc.CAND_DTYPE)
    s = orc.score(f, cfreq, foff, bw)
    assert list(s["is_physical"]) == [1, 0, 1, 1, 1, 1]
    assert list(s["is_adjacent"]) == [1, 1, 1, 0, 1, 0]
    assert s["ddm_count_ratio"][5] == np.float32(2.0 / 3.0)
    assert s["ddm_snr_ratio"][5] == np.float32(16.0) / np.float32(20.0)
    # foff < 0 (the usual SIGPROC sign) makes every candidate physical
    s2 = orc.score(f, cfreq, -foff, bw)
    assert np.all(s2["is_physical"][:2] == 1) and orc.score(f, cfreq, -foff, bw)["is_physical"].all()


def test_example_output_candidates_are_physical(orc):
    """overview.xml's 1,250 final candidates: foff = -8.5 MHz, all physical."""
    cands = np.load(os.path.join(GOLD, "overview_candidates.npy"))
    params = json.load(open(os.path.join(GOLD, "overview_search_params.json")))
    foff, nchans = params["foff"], params["nchans"]
    cfreq = np.float32(params["fch1"] + foff * nchans / 2)   # Filterbank::get_cfreq, foff < 0
    recs = np.zeros(len(cands), orc.CAND_DTYPE)
    recs["dm"] = cands["dm"]
    recs["freq"] = (1.0 / cands["period"]).astype(np.float32)
    recs["snr"] = cands["snr"]
    s = orc.score(recs, cfreq, foff, abs(foff) * nchans)
    assert np.array_equal(s["is_physical"], cands["is_physical"])


def test_malformed_forest_rejected(psx_host, pp):
    f = _flatten([_tree(1.0, 0, 0, 0, 9.0, 3.0, [_tree(1.0, 0, 0, 0, 8.0, 3.0)])], pp.CAND_DTYPE)
    f["nassoc"][0] = 5
    with pytest.raises(psx_host.PsxError):
        pp.finalize(f, pp.make_final_args())
    with pytest.raises(psx_host.PsxError):
        pp.write_candidates("/tmp/never_written.peasoup", f)


# ----------------------------------------------------------------------------
# candidates.peasoup
# ----------------------------------------------------------------------------

def _forest_from_file_records(recs, dtype):
    """Top-level candidate + its detections as direct associates (the file
    stores pre-order detections; the nesting is not recorded)."""
    parts = []
    for _off, _fold, hits in recs:
        f = np.zeros(len(hits), dtype)
        for name in ["dm", "dm_idx", "acc", "nh", "snr", "freq"]:
            f[name] = hits[name]
        f["nassoc"][0] = len(hits) - 1
        parts.append(f)
    return np.concatenate(parts)


def test_example_candidates_file_roundtrip(psx_host, pp, tmp_path):
    src = os.path.join(GOLD, "example_candidates.peasoup")
    recs = pp.read_candidates(src)
    assert len(recs) == 10 and all(f is not None and f.shape == (16, 64) for _, f, _ in recs)
    assert [len(h) for _, _, h in recs[:3]] == [156, 165, 39]
    forest = _forest_from_file_records(recs, pp.CAND_DTYPE)
    out = tmp_path / "candidates.peasoup"
    offs = pp.write_candidates(out, forest, folds=[f for _, f, _ in recs])
    assert list(offs) == [o for o, _, _ in recs]
    assert out.read_bytes() == open(src, "rb").read()
    # random access as the reference's tools do it
    with pp.CandidateFileParser(str(out)) as p:
        fold, hits = p.cand_from_offset(int(offs[3]))
        assert np.array_equal(fold, recs[3][1]) and np.array_equal(hits, recs[3][2])


def test_example_overview_byte_offsets(psx_host, pp, tmp_path):
    """overview.xml's byte_offset column is CandidateFileWriter::byte_mapping
    of a file without folds: offset[i+1] = offset[i] + 4 + 24 (1 + nassoc[i])."""
    cands = np.load(os.path.join(GOLD, "overview_candidates.npy"))
    parts = []
    for c in cands:
        f = np.zeros(1 + int(c["nassoc"]), pp.CAND_DTYPE)
        f["nassoc"][0] = int(c["nassoc"])
        f["dm"], f["snr"] = c["dm"], c["snr"]
        parts.append(f)
    forest = np.concatenate(parts)
    out = tmp_path / "candidates.peasoup"
    offs = pp.write_candidates(out, forest)
    assert np.array_equal(offs, cands["byte_offset"])
    assert os.path.getsize(out) == int(cands["byte_offset"][-1]) + 4 + 24 * (1 + int(cands["nassoc"][-1]))
    back = pp.read_candidates(str(out))
    assert [len(h) - 1 for _, _, h in back] == list(cands["nassoc"])


def test_empty_forest(psx_host, pp, tmp_path):
    out, sc = pp.finalize(np.zeros(0, pp.CAND_DTYPE), pp.make_final_args())
    assert len(out) == 0 and len(sc) == 0
    p = tmp_path / "empty.peasoup"
    assert len(pp.write_candidates(p, np.zeros(0, pp.CAND_DTYPE))) == 0
    assert os.path.getsize(p) == 0
