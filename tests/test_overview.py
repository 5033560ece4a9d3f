"""overview.xml writer against the reference's example overview.xml (CPU).

The fixture is rebuilt section by section from typed values - numbers parsed
back to the types the reference prints (float32 candidate fields, float32
acceleration trials from the pinned golden list) - and the writer's output
must equal the file byte for byte.
"""
import lzma
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def fixture_bytes():
    with open(os.path.join(GOLD, "example_overview.xml.xz"), "rb") as f:
        return lzma.decompress(f.read())


def _typed(text):
    """A header/parameter value as the C++ side held it: int, double or string."""
    if text is None:
        return ""
    if re.fullmatch(r"-?\d+", text):
        return int(text)
    try:
        return float(text)
    except ValueError:
        return text


def test_overview_byte_exact(fixture_bytes):
    from peasoup_amd.overview import OverviewWriter, HEADER_FIELDS, SEARCH_FIELDS
    root = ET.fromstring(fixture_bytes)
    w = OverviewWriter()
    misc = root.find("misc_info")
    w.add_misc_info(misc.find("local_datetime").text, misc.find("utc_datetime").text)
    hdr = root.find("header_parameters")
    w.add_header({k: _typed(hdr.find(k).text) for k in HEADER_FIELDS})
    sp = root.find("search_parameters")
    w.add_search_parameters({k: _typed(sp.find(k).text) for k in SEARCH_FIELDS})
    seg = root.find("segment_parameters")
    w.add_segment_parameters(int(seg.find("segment_start_sample").text), int(seg.find("segment_nsamples").text),
                             float(seg.find("segment_pepoch").text))
    w.add_dm_list([np.float32(t.text) for t in root.find("dedispersion_trials").findall("trial")])
    acc_el = root.find("acceleration_trials")
    accs = np.load(os.path.join(GOLD, "overview_accel_list.npy"))   # pinned by the accel-plan tests
    w.add_acc_list(accs, np.float32(acc_el.get("DM")))
    gpu = root.find("cuda_device_parameters")
    devs = [(int(d.get("id")), d.find("name").text, int(d.find("major_cc").text), int(d.find("minor_cc").text))
            for d in gpu.findall("cuda_device")]
    w.add_gpu_info(int(gpu.find("runtime").text), int(gpu.find("driver").text), devs)

    import peasoup_amd.workloads as W
    from peasoup_amd.postprocess import SCORE_DTYPE
    cands = root.find("candidates").findall("candidate")
    tops = np.zeros(len(cands), W.CAND_DTYPE)
    scores = np.zeros(len(cands), SCORE_DTYPE)
    offs = np.zeros(len(cands), np.int64)
    for i, c in enumerate(cands):
        g = lambda k: c.find(k).text
        tops[i]["freq"] = np.float32(1.0 / float(g("period")))
        tops[i]["dm"], tops[i]["acc"], tops[i]["snr"] = np.float32(g("dm")), np.float32(g("acc")), np.float32(g("snr"))
        tops[i]["nh"], tops[i]["nassoc"] = int(g("nh")), int(g("nassoc"))
        scores[i]["is_adjacent"], scores[i]["is_physical"] = int(g("is_adjacent")), int(g("is_physical"))
        scores[i]["ddm_count_ratio"] = np.float32(g("ddm_count_ratio"))
        scores[i]["ddm_snr_ratio"] = np.float32(g("ddm_snr_ratio"))
        offs[i] = int(g("byte_offset"))
    w.add_candidates(tops, scores, offs)
    w.add_timing_info({t.tag: np.float32(t.text) for t in root.find("execution_times")})
    got = w.to_string().encode("latin-1")
    if got != fixture_bytes:
        a, b = got.decode("latin-1").splitlines(), fixture_bytes.decode("latin-1").splitlines()
        bad = [(i, x, y) for i, (x, y) in enumerate(zip(a, b)) if x != y][:5]
        raise AssertionError("first differing lines: %r (lengths %d vs %d)" % (bad, len(a), len(b)))


def test_overview_written_from_a_search(psx_host, tmp_path):
    """The writer consumes psx_finalize / psx_write_candidates outputs."""
    import peasoup_amd.postprocess as pp
    import peasoup_amd.workloads as W
    from peasoup_amd.overview import OverviewWriter
    f = np.zeros(3, W.CAND_DTYPE)
    f["freq"] = [10.0, 10.0005, 33.3]
    f["dm_idx"] = [0, 1, 4]
    f["dm"] = [1.0, 2.0, 5.0]
    f["snr"] = [20.0, 15.0, 9.0]
    fin, sc = pp.finalize(f, pp.make_final_args(limit=10))
    offs = pp.write_candidates(tmp_path / "candidates.peasoup", fin)
    w = OverviewWriter()
    w.add_misc_info()
    w.add_candidates(W.top_level(fin), sc, offs)
    w.to_file(tmp_path / "overview.xml")
    root = ET.parse(tmp_path / "overview.xml").getroot()
    cs = root.find("candidates").findall("candidate")
    assert [int(c.find("byte_offset").text) for c in cs] == list(offs)
    assert [int(c.find("nassoc").text) for c in cs] == [1, 0]
    assert cs[0].find("period").text == "%.15g" % (1.0 / float(np.float32(10.0)))
