"""The drop-in C++ headers compile with a plain C++ compiler against libpsx
and their host-side parts run without a device (CPU): SigprocFilterbank and
Dedisperser::generate_dm_list on the tutorial file, the distillers and the
scorer on a handful of candidates, and CandidateCollection plumbing."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUTORIAL = os.path.join(ROOT, "tests", "golden", "tutorial.fil")

PROGRAM = r'''
#include <cstdio>
#include <vector>
#include "peasoup/candidates.hpp"
#include "peasoup/distiller.hpp"
#include "peasoup/filterbank.hpp"
#include "peasoup/scorer.hpp"
#include "peasoup/transforms.hpp"

int main(int argc, char** argv) {
  SigprocFilterbank fil(argv[1]);
  std::printf("nchans %u nbits %u nsamps %zu tsamp %.9g cfreq %.9g\n", fil.get_nchans(), fil.get_nbits(),
              fil.get_nsamps(), fil.get_tsamp(), fil.get_cfreq());
  Dedisperser dd(fil);
  dd.generate_dm_list(0.0f, 260.0f, 64.0f, 1.1f);
  std::vector<float> dms = dd.get_dm_list();
  std::printf("ndm %zu dm3 %.9g dm58 %.9g\n", dms.size(), dms[3], dms[58]);
  std::size_t got = fil.load_gulp(0, 10);
  std::printf("gulp %zu first %d\n", got, (int)fil.get_data()[0]);

  std::vector<Candidate> c;
  c.push_back(Candidate(10.0f, 5, 0.0f, 1, 20.0f, 10.0f));
  c.push_back(Candidate(11.0f, 6, 5.0f, 1, 15.0f, 10.0005f));
  c.push_back(Candidate(10.0f, 5, -5.0f, 2, 12.0f, 20.0f));
  c.push_back(Candidate(30.0f, 9, 0.0f, 0, 8.0f, 33.3f));
  DMDistiller dm_still(1e-4f, true);
  HarmonicDistiller harm_still(1e-4f, 16.0f, true, false);
  c = dm_still.distill(c);
  c = harm_still.distill(c);
  CandidateScorer scorer(64e-6f, 1400.0f, 0.390625f, 400.0f);
  scorer.score_all(c);
  for (const Candidate& x : c)
    std::printf("cand %.6g %d %d %d %.6g %.6g\n", x.freq, x.count_assoc(), (int)x.is_adjacent, (int)x.is_physical,
                x.ddm_count_ratio, x.ddm_snr_ratio);
  return 0;
}
'''


@pytest.fixture(scope="module")
def binary(tmp_path_factory, psx_host):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("hdr")
    src = d / "prog.cpp"
    src.write_text(PROGRAM)
    exe = d / "prog"
    lib = os.path.join(ROOT, "peasoup_amd")
    cmd = [cxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "peasoup_amd", "include"), str(src), "-L", lib, "-lpsx", "-Wl,-rpath," + lib,
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def test_headers_compile_and_run(binary):
    r = subprocess.run([binary, TUTORIAL], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "nchans 64 nbits 2 nsamps 187520 tsamp %.9g cfreq %.9g" % (np.float32(0.00032), np.float32(
        np.float32(1510.0) + np.float32(-1.09) * np.float32(64) / np.float32(2)))
    assert lines[1] == "ndm %d dm3 %.9g dm58 %.9g" % (int(lines[1].split()[1]), np.float32(9.908315),
                                                     float(lines[1].split()[5]))
    assert lines[2].startswith("gulp 10 ")
    cands = [l.split() for l in lines[3:]]
    # A absorbs B (one DM trial up) and C (its second harmonic); D stays alone
    assert [c[1:3] for c in cands] == [["10", "2"], ["33.3", "0"]]
    assert cands[0][3:] == ["1", "1", "1", "1"]


PROGRAM2 = r'''
#include <cstdio>
#include <string>
#include <vector>
#include "peasoup/candidates.hpp"
#include "peasoup/transforms.hpp"

int main(int argc, char** argv) {
  const std::string dir = argv[1];
  // AccelerationPlan at the example run's parameters (pipeline_multi.cu:412-421 units)
  const float fch1 = std::stof(argv[2]), foff = std::stof(argv[3]), tsamp = std::stof(argv[4]);
  const int nchans = std::stoi(argv[5]);
  const float cfreq = foff < 0 ? fch1 + foff * nchans / 2 : fch1 - foff * nchans / 2;
  AccelerationPlan plan(std::stof(argv[6]), std::stof(argv[7]), std::stof(argv[8]), std::stof(argv[9]) * 1e-6,
                        std::stoull(argv[10]), tsamp, cfreq * 1e6, foff * 1e6);
  std::vector<float> acc;
  plan.generate_accel_list(std::stof(argv[11]), std::stof(argv[12]), acc);
  FILE* fa = std::fopen((dir + "/accs.bin").c_str(), "wb");
  std::fwrite(acc.data(), sizeof(float), acc.size(), fa);
  std::fclose(fa);

  CandidateCollection cc;
  Candidate a(12.5f, 3, -7.25f, 2, 21.3f, 18.75f, 123.456f);
  Candidate b(12.5f, 4, 0.0f, 0, 9.1f, 61.728f);
  Candidate c(13.0f, 5, 1.5f, 1, 7.4f, 246.912f);
  b.append(c);
  a.append(b);
  a.is_adjacent = true;
  a.ddm_count_ratio = 0.5f;
  a.ddm_snr_ratio = 0.4271f;
  std::vector<float> fold(6);
  for (int i = 0; i < 6; ++i) fold[i] = 0.5f * i;
  a.set_fold(fold.data(), 3, 2);
  std::printf("fold %d %d %zu %g\n", a.nbins, a.nints, a.fold.size(), a.fold[5]);
  cc.cands.push_back(a);
  cc.cands.push_back(Candidate(30.0f, 9, 0.0f, 0, 8.0f, 33.3f));
  cc.print(stdout);
  cc.write_candidate_file(dir + "/candidates.txt");
  cc.generate_candidate_binaries(dir);
  return 0;
}
'''


@pytest.fixture(scope="module")
def binary2(tmp_path_factory, psx_host):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp("hdr2")
    src = d / "prog2.cpp"
    src.write_text(PROGRAM2)
    exe = d / "prog2"
    lib = os.path.join(ROOT, "peasoup_amd")
    cmd = [cxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "peasoup_amd", "include"), str(src), "-L", lib, "-lpsx", "-Wl,-rpath," + lib,
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def _cand_line(freq, opt_period, dm, acc, nh, snr, folded, adj, phys, r1, r2, nassoc):
    f = float(np.float32(freq))
    return "%.15f\t%.15f\t%.15f\t%.2f\t%.2f\t%d\t%.1f\t%.1f\t%d\t%d\t%.4f\t%.4f\t%d" % (
        1.0 / f, opt_period, f, float(np.float32(dm)), float(np.float32(acc)), nh, float(np.float32(snr)),
        float(np.float32(folded)), adj, phys, float(np.float32(r1)), float(np.float32(r2)), nassoc)


def test_product_accel_plan_and_candidate_writers(binary2, tmp_path):
    """The library's AccelerationPlan reproduces the example run's 10,758
    accelerations (example_output/overview.xml:91-10850) from the run's own
    parameters; Candidate::print, CandidateCollection::write_candidate_file
    and generate_candidate_binaries write the reference's text layout
    (include/data_types/candidates.hpp:86-148)."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "overview_search_params.json")) as f:
        p = json.load(f)
    gold = np.load(os.path.join(ROOT, "tests", "golden", "overview_accel_list.npy"))
    args = [str(tmp_path), repr(p["fch1"]), repr(p["foff"]), repr(p["tsamp"]), str(p["nchans"]),
            repr(p["acc_start"]), repr(p["acc_end"]), repr(p["acc_tol"]), repr(p["acc_pulse_width"]),
            str(p["size"]), repr(p["dm"]), repr(p["cdm"])]
    r = subprocess.run([binary2] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    acc = np.fromfile(str(tmp_path / "accs.bin"), np.float32)
    assert acc.size == gold.size == p["count"]
    assert np.array_equal(acc, gold)

    lines = r.stdout.splitlines()
    assert lines[0] == "fold 3 2 6 2.5"
    expect = [_cand_line(123.456, 0.0, 12.5, -7.25, 2, 21.3, 18.75, 1, 0, 0.5, 0.4271, 1),
              _cand_line(61.728, 0.0, 12.5, 0.0, 0, 9.1, 0.0, 0, 0, 0.0, 0.0, 1),
              _cand_line(246.912, 0.0, 13.0, 1.5, 1, 7.4, 0.0, 0, 0, 0.0, 0.0, 0),
              _cand_line(33.3, 0.0, 30.0, 0.0, 0, 8.0, 0.0, 0, 0, 0.0, 0.0, 0)]
    assert lines[1:] == expect
    txt = (tmp_path / "candidates.txt").read_text().splitlines()
    assert txt == (["#Period...Optimal period...Frequency...DM...Acceleration...Harmonic number...S/N...Folded S/N",
                    "#Candidate 0"] + expect[:3] + ["#Candidate 1", expect[3]])
    names = sorted(x.name for x in tmp_path.iterdir() if x.name.startswith("cand_"))
    want = ["cand_%04d_%.5f_%.1f_%.1f.peasoup" % (0, 1.0 / float(np.float32(123.456)), 12.5, -7.25),
            "cand_%04d_%.5f_%.1f_%.1f.peasoup" % (1, 1.0 / float(np.float32(33.3)), 30.0, 0.0)]
    assert names == sorted(want)
    assert (tmp_path / want[0]).read_text().splitlines() == expect[:3]
    assert (tmp_path / want[1]).read_text().splitlines() == expect[3:]
