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
