"""Extract golden vectors from the reference's shipped example output.

Reads /root/reference/example_output/overview.xml (a file of the reference
repository, example_output/overview.xml) and writes, next to this script:

  overview_accel_list.npy     float32 acceleration trials at DM = cdm = 54.5
                              (overview.xml:91-10850, 10,758 values)
  overview_search_params.json the search/header parameters that fix that list
                              (overview.xml:15-19,40-62)
  overview_candidates.npy     the 1,250 final candidates' fields
                              (overview.xml:10860-29611): period, dm, acc,
                              nh, snr, scores, nassoc, byte_offset
  example_candidates.peasoup  the reference's example candidates.peasoup
                              (example_output/candidates.peasoup, a data
                              file: 10 FOLD + detection records), copied
  example_overview.xml.xz     example_output/overview.xml, xz-compressed
                              (the byte-exact format fixture of overview.xml)
  tutorial.fil                example_data/tutorial.fil, copied: the SIGPROC
                              filterbank (64 channels, 2 bits, 320 us; a
                              250 ms pulsar at DM 30) the example
                              candidates.peasoup was searched from

Run once in a container that has the reference mounted; the outputs are
committed so the tests do not need the reference.
"""
import json
import os
import re
import sys

import numpy as np

SRC = "/root/reference/example_output/overview.xml"
CANDS = "/root/reference/example_output/candidates.peasoup"
CAND_FIELDS = [("period", "<f8"), ("opt_period", "<f8"), ("dm", "<f8"), ("acc", "<f8"), ("nh", "<i4"),
               ("snr", "<f8"), ("folded_snr", "<f8"), ("is_adjacent", "<i4"), ("is_physical", "<i4"),
               ("ddm_count_ratio", "<f8"), ("ddm_snr_ratio", "<f8"), ("nassoc", "<i4"), ("byte_offset", "<i8")]
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    txt = open(SRC, encoding="latin-1").read()
    m = re.search(r"<acceleration_trials DM='([0-9.]+)' count='(\d+)'>(.*?)</acceleration_trials>", txt, re.S)
    vals = [float(v) for v in re.findall(r"<trial id='\d+'>([^<]+)</trial>", m.group(3))]
    assert len(vals) == int(m.group(2))
    np.save(os.path.join(HERE, "overview_accel_list.npy"), np.array(vals, dtype=np.float64).astype(np.float32))

    def tag(name):
        return re.search(r"<%s>([^<]*)</%s>" % (name, name), txt).group(1)

    params = {
        "dm": float(m.group(1)),
        "count": int(m.group(2)),
        "tsamp": float(tag("tsamp")),
        "fch1": float(tag("fch1")),
        "foff": float(tag("foff")),
        "nchans": int(tag("nchans")),
        "size": int(tag("size")),
        "cdm": float(tag("cdm")),
        "acc_start": float(tag("acc_start")),
        "acc_end": float(tag("acc_end")),
        "acc_tol": float(tag("acc_tol")),
        "acc_pulse_width": float(tag("acc_pulse_width")),
        "nharmonics": int(tag("nharmonics")),
        "min_snr": float(tag("min_snr")),
        "min_freq": float(tag("min_freq")),
        "max_freq": float(tag("max_freq")),
        "freq_tol": float(tag("freq_tol")),
        "max_harm": float(tag("max_harm")),
    }
    with open(os.path.join(HERE, "overview_search_params.json"), "w") as f:
        json.dump(params, f, indent=1, sort_keys=True)
    print("wrote %d accelerations" % len(vals), file=sys.stderr)

    block = re.search(r"<candidates>(.*)</candidates>", txt, re.S).group(1)
    rows = []
    for m in re.finditer(r"<candidate id='(\d+)'>(.*?)</candidate>", block, re.S):
        assert int(m.group(1)) == len(rows)
        body = m.group(2)
        rows.append(tuple(float(re.search(r"<%s>([^<]*)</%s>" % (k, k), body).group(1)) for k, _ in CAND_FIELDS))
    np.save(os.path.join(HERE, "overview_candidates.npy"), np.array(rows, dtype=CAND_FIELDS))
    with open(CANDS, "rb") as f:
        data = f.read()
    with open(os.path.join(HERE, "example_candidates.peasoup"), "wb") as f:
        f.write(data)
    print("wrote %d candidates, %d bytes of candidates.peasoup" % (len(rows), len(data)), file=sys.stderr)
    with open("/root/reference/example_data/tutorial.fil", "rb") as f, \
            open(os.path.join(HERE, "tutorial.fil"), "wb") as g:
        g.write(f.read())
    import lzma
    with open(SRC, "rb") as f, open(os.path.join(HERE, "example_overview.xml.xz"), "wb") as g:
        g.write(lzma.compress(f.read(), preset=9))


if __name__ == "__main__":
    main()
