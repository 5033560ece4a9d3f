"""Extract golden vectors from the reference's shipped example output.

Reads /root/reference/example_output/overview.xml (a file of the reference
repository, example_output/overview.xml) and writes, next to this script:

  overview_accel_list.npy     float32 acceleration trials at DM = cdm = 54.5
                              (overview.xml:91-10850, 10,758 values)
  overview_search_params.json the search/header parameters that fix that list
                              (overview.xml:15-19,40-62)

Run once in a container that has the reference mounted; the outputs are
committed so the tests do not need the reference.
"""
import json
import os
import re
import sys

import numpy as np

SRC = "/root/reference/example_output/overview.xml"
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


if __name__ == "__main__":
    main()
