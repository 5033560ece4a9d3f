"""overview.xml, the search summary the reference writes beside
candidates.peasoup (OutputFileWriter, include/utils/output_stats.hpp:20-243,
with the element formatting of include/utils/xml_util.hpp).

Formatting follows the reference byte for byte, since downstream tools parse
it: numbers as a C++ stream with precision 15 prints them (printf "%.15g";
floats widened to double first), booleans as 0/1, attributes in key order
and quoted with single quotes, two spaces of indentation per level, an
ISO-8859-1 XML declaration.  DM trials and the centre DM are printed with
four decimals (output_stats.hpp:101,163-167,176).

The device section keeps the reference's element names
(cuda_device_parameters / cuda_device / runtime / driver / name / major_cc /
minor_cc) so existing parsers keep working; it carries the HIP runtime and
driver versions and the GPU's name and gfx major/minor.
"""
import time

import numpy as np


def fmt(v):
    """Value as `stream << std::setprecision(15) << v` prints it."""
    if isinstance(v, (bool, np.bool_)):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        s = "%.15g" % float(v)
        return s
    return str(v)


def fixed4(v):
    """std::fixed << std::setprecision(4) of a float."""
    return "%.4f" % float(np.float32(v))


class Element:
    """One XML element: text or children, attributes printed in key order."""

    def __init__(self, name, value=None):
        self.name = name
        self.attributes = {}
        self.children = []
        self.text = "" if value is None else fmt(value)

    def append(self, child):
        self.children.append(child)
        return child

    def add_attribute(self, key, value):
        self.attributes[key] = "'%s'" % fmt(value)

    def set_text(self, value):
        self.text = fmt(value)

    def to_string(self, header=False, level=0):
        out = []
        if header:
            out.append("<?xml version='1.0' encoding='ISO-8859-1'?>\n")
        out.append("  " * level + "<" + self.name)
        for k in sorted(self.attributes):
            out.append(" %s=%s" % (k, self.attributes[k]))
        out.append(">")
        if not self.children:
            out.append(self.text)
        else:
            out.append("\n")
            out.extend(c.to_string(False, level + 1) for c in self.children)
            out.append("  " * level)
        out.append("</%s>\n" % self.name)
        return "".join(out)


# Field order of each section (output_stats.hpp:47-75, 85-119)
HEADER_FIELDS = ["source_name", "rawdatafile", "az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp",
                 "period", "fch1", "foff", "nchans", "telescope_id", "machine_id", "data_type", "ibeam", "nbeams",
                 "nbits", "barycentric", "pulsarcentric", "nbins", "nsamples", "nifs", "npuls", "refdm", "signed"]
SEARCH_FIELDS = ["infilename", "outdir", "killfilename", "zapfilename", "max_num_threads", "size", "dmfilename",
                 "cdm", "dm_start", "dm_end", "dm_tol", "dm_pulse_width", "acc_start", "acc_end", "acc_tol",
                 "acc_pulse_width", "boundary_5_freq", "boundary_25_freq", "nharmonics", "npdmp", "min_snr",
                 "min_freq", "max_freq", "max_harm", "freq_tol", "verbose", "progress_bar"]
CANDIDATE_FIELDS = ["period", "opt_period", "dm", "acc", "nh", "snr", "folded_snr", "is_adjacent", "is_physical",
                    "ddm_count_ratio", "ddm_snr_ratio", "nassoc", "byte_offset"]


class OverviewWriter:
    """Builds overview.xml section by section, in the order main() adds them
    (src/pipeline_multi.cu:563-580)."""

    def __init__(self):
        self.root = Element("peasoup_search")

    def add_misc_info(self, local_datetime=None, utc_datetime=None):
        t = time.time()
        info = self.root.append(Element("misc_info"))
        info.append(Element("local_datetime",
                            local_datetime or time.strftime("%Y-%m-%d-%H:%M", time.localtime(t))))
        info.append(Element("utc_datetime", utc_datetime or time.strftime("%Y-%m-%d-%H:%M", time.gmtime(t))))

    def add_header(self, header):
        """header: mapping with the SIGPROC header fields of HEADER_FIELDS."""
        h = self.root.append(Element("header_parameters"))
        for k in HEADER_FIELDS:
            h.append(Element(k, header.get(k, 0 if k not in ("source_name", "rawdatafile") else "")))

    def add_search_parameters(self, params):
        """params: the command-line options of SEARCH_FIELDS (cdm printed with
        four decimals)."""
        s = self.root.append(Element("search_parameters"))
        for k in SEARCH_FIELDS:
            v = params.get(k, "" if k in ("infilename", "outdir", "killfilename", "zapfilename", "dmfilename")
                           else 0)
            s.append(Element(k, fixed4(v) if k == "cdm" else v))

    def add_segment_parameters(self, start_sample, nsamples, pepoch):
        s = self.root.append(Element("segment_parameters"))
        s.append(Element("segment_start_sample", int(start_sample)))
        s.append(Element("segment_nsamples", int(nsamples)))
        s.append(Element("segment_pepoch", float(pepoch)))

    def add_dm_list(self, dms):
        e = self.root.append(Element("dedispersion_trials"))
        e.add_attribute("count", len(dms))
        for i, dm in enumerate(dms):
            t = e.append(Element("trial"))
            t.add_attribute("id", i)
            t.text = fixed4(dm)

    def add_acc_list(self, accs, cdm):
        e = self.root.append(Element("acceleration_trials"))
        e.add_attribute("count", len(accs))
        e.attributes["DM"] = "'%s'" % fixed4(cdm)
        for i, a in enumerate(np.asarray(accs, np.float32)):
            t = e.append(Element("trial"))
            t.add_attribute("id", i)
            t.set_text(np.float32(a))

    def add_gpu_info(self, runtime, driver, devices):
        """devices: [(id, name, major, minor), ...]."""
        g = self.root.append(Element("cuda_device_parameters"))
        g.append(Element("runtime", int(runtime)))
        g.append(Element("driver", int(driver)))
        for idx, name, major, minor in devices:
            d = g.append(Element("cuda_device"))
            d.add_attribute("id", int(idx))
            d.append(Element("name", name))
            d.append(Element("major_cc", int(major)))
            d.append(Element("minor_cc", int(minor)))

    def add_candidates(self, tops, scores, byte_offsets):
        """tops: top-level final candidates (CAND_DTYPE rows), scores: their
        SCORE_DTYPE rows (psx_finalize), byte_offsets: their offsets in
        candidates.peasoup (psx_write_candidates)."""
        e = self.root.append(Element("candidates"))
        for i in range(len(tops)):
            c, s = tops[i], scores[i]
            freq = np.float32(c["freq"])
            vals = {
                "period": 1.0 / float(freq), "opt_period": 0.0, "dm": np.float32(c["dm"]),
                "acc": np.float32(c["acc"]), "nh": int(c["nh"]), "snr": np.float32(c["snr"]), "folded_snr": 0.0,
                "is_adjacent": bool(s["is_adjacent"]), "is_physical": bool(s["is_physical"]),
                "ddm_count_ratio": np.float32(s["ddm_count_ratio"]), "ddm_snr_ratio": np.float32(s["ddm_snr_ratio"]),
                "nassoc": int(c["nassoc"]), "byte_offset": int(byte_offsets[i]),
            }
            ce = e.append(Element("candidate"))
            ce.add_attribute("id", i)
            for k in CANDIDATE_FIELDS:
                ce.append(Element(k, vals[k]))

    def add_timing_info(self, times):
        """times: {name: seconds}; printed in key order (a std::map)."""
        e = self.root.append(Element("execution_times"))
        for k in sorted(times):
            e.append(Element(k, np.float32(times[k])))

    def to_string(self):
        return self.root.to_string(True)

    def to_file(self, path):
        with open(path, "wb") as f:
            f.write(self.to_string().encode("latin-1"))
