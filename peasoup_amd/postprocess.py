"""Program-level candidate post-processing and candidates.peasoup I/O.

``finalize`` is main()'s chain after the searches
(src/pipeline_multi.cu:406-407,547-558): DM distilling and harmonic
distilling with related candidates kept, scoring (CandidateScorer,
include/transforms/scorer.hpp) and the --limit cut, computed by libpsx.

``write_candidates`` writes candidates.peasoup as CandidateFileWriter::
write_binary does (include/utils/output_stats.hpp:261-298);
``CandidateFileParser`` reads it back the way the reference's tools do
(tools/peasoup_tools.py:46-80), and ``read_candidates`` walks a whole file.
"""
import ctypes as C
import struct

import numpy as np

import peasoup_amd as P
from peasoup_amd.workloads import CAND_DTYPE, top_level

# CandidatePOD (include/data_types/candidates.hpp:10-17), 24 bytes
POD_DTYPE = np.dtype([("dm", "<f4"), ("dm_idx", "<i4"), ("acc", "<f4"), ("nh", "<i4"), ("snr", "<f4"),
                      ("freq", "<f4")])
SCORE_DTYPE = np.dtype([("ddm_count_ratio", "<f4"), ("ddm_snr_ratio", "<f4"), ("is_adjacent", "<i4"),
                        ("is_physical", "<i4"), ("nassoc", "<i4"), ("ndirect", "<i4")])
FOLD_MAGIC = b"FOLD"


class FinalArgs(C.Structure):
    _fields_ = [("freq_tol", C.c_float), ("max_harm", C.c_float), ("tsamp", C.c_float), ("cfreq", C.c_float),
                ("foff", C.c_float), ("bw", C.c_float), ("limit", C.c_int32), ("reserved", C.c_int32)]


class Fold(C.Structure):
    _fields_ = [("nbins", C.c_int32), ("nints", C.c_int32), ("data", C.c_void_p)]


P.register_signatures({
    "psx_finalize": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(FinalArgs), C.c_void_p, C.c_uint32,
                               C.POINTER(C.c_uint32), C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "psx_write_candidates": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                       C.POINTER(C.c_uint32)]),
})


def make_final_args(freq_tol=1e-4, max_harm=16.0, tsamp=64e-6, cfreq=1400.0, foff=-0.390625, nchans=1024,
                    limit=1000):
    """Arguments as main() builds them (bw = |foff| * nchans, MHz)."""
    return FinalArgs(freq_tol, max_harm, tsamp, cfreq, foff, abs(foff) * nchans, limit, 0)


def finalize(forest, args, cap=None):
    """Final candidates of a search (forest in, forest out) and their scores
    (one SCORE_DTYPE row per top-level candidate)."""
    forest = np.ascontiguousarray(forest, CAND_DTYPE)
    cap = cap or max(4 * len(forest), 1024)
    out = np.zeros(cap, CAND_DTYPE)
    scores = np.zeros(max(len(forest), 1), SCORE_DTYPE)
    cnt, nfin = C.c_uint32(), C.c_uint32()
    P.check(P.lib().psx_finalize(forest.ctypes.data, len(forest), C.byref(args), out.ctypes.data, cap,
                                 C.byref(cnt), scores.ctypes.data, len(scores), C.byref(nfin)))
    return out[:cnt.value], scores[:nfin.value]


def write_candidates(path, forest, folds=None):
    """Write candidates.peasoup; folds: None or one (nints, nbins) float32
    array (or None) per top-level candidate.  Returns the byte offsets of the
    top-level candidates (CandidateFileWriter::byte_mapping)."""
    forest = np.ascontiguousarray(forest, CAND_DTYPE)
    ntop = len(top_level(forest)) if len(forest) else 0
    offsets = np.zeros(max(ntop, 1), np.int64)
    keep = []
    fold_arr = None
    if folds is not None:
        if len(folds) != ntop:
            raise ValueError("need one fold per top-level candidate (%d), got %d" % (ntop, len(folds)))
        fold_arr = (Fold * max(ntop, 1))()
        for t, f in enumerate(folds):
            if f is None:
                fold_arr[t] = Fold(0, 0, None)
                continue
            f = np.ascontiguousarray(f, np.float32)
            keep.append(f)
            fold_arr[t] = Fold(f.shape[1], f.shape[0], f.ctypes.data)
    n = C.c_uint32()
    P.check(P.lib().psx_write_candidates(str(path).encode(), forest.ctypes.data, len(forest),
                                         C.cast(fold_arr, C.c_void_p) if fold_arr is not None else None,
                                         offsets.ctypes.data, max(ntop, 1), C.byref(n)))
    return offsets[:n.value]


class CandidateFileParser:
    """Random access by byte offset (tools/peasoup_tools.py:46-80): returns
    (fold or None, detections) where fold is (nints, nbins) and detections a
    POD_DTYPE array (the candidate first, then its associates)."""

    def __init__(self, path):
        self._f = open(path, "rb")

    def _read_fold(self):
        nbins, nints = struct.unpack("<ii", self._f.read(8))
        fold = np.fromfile(self._f, dtype="<f4", count=nbins * nints)
        return fold.reshape(nints, nbins)

    def _read_hits(self):
        count, = struct.unpack("<i", self._f.read(4))
        return np.fromfile(self._f, dtype=POD_DTYPE, count=count)

    def cand_from_offset(self, offset):
        self._f.seek(offset)
        if self._f.read(4) == FOLD_MAGIC:
            fold = self._read_fold()
            return fold, self._read_hits()
        self._f.seek(offset)
        return None, self._read_hits()

    def tell(self):
        return self._f.tell()

    def close(self):
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_candidates(path):
    """Every record of a candidates.peasoup file in order:
    [(offset, fold or None, detections), ...]."""
    import os
    size = os.path.getsize(path)
    out = []
    with CandidateFileParser(path) as f:
        off = 0
        while off < size:
            fold, hits = f.cand_from_offset(off)
            out.append((off, fold, hits))
            off = f.tell()
    return out
