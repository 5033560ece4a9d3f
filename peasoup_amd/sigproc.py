"""SIGPROC filterbanks and dedispersion: the input side of the search.

ctypes client of libpsx's psx_sigproc_* and psx_dedisp* entry points
(include/data_types/header.hpp, include/data_types/filterbank.hpp and
include/transforms/dedisperser.hpp in the reference; libdedisp's algorithm
restated on the GPU).  ``Filterbank`` reads a file's header and samples;
``Dedisperser`` mirrors the reference class (generate_dm_list, set_dm_list,
set_killmask, dedisperse) and returns the dispersion trials resident in HBM.
"""
import ctypes as C

import numpy as np

import peasoup_amd as P


class SigprocHeader(C.Structure):
    _fields_ = [("source_name", C.c_char * 80), ("rawdatafile", C.c_char * 80)] + \
               [(k, C.c_double) for k in ("az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp", "period",
                                          "fch1", "foff", "refdm")] + \
               [(k, C.c_int32) for k in ("nchans", "telescope_id", "machine_id", "data_type", "ibeam", "nbeams",
                                         "nbits", "barycentric", "pulsarcentric", "nbins", "nsamples", "nifs",
                                         "npuls")] + \
               [("signed_data", C.c_uint8), ("reserved", C.c_uint8 * 3), ("size", C.c_uint32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("reserved",)}
        d["source_name"] = self.source_name.decode("latin-1")
        d["rawdatafile"] = self.rawdatafile.decode("latin-1")
        d["signed"] = d.pop("signed_data")
        return d


vp, u64, u32, f32, i32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_float, C.c_int
P.register_signatures({
    "psx_sigproc_read_header": (i32, [C.c_char_p, C.POINTER(SigprocHeader)]),
    "psx_sigproc_read_samples": (i32, [C.c_char_p, C.POINTER(SigprocHeader), u64, u64, vp, C.POINTER(u64)]),
    "psx_sigproc_cfreq": (f32, [C.POINTER(SigprocHeader)]),
    "psx_dedisp_dm_list": (i32, [f32, f32, f32, f32, f32, f32, u32, f32, vp, u32, C.POINTER(u32)]),
    "psx_dedisp_delay_table": (i32, [u32, f32, f32, f32, vp]),
    "psx_dedisp_max_delay": (u64, [vp, u32, vp, u32]),
    "psx_dedisp_work_bytes": (u64, [u64, u32, u32]),
    "psx_dedisperse": (i32, [vp, u64, u32, u32, vp, u32, vp, f32, f32, f32, vp, u64, vp, vp]),
})


class Filterbank:
    """A SIGPROC filterbank file (SigprocFilterbank, filterbank.hpp:248-333)."""

    def __init__(self, path):
        self.path = str(path)
        self.header = SigprocHeader()
        P.check(P.lib().psx_sigproc_read_header(self.path.encode(), C.byref(self.header)))

    @property
    def nchans(self):
        return self.header.nchans

    @property
    def nbits(self):
        return self.header.nbits

    @property
    def nsamps(self):
        return self.header.nsamples

    def get_tsamp(self):
        return np.float32(self.header.tsamp)

    def get_fch1(self):
        return np.float32(self.header.fch1)

    def get_foff(self):
        return np.float32(self.header.foff)

    def get_cfreq(self):
        return np.float32(P.lib().psx_sigproc_cfreq(C.byref(self.header)))

    def load(self, start=0, nsamps=None):
        """Packed bytes of samples [start, start + nsamps) as a uint8 array."""
        n = self.nsamps - start if nsamps is None else nsamps
        bps = self.nchans * self.nbits // 8
        buf = np.zeros(n * bps, np.uint8)
        got = C.c_uint64()
        P.check(P.lib().psx_sigproc_read_samples(self.path.encode(), C.byref(self.header), start, n,
                                                 buf.ctypes.data, C.byref(got)))
        return buf[:got.value * bps], got.value


def generate_dm_list(dm_start, dm_end, tsamp, pulse_width_us, f0, df, nchans, tol):
    cnt = C.c_uint32()
    P.check(P.lib().psx_dedisp_dm_list(dm_start, dm_end, tsamp, pulse_width_us, f0, df, nchans, tol, None, 0,
                                       C.byref(cnt)))
    out = np.zeros(cnt.value, np.float32)
    P.check(P.lib().psx_dedisp_dm_list(dm_start, dm_end, tsamp, pulse_width_us, f0, df, nchans, tol,
                                       out.ctypes.data, out.size, C.byref(cnt)))
    return out


def delay_table(nchans, tsamp, f0, df):
    t = np.zeros(nchans, np.float32)
    P.check(P.lib().psx_dedisp_delay_table(nchans, tsamp, f0, df, t.ctypes.data))
    return t


class Dedisperser:
    """The reference's Dedisperser (dedisperser.hpp:13-186) over libpsx."""

    def __init__(self, fil):
        self.fil = fil
        self.dm_list = np.zeros(0, np.float32)
        self.killmask = None

    def generate_dm_list(self, dm_start, dm_end, width_us, tol):
        self.dm_list = generate_dm_list(dm_start, dm_end, self.fil.get_tsamp(), width_us, self.fil.get_fch1(),
                                        self.fil.get_foff(), self.fil.nchans, tol)
        return self.dm_list

    def set_dm_list(self, dms):
        self.dm_list = np.ascontiguousarray(dms, np.float32)

    def get_dm_list(self):
        return self.dm_list

    def set_killmask(self, mask):
        m = np.ascontiguousarray(mask, np.int32)
        if m.size != self.fil.nchans:
            raise ValueError("killmask has %d entries for %d channels" % (m.size, self.fil.nchans))
        self.killmask = m

    def max_delay(self):
        t = delay_table(self.fil.nchans, self.fil.get_tsamp(), self.fil.get_fch1(), self.fil.get_foff())
        return int(P.lib().psx_dedisp_max_delay(self.dm_list.ctypes.data, self.dm_list.size, t.ctypes.data,
                                                t.size))

    def dedisperse(self, start=0, end=None, stream=None):
        """Dispersion trials of samples [start, end): (DeviceArray [ndm, nout], nout)."""
        end = self.fil.nsamps if end is None else end
        packed, n = self.fil.load(start, end - start)
        nout = n - self.max_delay()
        if nout <= 0:
            raise ValueError("%d samples do not cover the dispersion delay" % n)
        d_in = P.DeviceArray.from_numpy(packed)
        out = P.DeviceArray((self.dm_list.size, nout), np.float32)
        work = P.DeviceArray(int(P.lib().psx_dedisp_work_bytes(n, self.fil.nchans, self.dm_list.size)), np.uint8)
        km = None if self.killmask is None else self.killmask.ctypes.data
        P.check(P.lib().psx_dedisperse(d_in.ptr, n, self.fil.nchans, self.fil.nbits, self.dm_list.ctypes.data,
                                       self.dm_list.size, km, self.fil.get_tsamp(), self.fil.get_fch1(),
                                       self.fil.get_foff(), out.ptr, nout, work.ptr, stream))
        return out, nout
