"""peasoup_amd - MI355X-native acceleration-search core (Python bindings).

The product is ``libpsx.so`` (HIP kernels for gfx950 behind the C ABI in
``include/psx.h``) plus the C++ transform classes in
``peasoup_amd/include/peasoup``.  This module is a thin ctypes layer over
the C ABI for tests, benchmarks and Python hosts.  It never falls back to a
CPU implementation: if the library is missing, importing it raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsx.so")

PSX_OK = 0
PSX_FFT_R2C, PSX_FFT_C2R = 0, 1
PSX_U8, PSX_I8, PSX_U32, PSX_F32 = 0, 1, 2, 3


class PsxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("psx error %d: %s" % (code, msg))
        self.code = code


class Pulsar(C.Structure):
    _fields_ = [("period", C.c_double), ("accel", C.c_double), ("snr", C.c_double),
                ("phase0", C.c_double), ("duty", C.c_double)]


_lib = None


def lib():
    """Load libpsx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libpsx.so not built at %s: run __graft_entry__.build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, u64, u32, i64, f32, f64, i32 = (C.c_void_p, C.c_uint64, C.c_uint32, C.c_int64, C.c_float,
                                        C.c_double, C.c_int)
    sigs = {
        "psx_last_error_string": (C.c_char_p, []),
        "psx_version": (i32, []),
        "psx_device_count": (i32, [C.POINTER(i32)]),
        "psx_set_device": (i32, [i32]),
        "psx_get_device": (i32, [C.POINTER(i32)]),
        "psx_device_info": (i32, [i32, C.c_char_p, C.c_size_t, C.POINTER(u64), C.POINTER(i32), C.POINTER(i32)]),
        "psx_malloc": (i32, [C.POINTER(vp), u64]),
        "psx_free": (i32, [vp]),
        "psx_host_malloc": (i32, [C.POINTER(vp), u64]),
        "psx_host_free": (i32, [vp]),
        "psx_memcpy_h2d": (i32, [vp, vp, u64, vp]),
        "psx_memcpy_d2h": (i32, [vp, vp, u64, vp]),
        "psx_memcpy_d2d": (i32, [vp, vp, u64, vp]),
        "psx_stream_create": (i32, [C.POINTER(vp)]),
        "psx_stream_destroy": (i32, [vp]),
        "psx_stream_synchronize": (i32, [vp]),
        "psx_device_synchronize": (i32, []),
        "psx_event_create": (i32, [C.POINTER(vp)]),
        "psx_event_destroy": (i32, [vp]),
        "psx_event_record": (i32, [vp, vp]),
        "psx_event_synchronize": (i32, [vp]),
        "psx_event_elapsed_ms": (i32, [vp, vp, C.POINTER(f32)]),
        "psx_mem_info": (i32, [C.POINTER(u64), C.POINTER(u64)]),
        "psx_resample": (i32, [vp, vp, u64, f32, f32, vp]),
        "psx_resample_v1": (i32, [vp, vp, u64, f32, f32, vp]),
        "psx_form_amplitude": (i32, [vp, vp, u64, vp]),
        "psx_form_interbin": (i32, [vp, vp, u64, vp]),
        "psx_median_scrunch5": (i32, [vp, u64, vp, vp]),
        "psx_linear_stretch": (i32, [vp, u64, vp, u64, vp]),
        "psx_divide_c_by_f": (i32, [vp, vp, u64, vp]),
        "psx_zap_birdies": (i32, [vp, vp, vp, u32, f32, u64, vp]),
        "psx_deredden_work_floats": (u64, [u64]),
        "psx_deredden_median": (i32, [vp, u64, f64, f32, f32, vp, vp, vp]),
        "psx_reduce_work_bytes": (u64, []),
        "psx_mean_rms": (i32, [vp, u64, u64, C.POINTER(f32), C.POINTER(f32), vp, vp]),
        "psx_normalise": (i32, [vp, f32, f32, u64, vp]),
        "psx_remove_baseline": (i32, [vp, u64, vp, vp]),
        "psx_fill": (i32, [vp, u64, u64, f32, vp]),
        "psx_convert_to_f32": (i32, [vp, i32, vp, u64, vp]),
        "psx_harmonic_sum": (i32, [vp, C.POINTER(vp), u32, u64, vp]),
        "psx_find_peaks_work_bytes": (u64, [u64]),
        "psx_find_peaks": (i32, [vp, i64, i64, f32, vp, vp, u32, vp, vp, vp]),
        "psx_fft_plan_create": (i32, [u64, i32, u32, C.POINTER(vp)]),
        "psx_fft_plan_create_decomp": (i32, [u64, i32, u32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(vp)]),
        "psx_fft_plan_destroy": (i32, [vp]),
        "psx_fft_exec": (i32, [vp, vp, vp, vp]),
        "psx_fft_plan_describe": (i32, [vp, C.c_char_p, C.c_size_t]),
        "psx_synth_work_bytes": (u64, [u64]),
        "psx_synth_series": (i32, [vp, u64, f64, u64, u32, f64, C.POINTER(Pulsar), i32, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    _register_extras(L)
    return L


_extra_sigs = []


def register_signatures(sigs):
    """Add ctypes signatures for entry points defined in other modules."""
    _extra_sigs.append(sigs)
    if _lib is not None:
        _register_extras(_lib)


def _register_extras(L):
    for sigs in _extra_sigs:
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args


def check(rc):
    if rc != PSX_OK:
        raise PsxError(rc, lib().psx_last_error_string().decode())
    return rc


def gpu_available():
    """True when a HIP device is visible (does not initialise a context)."""
    try:
        n = C.c_int(0)
        return lib().psx_device_count(C.byref(n)) == 0 and n.value > 0
    except (ImportError, OSError):
        return False


def device_info(dev=0):
    name = C.create_string_buffer(256)
    hbm, cus, arch = C.c_uint64(), C.c_int(), C.c_int()
    check(lib().psx_device_info(dev, name, 256, C.byref(hbm), C.byref(cus), C.byref(arch)))
    gfx = "gfx%x" % arch.value
    nm = name.value.decode() or "%s-%dCU" % (gfx, cus.value)
    return {"name": nm, "hbm_bytes": hbm.value, "cu_count": cus.value, "gfx": gfx}


class DeviceArray:
    """Owning device buffer with a numpy dtype/shape (allocated by libpsx)."""

    def __init__(self, shape, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().psx_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        check(lib().psx_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream))

    def numpy(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        check(lib().psx_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream))
        return out

    def offset(self, elems):
        return self.ptr + int(elems) * self.dtype.itemsize

    def view(self, elem_offset, shape):
        """Non-owning window of `shape` elements starting at `elem_offset`."""
        return DeviceView(self, elem_offset, shape)

    def free(self):
        if getattr(self, "ptr", None):
            lib().psx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView(DeviceArray):
    """Window into a DeviceArray; keeps its owner alive and frees nothing."""

    def __init__(self, owner, elem_offset, shape):
        self.owner = owner
        self.dtype = owner.dtype
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        if elem_offset < 0 or elem_offset + self.size > owner.size:
            raise ValueError("view [%d, %d) outside a buffer of %d elements" %
                             (elem_offset, elem_offset + self.size, owner.size))
        self.ptr = owner.offset(elem_offset)

    def free(self):
        self.ptr = None


def synchronize():
    check(lib().psx_device_synchronize())


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib().psx_stream_create(C.byref(s)))
        self.handle = s.value

    def synchronize(self):
        check(lib().psx_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().psx_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib().psx_event_create(C.byref(e)))
        self.handle = e.value

    def record(self, stream=None):
        check(lib().psx_event_record(self.handle, stream.handle if isinstance(stream, Stream) else stream))

    def synchronize(self):
        check(lib().psx_event_synchronize(self.handle))

    def elapsed_ms(self, later):
        ms = C.c_float()
        check(lib().psx_event_elapsed_ms(self.handle, later.handle, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().psx_event_destroy(self.handle)
        except Exception:
            pass


class FFTPlan:
    """R2C / C2R plan (CuFFTerR2C / CuFFTerC2R, ffter.hpp:54-110)."""

    def __init__(self, n, kind=PSX_FFT_R2C, batch=1, decomp=None):
        """decomp: optional [(length, column tile), ...] per pass (the
        autotuner's schedule form); None picks the default."""
        p = C.c_void_p()
        if decomp:
            lens = (C.c_int * 3)(*[int(L) for L, _ in decomp])
            cols = (C.c_int * 3)(*[int(c or 0) for _, c in decomp])
            check(lib().psx_fft_plan_create_decomp(n, kind, batch, len(decomp), lens, cols, C.byref(p)))
        else:
            check(lib().psx_fft_plan_create(n, kind, batch, C.byref(p)))
        self.handle = p.value
        self.n, self.kind, self.batch = n, kind, batch

    def execute(self, d_in, d_out, stream=None):
        check(lib().psx_fft_exec(self.handle, _ptr(d_in), _ptr(d_out), stream))

    def describe(self):
        b = C.create_string_buffer(256)
        check(lib().psx_fft_plan_describe(self.handle, b, 256))
        return b.value.decode()

    def __del__(self):
        try:
            lib().psx_fft_plan_destroy(self.handle)
        except Exception:
            pass


def _ptr(x):
    return x.ptr if isinstance(x, DeviceArray) else x
