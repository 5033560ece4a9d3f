"""The C ABI library loads and exports every entry point include/psx.h declares
(no device calls: runs without a GPU)."""
import ctypes as C
import os
import re

import peasoup_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if hdr.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", hdr)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(psx_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(P.LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_error_reporting_without_device():
    L = P.lib()
    assert L.psx_version() >= 10000
    # argument validation happens before any device work
    assert L.psx_resample(None, None, 0, 0.0, 1.0, None) == 1
    assert b"bad arguments" in L.psx_last_error_string()
    assert L.psx_harmonic_sum(None, None, 3, 10, None) == 1


def test_fft_plan_rejects_unsupported_sizes_before_device_work():
    L = P.lib()
    p = C.c_void_p()
    assert L.psx_fft_plan_create(1000, 0, 1, C.byref(p)) == 1
    assert L.psx_fft_plan_create(1 << 29, 0, 1, C.byref(p)) == 1
    assert L.psx_fft_plan_create(4096, 7, 1, C.byref(p)) == 1


def test_server_rejects_bad_schedule_without_device():
    import ctypes as C
    import peasoup_amd as P
    from peasoup_amd.serve import Schedule, ServerConfig
    arr = (Schedule * 1)()
    arr[0].size = 1 << 20
    arr[0].fft_npass = 7
    cfg = ServerConfig(1 << 30, 16, 0, 1, 1, C.cast(arr, C.POINTER(Schedule)))
    h = C.c_void_p()
    assert P.lib().psx_server_create(C.byref(cfg), C.byref(h)) == 1
    assert b"bad schedule" in P.lib().psx_last_error_string()
