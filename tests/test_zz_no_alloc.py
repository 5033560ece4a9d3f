"""psx_deredden_median_noalloc (the running median with its scratch in a
static device array) returns exactly what psx_deredden_median returns with a
caller workspace, call after call on one device (the second call smaller
than the first, so stale scratch contents of a larger call cannot leak into
a smaller one), and rejects lengths beyond its scratch on the host."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind(P):
    L = P.lib()
    L.psx_deredden_median_noalloc.restype = C.c_int
    L.psx_deredden_median_noalloc.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_float, C.c_float,
                                              C.c_void_p, C.c_void_p]
    L.psx_deredden_noalloc_max_bins.restype = C.c_uint64
    L.psx_deredden_noalloc_max_bins.argtypes = []
    return L


def _spectrum(nb, seed):
    rng = np.random.default_rng(seed)
    # red-noise-like amplitude spectrum: a falling baseline times |N(0,1)|
    k = np.arange(nb, dtype=np.float64)
    return (np.abs(rng.standard_normal(nb)) * (1.0 + 50.0 / (1.0 + k / 200.0))).astype(np.float32)


@pytest.mark.parametrize("sizes", [((1 << 26, 64e-6), (1 << 23, 76e-6), (1 << 17, 64e-6))])
def test_noalloc_equals_workspace_entry_point(psx, sizes):
    L = _bind(psx)
    assert L.psx_deredden_noalloc_max_bins() == (1 << 25) + 1
    for call, (n, tsamp) in enumerate(sizes):
        nb = n // 2 + 1
        print("noalloc call %d: n=%d nb=%d" % (call, n, nb), flush=True)
        A = _spectrum(nb, 100 + call)
        bw = float(np.float32(1.0 / np.float32(np.float32(n) * np.float32(tsamp))))
        d_A = psx.DeviceArray.from_numpy(A)
        d_ref = psx.DeviceArray(nb)
        d_got = psx.DeviceArray(nb)
        d_w = psx.DeviceArray(psx.lib().psx_deredden_work_floats(nb))
        psx.check(psx.lib().psx_deredden_median(d_A.ptr, nb, bw, 0.05, 0.5, d_ref.ptr, d_w.ptr, None))
        psx.check(L.psx_deredden_median_noalloc(d_A.ptr, nb, bw, 0.05, 0.5, d_got.ptr, None))
        psx.synchronize()
        ref, got = d_ref.numpy(), d_got.numpy()
        assert np.array_equal(got, ref), (n, int(np.count_nonzero(got != ref)))
        assert np.all(np.isfinite(got)) and np.count_nonzero(got) > nb // 2


def test_noalloc_rejects_lengths_beyond_its_scratch(psx):
    L = _bind(psx)
    d = psx.DeviceArray(64)
    too_many = int(L.psx_deredden_noalloc_max_bins()) + 1
    rc = L.psx_deredden_median_noalloc(d.ptr, too_many, 0.01, 0.05, 0.5, d.ptr, None)
    assert rc == 1  # PSX_ERR_ARG
    assert "exceed the static scratch" in psx.lib().psx_last_error_string().decode()
