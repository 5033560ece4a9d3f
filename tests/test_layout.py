"""Host-side bounds validation of every FFT launch geometry (no GPU needed).

psx_fft_layout_check builds the pass launches of a plan without device
memory and checks each pass's largest read and write index against its
buffer, for the plain R2C/C2R transform and for the fused batched search
(whose half-spectra are stored back to back with stride n/2)."""
import ctypes as C

import pytest

import peasoup_amd as P
from peasoup_amd import tuning

SIZES = [16, 32, 64, 96, 1024, 6144, 1 << 13, 1 << 14, 1 << 16, 1 << 17, 3 << 17, 1 << 20, 1 << 22,
         3 << 21, 1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27, 1 << 28, 3 << 23, 3 << 24, 3 << 25]


def check(n, kind, batch, dec=None):
    L = P.lib()
    L.psx_fft_layout_check.restype = C.c_int
    L.psx_fft_layout_check.argtypes = [C.c_uint64, C.c_int, C.c_uint32, C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]
    if dec:
        lens = (C.c_int * 3)(*[d[0] for d in dec])
        cols = (C.c_int * 3)(*[d[1] for d in dec])
        return L.psx_fft_layout_check(n, kind, batch, len(dec), lens, cols)
    return L.psx_fft_layout_check(n, kind, batch, 0, None, None)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("batch", [1, 3, 16, 64])
def test_default_plans_in_bounds(n, batch):
    for kind in (P.PSX_FFT_R2C, P.PSX_FFT_C2R):
        rc = check(n, kind, batch)
        assert rc == 0, (n, kind, batch, P.lib().psx_last_error_string())


@pytest.mark.parametrize("n", [1 << 17, 1 << 20, 1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26, 3 << 24])
def test_tuner_candidates_in_bounds(n):
    # the host tables of the longest lengths take a while to build: fewer batches there
    for dec in tuning.candidates(n, limit=100):
        for batch in ((1, 8, 16, 32) if n <= 1 << 24 else (1, 64)):
            assert check(n, P.PSX_FFT_R2C, batch, dec) == 0, (n, dec, batch, P.lib().psx_last_error_string())


def test_bad_decomposition_rejected():
    assert check(1 << 20, P.PSX_FFT_R2C, 1, [(1024, 4), (256, 16)]) != 0
