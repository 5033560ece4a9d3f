"""Birdie zapping (Zapper, include/transforms/birdiezapper.hpp; zap_birdies_kernel,
src/kernels.cu:1063-1084): oracle pinned by hand-computed bin ranges on exact
binary fractions, and the zap-list reader (CPU)."""
import numpy as np
import pytest


def _zapped(X):
    return np.flatnonzero((X.real == 1.0) & (X.imag == 0.0))


def test_zap_bin_ranges_by_hand(orc):
    nb = 300
    X = (np.arange(nb) + 2.0 + 3.0j).astype(np.complex64)
    # bin width 0.25 Hz: 50 +- 0.5 Hz -> [floor(198), ceil(202)) = 198..201
    assert list(_zapped(orc.zap_birdies(X, 0.25, [(50.0, 0.5)]))) == [198, 199, 200, 201]
    # 0.1 +- 0.5 Hz: floor(-1.6) = -2 raised to 0; ceil(2.4) = 3 -> 0..2
    assert list(_zapped(orc.zap_birdies(X, 0.25, [(0.1, 0.5)]))) == [0, 1, 2]
    # near the top: floor(297.6) = 297, ceil(301.6) = 302 lowered to nb-1 = 299 -> 297, 298
    assert list(_zapped(orc.zap_birdies(X, 0.25, [(74.9, 0.5)]))) == [297, 298]
    # starting at or past nb: nothing
    assert list(_zapped(orc.zap_birdies(X, 0.25, [(75.5, 0.25)]))) == []
    # zero width on a bin edge: [200, 200) is empty; off the edge one bin
    assert list(_zapped(orc.zap_birdies(X, 0.25, [(50.0, 0.0)]))) == []
    assert list(_zapped(orc.zap_birdies(X, 0.25, [(50.1, 0.0)]))) == [200]
    # several birdies, overlapping; the rest of the spectrum untouched
    Y = orc.zap_birdies(X, 0.25, [(50.0, 0.5), (50.25, 0.5), (10.0, 0.125)])
    z = list(_zapped(Y))
    assert z == [39, 40, 198, 199, 200, 201, 202]
    keep = np.setdiff1d(np.arange(nb), z)
    assert np.array_equal(Y[keep], X[keep])


def test_read_zaplist(tmp_path):
    import peasoup_amd.workloads as W
    p = tmp_path / "zap.txt"
    p.write_text("50.000 0.100\n\n100.00 0.15\n   \n150.00 0.15 extra\n")
    z = W.read_zaplist(str(p))
    assert z.shape == (3, 2) and np.allclose(z[:, 0], [50, 100, 150]) and np.allclose(z[:, 1], [0.1, 0.15, 0.15])
    p.write_text("50.0\n")
    with pytest.raises(ValueError):
        W.read_zaplist(str(p))
