"""The library's host-side dielectric function on the grid (librpa_amd_interpolate_dielec_func),
checked without the reference tree (CPU only; the comparison with LibRPA's own compiled routines
is tests/test_oracle_ref_dielec.py):

* option 0 returns the values as given, and needs one per target frequency;
* option 1 equals the oracle's restatement of LibRPA's spline and passes through the data;
* option 2 reproduces data the model holds exactly (two equal Havriliak-Negami terms:
  1 + A / (B + u^2)) and equals the oracle's restatement of the fit;
* option 3 (head and wings) and malformed inputs raise.
"""
import numpy as np
import pytest

from oracle import dielec


def test_dielectric_function_on_the_grid(lib):
    import librpa_amd as L
    xs = np.geomspace(0.02, 20.0, 24)
    ys = 1 + 8.0 / (0.5 + xs ** 2)
    tg = np.array([0.01, 0.1, 0.7, 3.0, 15.0, 40.0])
    assert np.array_equal(L.interpolate_dielec_func(0, tg, ys[:6], tg), ys[:6])
    s1 = L.interpolate_dielec_func(1, xs, ys, tg)
    assert np.abs(s1 - dielec.spline(list(xs), list(ys), list(tg))).max() <= 1e-13 * np.abs(s1).max()
    assert np.abs(L.interpolate_dielec_func(1, xs, ys, xs) - ys).max() <= 1e-13 * ys.max()
    s2 = L.interpolate_dielec_func(2, xs, ys, tg)
    assert np.abs(s2 - (1 + 8.0 / (0.5 + tg ** 2))).max() <= 1e-6 * s2.max()
    assert np.abs(s2 - np.array(dielec.interpolate(2, xs, ys, tg))).max() <= 1e-13 * s2.max()
    for bad in ((3, xs, ys, tg), (0, xs, ys, tg), (1, xs[:1], ys[:1], tg), (7, xs, ys, tg)):
        with pytest.raises(L.LibrpaError):
            L.interpolate_dielec_func(*bad)
