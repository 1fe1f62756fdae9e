"""The dielectric function on the grid frequencies against the reference's own code (CPU only).

``oracle/_ref/ref_dielec`` (built by ``make -C oracle`` where the reference tree is present) runs
LibRPA's cubic spline (src/math/interpolate.cpp) and Levenberg-Marquardt fit (src/math/fitting.cpp)
compiled from its sources, as interpolate_dielec_func (src/core/dielecmodel.cpp:65-104) calls them
for option_dielect_func 1 and 2. The oracle (``oracle/dielec.py``) and the library's host code
(``librpa_amd_interpolate_dielec_func``) must give the same values on minimax frequencies,
including those outside the input range. The arithmetic is done in the same order, so the
comparisons are to 1e-13.

What the mathematics fixes: the spline passes through the data with continuous value and second
derivative, and the fit reproduces data of the form 1 + A / (B + u^2), which the model with two
equal terms holds exactly. A textbook natural spline is not the reference's: it must fail the
comparison.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import dielec, tfgrids

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "oracle", "_ref", "ref_dielec")

pytestmark = pytest.mark.skipif(
    not os.access(REF, os.X_OK),
    reason="oracle/_ref/ref_dielec is not built (make -C oracle needs the reference tree)")


def _run_ref(cases, tmp_path):
    fin, fout = str(tmp_path / "dielec_in.bin"), str(tmp_path / "dielec_out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack("<i", len(cases)))
        for opt, xs, ys, xn in cases:
            f.write(struct.pack("<3i", opt, len(xs), len(xn)))
            for a in (xs, ys, xn):
                f.write(np.asarray(a, dtype="<f8").tobytes())
    subprocess.run([REF, fin, fout], check=True, timeout=120,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    raw, off, out = open(fout, "rb").read(), 0, []
    for opt, xs, ys, xn in cases:
        v = np.frombuffer(raw, "<f8", len(xn), off).copy()
        off += 8 * len(xn)
        pars = None
        if opt == 2:
            pars = np.frombuffer(raw, "<f8", 8, off).copy()
            off += 64
        out.append((v, pars))
    assert off == len(raw)
    return out


def _grid(n, emin, emax):
    return tfgrids.minimax(n, emin, emax)["omega"]


# eps_M(i u) data on a caller's own frequencies: a Drude-Lorentz-like insulator, data from the
# model itself, and a slowly converging one; targets are minimax frequencies that reach below
# and above the data
XS = np.geomspace(0.02, 20.0, 30)
TRUE_DHN = [3.2, 0.9, 1.1, 2.0, 1.8, 0.8, 1.0, 0.15]
DATA = {
    "lorentz": 1 + 8.0 / (0.5 + XS ** 2),
    "dhn": np.array([dielec.dhn(x, TRUE_DHN) for x in XS]),
    "slow": 1 + 5.0 / np.sqrt(1 + XS),
}
TARGETS = {16: _grid(16, 0.1, 30.0), 24: _grid(24, 0.05, 100.0), 6: _grid(6, 0.3, 10.0)}


def _cases():
    out = []
    for opt in (1, 2):
        for name, ys in DATA.items():
            for n, tg in TARGETS.items():
                out.append((opt, XS, ys, tg))
    # short and uneven inputs
    xs = np.array([0.05, 0.3, 0.31, 2.0, 9.0])
    out.append((1, xs, 1 + 3.0 / (1 + xs ** 2), TARGETS[16]))
    out.append((1, xs[:2], np.array([4.0, 2.5]), TARGETS[6]))
    out.append((2, xs, 1 + 3.0 / (1 + xs ** 2), TARGETS[16]))
    return out


@pytest.fixture(scope="module")
def reference(tmp_path_factory):
    cases = _cases()
    return cases, _run_ref(cases, tmp_path_factory.mktemp("dielec"))


def test_oracle_equals_reference(reference):
    cases, ref = reference
    for i, ((opt, xs, ys, tg), (rv, rp)) in enumerate(zip(cases, ref)):
        mine = np.array(dielec.interpolate(opt, xs, ys, tg))
        assert np.abs(mine - rv).max() <= 1e-13 * np.abs(rv).max(), (i, opt, np.abs(mine - rv).max())
        if opt == 2:
            p0 = [1.0] * 8
            p0[0] = p0[4] = ys[0]
            p = dielec.levmarq_fit(p0, list(xs), list(ys), dielec.dhn, dielec.dhn_grad)
            assert np.abs(np.array(p) - rp).max() <= 1e-12 * np.abs(rp).max(), (i, p, rp)


def test_library_equals_reference(reference, lib):
    import librpa_amd as L
    cases, ref = reference
    for i, ((opt, xs, ys, tg), (rv, _)) in enumerate(zip(cases, ref)):
        got = L.interpolate_dielec_func(opt, xs, ys, tg)
        assert np.abs(got - rv).max() <= 1e-13 * np.abs(rv).max(), (i, opt, np.abs(got - rv).max())
    # option 0 copies; the head-and-wing model (3) is not in this build
    tg = TARGETS[6]
    vals = np.linspace(5.0, 1.2, len(tg))
    assert np.array_equal(L.interpolate_dielec_func(0, tg, vals, tg), vals)
    with pytest.raises(L.LibrpaError):
        L.interpolate_dielec_func(3, XS, DATA["lorentz"], tg)
    with pytest.raises(L.LibrpaError):
        L.interpolate_dielec_func(0, XS, DATA["lorentz"], tg)


def test_what_the_mathematics_fixes():
    # Whatever its tridiagonal solve gives, the spline's pieces pass through the data and join
    # with continuous value and second derivative (b and d are built from c to make it so).
    # Its first derivative jumps at the knots, so it does not reproduce even a straight line:
    xs = [0.1, 0.4, 0.5, 1.3, 2.0, 3.7]
    ys = [2.0 - 0.3 * x for x in xs]
    assert np.abs(np.array(dielec.spline(xs, ys, xs)) - ys).max() <= 1e-15
    eps = 1e-7
    for x in xs[1:-1]:
        lo, mid, hi = dielec.spline(xs, ys, [x - eps, x, x + eps])
        assert abs(lo - mid) < 1e-6 and abs(hi - mid) < 1e-6
    assert abs(dielec.spline(xs, ys, [0.2])[0] - (2.0 - 0.3 * 0.2)) > 1e-3
    # the model with two equal terms holds 1 + A / (B + u^2) exactly: 2 (p0 - 1) = A, p3^2 = 1 / B
    got = dielec.interpolate(2, XS, DATA["lorentz"], TARGETS[16])
    want = 1 + 8.0 / (0.5 + TARGETS[16] ** 2)
    assert np.abs(np.array(got) - want).max() <= 1e-6 * want.max()


def test_comparison_detects_a_textbook_spline(reference):
    # scipy's natural cubic spline is the textbook algorithm, not LibRPA's
    from scipy.interpolate import CubicSpline
    cases, ref = reference
    opt, xs, ys, tg = cases[0]
    inside = (tg >= xs[0]) & (tg <= xs[-1])
    textbook = CubicSpline(xs, ys, bc_type="natural")(tg[inside])
    assert np.abs(textbook - ref[0][0][inside]).max() > 1e-4 * np.abs(ref[0][0]).max()
