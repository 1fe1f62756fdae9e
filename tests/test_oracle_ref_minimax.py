"""Minimax grids and transform matrices against the reference's own greenX (CPU only).

``oracle/_ref/ref_minimax`` (built by ``make -C oracle`` where the reference tree is present) runs
greenX's ``gx_minimax_grid`` (thirdparty/greenX/GX-TimeFrequency/src/minimax_grids.F90:53-237),
compiled from its sources, as LibRPA calls it (src/core/timefreq.cpp:193-219 through
src/interface/gx_minimax_wrp.f90:94-180). For every one of the 15 sizes and several energy
ranges, and for the four regression cases' own ranges:

* the tau / omega nodes and weights of the oracle and of the library equal greenX's (1e-14);
* the cosine / sine transform matrices equal greenX's up to what the least-squares problem's
  conditioning allows. greenX takes the plain pseudo-inverse V S^-1 U^T psi of each fit
  (minimax_grids.F90:366-395) with no cut of small singular values, so a fit matrix with
  condition number kappa fixes the weights only to about eps * kappa. Where eps * kappa < 1e-2
  the matrices must agree to max(2e-13, eps * kappa); the differences seen are below
  0.1 eps * kappa;
* everywhere, each transform reproduces greenX's target functions on greenX's x nodes as well as
  greenX's own matrices do (this residual is the max_errors greenX reports): at most twice
  greenX's residual for the well-posed fits. In the numerically singular fits (eps * kappa >=
  1e-2: large n on a narrow energy range) the weights reach 1e8 and beyond in greenX too, both
  residuals are set by rounding, and they differ by up to 3.4x; the bound there is 4x.

A planted mistake in the fit's x-node count fails the matrix comparison. With the Tikhonov
regularisation LibRPA passes on (minimax_regulation), the fits that are singular without it agree
entry by entry (1e-8 at 1e-6, 1e-5 at 1e-8).
"""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import tfgrids

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_minimax")
GOLD = json.load(open(os.path.join(HERE, "golden", "minimax_goldens.json")))
EPS = np.finfo(float).eps

pytestmark = pytest.mark.skipif(
    not os.access(REF, os.X_OK),
    reason="oracle/_ref/ref_minimax is not built (make -C oracle needs the reference tree)")

SIZES = (6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34)
RANGES = ((0.35, 60.0), (0.05, 300.0), (0.5, 8.0), (0.25, 3.6))
KINDS = ("cos_t2f", "cos_f2t", "sin_t2f")


def _gold_range(r):
    return tfgrids.e_min_max(np.array(r["eps"]), np.array(r["occ"]), r["nspin"], r["nk"])


CASES = [(n, a, b) for n in SIZES for (a, b) in RANGES] + \
        [(GOLD[k]["n"],) + tuple(_gold_range(GOLD[k])) for k in sorted(GOLD)]


def _run_greenx(cases, tmp, reg=None):
    fin, fout = os.path.join(tmp, "mm_in.bin"), os.path.join(tmp, "mm_out.bin")
    regs = reg if reg is not None else [0.0] * len(cases)
    with open(fin, "wb") as f:
        f.write(struct.pack("<i", len(cases)))
        for (n, a, b), r in zip(cases, regs):
            f.write(struct.pack("<iddd", n, a, b, r))
    env = dict(os.environ, MKL_NUM_THREADS="2", OMP_NUM_THREADS="2")
    subprocess.run([REF, fin, fout], check=True, timeout=300, env=env,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    raw, off, out = open(fout, "rb").read(), 0, []
    for (n, a, b) in cases:
        (ierr,) = struct.unpack_from("<i", raw, off)
        off += 4
        assert ierr == 0, (n, a, b, ierr)

        def take(k):
            nonlocal off
            v = np.frombuffer(raw, "<f8", k, off).copy()
            off += 8 * k
            return v
        g = dict(tau=take(n), tau_w=take(n), omega=take(n), omega_w=take(n))
        for k in KINDS:   # Fortran order: element (i, j) is row i of LibRPA's matrix
            g[k] = take(n * n).reshape(n, n, order="F")
        g["max_errors"] = take(3)
        take(1)
        out.append(g)
    assert off == len(raw)
    return out


@pytest.fixture(scope="module")
def greenx(tmp_path_factory):
    return dict(zip(CASES, _run_greenx(CASES, str(tmp_path_factory.mktemp("greenx")))))


def _x_nodes(n, emin, emax):
    """greenX's log-spaced fit nodes (minimax_grids.F90:333-357)."""
    nx = max((int(np.log10(emax / emin)) + 1) * 200, n)
    return emin * ((emax / emin) ** (1.0 / (nx - 1.0))) ** np.arange(nx)


def _targets(g, x):
    t, w = g["tau"], g["omega"]
    e = np.exp(-np.outer(x, t))                       # (x, tau)
    lor = 2 * x[:, None] / (x[:, None] ** 2 + w[None, :] ** 2)
    return {"cos_t2f": (e, lor), "cos_f2t": (lor, e),
            "sin_t2f": (e, 2 * w[None, :] / (x[:, None] ** 2 + w[None, :] ** 2))}


def _residuals(g, emin, emax):
    """max_x,i |sum_j M_ij f_j(x) - psi_i(x)| for each transform M."""
    x = _x_nodes(len(g["tau"]), emin, emax)
    return {k: float(np.abs(basis @ g[k].T - psi).max()) for k, (basis, psi) in _targets(g, x).items()}


_KAPPA = {}


def _kappas(g, emin, emax):
    """Largest condition number over the rows of each least-squares fit (the matrix A of
    calculate_psi_and_mat_A, minimax_grids.F90:420-500), from greenX's nodes ``g``."""
    key = (len(g["tau"]), emin, emax)
    if key not in _KAPPA:
        _KAPPA[key] = _kappas_uncached(g, emin, emax)
    return _KAPPA[key]


def _kappas_uncached(g, emin, emax):
    x = _x_nodes(len(g["tau"]), emin, emax)
    t, w = g["tau"], g["omega"]
    out = {}
    for k in KINDS:
        worst = 0.0
        for i in range(len(t)):
            if k == "cos_t2f":
                A = np.cos(w[i] * t)[None, :] * np.exp(-np.outer(x, t))
            elif k == "cos_f2t":
                A = np.cos(t[i] * w)[None, :] * 2 * x[:, None] / (x[:, None] ** 2 + w[None, :] ** 2)
            else:
                A = np.sin(w[i] * t)[None, :] * np.exp(-np.outer(x, t))
            s = np.linalg.svd(A, compute_uv=False)
            worst = max(worst, np.inf if s[-1] == 0 else s[0] / s[-1])
        out[k] = worst
    return out


def _check(mine, ref, emin, emax, label):
    for k in ("tau", "tau_w", "omega", "omega_w"):
        assert np.abs(mine[k] - ref[k]).max() <= 1e-14 * np.abs(ref[k]).max(), (label, k)
    kap = _kappas(ref, emin, emax)
    res_mine, res_ref = _residuals(mine, emin, emax), _residuals(ref, emin, emax)
    compared = 0
    for k in KINDS:
        bound = EPS * kap[k]
        factor = 2.0 if bound < 1e-2 else 4.0
        assert res_mine[k] <= factor * res_ref[k] + 1e-14, (label, k, res_mine[k], res_ref[k])
        if bound < 1e-2:
            d = np.abs(mine[k] - ref[k]).max() / np.abs(ref[k]).max()
            assert d <= max(2e-13, bound), (label, k, d, bound)
            compared += 1
    return compared


def test_greenx_residual_metric_is_its_max_errors(greenx):
    # the same quantity as greenX's calculate_max_error (minimax_grids.F90:540-640); in the
    # singular fits the sum cancels and the two roundings differ by up to about 10 %
    for (n, a, b), g in greenx.items():
        r = _residuals(g, a, b)
        assert np.allclose([r[k] for k in KINDS], g["max_errors"], rtol=0.25, atol=1e-15), (n, a, b)


def test_oracle_grids_equal_greenx(greenx):
    compared = 0
    for (n, a, b), ref in greenx.items():
        compared += _check(tfgrids.minimax(n, a, b), ref, a, b, (n, a, b))
    # more than half of the fits are well posed and compared entry by entry
    assert compared >= 0.5 * 3 * len(greenx)


def test_matrix_comparison_detects_planted_x_nodes(greenx, monkeypatch):
    # the fit on 200 fewer x nodes (int(log10(range)) * 200): a plausible off-by-one
    orig = tfgrids._fit

    def fewer_nodes(n, tau, omega, e_min, e_max, ttype, reg=0.0):
        return orig(n, tau, omega, e_min, e_min * (e_max / e_min) ** 0.999, ttype, reg)
    monkeypatch.setattr(tfgrids, "_fit", fewer_nodes)
    n, a, b = 16, 0.05, 300.0
    with pytest.raises(AssertionError):
        _check(tfgrids.minimax(n, a, b), greenx[(n, a, b)], a, b, (n, a, b))


def _library_grid(n, emin, emax, reg=0.0):
    import librpa_amd as L
    # one k, six bands: gap emin across the Fermi level, band width emax
    eps = np.array([-0.6, -emin / 2 - 0.05, -emin / 2, emin / 2, 1.0, -0.6 + emax])
    assert np.all(np.diff(eps) > 0)
    occ = np.array([2.0, 2.0, 2.0, 0.0, 0.0, 0.0])
    h = L.Handler()
    h.set_scf_dimension(1, 1, 6, 6)
    h.set_wg_ekb_efermi(1, 1, 6, occ, eps, 0.0)
    return h.get_tfgrids(L.Options(nfreq=n, tfgrids_type=L.MINIMAX, minimax_regulation=reg))


LIBRARY_RANGES = RANGES[:1] + RANGES[2:3]   # a wide and a narrow range: every size, well posed and singular


def test_library_grids_equal_greenx(greenx, lib):
    import librpa_amd as L
    compared = 0
    for (n, a, b), ref in greenx.items():
        if (a, b) not in LIBRARY_RANGES:
            continue
        g = _library_grid(n, a, b)
        compared += _check(g, ref, a, b, ("library", n, a, b))
    assert compared >= 0.5 * 3 * len(SIZES) * len(LIBRARY_RANGES)
    # the regression cases' own eigenvalues
    for key in sorted(GOLD):
        r = GOLD[key]
        eps, occ = np.array(r["eps"]), np.array(r["occ"])
        ns, nk, nb = eps.shape
        h = L.Handler()
        h.set_scf_dimension(ns, nk, nb, nb)
        h.set_wg_ekb_efermi(ns, nk, nb, occ.reshape(-1), eps.reshape(-1), r["efermi"])
        g = h.get_tfgrids(L.Options(nfreq=r["n"], tfgrids_type=L.MINIMAX))
        a, b = _gold_range(r)
        _check(g, greenx[(r["n"], a, b)], a, b, ("library", key))


# Tikhonov-regularised fits (minimax_regulation, passed to greenX by LibRPA's
# src/api/dataset_helper.cpp:20-28): the filter s / (reg^2 + s^2) bounds the weights, so the
# fits that are singular without it agree entry by entry; 1e-12 and the 32-point H2O range are
# the reference's own unit test (src/test/test_timefreq.cpp:124-137), still singular there.
EMIN_H2O = 0.302066195818780992e-2 + 0.260399583086844855e0
EMAX_H2O = 0.322960860707412234e1 + 0.187886173733204913e2
REG_CASES = [(32, EMIN_H2O, EMAX_H2O, 1e-6), (34, 0.25, 3.6, 1e-6), (24, 0.35, 60.0, 1e-6),
             (32, EMIN_H2O, EMAX_H2O, 1e-8), (16, 0.5, 8.0, 1e-8), (32, EMIN_H2O, EMAX_H2O, 1e-12)]
REG_TOL = {1e-6: 1e-8, 1e-8: 1e-5}


def test_regularised_fits_equal_greenx(tmp_path, lib):
    refs = _run_greenx([c[:3] for c in REG_CASES], str(tmp_path), reg=[c[3] for c in REG_CASES])
    for (n, a, b, reg), ref in zip(REG_CASES, refs):
        for who, mine in (("oracle", tfgrids.minimax(n, a, b, reg)), ("library", _library_grid(n, a, b, reg))):
            for k in ("tau", "tau_w", "omega", "omega_w"):
                assert np.abs(mine[k] - ref[k]).max() <= 1e-14 * np.abs(ref[k]).max(), (who, n, reg, k)
            res_mine, res_ref = _residuals(mine, a, b), _residuals(ref, a, b)
            for k in KINDS:
                assert res_mine[k] <= 4.0 * res_ref[k] + 1e-14, (who, n, reg, k, res_mine[k], res_ref[k])
                if reg in REG_TOL:
                    d = np.abs(mine[k] - ref[k]).max() / np.abs(ref[k]).max()
                    assert d <= REG_TOL[reg], (who, n, reg, k, d)
