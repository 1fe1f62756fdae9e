"""Minimax grid pins: LibRPA's printed grids for four regression cases.

Fixture: tests/golden/minimax_goldens.json (made by make_minimax_goldens.py from
regression_tests/refs/*/librpa.out and the cases' band_out files).
"""
import json
import os

import numpy as np
import pytest

from oracle import tfgrids

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "minimax_goldens.json")))


def _rel(a, b, floor=0.0):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


@pytest.mark.parametrize("case", sorted(GOLD))
def test_oracle_nodes_match_reference(case):
    r = GOLD[case]
    emin, emax = tfgrids.e_min_max(np.array(r["eps"]), np.array(r["occ"]), r["nspin"], r["nk"])
    g = tfgrids.minimax(r["n"], emin, emax)
    for k in ("omega", "omega_w", "tau", "tau_w"):
        assert _rel(g[k], r[k]) < 1e-12, k


@pytest.mark.parametrize("case", ["h2o_rpa", "mgo_g0w0", "li_g0w0"])
def test_oracle_transforms_match_reference(case):
    # printed with 6 significant digits; n = 32 (h2_g0w0) is numerically singular in the reference too
    r = GOLD[case]
    emin, emax = tfgrids.e_min_max(np.array(r["eps"]), np.array(r["occ"]), r["nspin"], r["nk"])
    g = tfgrids.minimax(r["n"], emin, emax)
    tol = 2e-4 if case == "li_g0w0" else 1e-5
    for k in ("cos_t2f", "sin_t2f", "cos_f2t"):
        assert _rel(g[k], r[k], 1e-3) < tol, k


def test_bin_selection_and_ratio_heuristic():
    # n = 32 in the first bin uses the e_ratio rescaling (minimax_utils.F90:74-80)
    r = GOLD["h2_g0w0"]
    emin, emax = tfgrids.e_min_max(np.array(r["eps"]), np.array(r["occ"]), r["nspin"], r["nk"])
    g = tfgrids.minimax(32, emin, emax)
    assert _rel(g["tau"], r["tau"]) < 1e-12
    # a plausible mistake (no rescaling) is caught
    edges, cols = tfgrids.load_table()[("tau", 32)]
    raw = np.array(cols[0][:32]) / (2 * emin)
    assert _rel(raw, r["tau"]) > 1e-3


@pytest.mark.parametrize("case", sorted(GOLD))
def test_library_grid_matches_reference(case, lib):
    import librpa_amd as L
    r = GOLD[case]
    eps, occ = np.array(r["eps"]), np.array(r["occ"])
    ns, nk, nb = eps.shape
    h = L.Handler()
    h.set_scf_dimension(ns, nk, nb, nb)
    h.set_wg_ekb_efermi(ns, nk, nb, occ.reshape(-1), eps.reshape(-1), r["efermi"])
    g = h.get_tfgrids(L.Options(nfreq=r["n"], tfgrids_type=L.MINIMAX))
    for k in ("omega", "omega_w", "tau", "tau_w"):
        assert _rel(g[k], r[k]) < 1e-12, k
    if r["n"] <= 16:
        tol = 2e-4 if case == "li_g0w0" else 1e-5
        for k in ("cos_t2f", "sin_t2f", "cos_f2t"):
            assert _rel(g[k], r[k], 1e-3) < tol, k
    om, w = h.get_imaginary_frequency_grids(L.Options(nfreq=r["n"], tfgrids_type=L.MINIMAX))
    assert _rel(om, r["omega"]) < 1e-12 and _rel(w, r["omega_w"]) < 1e-12


# ---------------------------------------------------------------------------------------------
# Every greenX size (6, 8, ..., 34; minimax_tau.F90:25-26) against what the mathematics fixes.
# The tau grid is a minimax fit of 1/x by exponentials (minimax_tau.F90:6-7), so after the
# scaling tau = a / (2 e_min) (minimax_grids.F90:142-145) sum_j w_j exp(-2 E tau_j) = 1 / (2E)
# on [e_min, e_max]; the omega grid integrates (4E/pi) E^2 / (E^2 + w^2)^2 dw = 1.  The
# transforms are least-squares fits (minimax_grids.F90:286-406) of the exact Fourier pairs
#   int cos(w t) exp(-E|t|) dt = 2E / (E^2 + w^2),  int sin(w t) exp(-E|t|) sgn(t) dt = 2w / (E^2 + w^2),
# so each row must reproduce its target to the fit error on the whole energy range.
# ---------------------------------------------------------------------------------------------
SIZES = list(range(6, 36, 2))
# bounds on the relative quadrature error for e_max / e_min = 14.4 and 300 (the minimax error
# falls roughly tenfold per two points until it saturates near 1e-11)
QUAD_TOL = {14.4: {6: 1e-4, 8: 1e-6, 10: 3e-8, 12: 3e-9}, 300.0: {6: 0.1, 8: 2e-2, 10: 2e-3, 12: 2e-4,
                                                           14: 2e-5, 16: 2e-6, 18: 1e-7, 20: 1e-8}}


def _quad_tol(n, er):
    return QUAD_TOL[er].get(n, 1e-9)


# bounds on the transform fit errors (the least-squares fits stop improving near 1e-8)
FIT_TOL = {14.4: {6: 3e-4, 8: 3e-5, 10: 3e-6, 12: 3e-7}, 300.0: {6: 2e-2, 8: 3e-3, 10: 5e-4, 12: 5e-5,
                                                         14: 2e-5, 16: 1e-5, 18: 1e-6, 20: 3e-7}}


def _fit_tol(n, er):
    return FIT_TOL[er].get(n, 5e-7 if er < 100 else 1e-7)


def grid_identity_errors(g, emin, emax):
    x = np.geomspace(emin, emax, 1201)
    t, tw, om, ow = (np.asarray(g[k], dtype=float) for k in ("tau", "tau_w", "omega", "omega_w"))
    X, O = x[:, None], om[None, :]
    lor = 2 * X / (X ** 2 + O ** 2)
    return {
        "tau_quad": abs(2 * x * (np.exp(-2 * np.outer(x, t)) @ tw) - 1).max(),
        "omega_quad": abs(4 * x / np.pi * ((X ** 2 / (X ** 2 + O ** 2) ** 2) @ ow) - 1).max(),
        # absolute errors on the scale of each target's maximum (2 / e_min, 1 / e_min, 1)
        "cos_t2f": abs(np.exp(-np.outer(x, t)) @ np.asarray(g["cos_t2f"]).T - lor).max() * emin / 2,
        "sin_t2f": abs(np.exp(-np.outer(x, t)) @ np.asarray(g["sin_t2f"]).T
                       - 2 * O / (X ** 2 + O ** 2)).max() * emin,
        "cos_f2t": abs(lor @ np.asarray(g["cos_f2t"]).T - np.exp(-np.outer(x, t))).max(),
    }


def _check_identities(g, n, er, emin):
    e = grid_identity_errors(g, emin, emin * er)
    assert e["tau_quad"] < _quad_tol(n, er), e
    assert e["omega_quad"] < _quad_tol(n, er), e
    for k in ("cos_t2f", "cos_f2t"):
        assert e[k] < _fit_tol(n, er), (k, e)
    assert e["sin_t2f"] < 50 * _fit_tol(n, er), e


@pytest.mark.parametrize("er", [14.4, 300.0])
@pytest.mark.parametrize("n", SIZES)
def test_oracle_grid_identities_every_size(n, er):
    emin = 0.25
    _check_identities(tfgrids.minimax(n, emin, emin * er), n, er, emin)


def test_identities_catch_planted_grid_mistakes():
    emin, er, n = 0.25, 14.4, 24
    g = tfgrids.minimax(n, emin, emin * er)
    e = grid_identity_errors(g, emin, emin * er)
    # cosine matrix stored with tau rows instead of omega rows
    bad = dict(g, cos_t2f=np.asarray(g["cos_t2f"]).T)
    assert grid_identity_errors(bad, emin, emin * er)["cos_t2f"] > 1e3 * e["cos_t2f"]
    # tau nodes scaled by 1 / e_min instead of 1 / (2 e_min) (weights right); scaling nodes and
    # weights together leaves the 1 / x fit intact on a shifted range, so that is not a mistake here
    bad = dict(g, tau=2 * np.asarray(g["tau"]))
    assert grid_identity_errors(bad, emin, emin * er)["tau_quad"] > 0.1
    bad = dict(g, omega_w=0.5 * np.asarray(g["omega_w"]))
    assert grid_identity_errors(bad, emin, emin * er)["omega_quad"] > 0.1
    # sine and cosine t->f swapped
    bad = dict(g, cos_t2f=g["sin_t2f"])
    assert grid_identity_errors(bad, emin, emin * er)["cos_t2f"] > 1e-2
    # the bin below the selected one (a range too short for e_max / e_min) for the small sizes,
    # whose bins are far apart
    orig = tfgrids._coeffs

    def lower_bin(kind, nn, e_range):
        edges, cols = tfgrids.load_table()[(kind, nn)]
        c, r = orig(kind, nn, e_range)
        idx = next(i for i in range(len(cols)) if np.array_equal(cols[i], c))
        return np.array(cols[max(idx - 1, 0)]), r

    for nn in (6, 8, 10, 12, 14):
        good = grid_identity_errors(tfgrids.minimax(nn, emin, emin * er), emin, emin * er)["tau_quad"]
        tfgrids._coeffs = lower_bin
        try:
            wrong = grid_identity_errors(tfgrids.minimax(nn, emin, emin * er), emin, emin * er)["tau_quad"]
        finally:
            tfgrids._coeffs = orig
        assert wrong > 30 * good and wrong > _quad_tol(nn, er), (nn, good, wrong)


@pytest.mark.parametrize("n", SIZES)
def test_library_grid_every_size(n, lib):
    # the library's nodes equal the oracle's; its transforms (own SVD) satisfy the same identities
    import librpa_amd as L
    emin, er = 0.25, 14.4
    nb = 6
    eps = np.array([[[-0.6, -emin / 2 - 0.05, -emin / 2, emin / 2, 1.0, -0.6 + emin * er]]])
    occ = np.array([[[2.0, 2.0, 2.0, 0.0, 0.0, 0.0]]])
    h = L.Handler()
    h.set_scf_dimension(1, 1, nb, nb)
    h.set_wg_ekb_efermi(1, 1, nb, occ.reshape(-1), eps.reshape(-1), 0.0)
    g = h.get_tfgrids(L.Options(nfreq=n, tfgrids_type=L.MINIMAX))
    ref = tfgrids.minimax(n, emin, emin * er)
    for k in ("omega", "omega_w", "tau", "tau_w"):
        assert _rel(g[k], ref[k]) < 1e-12, k
    _check_identities(g, n, er, emin)
