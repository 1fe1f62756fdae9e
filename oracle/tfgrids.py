"""Minimax time/frequency grids (oracle restatement).

Follows LibRPA ``TFGrids::generate_minimax`` (src/core/timefreq.cpp:193-236)
which calls greenX ``gx_minimax_grid`` (GX-TimeFrequency/src/minimax_grids.F90:53-237):

* bin selection ``find_erange`` and the e_ratio heuristic for n > 20 in the
  first bin (minimax_utils.F90:38-86);
* scaling: omega = x * e_min, tau = x / (2 e_min) (minimax_grids.F90:130-147);
* transform weights by least squares on log-spaced x nodes
  (``get_transformation_weights``, minimax_grids.F90:286-406 and
  ``calculate_psi_and_mat_A`` :420-500), then multiplied by cos/sin(omega tau)
  (:182-206).

e_min is the band gap and e_max = eps_max - eps_min
(``MeanField::get_E_min_max``, src/core/meanfield.cpp:151-198).

Pins: tests/golden/h2o_minimax16.json holds the 16-point grid and cosine /
sine transform matrices LibRPA printed for the H2O regression case
(regression_tests/refs/rpa_aims_mole_H2O_libri/librpa/librpa.out:129-224)
together with that case's eigenvalues; test_timefreq.cpp:54-57 values (n=6).
tests/test_oracle_ref_minimax.py: against the reference's own greenX gx_minimax_grid
(oracle/ref_minimax.f90, compiled from the greenX sources) for all 15 sizes on four energy
ranges and the regression cases' ranges: nodes and weights to 1e-14 (they come out identical),
transform matrices to max(2e-13, eps * kappa) wherever the least-squares fit is well posed,
and fit residuals as small as greenX's everywhere.
"""
from __future__ import annotations

import os

import numpy as np

_TABLE = os.path.join(os.path.dirname(__file__), os.pardir, "librpa_amd", "data", "minimax_coeffs.txt")
_cache = None


def load_table(path=_TABLE):
    global _cache
    if _cache is not None and path == _TABLE:
        return _cache
    tab = {}
    with open(path) as f:
        lines = [l for l in f if not l.startswith('#')]
    i = 0
    while i < len(lines):
        kind, n, nb = lines[i].split()
        n, nb = int(n), int(nb)
        edges = np.array([float(x) for x in lines[i + 1].split()])
        cols = [np.array([float(x) for x in lines[i + 2 + c].split()]) for c in range(nb + 1)]
        tab[(kind, n)] = (edges, cols)
        i += 2 + nb + 1
    if path == _TABLE:
        _cache = tab
    return tab


def _coeffs(kind, n, e_range):
    edges, cols = load_table()[(kind, n)]
    # find_erange: smallest edge strictly greater than e_range, else the overflow column
    idx = len(edges)
    best = np.inf
    for j, e in enumerate(edges):
        if e_range < e and e < best:
            idx, best = j, e
    e_ratio = 1.0
    if idx == 0 and n > 20:
        e_ratio = edges[0] / e_range
        if e_ratio > 1.5:
            e_ratio = e_ratio / 1.5
    return np.array(cols[idx]), e_ratio


def _fit(n, tau, omega, e_min, e_max, ttype, reg=0.0):
    nx = (int(np.log10(e_max / e_min)) + 1) * 200
    nx = max(nx, n)
    fac = (e_max / e_min) ** (1.0 / (nx - 1.0))
    x = e_min * fac ** np.arange(nx)
    W = np.empty((n, n))
    for ip in range(n):
        if ttype == 1:      # cos t -> w
            om = omega[ip]
            psi = 2.0 * x / (x ** 2 + om ** 2)
            A = np.cos(om * tau)[None, :] * np.exp(-np.outer(x, tau))
        elif ttype == 2:    # cos w -> t
            t = tau[ip]
            psi = np.exp(-x * t)
            A = np.cos(t * omega)[None, :] * 2.0 * x[:, None] / (x[:, None] ** 2 + omega[None, :] ** 2)
        elif ttype == 3:    # sin t -> w
            om = omega[ip]
            psi = 2.0 * om / (x ** 2 + om ** 2)
            A = np.sin(om * tau)[None, :] * np.exp(-np.outer(x, tau))
        else:               # sin w -> t
            t = tau[ip]
            psi = np.exp(-x * t)
            A = np.sin(t * omega)[None, :] * 2.0 * omega[None, :] / (x[:, None] ** 2 + omega[None, :] ** 2)
        U, s, Vt = np.linalg.svd(A, full_matrices=False)
        W[ip] = Vt.T @ ((s / (reg ** 2 + s ** 2)) * (U.T @ psi))
    return W


def minimax(n, e_min, e_max, reg=0.0):
    """Return dict with tau, tau_w, omega, omega_w, cos_t2f (omega x tau),
    cos_f2t (tau x omega), sin_t2f (omega x tau)."""
    e_range = e_max / e_min
    xo, r = _coeffs('omega', n, e_range)
    xo = xo / r
    omega, omega_w = xo[:n] * e_min, xo[n:] * e_min
    xt, r = _coeffs('tau', n, e_range)
    xt = xt * r
    tau, tau_w = xt[:n] / (2 * e_min), xt[n:] / (2 * e_min)
    ct2f = _fit(n, tau, omega, e_min, e_max, 1, reg) * np.cos(np.outer(omega, tau))
    cf2t = _fit(n, tau, omega, e_min, e_max, 2, reg) * np.cos(np.outer(tau, omega))
    st2f = _fit(n, tau, omega, e_min, e_max, 3, reg) * np.sin(np.outer(omega, tau))
    return dict(tau=tau, tau_w=tau_w, omega=omega, omega_w=omega_w, cos_t2f=ct2f, cos_f2t=cf2t,
                sin_t2f=st2f)


def e_min_max(eps, occ, nspin, nk, nspinor=1):
    """MeanField::get_E_min_max (src/core/meanfield.cpp:151-198); occ = API occupations (0..2)."""
    wg = occ / nk
    midpoint = 1.0 / (nspin * nk * nspinor)
    homo, lumo = -1e6, 1e6
    for s in range(eps.shape[0]):
        for k in range(eps.shape[1]):
            hl = -1
            for b in range(eps.shape[2]):
                if wg[s, k, b] >= midpoint:
                    hl = b
            lumo = min(lumo, eps[s, k, hl + 1])
            if hl != -1:
                homo = max(homo, eps[s, k, hl])
    lb = eps[:, :, 0].min()
    ub = max(np.finfo(float).tiny, eps[:, :, -1].max())
    return lumo - homo, ub - lb
