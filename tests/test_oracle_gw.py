"""Pins for the G0W0 oracle (oracle/gw.py): closed forms and independent restatements."""
import numpy as np
import pytest

from librpa_amd import synth
from oracle import gw, rpa, tfgrids


def _tiny(case):
    if case == "k2":
        return synth.tiny(natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0, nfreq=6)
    if case == "k3":
        return synth.tiny(natom=2, period=(3, 1, 1), nao=(3, 2), nabf=(5, 3), nocc=(1, 1), a=7.0, seed=9, nfreq=6)
    if case == "gamma3":
        return synth.tiny(natom=3, period=(1, 1, 1), nao=(3, 4), nabf=(5, 6), nocc=(1, 2), a=9.0, seed=5, nfreq=6)
    raise KeyError(case)


def _random_inputs(sys, seed=0):
    """Random G(R) and symmetric-in-R W(R) of the right shapes for one tau."""
    rng = np.random.default_rng(seed)
    Rs = rpa.r_grid(sys.period)
    n, m = sys.nbasis, sys.nabf_tot
    G, W = {}, {}
    for R in Rs:
        G[R] = rng.standard_normal((n, n))
        W[R] = rng.standard_normal((m, m))
    # W(R) = W(-R)^T and G(R) = G(-R)^T, as the real parts of Hermitian k-space matrices are
    for R in Rs:
        mR = rpa.fold_R(-np.array(R), sys.period)
        if R == mR:
            W[R] = 0.5 * (W[R] + W[R].T)
            G[R] = 0.5 * (G[R] + G[R].T)
        elif R < mR:
            W[mR] = W[R].T.copy()
            G[mR] = G[R].T.copy()
    return G, W


@pytest.mark.parametrize("case", ["k2", "k3", "gamma3"])
def test_sigma_labels_vs_supercell(case):
    s = _tiny(case)
    cs = rpa.fold_cs(s)
    G, W = _random_inputs(s)
    a = gw.sigma_labels(s, cs, G, W)
    b = gw.sigma_supercell(s, cs, G, W)
    scale = max(np.abs(v).max() for v in b.values())
    for R in a:
        assert np.abs(a[R] - b[R]).max() / scale < 1e-12


def test_sigma_labels_planted_mistake():
    """A wrong cell in one label (+Ra for -Ra) must fail the supercell comparison."""
    s = _tiny("k3")
    cs = rpa.fold_cs(s)
    G, W = _random_inputs(s, 3)
    a = gw.sigma_labels(s, cs, G, W, mistake='a1b1_cell')
    b = gw.sigma_supercell(s, cs, G, W)
    scale = max(np.abs(v).max() for v in b.values())
    assert max(np.abs(a[R] - b[R]).max() for R in a) / scale > 1e-3


@pytest.mark.parametrize("case", ["k2", "k3", "gamma3"])
def test_sigma_factorized_vs_labels(case):
    s = _tiny(case)
    cs = rpa.fold_cs(s)
    G, W = _random_inputs(s, 1)
    a = gw.sigma_labels(s, cs, G, W)
    f = gw.sigma_factorized(s, cs, G, W)
    scale = max(np.abs(v).max() for v in a.values())
    for R in a:
        assert np.abs(a[R] - f[R]).max() / scale < 1e-12
        # Sigma_XY(R) = Sigma_YX(-R)^T for transpose-symmetric G and W
        mR = rpa.fold_R(-np.array(R), s.period)
        assert np.abs(a[R] - a[mR].T).max() / scale < 1e-12


def _with_real_gamma(s):
    for iq, q in enumerate(s.kfrac):
        if gw.is_gamma(q):
            s.vq[iq] = s.vq[iq].real.astype(np.complex128)
    return s


@pytest.mark.parametrize("case", ["k2", "gamma3"])
def test_w_closed_form(case):
    """v^1/2 [(1 - v^1/2 chi v^1/2)^-1 - 1] v^1/2 = v (1 - chi v)^-1 - v for v > 0, v_c = v."""
    s = _with_real_gamma(_tiny(case))
    emin, emax = tfgrids.e_min_max(s.eps, s.occ, s.nspin, s.nk)
    grid = tfgrids.minimax(s.nfreq, emin, emax)
    chi0s = rpa.chi0s_all_tau(s, grid)
    chi0q = rpa.chi0_q_omega(s, grid, chi0s)
    wq = gw.wc_q_omega(s, grid, chi0q)
    abf = rpa.Basis(s.nabf)
    for (iw, iq), w in wq.items():
        chi = rpa.full_matrix(chi0q, abf, s.natom, lambda I, J: chi0q[(iw, iq, I, J)])
        ref = gw.w_closed_form(s.vq[iq], chi)
        assert np.abs(w - ref).max() / np.abs(ref).max() < 1e-10
        assert np.abs(w - w.conj().T).max() / np.abs(w).max() < 1e-10


def test_w_gamma_uses_real_v_for_sqrt():
    """At q = 0 the square root of v comes from Re v while v_c stays complex."""
    s = _tiny("gamma3")
    emin, emax = tfgrids.e_min_max(s.eps, s.occ, s.nspin, s.nk)
    grid = tfgrids.minimax(s.nfreq, emin, emax)
    chi0q = rpa.chi0_q_omega(s, grid, rpa.chi0s_all_tau(s, grid))
    wq = gw.wc_q_omega(s, grid, chi0q)
    abf = rpa.Basis(s.nabf)
    chi = rpa.full_matrix(chi0q, abf, s.natom, lambda I, J: chi0q[(0, 0, I, J)])
    V = s.vq[0]
    sv = gw.power_hemat(V.real.astype(complex), 0.5)
    svc = gw.power_hemat(V, 0.5)
    n = V.shape[0]
    ref = svc @ (np.linalg.inv(np.eye(n) - sv @ chi @ sv) - np.eye(n)) @ svc
    assert np.abs(wq[(0, 0)] - ref).max() < 1e-12 * np.abs(ref).max()
    # and it differs from the all-complex closed form when Im v(0) != 0
    assert np.abs(wq[(0, 0)] - gw.w_closed_form(V, chi)).max() > 1e-8 * np.abs(ref).max()


def test_pade_closed_form():
    """src/test/test_analycont.cpp:15-46: -1/(x - x0) through 6 points on the imaginary axis."""
    x0 = 2.0 + 2.0j
    xs = [1j * (i + 1) for i in range(6)]
    data = [-1.0 / (x - x0) for x in xs]
    p = gw.Pade(6, xs, data)
    tx = 1.0 + 1.0j
    assert abs(p(tx) - (-1.0 / (tx - x0))) < 1e-12
    # subsampled fit (fewer parameters than points) of a two-pole function stays accurate near the axis
    f = lambda z: 0.3 / (z - (0.5 - 0.4j)) + 0.2 / (z - (-0.7 + 0.3j))
    xs = [1j * w for w in np.linspace(0.05, 5, 16)]
    p = gw.Pade(9, xs, [f(x) for x in xs])
    assert p.n == 9
    assert abs(p(0.4j) - f(0.4j)) < 1e-6


def test_qpe_pole_model():
    """QP equation e = c + a / (e - mu - b) with a single-pole Sigma: closed-form root."""
    a, b, mu = 0.02, 0.9, -0.1
    xs = [1j * w for w in np.linspace(0.1, 3.0, 8)]
    p = gw.Pade(-1, xs, [a / (x - b) for x in xs])
    e_mf, vxc, sigx = -0.35, -0.4, -0.5
    c = e_mf - vxc + sigx
    info, e_qp, sig = gw.qpe_solve(p, e_mf, mu, vxc, sigx)
    # (e - c)(e - mu - b) = a, root nearest c
    roots = np.roots([1.0, -(c + mu + b), c * (mu + b) - a])
    ref = roots[np.argmin(np.abs(roots - c))].real
    assert info == 0
    assert abs(e_qp - ref) < 1e-5
    assert abs(sig.real - a / (e_qp - mu - b)) < 1e-6


def test_ks_nearest_image_phase_on_grid_k():
    """On k points of the mesh the nearest-image cell changes no phase (k . period is integral)."""
    s = _tiny("k2")
    Rs = rpa.r_grid(s.period)
    rng = np.random.default_rng(4)
    sig = {(0, R): rng.standard_normal((s.nbasis, s.nbasis)) + 1j * rng.standard_normal((s.nbasis, s.nbasis))
           for R in Rs}
    frac = s.positions @ np.linalg.inv(s.latvec)
    a = gw.sigma_ks_diag(s, sig, [0, 1], [0, 2], frac=frac)
    b = gw.sigma_ks_diag(s, sig, [0, 1], [0, 2], frac=None)
    for k in a:
        assert np.abs(a[k] - b[k]).max() < 1e-12


def test_head_replacement_and_root_filter_change_w():
    # the oracle's head replacement (eps_head_gamma) and eigenvalue filter (power_hemat with a
    # threshold) are not no-ops: each moves W of a Gamma-only system by more than 1e-3 of its
    # scale; and the head form equals the rank-1 update eps + (eps_mac - u^T eps u) u u^T with u
    # the leading eigenvector of Re v (the library's form)
    from librpa_amd import synth
    from oracle import tfgrids
    s = synth.tiny(natom=3, period=(1, 1, 1), nao=(3, 4), nabf=(5, 6), nocc=(1, 2), a=9.0, seed=5, nfreq=6,
                   real_gamma_v=True)
    emin, emax = tfgrids.e_min_max(s.eps, s.occ, s.nspin, s.nk)
    g = tfgrids.minimax(s.nfreq, emin, emax)
    chi0q = rpa.chi0_q_omega(s, g, rpa.chi0s_all_tau(s, g))
    eps_mac = 1.0 + 6.0 / (1.0 + (g["omega"] / 0.5) ** 2)
    base = gw.wc_q_omega(s, g, chi0q)
    head = gw.wc_q_omega(s, g, chi0q, head=eps_mac)
    thr = float(np.quantile(np.linalg.eigvalsh(s.vq[0]), 0.3))
    filt = gw.wc_q_omega(s, g, chi0q, threshold=thr)
    scale = max(np.abs(v).max() for v in base.values())
    assert max(np.abs(head[k] - base[k]).max() for k in base) > 1e-3 * scale
    assert max(np.abs(filt[k] - base[k]).max() for k in base) > 1e-3 * scale
    abf = rpa.Basis(s.nabf)
    V = s.vq[0]
    w, U = np.linalg.eigh(V.real)
    u = U[:, -1]
    sv = gw.power_hemat(V.real.astype(np.complex128), 0.5)
    for iw in range(s.nfreq):
        chi = rpa.full_matrix(chi0q, abf, s.natom, lambda I, J: chi0q[(iw, 0, I, J)])
        eps = np.eye(abf.total) - sv @ chi @ sv
        rank1 = eps + (eps_mac[iw] - u @ eps @ u) * np.outer(u, u)
        assert np.abs(rank1 - gw.eps_head_gamma(V, chi, eps_mac[iw])).max() < 1e-12 * np.abs(eps).max()
