"""G0W0 inputs that take W off its fast paths, through the C-ABI against the FP64 oracle:

* a Coulomb matrix with negative eigenvalues (v_c = v): the square root drops them
  (power_hemat, src/math/matrix_m.h:916-970, sqrt_coulomb_threshold = 0), so the library's
  Cholesky and Newton-Schulz roots give way to an eigen-decomposition;
* a positive sqrt_coulomb_threshold with a truncated Coulomb matrix that differs from the bare
  one: eigenvalues of both below the threshold are dropped;
* the head of the dielectric matrix at q = 0 replaced by a macroscopic dielectric function
  (replace_w_head; epsilon.cpp:2106-2181), Gamma-only and with k points (where q = 0 is handled
  in the real embedding of the complex path): given on the grid frequencies (option_dielect_func
  0), or on the caller's own frequencies and put on the grid by LibRPA's cubic spline (1) or
  Havriliak-Negami fit (2) (interpolate_dielec_func, dielecmodel.cpp:65-104; the oracle's
  restatement is held against the reference's compiled routines on the CPU);
* an IBZ mapping that pairs each k with -k on time-reversal-symmetric inputs: W is formed at the
  irreducible q only (FT_Wc_freq_q, epsilon.cpp:2430-2452); the oracle forms it on the full grid.

Bars as in test_gpu_gw.py: W(R, tau) and Sigma_c(R, i omega) to 1e-10 of their largest element.
"""
import dataclasses

import numpy as np
import pytest

import librpa_amd as L
from librpa_amd import synth
from oracle import dielec, gw, rpa

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _k2():
    return synth.tiny(natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0, nfreq=6,
                      real_gamma_v=True)


def _gamma():
    return synth.tiny(natom=3, period=(1, 1, 1), nao=(3, 4), nabf=(5, 6), nocc=(1, 2), a=9.0, seed=5, nfreq=6,
                      real_gamma_v=True)


def _k3_tr():
    return synth.time_reversal_symmetric(
        synth.tiny(natom=2, period=(3, 2, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0, seed=3, nfreq=6,
                   real_gamma_v=True))


def _indefinite(s, shift=0.05, nneg=2):
    vq = s.vq.copy()
    for iq in range(s.nk):
        w, u = np.linalg.eigh(vq[iq])
        w = w.copy()
        w[:nneg] = -shift * w[-1] * (1.0 + 0.3 * np.arange(nneg))
        vq[iq] = (u * w[None, :]) @ u.conj().T
        vq[iq] = 0.5 * (vq[iq] + vq[iq].conj().T)
        if np.all(s.kfrac[iq] == 0):
            vq[iq] = vq[iq].real.astype(np.complex128)
    return dataclasses.replace(s, vq=vq)


def _cut(s):
    # a truncated Coulomb matrix that differs from v: 0.6 v + 0.05 v^2 / |v| (still Hermitian)
    vc = np.array([0.6 * v + 0.05 * (v @ v) / np.abs(v).max() for v in s.vq])
    for iq in range(s.nk):
        if np.all(s.kfrac[iq] == 0):
            vc[iq] = vc[iq].real.astype(np.complex128)
    return vc


def _threshold(s, frac=0.3):
    # a threshold between the eigenvalues, dropping about the lowest `frac` of every v and v_c
    ws = np.concatenate([np.linalg.eigvalsh(v) for v in s.vq] + [np.linalg.eigvalsh(v) for v in _cut(s)])
    return float(np.quantile(ws, frac))


CASES = {
    "indefinite_k2": dict(sys=lambda: _indefinite(_k2())),
    "indefinite_gamma": dict(sys=lambda: _indefinite(_gamma(), nneg=1)),
    "threshold_k2": dict(sys=_k2, cut=True, thr=True),
    "head_gamma": dict(sys=_gamma, head=0),
    "head_k2": dict(sys=_k2, head=0, cut=True),
    "head_spline_k2": dict(sys=_k2, head=1, cut=True),
    "head_fit_gamma": dict(sys=_gamma, head=2),
    "ibz_k3x2": dict(sys=_k3_tr, ibz=True),
}


def _eps_mac(omega):
    return 1.0 + 6.0 / (1.0 + (omega / 0.5) ** 2)


@pytest.mark.parametrize("case", sorted(CASES))
def test_w_and_sigma_special_inputs(case):
    _say(f"g0w0 special input case={case}")
    c = CASES[case]
    s = c["sys"]()
    vc = _cut(s) if c.get("cut") else s.vq
    thr = _threshold(s) if c.get("thr") else 0.0
    h = L.Handler()
    L.load_system(h, s, vq_cut=vc if c.get("cut") else True)
    ibz = None
    if c.get("ibz"):
        mk = synth._minus_k(s.kfrac, s.period)
        ibz = np.array([min(k, m) for k, m in enumerate(mk)])
        assert len(set(ibz.tolist())) < s.nk
        h.set_ibz_mapping(ibz)
    o = L.Options(nfreq=s.nfreq, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI, sqrt_coulomb_threshold=thr)
    grid = h.get_tfgrids(o)
    head = None
    if "head" in c:
        opt = c["head"]
        # option 0: values at the grid frequencies; 1, 2: on the caller's own frequencies, which
        # the grid reaches beyond at both ends
        om_in = grid["omega"] if opt == 0 else np.geomspace(0.02, 20.0, 24)
        df_in = _eps_mac(om_in)
        h.set_dielect_func_imagfreq(om_in, df_in)
        o.replace_w_head = 1
        o.option_dielect_func = opt
        head = np.array(dielec.interpolate(opt, om_in, df_in, grid["omega"]))
        _say(f"  head option {opt}: eps_mac on the grid {np.round(head, 4).tolist()}")
    h.build_g0w0_sigma(o)
    chi0q = rpa.chi0_q_omega(s, grid, rpa.chi0s_all_tau(s, grid))
    wq = gw.wc_q_omega(s, grid, chi0q, vq=s.vq, vq_cut=vc, threshold=thr, head=head, ibz=ibz)
    wrt = gw.w_r_tau(s, grid, wq, ibz=ibz)
    if ibz is not None:
        # the full grid gives the same W on time-reversal-symmetric inputs
        wfull = gw.w_r_tau(s, grid, gw.wc_q_omega(s, grid, chi0q, vq=s.vq, vq_cut=vc))
        dfull = max(np.abs(wfull[k] - wrt[k]).max() for k in wrt) / max(np.abs(wrt[k]).max() for k in wrt)
        _say(f"  oracle: IBZ W vs full-grid W rel diff {dfull:.1e}")
        assert dfull < 1e-11
    sig = gw.sigma_all(s, grid, wrt)
    abf, ao = rpa.Basis(s.nabf), rpa.Basis(s.nao)
    Rs = rpa.r_grid(s.period)
    num = den = 0.0
    for it in range(s.nfreq):
        for R in Rs:
            ref = wrt[(it, R)]
            for I in range(s.natom):
                for J in range(s.natom):
                    got = h.get_wc_block(it, I, J, R, int(s.nabf[I]), int(s.nabf[J]))
                    num = max(num, np.abs(got - ref[abf.sl(I), abf.sl(J)]).max())
            den = max(den, np.abs(ref).max())
    _say(f"  W(R,tau) rel err {num / den:.3e}")
    assert num / den < 1e-10
    num = den = 0.0
    for iw in range(s.nfreq):
        for R in Rs:
            ref = sig[(iw, R)]
            for I in range(s.natom):
                for J in range(s.natom):
                    got = h.get_sigma_block(0, iw, I, J, R, int(s.nao[I]), int(s.nao[J]))
                    num = max(num, np.abs(got - ref[ao.sl(I), ao.sl(J)]).max())
            den = max(den, np.abs(ref).max())
    _say(f"  Sigma(R,iw) rel err {num / den:.3e}")
    assert num / den < 1e-10
