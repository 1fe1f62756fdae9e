"""GPU checks of the boundary's conventions through the C-ABI.

* k points: the library's Ec of a time-reversal-symmetric k-mesh system equals 1 / N_k of the
  library's Ec of its Gamma-only BvK supercell (built by the Bloch convention in
  ``synth.bvk_supercell``), and both equal the oracle's.
* routing: every ``parallel_routing`` value (RTAU, ATOMPAIR, LIBRI, AUTO; librpa_enums.h) gives
  the LibRI result, for the routing passed to ``librpa_set_lri_coeff`` and in the options.
* a Coulomb matrix that is not positive definite: Ec equals the oracle's LU-based
  ln|det(1 - V chi0)| (epsilon.cpp:600-693).  (W from such a matrix drops its negative
  eigenvalues in v^{1/2}; test_gpu_gw_special.py covers it.)
* split-K: the contraction's atomically accumulated K parts agree with one plain write to
  1e-13 in FP64, and one plain write is bitwise repeatable.
* the host-staged chi0 store gives the device store's Ec and blocks; Ec with several q per pass
  gives the per-q contributions of one q per pass.
"""
import numpy as np
import pytest

import librpa_amd as L
from librpa_amd import synth
from oracle import rpa

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ec(s, routing=L.LIBRI, opt_routing=L.LIBRI, keep=False, prec=L.FP64):
    h = L.Handler()
    L.load_system(h, s, routing=routing)
    h.set_precision(prec, L.FP64 if prec == L.MIXED else prec)
    if keep:
        h.set_keep_chi0(True)
    o = L.Options(nfreq=s.nfreq, tfgrids_type=L.MINIMAX, parallel_routing=opt_routing)
    ec, contrib = h.get_rpa_correlation_energy(o, s.nk)
    return h, o, ec, contrib


def _oracle_ec(s, h, o):
    grid = h.get_tfgrids(o)
    return rpa.rpa_ec(s, grid, rpa.chi0_q_omega(s, grid, rpa.chi0s_all_tau(s, grid)))


KCASES = {
    "k3x2_mixed": dict(natom=2, period=(3, 2, 1), nao=(2, 3), nabf=(3, 4), nocc=(1, 1), a=6.5, seed=11, nfreq=6),
    "k1x3x3": dict(natom=1, period=(1, 3, 3), nao=(4,), nabf=(6,), nocc=(2,), a=6.0, seed=5, nfreq=8),
    "k3x3_larger": dict(natom=2, period=(3, 3, 1), nao=(5, 4), nabf=(12, 9), nocc=(2, 1), a=7.0, seed=21,
                        nfreq=16),
}


@pytest.mark.parametrize("case", sorted(KCASES))
def test_kmesh_equals_bvk_supercell(case):
    _say(f"k mesh vs BvK supercell case={case}")
    s = synth.time_reversal_symmetric(synth.tiny(**KCASES[case]))
    sc = synth.bvk_supercell(s)
    h, o, ec_k, _ = _ec(s)
    hs, os_, ec_sc, _ = _ec(sc)
    _say(f"  Ec k-mesh {ec_k!r}  supercell / N {ec_sc / s.nk!r}")
    assert abs(ec_k - ec_sc / s.nk) < 1e-10 * abs(ec_k)
    if case != "k3x3_larger":
        ec_ref, _ = _oracle_ec(s, h, o)
        _say(f"  oracle {ec_ref!r}")
        assert abs(ec_k - ec_ref) < 1e-10 * abs(ec_ref)


@pytest.mark.parametrize("routing", [L.RTAU, L.ATOMPAIR, L.AUTO])
def test_every_routing_gives_the_libri_result(routing):
    _say(f"routing {routing}")
    s = synth.tiny(natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0, nfreq=6)
    _, _, ec_libri, c_libri = _ec(s)
    _, _, ec1, c1 = _ec(s, routing=routing, opt_routing=routing)
    _, _, ec2, c2 = _ec(s, routing=routing, opt_routing=L.LIBRI)
    _, _, ec3, c3 = _ec(s, routing=L.LIBRI, opt_routing=routing)
    for e, c in ((ec1, c1), (ec2, c2), (ec3, c3)):
        assert abs(e - ec_libri) <= 1e-13 * abs(ec_libri)
        assert np.abs(c - c_libri).max() <= 1e-13 * abs(ec_libri)


def _indefinite(s, shift=0.02, nneg=1):
    """The same system with each v(q) moved so that its nneg lowest eigenvalues are -shift * (its
    largest): Hermitian, invertible and not positive definite."""
    vq = s.vq.copy()
    for iq in range(s.nk):
        w, u = np.linalg.eigh(vq[iq])
        w = w.copy()
        w[:nneg] = -shift * w[-1] * (1.0 + 0.1 * np.arange(nneg))
        vq[iq] = (u * w[None, :]) @ u.conj().T
        vq[iq] = 0.5 * (vq[iq] + vq[iq].conj().T)
        if np.all(s.kfrac[iq] == 0):
            vq[iq] = vq[iq].real.astype(np.complex128)
    import dataclasses
    return dataclasses.replace(s, vq=vq)


@pytest.mark.parametrize("kw", [
    dict(natom=3, period=(1, 1, 1), nao=(3, 4), nabf=(5, 6), nocc=(1, 2), a=9.0, seed=5, nfreq=6),
    dict(natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0, nfreq=6),
    "cell2_gamma",   # n_abf = 120: the negative pivots fall in different 64-column blocks
])
def test_indefinite_coulomb_matrix_ec(kw):
    base = synth.cell2(period=(1, 1, 1), nfreq=8) if kw == "cell2_gamma" else synth.tiny(**kw)
    s = _indefinite(base, nneg=3 if kw == "cell2_gamma" else 1)
    assert min(np.linalg.eigvalsh(v).min() for v in s.vq) < 0
    _say(f"indefinite v: nk={s.nk}")
    h, o, ec, contrib = _ec(s)
    ec_ref, c_ref = _oracle_ec(s, h, o)
    _say(f"  Ec lib={ec!r} oracle={ec_ref!r}")
    assert abs(ec - ec_ref) < 1e-10 * max(1.0, abs(ec_ref))
    assert np.abs(contrib.real - c_ref.real).max() < 1e-10 * max(1.0, abs(ec_ref))


def _chi0_blocks(h, s, taus):
    out = {}
    for it in taus:
        for I in range(s.natom):
            for J in range(I, s.natom):
                for R in rpa.r_grid(s.period):
                    out[(it, I, J, R)] = h.get_chi0s_block(it, I, J, R, int(s.nabf[I]), int(s.nabf[J]))
    return out


def test_split_k_parts_agree_and_single_part_is_bitwise_repeatable(monkeypatch):
    _say("split-K reproducibility cell2")
    s = synth.cell2()
    taus = (0, 7, 15)
    monkeypatch.setenv("LIBRPA_AMD_SPLITK", "1")
    h1, _, ec1, _ = _ec(s, keep=True)
    b1 = _chi0_blocks(h1, s, taus)
    h1b, _, ec1b, _ = _ec(s, keep=True)
    b1b = _chi0_blocks(h1b, s, taus)
    assert ec1 == ec1b
    assert all(np.array_equal(b1[k], b1b[k]) for k in b1)
    monkeypatch.setenv("LIBRPA_AMD_SPLITK", "4")
    h4, _, ec4, _ = _ec(s, keep=True)
    b4 = _chi0_blocks(h4, s, taus)
    worst = max(np.abs(b4[k] - b1[k]).max() / max(np.abs(b1[k]).max(), 1e-300) for k in b1)
    _say(f"  Ec split 1 {ec1!r}  split 4 {ec4!r}  chi0 rel diff {worst:.1e}")
    assert worst < 1e-13
    assert abs(ec4 - ec1) < 1e-12 * abs(ec1)


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_host_staged_chi0_store_gives_the_same_result(monkeypatch, prec):
    # chi0(R, tau) kept in host memory slice by slice (what a store too large for the device
    # does) and streamed back into chi0(R, omega) for Ec: the same Ec and blocks
    _say(f"host-staged chi0 store, prec {prec}")
    s = synth.cell2()
    taus = (0, 7, 15)
    monkeypatch.setenv("LIBRPA_AMD_SPLITK", "1")
    h0, _, ec0, c0 = _ec(s, keep=True, prec=prec)
    b0 = _chi0_blocks(h0, s, taus)
    monkeypatch.setenv("LIBRPA_AMD_HOST_CHI0", "1")
    h1, _, ec1, c1 = _ec(s, keep=True, prec=prec)
    b1 = _chi0_blocks(h1, s, taus)
    worst = max(np.abs(b1[k] - b0[k]).max() / max(np.abs(b0[k]).max(), 1e-300) for k in b0)
    _say(f"  Ec device store {ec0!r}, host-staged {ec1!r}; chi0 blocks max rel diff {worst:.1e}")
    assert worst == 0.0
    assert abs(ec1 - ec0) <= 1e-13 * abs(ec0)
    assert np.abs(c1 - c0).max() <= 1e-13 * abs(ec0)


def test_ec_q_grouping_gives_the_same_result(monkeypatch):
    # Ec factorises several q per pass on k-point workloads (batched V Cholesky, GEMMs and
    # log-det; chi0(R, omega) read once per 8 q): one q per pass, and passes of 4 + 4 + 1 q, give
    # the same per-q contributions as all 9 q in one pass
    _say("Ec q grouping, 3x3 k")
    s = synth.tiny(**KCASES["k3x3_larger"])
    monkeypatch.setenv("LIBRPA_AMD_SPLITK", "1")
    _, _, ec_all, c_all = _ec(s)
    res = {}
    for qg in ("1", "4"):
        monkeypatch.setenv("LIBRPA_AMD_EC_QGROUP", qg)
        _, _, ec, c = _ec(s)
        res[qg] = (ec, c)
        _say(f"  q per pass {qg}: Ec {ec!r} (all in one pass {ec_all!r})")
        assert abs(ec - ec_all) <= 1e-12 * abs(ec_all)
        assert np.abs(c - c_all).max() <= 1e-12 * abs(ec_all)
