"""The oracle's label contractions against the reference's own LibRI code (CPU only).

``oracle/_ref/ref_libri`` (built by ``make -C oracle`` where the reference tree is present) runs
LibRI on one tau point of a small system and writes every block it returns:

* ``chi0s``: ``RPA::cal_chi0s`` (thirdparty/LibRI/include/RI/physics/RPA.hpp:104-138), fed as
  LibRPA feeds it (src/core/chi0.cpp:956-1066): raw Cs tensors (mu, i, j) with their unfolded
  cells (src/api/input.cpp:525-540), G(R, +tau) and G(R, -tau) blocks (i, j) on the BvK grid
  (src/core/chi0.cpp:395-419). Every block must equal ``oracle.rpa.chi0s_bruteforce``.
* ``sigmas``: ``GW::cal_Sigmas`` (thirdparty/LibRI/include/RI/physics/GW.hpp:99-122), fed as
  src/core/gw.cpp:425-470 and :616-618 feed it: Cs, W(R, tau) blocks (mu, nu), G(R, tau) of one
  sign. Every block must equal ``oracle.gw.sigma_labels``.

This pins the oracle's label sums, cell folding (LibRI's ``cal_period``) and R sign conventions on
k-point systems against the reference's code, not against a restatement of it. Planted cell-index
mistakes in the oracle must fail the comparisons.
"""
import numpy as np
import pytest

from librpa_amd import synth
from oracle import gw as ogw
from oracle import ref_io
from oracle import rpa

pytestmark = pytest.mark.skipif(
    not ref_io.available(),
    reason="oracle/_ref/ref_libri is not built (make -C oracle needs the reference tree)")


CASES = [
    dict(natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0),
    dict(natom=2, period=(3, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0),
    dict(natom=3, period=(1, 2, 2), nao=(2, 3), nabf=(3, 4), nocc=(1, 1), a=6.5, seed=11),
    dict(natom=2, period=(3, 2, 1), nao=(2, 3), nabf=(3, 2), nocc=(1, 1), a=5.5, seed=3),
    dict(natom=2, period=(1, 1, 1), nao=(3, 3), nabf=(4, 4), nocc=(1, 2), a=9.0, seed=5),
]


def _inputs(kw, tau=0.37):
    s = synth.tiny(**kw)
    Rs = rpa.r_grid(s.period)
    return s, rpa.fold_cs(s), Rs, rpa.gf_R(s, tau, Rs), rpa.gf_R(s, -tau, Rs)


def _libri_chi0s(tmp_path, s, Gp, Gn):
    return ref_io.chi0s(s, Gp, Gn, timeout=120)[0]


def _w_of_r(s, seed=0):
    """A real W(R) with the symmetry of a screened interaction, W(-R) = W(R)^T."""
    rng = np.random.default_rng(seed + 31)
    m = rpa.Basis(s.nabf).total
    A = {R: rng.standard_normal((m, m)) for R in rpa.r_grid(s.period)}
    return {R: A[R] + A[rpa.fold_R(tuple(-x for x in R), s.period)].T for R in A}


def _libri_sigmas(tmp_path, s, G, W):
    return ref_io.sigmas(s, G, W, timeout=120)


def _compare(ref, mine_of_key, Rs, what):
    assert ref, f"LibRI returned no {what} blocks"
    scale = max(np.abs(v).max() for v in ref.values())
    for key, v in ref.items():
        assert key[2] in Rs, key
        assert np.abs(mine_of_key(key) - v).max() <= 1e-12 * scale, key
    return scale


@pytest.mark.parametrize("kw", CASES)
def test_oracle_chi0s_equal_libri_cal_chi0s(tmp_path, kw):
    s, cs, Rs, Gp, Gn = _inputs(kw)
    ref = _libri_chi0s(tmp_path, s, Gp, Gn)
    pairs = sorted({(I, J) for (I, J, _) in ref})
    mine = rpa.chi0s_bruteforce(s, cs, Gp, Gn, pairs, Rs)
    scale = _compare(ref, lambda k: mine[k], Rs, "chi0s")
    # every atom pair and cell comes back
    assert len(ref) == s.natom ** 2 * len(Rs)
    for key, v in mine.items():
        if key not in ref:
            assert np.abs(v).max() <= 1e-12 * scale, key


def test_chi0s_comparison_detects_planted_cell_sign(tmp_path):
    # the oracle with G taken at -R instead of R (period 3, where -R and R are different cells)
    s, cs, Rs, Gp, Gn = _inputs(CASES[1])
    ref = _libri_chi0s(tmp_path, s, Gp, Gn)
    pairs = sorted({(I, J) for (I, J, _) in ref})
    orig = rpa._gblk
    try:
        rpa._gblk = lambda G, ao, I, J, R, period: orig(G, ao, I, J, -np.asarray(R), period)
        wrong = rpa.chi0s_bruteforce(s, cs, Gp, Gn, pairs, Rs)
    finally:
        rpa._gblk = orig
    scale = max(np.abs(v).max() for v in ref.values())
    assert max(np.abs(wrong[k] - v).max() for k, v in ref.items()) > 1e-6 * scale


@pytest.mark.parametrize("sign", [1, -1])
@pytest.mark.parametrize("kw", CASES)
def test_oracle_sigma_labels_equal_libri_cal_sigmas(tmp_path, kw, sign):
    s, cs, Rs, Gp, Gn = _inputs(kw)
    G = Gp if sign > 0 else Gn
    W = _w_of_r(s)
    ref = _libri_sigmas(tmp_path, s, G, W)
    ao = rpa.Basis(s.nao)
    mine = ogw.sigma_labels(s, cs, G, W)
    _compare(ref, lambda k: mine[k[2]][ao.sl(k[0]), ao.sl(k[1])], Rs, "Sigma")
    # blocks LibRI does not return vanish in the oracle
    got = {(I, J, R) for (I, J, R) in ref}
    for R in Rs:
        for I in range(s.natom):
            for J in range(s.natom):
                if (I, J, R) not in got:
                    assert np.abs(mine[R][ao.sl(I), ao.sl(J)]).max() <= 1e-12, (I, J, R)


def test_sigma_comparison_detects_planted_cell_mistake(tmp_path):
    # the a0b0_a1b1 term written to cell R + Rb + Ra instead of R + Rb - Ra
    s, cs, Rs, Gp, Gn = _inputs(CASES[1])
    W = _w_of_r(s)
    ref = _libri_sigmas(tmp_path, s, Gp, W)
    ao = rpa.Basis(s.nao)
    wrong = ogw.sigma_labels(s, cs, Gp, W, mistake="a1b1_cell")
    scale = max(np.abs(v).max() for v in ref.values())
    err = max(np.abs(wrong[k[2]][ao.sl(k[0]), ao.sl(k[1])] - v).max() for k, v in ref.items())
    assert err > 1e-6 * scale


def test_bench_reference_cpu_baseline_runs():
    # the bench's CPU baseline leg drives the same checker on atom samples of its workload
    import bench
    from oracle import tfgrids
    s = synth.tiny(natom=8, period=(1, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=9.0, seed=2)
    emin, emax = tfgrids.e_min_max(s.eps, s.occ, s.nspin, s.nk)
    grid = tfgrids.minimax(s.nfreq, emin, emax)
    r = bench.cpu_baseline_reference(s, grid, s.nfreq)
    assert r["kind"] == "reference" and r["value"] > 0
    assert set(r["parts"]["libri_samples"]) == {"2", "4"}
    # samples keep the mix of atom kinds (n_ao, n_abf): here 4 + 4 atoms of two kinds
    pick = bench._balanced_atoms(s, 4)
    assert len(pick) == 4 and len(set(pick)) == 4
    assert sorted(int(s.nao[i]) for i in pick) == [2, 2, 3, 3]
