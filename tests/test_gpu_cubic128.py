"""The 128-atom two-species cubic cell (Gamma only, nfreq 16, n_abf 11 520) at full size.

* FP64 chi0(R = 0, tau) blocks at two minimax times (the smallest tau, whose cosine-transform
  weights are largest, and a late one) against the FP64 oracle for sample output blocks that
  cover every species pairing and a diagonal block: 1e-9 of the largest sampled block.  The
  oracle builds G(+-tau) of all 2560 basis functions and the W(K, J) slabs of the sampled J.
* Size-independent checks of the whole build: diagonal blocks chi0s(I, I) are symmetric
  (G(-R) = G(R)^T at Gamma) and Ec < 0.
* The mixed precision mode: |Ec(mixed) - Ec(FP64)| <= 1e-5 Ha/atom, with FP32 at some tau points.
"""
import numpy as np
import pytest

import librpa_amd as L
from librpa_amd import synth
from oracle import rpa

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def system():
    _say("cubic128: generating the system")
    import torch
    return synth.cubic128(vq_device="cuda" if torch.cuda.is_available() else None)


def _handler(s):
    h = L.Handler()
    L.load_system(h, s)
    h.set_keep_chi0(True)
    return h


def test_cubic128_fp64_blocks_and_mixed_precision(system):
    s = system
    o = L.Options(nfreq=s.nfreq, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI)
    h = _handler(s)
    h.set_precision(L.FP64, L.FP64)
    ec64, _ = h.get_rpa_correlation_energy(o, s.nk)
    grid = h.get_tfgrids(o)
    _say(f"  FP64 Ec {ec64!r} Ha ({ec64 / s.natom:.6e} Ha/atom)")
    assert np.isfinite(ec64) and ec64 < 0

    types = np.asarray(s.types)
    a_atoms = [int(i) for i in np.nonzero(types == 0)[0][:3]]
    b_atoms = [int(i) for i in np.nonzero(types == 1)[0][:3]]
    pairs = sorted({tuple(sorted(p)) for p in [(a_atoms[0], a_atoms[1]), (b_atoms[0], b_atoms[1]),
                                               (a_atoms[2], b_atoms[2]), (b_atoms[0], a_atoms[0]),
                                               (a_atoms[1], a_atoms[1]), (b_atoms[2], b_atoms[2])]})
    cs = rpa.fold_cs(s)
    Rs = rpa.r_grid(s.period)
    for it in (0, 12):
        tau = grid["tau"][it]
        Gp = rpa.gf_R(s, tau, Rs)
        Gn = rpa.gf_R(s, -tau, Rs)
        ref = rpa.chi0s_factorized(s, cs, Gp, Gn, pairs, Rs)
        scale = max(np.abs(b).max() for b in ref.values())
        worst = 0.0
        for (I, J, R), r in ref.items():
            got = h.get_chi0s_block(it, I, J, R, int(s.nabf[I]), int(s.nabf[J]))
            worst = max(worst, np.abs(got - r).max() / scale)
        _say(f"  tau[{it}] = {tau:.4g}: {len(ref)} sampled blocks, max |lib - oracle| / max|oracle| = {worst:.2e}")
        assert worst < 1e-9
    for I in (a_atoms[0], b_atoms[0]):
        d = h.get_chi0s_block(0, I, I, (0, 0, 0), int(s.nabf[I]), int(s.nabf[I]))
        assert np.abs(d - d.T).max() <= 1e-12 * np.abs(d).max()
    del h

    hm = _handler(s)
    hm.set_precision(L.MIXED, L.FP32)
    ecm, _ = hm.get_rpa_correlation_energy(o, s.nk)
    st = hm.get_stats()
    _say(f"  mixed (FP32 at {int(st['n_fp32_tau'])} of {s.nfreq} tau, FP32 Ec transforms): Ec {ecm!r}, "
         f"dEc/atom {(ecm - ec64) / s.natom:.2e} Ha")
    assert 0 < st["n_fp32_tau"] < s.nfreq
    assert abs(ecm - ec64) / s.natom <= 1e-5
