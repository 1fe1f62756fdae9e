"""The 64-atom G0W0 workload at full size: 4x4x4 jittered lattice, two species, 2x2x2 k,
nfreq 24, n_abf 5760 (the chi0 -> W -> Sigma_c -> Pade -> QP chain of
librpa_get_g0w0_sigc_kgrid, compute_g0w0.cpp:162-260).

Checks that do not depend on the size (the chain is pinned against the oracle on small k-point
systems in test_gpu_gw.py and on the reference's H2O table):
* W(R, tau) keeps the symmetry of a Hermitian W(q): W_IJ(R) = W_JI(-R)^T (real parts kept);
* the quasi-particle energies of the 8 states around the gap at all 8 k are finite;
* with the chi0 and Sigma_c contractions in FP32 (W, the quadratures, Pade and the QP solve in
  FP64) they stay within 5 meV of the FP64 run (measured: 0.7 meV; the mixed and BF16x3 modes
  are reported by ``bench.py --task g0w0``).
"""
import numpy as np
import pytest

import librpa_amd as L
from librpa_amd import synth
from oracle import rpa

pytestmark = pytest.mark.gpu

HA2EV = 27.211386245988


def _say(msg):
    print(msg, flush=True)


@pytest.mark.timeout(1200)
def test_cubic64_g0w0_qp_fp64_and_fp32(capfd):
    import torch
    _say("cubic64 G0W0: generating the system")
    s = synth.cubic64(vq_device="cuda" if torch.cuda.is_available() else None)
    h = L.Handler()
    L.load_system(h, s, vq_cut=True)
    # debug output level: the library reports each q of W and each tau of Sigma_c as it goes (a
    # build takes minutes; the lines pass pytest's capture)
    o = L.Options(nfreq=s.nfreq, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI, output_level=4)
    nocc = int(round(s.occ[0, 0].sum() / 2.0))
    lo, hi = nocc - 4, nocc + 4
    iks = list(range(s.nk))
    zeros = np.zeros(s.nk * (hi - lo))
    e_ks = s.eps[0][:, lo:hi].reshape(-1)
    eqp = {}
    for name, prec in (("fp64", L.FP64), ("fp32", L.FP32)):
        h.set_precision(prec, prec)
        with capfd.disabled():
            _say(f"  {name}: building W and Sigma_c")
            h.build_g0w0_sigma(o)
        sc = h.get_g0w0_sigc_kgrid(o, 1, iks, lo, hi, zeros, zeros)
        eqp[name] = (e_ks + sc.real) * HA2EV
        assert np.isfinite(eqp[name]).all(), name
        if name == "fp64":
            worst = 0.0
            for (I, J, R) in ((0, 1, (0, 0, 0)), (2, 7, (1, 0, -1)), (5, 5, (0, 1, 0)), (11, 3, (-1, -1, 0))):
                mR = rpa.fold_R(tuple(-v for v in R), s.period)
                a = h.get_wc_block(0, I, J, R, int(s.nabf[I]), int(s.nabf[J]))
                b = h.get_wc_block(0, J, I, mR, int(s.nabf[J]), int(s.nabf[I]))
                worst = max(worst, np.abs(a - b.T).max() / max(np.abs(a).max(), 1e-300))
            _say(f"  W_IJ(R) vs W_JI(-R)^T at tau 0: max rel diff {worst:.1e}")
            assert worst < 1e-10
    d = np.abs(eqp["fp32"] - eqp["fp64"]) * 1e3
    _say(f"  QP states {lo}..{hi - 1} at {s.nk} k: FP32 vs FP64 max |dE_qp| {d.max():.3f} meV "
         f"(HOMO {d.reshape(s.nk, -1)[:, 3].max():.3f}, LUMO {d.reshape(s.nk, -1)[:, 4].max():.3f})")
    assert d.max() < 5.0
