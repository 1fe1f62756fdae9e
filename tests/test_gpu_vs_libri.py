"""The library's chi0(R, tau) against the reference's own LibRI contraction, through the C-ABI.

``oracle/_ref/ref_libri`` runs LibRI's ``RPA::cal_chi0s`` (thirdparty/LibRI/include/RI/physics/
RPA.hpp:104-138), compiled from the reference's sources, on the same Cs and on G(R, +-tau) at the
library's own minimax tau points (the oracle's G, which the FP64 parity tests hold the library's
to). Every block the library keeps (I <= J, all cells) must equal LibRI's to 1e-11 of the largest,
on k-point systems with periods up to 3 and at Gamma, at the first, a middle and the last tau.
The test is skipped where the checker is not built.
"""
import numpy as np
import pytest

import librpa_amd as L
from librpa_amd import synth
from oracle import ref_io, rpa

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref_io.available(), reason="oracle/_ref/ref_libri is not built")]

CASES = {
    "tiny_k2_mixed": lambda: synth.tiny(natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=7.0,
                                        nfreq=6),
    "tiny_k3_odd": lambda: synth.tiny(natom=2, period=(3, 1, 1), nao=(3, 2), nabf=(5, 3), nocc=(1, 1), a=7.0,
                                      seed=9, nfreq=6),
    "tiny_k4_3atoms": lambda: synth.tiny(natom=3, period=(1, 2, 2), nao=(2, 3), nabf=(3, 4), nocc=(1, 1), a=6.5,
                                         seed=11, nfreq=8),
    "tiny_gamma": lambda: synth.tiny(natom=3, period=(1, 1, 1), nao=(3, 4), nabf=(5, 6), nocc=(1, 2), a=9.0,
                                     seed=5, nfreq=6),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_library_chi0_equals_libri(case):
    print(f"library chi0 vs LibRI case={case}", flush=True)
    s = CASES[case]()
    assert s.nspin == 1
    h = L.Handler()
    L.load_system(h, s)
    h.set_precision(L.FP64, L.FP64)
    h.set_keep_chi0(True)
    o = L.Options(nfreq=s.nfreq, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI)
    h.get_rpa_correlation_energy(o, s.nk)
    grid = h.get_tfgrids(o)
    Rs = rpa.r_grid(s.period)
    for it in (0, s.nfreq // 2, s.nfreq - 1):
        tau = grid["tau"][it]
        ref, _ = ref_io.chi0s(s, rpa.gf_R(s, tau, Rs), rpa.gf_R(s, -tau, Rs), timeout=120)
        upper = {k: v for k, v in ref.items() if k[0] <= k[1]}
        assert len(upper) == s.natom * (s.natom + 1) // 2 * len(Rs)
        scale = max(np.abs(v).max() for v in upper.values())
        err = 0.0
        for (I, J, R), v in upper.items():
            got = h.get_chi0s_block(it, I, J, R, int(s.nabf[I]), int(s.nabf[J]))
            err = max(err, np.abs(got - v).max())
        print(f"  tau {it}: max |library - LibRI| / max |chi0s| = {err / scale:.2e}", flush=True)
        assert err <= 1e-11 * scale, (it, err / scale)
