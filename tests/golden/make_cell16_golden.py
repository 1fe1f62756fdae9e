"""Writes tests/golden/cell16/: FP64 oracle results for the 16-atom, 4x4x4-k, nfreq=24 workload.

The oracle needs about 70 s of CPU per tau point on this system (24 of them), so the test suite
reads these results instead of recomputing them.  Contents:

* expected.json: Ec and the per-k contributions (``oracle.rpa.rpa_ec_streamed``), the minimax
  nodes, and per tau the sum of squares of every chi0s(I <= J, R) block (a checksum of the whole
  chi0(R, tau) tensor);
* blocks.npz: a few whole chi0s(R, tau) blocks (``t{it}_{I}_{J}_{x}_{y}_{z}``).

Run from the repository root:  python tests/golden/make_cell16_golden.py [--workers 4]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

SAMPLE_TAU = (0, 23)
SAMPLE_BLOCKS = ((0, 0, (0, 0, 0)), (3, 12, (1, -1, 0)))


def _one_tau(it):
    import numpy as np
    from librpa_amd import synth
    from oracle import rpa, tfgrids
    s = synth.cell16()
    emin, emax = tfgrids.e_min_max(s.eps, s.occ, s.nspin, s.nk)
    g = tfgrids.minimax(s.nfreq, emin, emax)
    Rs = rpa.r_grid(s.period)
    tau = g['tau'][it]
    c = rpa.chi0s_factorized(s, rpa.fold_cs(s), rpa.gf_R(s, tau, Rs), rpa.gf_R(s, -tau, Rs),
                             rpa.upper_pairs(s.natom), Rs)
    return it, c


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=4)
    args = ap.parse_args()
    os.environ.setdefault("OMP_NUM_THREADS", "2")
    import multiprocessing as mp
    import numpy as np
    from librpa_amd import synth
    from oracle import rpa, tfgrids

    s = synth.cell16()
    emin, emax = tfgrids.e_min_max(s.eps, s.occ, s.nspin, s.nk)
    g = tfgrids.minimax(s.nfreq, emin, emax)
    acc, sumsq, samples = {}, [0.0] * s.nfreq, {}
    t0 = time.time()
    with mp.get_context("spawn").Pool(args.workers) as pool:
        for it, c in pool.imap_unordered(_one_tau, range(s.nfreq)):
            rpa.freq_accumulate(s, g, c, it, acc)
            sumsq[it] = float(sum(float(np.sum(b * b)) for b in c.values()))
            if it in SAMPLE_TAU:
                for (I, J, R) in SAMPLE_BLOCKS:
                    samples["t%d_%d_%d_%d_%d_%d" % (it, I, J, *R)] = c[(I, J, R)]
            print("tau %d done at %.0f s" % (it, time.time() - t0), flush=True)
    ec, contrib = rpa.rpa_ec_streamed(s, g, acc)
    print("Ec %.12f Ha at %.0f s" % (ec, time.time() - t0), flush=True)
    out = os.path.join(ROOT, "tests", "golden", "cell16")
    os.makedirs(out, exist_ok=True)
    json.dump(dict(system="cell16", seed=1, nfreq=s.nfreq, natom=s.natom, nk=s.nk,
                   ec_ha=ec, contrib_re=[float(x.real) for x in contrib],
                   contrib_im=[float(x.imag) for x in contrib],
                   tau=[float(x) for x in g['tau']], omega=[float(x) for x in g['omega']],
                   chi0s_sumsq=sumsq, sample_tau=list(SAMPLE_TAU),
                   sample_blocks=[[I, J, list(R)] for (I, J, R) in SAMPLE_BLOCKS]),
              open(os.path.join(out, "expected.json"), "w"), indent=1)
    np.savez_compressed(os.path.join(out, "blocks.npz"), **samples)


if __name__ == "__main__":
    main()
