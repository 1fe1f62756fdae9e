"""End-to-end check on LibRPA's regression case g0w0_aims_mole_H2O_libri.

Inputs: the H2O files (band, structure, basis, eigenvectors and Cs shared with the RPA case in
tests/golden/h2o_rpa; bare and truncated Coulomb matrices and v_xc in tests/golden/h2o_g0w0).
Expected: the reference's quasi-particle table (librpa.out:480-505, eV with 5 decimals; the
reference's own test allows 1e-4 eV).  v_exx, which this build does not compute, is taken from
that table.  The truncated Coulomb matrix differs from the bare one here, so W goes through the
square roots of both (power_hemat semantics).

* CPU: the FP64 oracle reproduces ReSigc, ImSigc and e_qp of all 24 states to 2e-5 eV (print
  precision 5e-6 eV per value, plus the rounding of the tabulated v_exx that is fed back in).
* GPU: the library through the C API reproduces the same table, and the oracle to 1e-7 eV.
"""
import json
import os

import numpy as np
import pytest

from librpa_amd import driver

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EXP = json.load(open(os.path.join(G, "h2o_g0w0", "expected.json")))
H = EXP["ha2ev"]


def _inputs():
    ds = driver.read_dataset(os.path.join(G, "h2o_rpa") + "/", 1e-6)
    ds["vq"] = driver.read_coulomb(os.path.join(G, "h2o_g0w0") + "/", 125, 1, "coulomb_mat")
    vc = driver.read_coulomb(os.path.join(G, "h2o_g0w0") + "/", 125, 1, "coulomb_cut")
    vxc = driver.read_vxc(os.path.join(G, "h2o_g0w0", "vxc_out"))
    return ds, vc, vxc


def _oracle_table(ds, vc, vxc):
    from oracle import gw, rpa, tfgrids
    s = driver.to_system(ds, 16)
    ds2 = dict(ds)
    ds2["vq"] = vc
    vcf = driver.to_system(ds2, 16).vq
    emin, emax = tfgrids.e_min_max(s.eps, s.occ, s.nspin, s.nk)
    grid = tfgrids.minimax(16, emin, emax)
    chi0q = rpa.chi0_q_omega(s, grid, rpa.chi0s_all_tau(s, grid))
    wrt = gw.w_r_tau(s, grid, gw.wc_q_omega(s, grid, chi0q, vq=s.vq, vq_cut=vcf))
    sig = gw.sigma_all(s, grid, wrt)
    frac = s.positions @ np.linalg.inv(s.latvec)
    ks = gw.sigma_ks_diag(s, sig, [0], list(range(24)), frac=frac)
    out = []
    for r in EXP["rows"]:
        st = r[0] - 1
        p = gw.Pade(EXP["options"]["n_params_anacon"], 1j * grid["omega"], ks[(0, st)])
        info, eqp, sc = gw.qpe_solve(p, s.eps[0, 0, st], s.efermi, vxc[0, 0, st], r[4] / H)
        assert info == 0
        out.append((sc.real * H, sc.imag * H, eqp * H))
    return np.array(out)


def test_h2o_g0w0_oracle_matches_reference_table():
    ds, vc, vxc = _inputs()
    assert np.abs(vc - ds["vq"]).max() > 1e-4  # the truncated matrix really differs
    got = _oracle_table(ds, vc, vxc)
    ref = np.array([[r[5], r[6], r[7]] for r in EXP["rows"]])
    assert np.abs(vxc[0, 0] * H - np.array([r[3] for r in EXP["rows"]])).max() < 6e-6
    assert np.abs(got - ref).max() < 2e-5, np.abs(got - ref).max()


@pytest.mark.gpu
def test_h2o_g0w0_library_matches_reference_table():
    import librpa_amd as L
    ds, vc, vxc = _inputs()
    h = L.Handler()
    driver.push(h, ds, L.LIBRI, vq_cut=vc)
    o = L.Options(nfreq=16, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI,
                  n_params_anacon=EXP["options"]["n_params_anacon"])
    vexx = np.array([r[4] / H for r in EXP["rows"]])
    sc = h.get_g0w0_sigc_kgrid(o, 1, [0], 0, 24, vxc[0, 0], vexx)
    eqp = ds["band"]["eps"][0, 0] - vxc[0, 0] + vexx + sc.real
    got = np.stack([sc.real * H, sc.imag * H, eqp * H], axis=1)
    ref = np.array([[r[5], r[6], r[7]] for r in EXP["rows"]])
    err = np.abs(got - ref).max()
    print(f"H2O G0W0 library vs reference table: max |d| = {err:.2e} eV", flush=True)
    assert err < 2e-5
    oracle = _oracle_table(ds, vc, vxc)
    assert np.abs(got - oracle).max() < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("prec_name,bar_mev", [("fp32", 1.0), ("bf16x3", 25.0), ("mixed", 1.0)])
def test_h2o_g0w0_reduced_precision_qp_error(prec_name, bar_mev):
    # chi0 and Sigma_c contractions in FP32 / BF16x3 (W inversion, quadratures, Pade and the QP
    # solve stay FP64), or chi0 in FP32 at the tau points of small cosine-transform weight only
    # (mixed): quasi-particle energies of all 24 states against the FP64 library path, which
    # matches the reference table above
    import librpa_amd as L
    ds, vc, vxc = _inputs()
    vexx = np.array([r[4] / H for r in EXP["rows"]])
    eqp = {}
    for name in ("fp64", prec_name):
        h = L.Handler()
        driver.push(h, ds, L.LIBRI, vq_cut=vc)
        h.set_precision(L.PREC[name], L.FP64 if name == "mixed" else L.PREC[name])
        o = L.Options(nfreq=16, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI,
                      n_params_anacon=EXP["options"]["n_params_anacon"])
        sc = h.get_g0w0_sigc_kgrid(o, 1, [0], 0, 24, vxc[0, 0], vexx)
        eqp[name] = (ds["band"]["eps"][0, 0] - vxc[0, 0] + vexx + sc.real) * H
    d = np.abs(eqp[prec_name] - eqp["fp64"]) * 1e3
    # the 8 states around the gap (HOMO = state 4); states far from it go through the Pade
    # continuation's ill-conditioned tail and move by up to a few hundred meV (printed)
    gap8 = d[1:9]
    print(f"H2O G0W0 {prec_name}: |dE_qp| over states 1..8 max {gap8.max():.3f} meV "
          f"(HOMO {d[4]:.3f}, LUMO {d[5]:.3f}); all 24: " + " ".join(f"{x:.2g}" for x in d), flush=True)
    assert gap8.max() < bar_mev
