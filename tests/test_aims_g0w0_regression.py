"""End-to-end checks on three of LibRPA's G0W0 regression cases (FHI-aims inputs, the
reference's own dataset files and printed quasi-particle tables; tests/golden/<case>, written by
make_aims_g0w0_expected.py):

* ``mgo``: rock-salt MgO, a 3x3x3 k grid reduced to 14 irreducible points by time reversal,
  nfreq 6, the head of the dielectric matrix at q = 0 replaced by the macroscopic dielectric
  function (replace_w_head = t, option_dielect_func = 0), and a bare Coulomb matrix v(0) with a
  negative eigenvalue (-0.20) that v^{1/2} drops (power_hemat, sqrt_coulomb_threshold = 0).
  G(k) -> G(R), chi0(R) -> chi0(q), the IBZ star of W and Sigma(R) -> Sigma(k) all enter it:
  the k-point pin that the reference itself holds.  Bar: every state at every k.
* ``li_atom``: Li atom, two spin channels (W from the spin sum, Sigma per spin), nfreq 16.
  Bar: every state of both channels.
* ``h2``: H2 molecule, nfreq 32.  The 32-point cosine/sine transform matrices are numerically
  singular (in the reference too), so quantities that depend on their null space differ between
  least-squares solvers; the states around the gap do not: bar on states 1-5 (HOMO = 1), the
  others printed (measured: 1e-4 eV at state 6, up to 3 eV higher up, where the Pade
  continuation of the poorly determined Sigma(i omega) takes over; the reference prints NaN for
  state 10).

v_exx, which this build does not compute, is taken from the tables; v_xc from the dataset.

* CPU: the FP64 oracle reproduces ReSigc, ImSigc and e_qp within the bar to 2e-5 eV (MgO:
  without the head replacement it misses by more than 0.1 eV).
* GPU: the library through the C API reproduces the same, and the oracle to 1e-7 eV (Li: 1e-6;
  H2: the reference bar).
"""
import json
import os
import tarfile

import numpy as np
import pytest

from librpa_amd import driver

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# fixture directory, states under the bar (None: all), library-vs-oracle bar in eV.  The Pade
# continuation amplifies rounding-level differences of Sigma(i omega) (Li 1s at -51 eV: 1.3e-7
# eV measured); H2: the singular 32-point transforms leave a null-space part that the library's
# and the oracle's least-squares solvers fix differently, as the reference's does, so the two
# agree to the reference bar there.
CASES = {"mgo": ("mgo_g0w0", None, 1e-7), "li_atom": ("li_atom_g0w0", None, 1e-6), "h2": ("h2_g0w0", 5, 2e-5)}
BAR = 2e-5
_ORACLE = {}


def _exp(case):
    return json.load(open(os.path.join(G, CASES[case][0], "expected.json")))


@pytest.fixture(scope="module")
def datasets(tmp_path_factory):
    out = {}
    for case, (d, _, _) in CASES.items():
        t = tmp_path_factory.mktemp(case)
        with tarfile.open(os.path.join(G, d, "dataset.tar.gz")) as tf:
            tf.extractall(t, filter="data")
        p = str(t / "dataset") + "/"
        exp = _exp(case)
        ds = driver.read_dataset(p, exp["options"]["cs_threshold"])
        n = int(ds["nabf"].sum())
        vc = driver.read_coulomb(p, n, ds["band"]["nk"], "coulomb_cut")
        vxc = driver.read_vxc(p + "vxc_out")
        om, df = driver.read_dielec_func(p + "dielecfunc_out")
        tab = np.array([[[r[1:] for r in k] for k in s] for s in exp["tables"]])  # [spin][k][state][occ .. e_qp]
        out[case] = dict(ds=ds, vc=vc, vxc=vxc, om=om, df=df, tab=tab, exp=exp)
    return out


def _oracle_table(c, head=True):
    from oracle import gw, rpa, tfgrids
    ds, exp, tab = c["ds"], c["exp"], c["tab"]
    H, nf = exp["ha2ev"], exp["options"]["nfreq"]
    s = driver.to_system(ds, nf)
    ds2 = dict(ds)
    ds2["vq"] = c["vc"]
    vcf = driver.to_system(ds2, nf).vq
    ibz = ds["bz"]["ibz"]
    emin, emax = tfgrids.e_min_max(s.eps, s.occ, s.nspin, s.nk)
    grid = tfgrids.minimax(nf, emin, emax)
    use_head = head and exp["options"]["replace_w_head"]
    if use_head:
        assert np.abs(grid["omega"] - c["om"]).max() < 1e-8 * c["om"].max()  # option 0: values on the grid
    chi0q = rpa.chi0_q_omega(s, grid, rpa.chi0s_all_tau(s, grid))
    wq = gw.wc_q_omega(s, grid, chi0q, vq=s.vq, vq_cut=vcf, head=c["df"] if use_head else None, ibz=ibz)
    wrt = gw.w_r_tau(s, grid, wq, ibz=ibz)
    frac = s.positions @ np.linalg.inv(s.latvec)
    nst = tab.shape[2]
    out = np.zeros((s.nspin, s.nk, nst, 3))
    for sp in range(s.nspin):
        sig = gw.sigma_all(s, grid, wrt, spin=sp)
        ks = gw.sigma_ks_diag(s, sig, list(range(s.nk)), list(range(nst)), frac=frac, spin=sp)
        for ik in range(s.nk):
            for st in range(nst):
                p = gw.Pade(exp["options"]["n_params_anacon"], 1j * grid["omega"], ks[(ik, st)])
                info, eqp, sc = gw.qpe_solve(p, s.eps[sp, ik, st], s.efermi, c["vxc"][sp, ik, st],
                                             tab[sp, ik, st, 3] / H)
                assert info == 0
                out[sp, ik, st] = (sc.real * H, sc.imag * H, eqp * H)
    return out


def _oracle_cached(case, c):
    if case not in _ORACLE:
        _ORACLE[case] = _oracle_table(c)
    return _ORACLE[case]


def _err(case, got, tab):
    nb = CASES[case][1]
    d = np.abs(got - tab[..., 4:7]).max(axis=-1)  # [spin][k][state]
    return (d if nb is None else d[..., :nb]).max(), d


def test_aims_datasets_have_the_stated_features(datasets):
    mgo = datasets["mgo"]
    assert len(set(int(x) for x in mgo["ds"]["bz"]["ibz"])) == 14
    assert np.linalg.eigvalsh(mgo["ds"]["vq"][0].real).min() < -0.1  # v(0) is indefinite
    assert datasets["li_atom"]["ds"]["band"]["nspin"] == 2
    for case, c in datasets.items():
        H = c["exp"]["ha2ev"]
        assert np.abs(c["vxc"] * H - c["tab"][..., 2]).max() < 6e-6, case


@pytest.mark.parametrize("case", sorted(CASES))
def test_aims_g0w0_oracle_matches_reference_table(datasets, case):
    c = datasets[case]
    got = _oracle_cached(case, c)
    err, d = _err(case, got, c["tab"])
    print(f"{case} G0W0 oracle vs reference table: max |d| = {err:.2e} eV within the bar; all states: "
          + " ".join(f"{x:.1e}" for x in d.reshape(-1)[:24]), flush=True)
    assert err < BAR
    if case == "mgo":
        # planted mistake: the head replacement left out
        miss = np.abs(_oracle_table(c, head=False) - c["tab"][..., 4:7]).max()
        assert miss > 0.1, miss


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_aims_g0w0_library_matches_reference_table(datasets, case):
    import librpa_amd as L
    c = datasets[case]
    ds, exp, tab = c["ds"], c["exp"], c["tab"]
    H = exp["ha2ev"]
    ns, nk, nst = ds["band"]["nspin"], ds["band"]["nk"], tab.shape[2]
    h = L.Handler()
    driver.push(h, ds, L.LIBRI, vq_cut=c["vc"])
    if exp["options"]["replace_w_head"]:
        h.set_dielect_func_imagfreq(c["om"], c["df"])
    o = L.Options(nfreq=exp["options"]["nfreq"], tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI,
                  n_params_anacon=exp["options"]["n_params_anacon"],
                  replace_w_head=1 if exp["options"]["replace_w_head"] else 0,
                  option_dielect_func=exp["options"]["option_dielect_func"])
    vexx = tab[..., 3] / H
    vxc = c["vxc"][:, :, :nst]
    sc = h.get_g0w0_sigc_kgrid(o, ns, list(range(nk)), 0, nst, vxc.reshape(-1), vexx.reshape(-1))
    sc = sc.reshape(ns, nk, nst)
    eqp = ds["band"]["eps"][:, :, :nst] - vxc + vexx + sc.real
    got = np.stack([sc.real * H, sc.imag * H, eqp * H], axis=-1)
    err, _ = _err(case, got, tab)
    print(f"{case} G0W0 library vs reference table: max |d| = {err:.2e} eV within the bar", flush=True)
    assert err < BAR
    oracle = _oracle_cached(case, c)
    nb = CASES[case][1]
    d = np.abs(got - oracle)
    d = (d if nb is None else d[:, :, :nb]).max()
    print(f"{case} G0W0 library vs oracle: max |d| = {d:.2e} eV", flush=True)
    assert d < CASES[case][2]
