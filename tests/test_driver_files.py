"""Seeded synthetic systems written as LibRPA dataset files and read back through the driver.

* CPU: write -> read -> System round trip of the 2-atom 8-k workload and of a spin-polarised
  mixed-species system (energies, occupations, eigenvectors, Cs, Coulomb matrix, geometry, k).
* GPU: ``librpa_amd.driver.main`` (task = rpa, the driver's file path) on the written 2-atom files
  prints the Ec of the library's direct C-API run and of the oracle.
"""
import numpy as np
import pytest

from librpa_amd import driver, synth


def _roundtrip(s, tmp_path):
    d = str(tmp_path / s.name)
    driver.write_dataset(s, d)
    ds = driver.read_dataset(d + "/", 0.0)
    return driver.to_system(ds, s.nfreq), ds


@pytest.mark.parametrize("make", [
    lambda: synth.cell2(),
    lambda: synth.spin_polarised(synth.tiny(natom=3, period=(1, 2, 2), nao=(2, 3), nabf=(3, 4), nocc=(1, 1),
                                            a=6.5, seed=11, nfreq=8)),
])
def test_dataset_files_roundtrip(make, tmp_path):
    s = make()
    r, ds = _roundtrip(s, tmp_path)
    assert r.nspin == s.nspin and r.nk == s.nk and r.natom == s.natom
    assert np.abs(r.eps - s.eps).max() < 1e-15 and np.abs(r.occ - s.occ).max() == 0
    assert abs(r.efermi - s.efermi) < 1e-15
    assert np.abs(r.wfc - s.wfc).max() < 1e-15
    assert np.abs(r.kfrac - s.kfrac).max() < 1e-14
    assert np.abs(r.positions - s.positions).max() < 1e-14 and np.abs(r.latvec - s.latvec).max() == 0
    assert list(r.nao) == list(s.nao) and list(r.nabf) == list(s.nabf) and list(r.types) == list(s.types)
    assert len(r.cs) == len(s.cs)
    for (a, b) in zip(r.cs, s.cs):
        assert a[:3] == b[:3] and np.abs(a[3] - b[3]).max() < 1e-15
    assert np.abs(r.vq - s.vq).max() < 1e-15 * np.abs(s.vq).max()
    assert list(ds["bz"]["ibz"]) == list(range(s.nk))


@pytest.mark.gpu
def test_driver_task_rpa_on_written_files(tmp_path, capsys):
    import librpa_amd as L
    from oracle import rpa
    s = synth.cell2()
    d = str(tmp_path / "cell2")
    driver.write_dataset(s, d)
    assert driver.main(["driver", d + "/librpa.in"]) == 0
    out = capsys.readouterr().out
    line = [x for x in out.splitlines() if x.startswith("| Total EcRPA:")][0]
    ec_file = float(line.split(":")[1])
    h = L.Handler()
    L.load_system(h, s)
    o = L.Options(nfreq=s.nfreq, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI)
    ec_api, _ = h.get_rpa_correlation_energy(o, s.nk)
    grid = h.get_tfgrids(o)
    ec_ref, _ = rpa.rpa_ec(s, grid, rpa.chi0_q_omega(s, grid, rpa.chi0s_all_tau(s, grid)))
    print(f"driver files Ec {ec_file:.9f}, C API {ec_api:.9f}, oracle {ec_ref:.9f}", flush=True)
    assert abs(ec_file - ec_api) < 2e-9  # printed with 9 decimals
    assert abs(ec_api - ec_ref) < 1e-7
