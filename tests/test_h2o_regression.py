"""End-to-end check on LibRPA's own regression case rpa_aims_mole_H2O_libri.

Inputs are the FHI-aims H2O files of the reference's regression data (binary Cs and Coulomb
matrices, ASCII eigenvectors) in tests/golden/h2o_rpa; the expected Ec is the reference's
output, librpa.out:273, printed to 9 decimals (the reference's own test allows 1e-4).

* CPU: the files parse to the right dimensions and the FP64 oracle reproduces the reference
  Ec to 1e-9 Ha (the print precision), which pins the whole oracle chain (minimax grid, G,
  chi0, transforms, log-det) on real data.
* GPU: the library, fed through the driver's setter sequence, reproduces the same Ec.
"""
import json
import os

import numpy as np
import pytest

from librpa_amd import driver

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h2o_rpa")
EXPECTED = json.load(open(os.path.join(HERE, "expected.json")))


def _ds():
    return driver.read_dataset(HERE + "/", 1e-6)


def test_h2o_files_parse():
    ds = _ds()
    b = ds["band"]
    assert (b["nk"], b["nspin"], b["nstates"], b["nbasis"]) == (1, 1, 24, 24)
    assert list(ds["nao"]) == [14, 5, 5] and list(ds["nabf"]) == [89, 18, 18]
    assert ds["vq"].shape == (1, 125, 125)
    # eigenvectors of a non-orthogonal NAO basis: finite, full rank
    U = ds["wfc"][0, 0]
    assert np.all(np.isfinite(U)) and np.linalg.matrix_rank(U) == 24
    assert len(ds["cs"]) > 0


def test_h2o_oracle_matches_reference_ec():
    from oracle import rpa
    s = driver.to_system(_ds(), 16)
    r = rpa.rpa_from_system(s)
    assert abs(r["ec"] - EXPECTED["ec_ha"]) < 1e-9, r["ec"]


@pytest.mark.gpu
def test_h2o_library_matches_reference_ec():
    import librpa_amd as L
    from oracle import rpa
    ds = _ds()
    h = L.Handler()
    driver.push(h, ds, L.LIBRI)
    o = L.Options(nfreq=16, tfgrids_type=L.MINIMAX, parallel_routing=L.LIBRI)
    ec, contrib = h.get_rpa_correlation_energy(o, 1)
    print(f"H2O Ec library {ec!r} reference {EXPECTED['ec_ha']}", flush=True)
    assert abs(ec - EXPECTED["ec_ha"]) < 1e-9
    ref = rpa.rpa_from_system(driver.to_system(ds, 16))["ec"]
    assert abs(ec - ref) < 1e-11
