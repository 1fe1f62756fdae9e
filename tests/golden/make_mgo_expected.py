"""Writes the MgO G0W0 fixtures from LibRPA's regression case g0w0_aims_MgO_libri.

* ``dataset.tar.gz``: the case's input files as the reference holds them
  (regression_tests/testcases/g0w0_aims_MgO_libri/dataset.tar.gz, copied unchanged);
* ``expected.json``: the quasi-particle table the reference printed for it
  (regression_tests/refs/g0w0_aims_MgO_libri/librpa/librpa_omp1_np4.out, "Printing
  quasi-particle energy"): per k point, rows of state, occ, e_mf, v_xc, v_exx, ReSigc, ImSigc,
  e_qp in eV, and the run's options (librpa.in of the test case).

Run once in a checkout that has the reference at /root/reference.
"""
import json
import os
import re
import shutil

REF = "/root/reference/regression_tests"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mgo_g0w0")


def main():
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(os.path.join(REF, "testcases", "g0w0_aims_MgO_libri", "dataset.tar.gz"),
                    os.path.join(OUT, "dataset.tar.gz"))
    txt = open(os.path.join(REF, "refs", "g0w0_aims_MgO_libri", "librpa", "librpa_omp1_np4.out")).read()
    tables = {}
    for blk in re.split(r"spin\s+1, k-point\s+", txt)[1:]:
        ik = int(blk.split(":")[0])
        rows = []
        for line in blk.splitlines()[4:]:
            t = line.split()
            if len(t) != 8:
                break
            rows.append([int(t[0])] + [float(x) for x in t[1:]])
        tables[ik] = rows
    assert len(tables) == 27 and all(len(r) == 17 for r in tables.values())
    exp = {
        "source": "regression_tests/refs/g0w0_aims_MgO_libri/librpa/librpa_omp1_np4.out",
        "ha2ev": 27.211386245988,
        "options": {"nfreq": 6, "option_dielect_func": 0, "replace_w_head": True, "n_params_anacon": -1,
                    "cs_threshold": 1e-6},
        "columns": ["state", "occ", "e_mf", "v_xc", "v_exx", "ReSigc", "ImSigc", "e_qp"],
        "kpoints": [tables[k] for k in range(1, 28)],
    }
    json.dump(exp, open(os.path.join(OUT, "expected.json"), "w"), indent=0)


if __name__ == "__main__":
    main()
