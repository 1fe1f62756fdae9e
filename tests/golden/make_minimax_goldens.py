"""Write tests/golden/minimax_goldens.json from LibRPA's regression outputs.

For each regression case that prints its minimax grid, the fixture keeps

* the mean-field eigenvalues and occupations of that case (from the case's
  ``band_out`` inside ``dataset.tar.gz``), which fix e_min / e_max;
* the grid LibRPA printed: frequency/time nodes and weights (16 digits), the
  cosine and sine t->f matrices and the cosine f->t matrix (6 digits).

Run in a tree where the reference regression data are available:
    python tests/golden/make_minimax_goldens.py
"""
import io
import json
import os
import re
import tarfile

REF = '/root/reference/regression_tests'
CASES = [
    ('h2o_rpa', 'rpa_aims_mole_H2O_libri', 'librpa/librpa_omp1_np4.out'),
    ('mgo_g0w0', 'g0w0_aims_MgO_libri', 'librpa/librpa_omp1_np4.out'),
    ('h2_g0w0', 'g0w0_aims_mole_H2_libri', 'librpa/librpa_omp1_np4.out'),
    ('li_g0w0', 'g0w0_aims_Li_atom_libri', 'librpa/librpa_omp1_np1.out'),
]


def parse_band_out(text):
    tok = text.split()
    nk, ns, nb, nbasis = (int(x) for x in tok[:4])
    efermi = float(tok[4])
    pos = 5
    eps = [[[0.0] * nb for _ in range(nk)] for _ in range(ns)]
    occ = [[[0.0] * nb for _ in range(nk)] for _ in range(ns)]
    for _ in range(nk * ns):
        ik, isp = int(tok[pos]) - 1, int(tok[pos + 1]) - 1
        pos += 2
        for b in range(nb):
            occ[isp][ik][b] = float(tok[pos + 1])
            eps[isp][ik][b] = float(tok[pos + 2])
            pos += 4
    return dict(nk=nk, nspin=ns, nbands=nb, nbasis=nbasis, efermi=efermi, eps=eps, occ=occ)


def parse_grid(lines):
    i = next(k for k, l in enumerate(lines) if l.startswith('Grid size:'))
    n = int(lines[i].split(':')[1])

    def nodes(start):
        rows = [lines[start + 1 + r].split() for r in range(n)]
        return [float(r[1]) for r in rows], [float(r[2]) for r in rows]

    jf = next(k for k in range(i, len(lines)) if lines[k].startswith('Frequency node'))
    jt = next(k for k in range(i, len(lines)) if lines[k].startswith('Time node'))
    omega, omega_w = nodes(jf)
    tau, tau_w = nodes(jt)
    mats = []
    k = jt
    while len(mats) < 3:
        if lines[k].startswith('nr = '):
            mats.append([[float(x) for x in lines[k + 1 + r].split()] for r in range(n)])
            k += n
        k += 1
    return dict(n=n, omega=omega, omega_w=omega_w, tau=tau, tau_w=tau_w,
                cos_t2f=mats[0], sin_t2f=mats[1], cos_f2t=mats[2])


def main():
    out = {}
    for name, case, log in CASES:
        with tarfile.open(os.path.join(REF, 'testcases', case, 'dataset.tar.gz')) as tf:
            member = next(m for m in tf.getmembers() if m.name.endswith('band_out'))
            band = parse_band_out(tf.extractfile(member).read().decode())
        lines = open(os.path.join(REF, 'refs', case, log)).read().splitlines()
        rec = dict(case=case, log=log, **parse_grid(lines), **band)
        out[name] = rec
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'minimax_goldens.json')
    with open(path, 'w') as f:
        json.dump(out, f)


if __name__ == '__main__':
    main()
