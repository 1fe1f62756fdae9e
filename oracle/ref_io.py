"""I/O of the reference-side checker ``_ref/ref_libri`` (test infrastructure only).

``ref_libri`` runs the reference's own LibRI ``RPA::cal_chi0s`` / ``GW::cal_Sigmas`` on one tau
point (see ref_libri.cpp for the stream format). This module writes its input from a system of
``librpa_amd.synth`` and reads the blocks it returns. Only ``tests/`` and the ``cpu_baseline``
leg of ``bench.py`` use it.
"""
from __future__ import annotations

import os
import shutil
import struct
import subprocess
import tempfile
import time

import numpy as np

from .rpa import Basis, r_grid

REF_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref")
REF_LIBRI = os.path.join(REF_DIR, "ref_libri")


def available(path=REF_LIBRI):
    return os.access(path, os.X_OK)


def _pack_blocks(blocks):
    out = [struct.pack("<i", len(blocks))]
    for (I, J, R, data) in blocks:
        out.append(struct.pack("<5i", I, J, *R))
        out.append(np.ascontiguousarray(data, dtype="<f8").tobytes())
    return b"".join(out)


def dense_blocks(sys, basis, M):
    """Blocks (I, J, R, M[R][I, J]) of a dense per-cell matrix for every atom pair and BvK cell."""
    return [(I, J, R, M[R][basis.sl(I), basis.sl(J)])
            for I in range(sys.natom) for J in range(sys.natom) for R in r_grid(sys.period)]


def cs_blocks(sys, atoms=None):
    """The raw Cs blocks (I, J, R, (mu, i, j) tensor) of ``sys`` as LibRPA hands them to LibRI
    (src/api/input.cpp:525-540), optionally only those whose first atom is in ``atoms``."""
    out = []
    for (I, J, R, blk) in sys.cs:
        if atoms is not None and I not in atoms:
            continue
        ni, nj, nmu = int(sys.nao[I]), int(sys.nao[J]), int(sys.nabf[I])
        out.append((I, J, tuple(int(x) for x in R), blk.reshape(ni, nj, nmu).transpose(2, 0, 1)))
    return out


def write_input(path, sys, cs, second, third):
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", sys.natom, *sys.period))
        f.write(np.asarray(sys.latvec, dtype="<f8").tobytes())
        f.write(np.asarray(sys.positions, dtype="<f8").tobytes())
        f.write(np.asarray(sys.nao, dtype="<i4").tobytes())
        f.write(np.asarray(sys.nabf, dtype="<i4").tobytes())
        for blocks in (cs, second, third):
            f.write(_pack_blocks(blocks))


def read_output(path):
    raw = open(path, "rb").read()
    (count,), off = struct.unpack_from("<i", raw, 0), 4
    out = {}
    for _ in range(count):
        I, J, r0, r1, r2, nr, nc = struct.unpack_from("<7i", raw, off)
        off += 28
        out[(I, J, (r0, r1, r2))] = np.frombuffer(raw, "<f8", nr * nc, off).reshape(nr, nc).copy()
        off += 8 * nr * nc
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    return out


def run_libri(mode, sys, cs, second, third, threads=1, timeout=600, read=True):
    """Run ``ref_libri mode`` and return (blocks or None, seconds spent in the LibRI process)."""
    tmp = tempfile.mkdtemp(prefix="ref_libri_")
    try:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        write_input(fin, sys, cs, second, third)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.perf_counter()
        subprocess.run([REF_LIBRI, mode, fin, fout], check=True, timeout=timeout, env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE)
        dt = time.perf_counter() - t0
        return (read_output(fout) if read else None), dt
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def chi0s(sys, Gp, Gn, atoms=None, threads=1, timeout=600, read=True):
    """LibRI's chi0s at one tau from G(R, +tau) / G(R, -tau) (dense per cell)."""
    ao = Basis(sys.nao)
    return run_libri("chi0s", sys, cs_blocks(sys, atoms), dense_blocks(sys, ao, Gp),
                     dense_blocks(sys, ao, Gn), threads, timeout, read)


def sigmas(sys, G, W, threads=1, timeout=600):
    """LibRI's Sigma blocks at one tau from G(R, tau) of one sign and W(R, tau) (dense per cell)."""
    return run_libri("sigmas", sys, cs_blocks(sys), dense_blocks(sys, Basis(sys.nabf), W),
                     dense_blocks(sys, Basis(sys.nao), G), threads, timeout)[0]
