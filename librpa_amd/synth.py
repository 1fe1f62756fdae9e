"""Seeded synthetic inputs for the space-time chi0 -> RPA Ec (and G0W0 Sigma) path.

The generator produces exactly the arrays LibRPA's C API consumes
(``librpa_input.h``): mean-field dimensions, occupations/eigenvalues/Fermi
level, per-k eigenvectors, atomic basis sizes, lattice and k-grid, RI triple
coefficients ``Cs`` per (I, J, R) and the auxiliary Coulomb matrix per k.

Value generators (identical for all workloads):

* occupied eigenvalues ~ U[-0.6, -0.2] Ha, unoccupied ~ U[0.05, 3.0] Ha,
  sorted per k; occupation 2.0 for occupied states (one spin channel);
  Fermi level at mid-gap.
* eigenvectors: the unitary factor Q of QR(complex Gaussian), one per k,
  stored row-major ``nstates x nbasis`` as ``librpa_set_wfc`` expects
  (src/api/input.cpp:101-126).
* ``Cs`` for every ordered atom pair (I, J) and lattice vector R with
  d = |tau_J + R.A - tau_I| <= 12 bohr: N(0, 1) * exp(-d / 2), ABF index on
  atom I, stored as ``Cs_in[(i * n_j + j) * n_mu + mu]`` like
  ``librpa_set_lri_coeff`` (src/api/input.cpp:480-554).  When n_i == n_j the
  block is made symmetric in (i, j) by (X + X^T) / sqrt(2).
* Coulomb: v(k) = B B^H / n_abf + 1e-3 I with B complex Gaussian whose
  column j is scaled by 1 / (1 + j); an independent B per k.
"""
from __future__ import annotations

import dataclasses
import math
import sys
from typing import Dict, List, Optional, Tuple

import numpy as np

CS_CUTOFF = 12.0
# when true, the generator reports each k of the Coulomb matrices on stderr (large workloads)
VERBOSE = False


@dataclasses.dataclass
class Species:
    nao: int
    nabf: int
    nocc: int


SPEC_A = Species(13, 60, 2)
SPEC_B = Species(27, 120, 4)


@dataclasses.dataclass
class System:
    name: str
    latvec: np.ndarray            # (3, 3): row i = lattice vector a_i (bohr)
    positions: np.ndarray         # (natom, 3) Cartesian (bohr)
    types: np.ndarray             # (natom,) species index
    nao: np.ndarray               # (natom,) NAOs per atom
    nabf: np.ndarray              # (natom,) ABFs per atom
    period: Tuple[int, int, int]  # k mesh == BvK supercell
    kfrac: np.ndarray             # (nk, 3) fractional k
    eps: np.ndarray               # (nspin, nk, nbands)
    occ: np.ndarray               # (nspin, nk, nbands) occupations as passed to the API (0..2)
    efermi: float
    wfc: np.ndarray               # (nspin, nk, nbands, nbasis) complex
    cs: List[Tuple[int, int, Tuple[int, int, int], np.ndarray]]
    vq: Optional[np.ndarray]      # (nk, nabf_tot, nabf_tot) complex, Hermitian
    nfreq: int
    nspin: int = 1

    @property
    def natom(self) -> int:
        return len(self.nao)

    @property
    def nk(self) -> int:
        return len(self.kfrac)

    @property
    def nbasis(self) -> int:
        return int(self.nao.sum())

    @property
    def nbands(self) -> int:
        return self.eps.shape[2]

    @property
    def nabf_tot(self) -> int:
        return int(self.nabf.sum())

    @property
    def recvec(self) -> np.ndarray:
        """Rows b_i with a_i . b_j = 2 pi delta_ij (bohr^-1)."""
        return 2 * np.pi * np.linalg.inv(self.latvec).T

    @property
    def kcart(self) -> np.ndarray:
        return self.kfrac @ self.recvec


def _kmesh(period):
    n1, n2, n3 = period
    ks = []
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                ks.append((i / n1, j / n2, k / n3))
    return np.array(ks, dtype=np.float64)


def _random_unitary(rng, n):
    z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))[None, :]
    return q


def _cs_pairs(latvec, pos, cutoff=CS_CUTOFF):
    """All (I, J, R) with |pos_J + R.A - pos_I| <= cutoff, ordered by (I, J, R)."""
    nat = len(pos)
    lens = np.linalg.norm(latvec, axis=1)
    # generous search box; the distance test below is exact
    rmax = [int(math.ceil((cutoff + np.ptp(pos[:, a]) + 1.0) / lens[a])) + 1 for a in range(3)]
    rng_R = [range(-rmax[a], rmax[a] + 1) for a in range(3)]
    Rs = np.array([(x, y, z) for x in rng_R[0] for y in rng_R[1] for z in rng_R[2]], dtype=np.int64)
    shifts = Rs @ latvec
    out = []
    for I in range(nat):
        for J in range(nat):
            d = np.linalg.norm(pos[J][None, :] + shifts - pos[I][None, :], axis=1)
            for idx in np.nonzero(d <= cutoff)[0]:
                out.append((I, J, tuple(int(v) for v in Rs[idx]), float(d[idx])))
    return out


def build_system(name: str, latvec, positions, types, species: List[Species], period, nfreq,
                 seed: int, with_vq: bool = True, vq_device: Optional[str] = None,
                 real_gamma_v: bool = False) -> System:
    """``real_gamma_v`` keeps only Re v(k) at k = 0 (a time-reversal-symmetric crystal has a real
    Coulomb matrix there; the G0W0 path needs it)."""
    latvec = np.asarray(latvec, dtype=np.float64)
    positions = np.asarray(positions, dtype=np.float64)
    types = np.asarray(types, dtype=np.int64)
    nao = np.array([species[t].nao for t in types], dtype=np.int64)
    nabf = np.array([species[t].nabf for t in types], dtype=np.int64)
    nocc = int(sum(species[t].nocc for t in types))
    nbasis = int(nao.sum())
    nbands = nbasis
    kfrac = _kmesh(period)
    nk = len(kfrac)
    root = np.random.SeedSequence(seed)
    s_eps, s_wfc, s_cs, s_vq = root.spawn(4)

    rng = np.random.default_rng(s_eps)
    eps = np.empty((1, nk, nbands))
    occ = np.zeros((1, nk, nbands))
    for ik in range(nk):
        e_o = np.sort(rng.uniform(-0.6, -0.2, nocc))
        e_u = np.sort(rng.uniform(0.05, 3.0, nbands - nocc))
        eps[0, ik] = np.concatenate([e_o, e_u])
        occ[0, ik, :nocc] = 2.0
    homo = eps[0, :, nocc - 1].max()
    lumo = eps[0, :, nocc].min()
    efermi = 0.5 * (homo + lumo)

    rng = np.random.default_rng(s_wfc)
    wfc = np.empty((1, nk, nbands, nbasis), dtype=np.complex128)
    for ik in range(nk):
        q = _random_unitary(rng, nbasis)
        wfc[0, ik] = q.T  # rows = states

    rng = np.random.default_rng(s_cs)
    cs = []
    for (I, J, R, d) in _cs_pairs(latvec, positions):
        ni, nj, nmu = int(nao[I]), int(nao[J]), int(nabf[I])
        x = rng.standard_normal((ni, nj, nmu))
        if ni == nj:
            x = (x + x.transpose(1, 0, 2)) / math.sqrt(2.0)
        blk = (x * math.exp(-d / 2.0)).reshape(ni * nj, nmu)
        cs.append((I, J, R, np.ascontiguousarray(blk)))

    vq = None
    if with_vq:
        n = int(nabf.sum())
        vq = np.empty((nk, n, n), dtype=np.complex128)
        rng = np.random.default_rng(s_vq)
        scale = 1.0 / (1.0 + np.arange(n))
        for ik in range(nk):
            b = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * scale[None, :]
            vq[ik] = _bbh(b, vq_device, n, 1e-3)
            if real_gamma_v and np.all(kfrac[ik] == 0):
                vq[ik] = vq[ik].real.astype(np.complex128)
            if VERBOSE:
                print(f"[synth] {name}: Coulomb matrix of k {ik + 1} / {nk}", file=sys.stderr, flush=True)
    return System(name=name, latvec=latvec, positions=positions, types=types, nao=nao, nabf=nabf,
                  period=tuple(period), kfrac=kfrac, eps=eps, occ=occ, efermi=float(efermi), wfc=wfc,
                  cs=cs, vq=vq, nfreq=nfreq)


def _bbh(b, device, inv=1.0, shift=0.0):
    """(0.5 (B B^H + (B B^H)^H)) / inv + shift * I, in this order of operations on both paths."""
    if device is not None:
        import torch
        t = torch.from_numpy(b).to(device)
        r = t @ t.conj().T
        r = 0.5 * (r + r.conj().T)
        r = r / inv
        r.diagonal().add_(shift)
        out = r.cpu().numpy()
        del t, r
        return out
    r = b @ b.conj().T
    r = 0.5 * (r + r.conj().T)
    r = r / inv
    r[np.diag_indices_from(r)] += shift
    return r


# ----------------------------------------------------------------------------------------------
# The five workloads.
# ----------------------------------------------------------------------------------------------

def cell2(seed=0, with_vq=True, period=(2, 2, 2), nfreq=16, nao=13, nabf=60, nocc_per_atom=2):
    """2-atom cubic cell, a = 10.26 bohr, atoms at 0 and a/4 (1,1,1)."""
    a = 10.26
    sp = [Species(nao, nabf, nocc_per_atom)]
    return build_system("cell2", np.eye(3) * a, [[0, 0, 0], [a / 4, a / 4, a / 4]], [0, 0], sp,
                        period, nfreq, seed, with_vq)


def _jittered(nx, ny, nz, spacing, jitter, rng):
    pts = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                pts.append((i * spacing, j * spacing, k * spacing))
    pts = np.array(pts, dtype=np.float64)
    return pts + rng.uniform(-jitter, jitter, pts.shape)


def cell16(seed=1, with_vq=True, period=(4, 4, 4), nfreq=24):
    """2x2x2 supercell of the 2-atom cell (L = 20.52 bohr), +-0.2 bohr jitter."""
    a = 10.26
    rng = np.random.default_rng(seed + 1000)
    base = np.array([[0, 0, 0], [a / 4, a / 4, a / 4]])
    pos = []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                for b in base:
                    pos.append(b + a * np.array([i, j, k]))
    pos = np.array(pos) + rng.uniform(-0.2, 0.2, (16, 3))
    return build_system("cell16", np.eye(3) * 2 * a, pos, [0] * 16, [SPEC_A], period, nfreq, seed,
                        with_vq)


def cubic128(seed=2, with_vq=True, nfreq=16, vq_device=None):
    """4x4x8 simple-cubic lattice, spacing 4.5 bohr, +-0.3 jitter, species 50/50, Gamma only."""
    rng = np.random.default_rng(seed + 1000)
    pos = _jittered(4, 4, 8, 4.5, 0.3, rng)
    types = rng.permutation(np.array([0] * 64 + [1] * 64))
    return build_system("cubic128", np.diag([18.0, 18.0, 36.0]), pos, types, [SPEC_A, SPEC_B],
                        (1, 1, 1), nfreq, seed, with_vq, vq_device)


def slab256(seed=3, with_vq=True, nfreq=16, vq_device=None):
    """8x8x4 slab, spacing 4.5 bohr, 20 bohr vacuum along z, 2x2x1 k."""
    rng = np.random.default_rng(seed + 1000)
    pos = _jittered(8, 8, 4, 4.5, 0.3, rng)
    types = rng.permutation(np.array([0] * 128 + [1] * 128))
    return build_system("slab256", np.diag([36.0, 36.0, 3 * 4.5 + 20.0]), pos, types,
                        [SPEC_A, SPEC_B], (2, 2, 1), nfreq, seed, with_vq, vq_device)


def cubic64(seed=4, with_vq=True, nfreq=24, vq_device=None):
    """4x4x4 jittered lattice, spacing 4.5 bohr, species 50/50, 2x2x2 k (G0W0 workload; v(0) real)."""
    rng = np.random.default_rng(seed + 1000)
    pos = _jittered(4, 4, 4, 4.5, 0.3, rng)
    types = rng.permutation(np.array([0] * 32 + [1] * 32))
    return build_system("cubic64", np.diag([18.0, 18.0, 18.0]), pos, types, [SPEC_A, SPEC_B],
                        (2, 2, 2), nfreq, seed, with_vq, vq_device, real_gamma_v=True)


def tiny(seed=7, natom=2, period=(2, 1, 1), nao=(3, 2), nabf=(4, 3), nocc=(1, 1), a=6.0,
         nfreq=6, with_vq=True, real_gamma_v=False):
    """A few-function system for brute-force checks (mixed species, k points)."""
    species = [Species(nao[i % len(nao)], nabf[i % len(nabf)], nocc[i % len(nocc)])
               for i in range(natom)]
    rng = np.random.default_rng(seed + 1000)
    pos = rng.uniform(0, a, (natom, 3))
    lat = np.eye(3) * a
    return build_system("tiny", lat, pos, list(range(natom)), species, period, nfreq, seed, with_vq,
                        real_gamma_v=real_gamma_v)


def spin_polarised(s: System, seed: int = 0, shift: float = 0.03) -> System:
    """The same system with two spin channels: spin 0 keeps the bands of ``s``; spin 1 has band
    energies moved by U(-shift, shift) (re-sorted, the gap kept) and new random unitary
    eigenvectors.  Occupations are 1 per occupied state and spin (the API's range for nspin = 2)."""
    rng = np.random.default_rng(seed + 77)
    nocc = int(round(s.occ[0, 0].sum() / 2.0))
    nk, nb = s.eps.shape[1], s.eps.shape[2]
    eps1 = np.sort(s.eps[0] + rng.uniform(-shift, shift, s.eps[0].shape), axis=1)
    wfc1 = np.empty_like(s.wfc[0])
    for ik in range(nk):
        wfc1[ik] = _random_unitary(rng, s.wfc.shape[3]).T
    eps = np.stack([s.eps[0], eps1])
    occ = np.zeros((2, nk, nb))
    occ[:, :, :nocc] = 1.0
    homo = eps[:, :, nocc - 1].max()
    lumo = eps[:, :, nocc].min()
    if not homo < lumo:
        raise ValueError("spin_polarised: the shift closes the gap")
    return dataclasses.replace(s, name=s.name + "_spin", eps=eps, occ=occ, wfc=np.stack([s.wfc[0], wfc1]),
                               efermi=float(0.5 * (homo + lumo)), nspin=2)


def _minus_k(kfrac, period):
    """Index of -k (mod 1) for each k of a Gamma-centred mesh."""
    idx = {tuple(int(round(v * p)) % p for v, p in zip(k, period)): i for i, k in enumerate(kfrac)}
    return [idx[tuple((-int(round(v * p))) % p for v, p in zip(k, period))] for k in kfrac]


def time_reversal_symmetric(s: System, seed: int = 0) -> System:
    """The same system with time-reversal-symmetric k-space inputs: for every pair (k, -k)
    eps(-k) = eps(k), U(-k) = conj U(k) and v(-k) = conj v(k); at the k points equal to their own
    -k (components 0 or 1/2) U and v are real.  G(R) and the BvK-supercell Coulomb matrix are then
    real, so the Gamma-only path of the supercell (real parts only) loses nothing."""
    if s.nspin != 1:
        raise ValueError("time_reversal_symmetric: one spin channel")
    rng = np.random.default_rng(seed + 91)
    mk = _minus_k(s.kfrac, s.period)
    eps, occ, wfc = s.eps.copy(), s.occ.copy(), s.wfc.copy()
    vq = None if s.vq is None else s.vq.copy()
    nb = s.wfc.shape[3]
    for ik, jk in enumerate(mk):
        if jk < ik:
            continue
        if jk == ik:
            q, r = np.linalg.qr(rng.standard_normal((nb, nb)))
            wfc[0, ik] = (q * np.sign(np.diag(r))[None, :]).T
            if vq is not None:
                vq[ik] = vq[ik].real.astype(np.complex128)
        else:
            eps[0, jk], occ[0, jk] = eps[0, ik], occ[0, ik]
            wfc[0, jk] = wfc[0, ik].conj()
            if vq is not None:
                vq[jk] = vq[ik].conj()
    return dataclasses.replace(s, name=s.name + "_tr", eps=eps, occ=occ, wfc=wfc, vq=vq)


def bvk_supercell(s: System) -> System:
    """The Gamma-only BvK supercell (period x cell) of a k-mesh system, by the Bloch convention
    psi_nk = sum_R e^{i 2 pi k.R} sum_i U_k(n, i) phi_i(r - R) / sqrt(N) for the eigenvectors and
    chi(q) = sum_R e^{i 2 pi q.R} chi(0, R) for auxiliary-basis matrices:

    * supercell atom (r, I) sits at tau_I + R_r . A, R_r on the BvK grid (``r_grid`` order);
    * eigenvector of state (k, n) on basis function (r, I, i): e^{i 2 pi k.R_r} U_k(n, (I, i)) / sqrt(N);
      states sorted by energy;
    * Cs block (I, K, R1) becomes ((r, I), (r', K), L) with R_r + R1 = R_r' + period * L;
    * Coulomb matrix v((r, I, mu), (r', J, nu)) = (1 / N) sum_q e^{-i 2 pi q.(R_r' - R_r)} v(q)_{(I,mu),(J,nu)}.

    The RPA correlation energy of the supercell is then N times the per-cell Ec of the k mesh."""
    if s.nspin != 1:
        raise ValueError("bvk_supercell: one spin channel")
    per = np.array(s.period)
    N = int(per.prod())
    cells = [(x, y, z) for x in range(-(per[0] // 2), (per[0] - 1) // 2 + 1)
             for y in range(-(per[1] // 2), (per[1] - 1) // 2 + 1)
             for z in range(-(per[2] // 2), (per[2] - 1) // 2 + 1)]
    cidx = {c: i for i, c in enumerate(cells)}
    na = s.natom
    lat = s.latvec
    pos = np.concatenate([s.positions + np.array(c, dtype=float) @ lat for c in cells])
    types = np.concatenate([s.types for _ in cells])
    nao = np.concatenate([s.nao for _ in cells])
    nabf = np.concatenate([s.nabf for _ in cells])
    ph_k = np.exp(2j * np.pi * (s.kfrac @ np.array(cells, dtype=float).T))   # (nk, N)
    nbasis = s.nbasis
    # states (k, n) on basis (r, b)
    Z = (ph_k[:, None, :, None] * s.wfc[0][:, :, None, :]).reshape(s.nk * s.nbands, N * nbasis) / np.sqrt(N)
    e = s.eps[0].reshape(-1)
    o = s.occ[0].reshape(-1)
    order = np.lexsort((np.arange(e.size), e))
    eps = e[order][None, None, :]
    occ = o[order][None, None, :]
    wfc = Z[order][None, None]

    def fold(v):
        out = []
        for a in range(3):
            p = int(per[a])
            f = ((v[a] % p) + (3 * p) // 2) % p - p // 2
            out.append((f, (v[a] - f) // p))
        return tuple(x[0] for x in out), tuple(x[1] for x in out)

    cs = []
    for (I, K, R1, blk) in s.cs:
        for r, c in enumerate(cells):
            f, L = fold([c[a] + R1[a] for a in range(3)])
            cs.append((r * na + I, cidx[f] * na + K, L, blk.copy()))
    vq = None
    if s.vq is not None:
        n = s.nabf_tot
        ph_q = np.exp(-2j * np.pi * (s.kfrac @ np.array(cells, dtype=float).T))  # (nq, N)
        # V[(r, mu), (r', nu)] = (1/N) sum_q e^{-i 2 pi q.(R_r' - R_r)} v(q)
        V = np.einsum('qa,qb,qmn->ambn', ph_q.conj(), ph_q, s.vq).reshape(N * n, N * n) / N
        if np.abs(V.imag).max() > 1e-10 * np.abs(V.real).max():
            raise ValueError("bvk_supercell: v(q) is not time-reversal symmetric (complex supercell v)")
        vq = V.real.astype(np.complex128)[None]
    return System(name=s.name + "_bvk", latvec=lat * per[:, None], positions=pos, types=types, nao=nao,
                  nabf=nabf, period=(1, 1, 1), kfrac=np.zeros((1, 3)), eps=eps, occ=occ, efermi=s.efermi,
                  wfc=wfc, cs=cs, vq=vq, nfreq=s.nfreq)


WORKLOADS = {
    "cell2": cell2,
    "cell16": cell16,
    "cubic128": cubic128,
    "slab256": slab256,
    "cubic64": cubic64,
}
