"""FP64 oracle: G(R, +-tau), chi0(R, tau), tau->omega / R->q transforms, RPA Ec.

Restates LibRPA's LibRI-routing path:

* ``gf_R`` : ``build_gf_Rt_libri_serial`` (src/core/chi0.cpp:303-426), occupation
  weights for +-tau (:333-355), real part kept (:414), sign flip for tau < 0 (:392).
* ``fold_cs``: LibRI ``RI_Tools::cal_period`` sums Cs images that map to the same
  BvK cell (thirdparty/LibRI/include/RI/ri/RI_Tools.hpp:48-67) with the symmetric
  modulo of ``Array_Operator::operator%`` (global/Array_Operator.hpp:16-25).
* ``chi0s_bruteforce``: the two ``LRI::cal_loop3`` labels a1b1_a2b2 and a1b2_a2b1
  called twice with G+/G- swapped (physics/RPA.hpp:104-138), written as the
  6-index sums of unittests/ri/LRI-loop3-test.hpp:353-393.
* ``chi0s_factorized``: the same quantity through W(K,J,S) = sum_L G(S+R2) C^J and
  T = G_IJ W_KJ + G_KJ W_IJ, then C^I^T T (the factorisation the GPU path uses).
* ``chi0_q_omega``: ``chi_libri_ft_ct`` (src/core/chi0.cpp:507-596):
  chi0(q, w) += (2/N_s) gamma_c(w, tau) e^{i 2 pi q.R} chi0s(R, tau), I <= J blocks.
* ``rpa_ec``: ``compute_RPA_correlation_blacs_2d{,_gamma_only}``
  (src/core/epsilon.cpp:43-558): Hermitian fill from upper blocks
  (src/math/utils_matrix_m_mpi.h:430-441), Gamma path on real parts (:106,163),
  Ec = sum_q w_q sum_w w_w / 2pi [Tr Pi + ln|det(1 - Pi)|], Pi = V chi0.

Pins (tests/test_oracle_rpa.py): G(tau -> 0-) = -D (src/test/test_meanfield.cpp:43-56)
and an explicit per-k double loop; brute force vs factorised chi0 on mixed-species
k-point systems; transforms vs explicit loops; ln det vs eigenvalue sum; Ec <= 0.
tests/test_oracle_libri.py: every block of ``chi0s_bruteforce`` equals the reference's own
LibRI ``RPA::cal_chi0s`` (oracle/ref_libri.cpp, built from the reference tree by
oracle/Makefile) on k-point systems with periods up to 3 x 2 x 1, to 1e-12 relative.
"""
from __future__ import annotations

import itertools
from typing import Dict, List, Tuple

import numpy as np


def bvk_mod(v, n):
    """LibRI/LibRPA symmetric modulo, result in [-n/2, (n-1)/2]."""
    return ((v % n) + (3 * n) // 2) % n - n // 2


def fold_R(R, period):
    return tuple(int(bvk_mod(R[a], period[a])) for a in range(3))


def r_grid(period):
    """construct_R_grid (src/core/pbc.cpp:178-190)."""
    out = []
    for x in range(-(period[0] // 2), (period[0] - 1) // 2 + 1):
        for y in range(-(period[1] // 2), (period[1] - 1) // 2 + 1):
            for z in range(-(period[2] // 2), (period[2] - 1) // 2 + 1):
                out.append((x, y, z))
    return out


class Basis:
    def __init__(self, sizes):
        self.sizes = np.asarray(sizes, dtype=np.int64)
        self.off = np.concatenate([[0], np.cumsum(self.sizes)])

    def sl(self, a):
        return slice(int(self.off[a]), int(self.off[a + 1]))

    @property
    def total(self):
        return int(self.off[-1])


def fold_cs(sys) -> Dict[Tuple[int, int, Tuple[int, int, int]], np.ndarray]:
    """Cs blocks folded to the BvK cell, as (mu, i, k) tensors (input.cpp:528-540)."""
    out = {}
    for (I, J, R, blk) in sys.cs:
        ni, nj, nmu = int(sys.nao[I]), int(sys.nao[J]), int(sys.nabf[I])
        t = blk.reshape(ni, nj, nmu).transpose(2, 0, 1)
        key = (I, J, fold_R(R, sys.period))
        if key in out:
            out[key] = out[key] + t
        else:
            out[key] = t.copy()
    return out


def occ_weights(sys, tau, nbands_g=-1, ispin=0):
    """Band scale factors s_{n,k}(tau) of spin ispin (chi0.cpp:333-355)."""
    nk = sys.nk
    wg = sys.occ[ispin] / nk
    if tau > 0:
        w = 1.0 / nk - wg * sys.nspin / 2.0
        w = np.where(w < 0, 0.0, w)
    else:
        w = wg * 0.5 * sys.nspin
    sc = -tau * (sys.eps[ispin] - sys.efermi)
    sc = np.where(sc > 0, 0.0, sc)
    s = np.exp(sc) * w
    if nbands_g >= 0:
        s[:, nbands_g:] = 0.0
    return s


def gf_R(sys, tau, Rs, nbands_g=-1, ispin=0):
    """Dense G(R, tau) (nbasis x nbasis, real) of spin ispin for each R in Rs."""
    s = occ_weights(sys, tau, nbands_g, ispin)
    U = sys.wfc[ispin]
    M = np.stack([(U[k].T * s[k][None, :]) @ U[k].conj() for k in range(sys.nk)])
    out = {}
    for R in Rs:
        ph = np.exp(-2j * np.pi * (sys.kfrac @ np.asarray(R, dtype=np.float64)))
        g = np.tensordot(ph, M, axes=(0, 0)).real
        out[R] = -g if tau < 0 else g
    return out


def _gblk(G, ao, I, J, R, period):
    return G[fold_R(R, period)][ao.sl(I), ao.sl(J)]


def chi0s_bruteforce(sys, cs, Gp, Gn, pairs, Rs):
    """chi0s[(I, J, R)] (mu x nu) straight from the LibRI label definitions."""
    ao = Basis(sys.nao)
    per = sys.period
    nb = {}
    for (I, K, R1) in cs:
        nb.setdefault(I, []).append((K, R1))
    out = {}
    for (I, J) in pairs:
        for R in Rs:
            acc = np.zeros((int(sys.nabf[I]), int(sys.nabf[J])))
            for (K, R1) in nb.get(I, []):
                a = cs[(I, K, R1)]
                for (L, R2) in nb.get(J, []):
                    b = cs[(J, L, R2)]
                    Rv, R1v, R2v = np.array(R), np.array(R1), np.array(R2)
                    for G1, G2 in ((Gp, Gn), (Gn, Gp)):
                        g_ij = _gblk(G1, ao, I, J, Rv, per)
                        g_kl = _gblk(G2, ao, K, L, Rv + R2v - R1v, per)
                        g_il = _gblk(G1, ao, I, L, Rv + R2v, per)
                        g_kj = _gblk(G2, ao, K, J, Rv - R1v, per)
                        acc += np.einsum('aik,bjl,ij,kl->ab', a, b, g_ij, g_kl, optimize=True)
                        acc += np.einsum('aik,bjl,il,kj->ab', a, b, g_il, g_kj, optimize=True)
            out[(I, J, R)] = acc
    return out


def chi0s_factorized(sys, cs, Gp, Gn, pairs, Rs):
    """Same chi0s through W(K,J,S)[k, j, nu] = sum_{(L,R2) in N(J)} G(S+R2)_KL C^J(L,R2)."""
    ao = Basis(sys.nao)
    per = sys.period
    nb = {}
    for (I, K, R1) in cs:
        nb.setdefault(I, []).append((K, R1))
    Jset = sorted({J for (_, J) in pairs})
    cells = r_grid(per)
    W = {}
    for sgn, G in (('+', Gp), ('-', Gn)):
        for J in Jset:
            for S in cells:
                Sv = np.array(S)
                acc = np.zeros((ao.total, int(sys.nao[J]), int(sys.nabf[J])))
                for (L, R2) in nb.get(J, []):
                    g = G[fold_R(Sv + np.array(R2), per)][:, ao.sl(L)]
                    acc += np.tensordot(g, cs[(J, L, R2)], axes=([1], [2])).transpose(0, 2, 1)
                W[(sgn, J, S)] = acc
    out = {}
    for (I, J) in pairs:
        for R in Rs:
            Rv = np.array(R)
            acc = np.zeros((int(sys.nabf[I]), int(sys.nabf[J])))
            Wp_IJ = W[('+', J, fold_R(Rv, per))][ao.sl(I)]
            Wn_IJ = W[('-', J, fold_R(Rv, per))][ao.sl(I)]
            gp_ij = _gblk(Gp, ao, I, J, Rv, per)
            gn_ij = _gblk(Gn, ao, I, J, Rv, per)
            for (K, R1) in nb.get(I, []):
                S = fold_R(Rv - np.array(R1), per)
                Wp_KJ = W[('+', J, S)][ao.sl(K)]
                Wn_KJ = W[('-', J, S)][ao.sl(K)]
                gp_kj = _gblk(Gp, ao, K, J, Rv - np.array(R1), per)
                gn_kj = _gblk(Gn, ao, K, J, Rv - np.array(R1), per)
                # T[i, k, nu] = sum_j G_IJ[i, j] W_KJ[k, j, nu] + G_KJ[k, j] W_IJ[i, j, nu]
                TA = np.matmul(gp_ij[None], Wn_KJ) + np.matmul(gn_ij[None], Wp_KJ)   # (k, i, nu)
                TB = np.matmul(gn_kj[None], Wp_IJ) + np.matmul(gp_kj[None], Wn_IJ)   # (i, k, nu)
                T = TA.transpose(1, 0, 2) + TB
                c = cs[(I, K, R1)]  # (mu, i, k)
                acc += c.reshape(c.shape[0], -1) @ T.reshape(-1, T.shape[2])
            out[(I, J, R)] = acc
    return out


def upper_pairs(natom):
    return [(I, J) for I in range(natom) for J in range(I, natom)]


def chi0s_all_tau(sys, grid, nbands_g=-1, method='factorized'):
    """chi0s[it][(I, J, R)] for all minimax tau points, I <= J, R on the BvK grid, summed over
    the spin channels (Chi0::build loops over spins into one chi0s, chi0.cpp:57-135; the 2 / N_s
    factor is applied in the transform)."""
    cs = fold_cs(sys)
    Rs = r_grid(sys.period)
    pairs = upper_pairs(sys.natom)
    f = chi0s_factorized if method == 'factorized' else chi0s_bruteforce
    out = []
    for tau in grid['tau']:
        acc = None
        for isp in range(sys.nspin):
            Gp = gf_R(sys, tau, Rs, nbands_g, isp)
            Gn = gf_R(sys, -tau, Rs, nbands_g, isp)
            c = f(sys, cs, Gp, Gn, pairs, Rs)
            acc = c if acc is None else {k: acc[k] + c[k] for k in acc}
        out.append(acc)
    return out


def chi0_q_omega(sys, grid, chi0s):
    """chi0[(iw, iq, I, J)] complex blocks (chi_libri_ft_ct, chi0.cpp:507-596)."""
    nspin = sys.nspin
    out = {}
    for iw in range(len(grid['omega'])):
        for iq, q in enumerate(sys.kfrac):
            for it in range(len(grid['tau'])):
                trans = grid['cos_t2f'][iw, it]
                for (I, J, R), blk in chi0s[it].items():
                    ph = np.exp(2j * np.pi * float(np.dot(q, R)))
                    key = (iw, iq, I, J)
                    out[key] = out.get(key, 0) + (2.0 / nspin) * trans * ph * blk
    return out


def full_matrix(blocks, abf, natom, getter):
    n = abf.total
    M = np.zeros((n, n), dtype=np.complex128)
    for I in range(natom):
        for J in range(I, natom):
            b = getter(I, J)
            M[abf.sl(I), abf.sl(J)] = b
            if J != I:
                M[abf.sl(J), abf.sl(I)] = b.conj().T
    return M


def ln_det_ref(A):
    """ln det via LU diagonal: sum log(+-d) as in compute_pi_det_blacs_2d (epsilon.cpp:600-693)."""
    import scipy.linalg
    lu, piv = scipy.linalg.lu_factor(A)
    d = np.diag(lu)
    if np.isrealobj(d):
        return np.sum(np.log(np.abs(d)))
    return np.sum(np.where(d.real > 0, np.log(d.astype(complex)), np.log(-d.astype(complex))))


def ibz_weights(nk, ibz_map=None):
    """IBZ k list in order of first appearance and weights count / N_k
    (PeriodicBoundaryData::set_ibz_mapping, src/core/pbc.cpp:110-153)."""
    m = list(range(nk)) if ibz_map is None else [int(x) for x in ibz_map]
    ks, w = [], []
    for k in range(nk):
        if m[k] in ks:
            w[ks.index(m[k])] += 1.0 / nk
        else:
            ks.append(m[k])
            w.append(1.0 / nk)
    return ks, w


def rpa_ec(sys, grid, chi0q, vq=None, ibz_map=None):
    """Ec and the contributions of the IBZ k points (all k without a mapping), each weighted by
    its star's share of the k grid (epsilon.cpp:248, map_ibzk_weight)."""
    abf = Basis(sys.nabf)
    vq = sys.vq if vq is None else vq
    gamma_only = sys.nk == 1
    qs, wts = ibz_weights(sys.nk, ibz_map)
    contrib = np.zeros(len(qs), dtype=np.complex128)
    for a, iq in enumerate(qs):
        V = vq[iq]
        for iw in range(len(grid['omega'])):
            chi = full_matrix(chi0q, abf, sys.natom, lambda I, J: chi0q[(iw, iq, I, J)])
            if gamma_only:
                Pi = V.real @ chi.real
            else:
                Pi = V @ chi
            tr = np.trace(Pi)
            ld = ln_det_ref(np.eye(abf.total) - Pi)
            if gamma_only:
                val = complex(tr + ld)
            else:
                val = tr + ld
            contrib[a] += val * grid['omega_w'][iw] * wts[a] / (2 * np.pi)
    return float(contrib.sum().real), contrib


def freq_accumulate(sys, grid, chi0s_tau, it, acc):
    """acc[(I, J, R)][w] += (2/N_s) gamma_c(w, tau_it) chi0s(R, tau_it).

    The cosine half of ``chi_libri_ft_ct`` (src/core/chi0.cpp:507-596) applied one tau at a
    time, so chi0s of all tau need not be held at once (the R -> q half is ``rpa_ec_streamed``).
    """
    col = (2.0 / sys.nspin) * np.asarray(grid['cos_t2f'])[:, it]
    for key, blk in chi0s_tau.items():
        add = col[:, None, None] * blk[None, :, :]
        if key in acc:
            acc[key] += add
        else:
            acc[key] = add.copy()
    return acc


def rpa_ec_streamed(sys, grid, acc, vq=None):
    """``rpa_ec`` of the chi0(q, w) that ``chi0_q_omega`` would build from the same chi0s.

    ``acc`` is the output of ``freq_accumulate`` over all tau.  Per q the R -> q phase sum
    e^{i 2 pi q.R} (chi0.cpp:587-593) forms every omega at once; the Hermitian fill, the Gamma
    real-part rule and the log-det are those of ``rpa_ec``.
    """
    abf = Basis(sys.nabf)
    vq = sys.vq if vq is None else vq
    gamma_only = sys.nk == 1
    nq, nw, n = sys.nk, len(grid['omega']), abf.total
    by_pair: Dict[Tuple[int, int], List] = {}
    for (I, J, R), a in acc.items():
        by_pair.setdefault((I, J), []).append((R, a))
    contrib = np.zeros(nq, dtype=np.complex128)
    for iq in range(nq):
        q = sys.kfrac[iq]
        chi = np.zeros((nw, n, n), dtype=np.complex128)
        for (I, J), lst in by_pair.items():
            blk = 0
            for R, a in lst:
                blk = blk + np.exp(2j * np.pi * float(np.dot(q, R))) * a
            chi[:, abf.sl(I), abf.sl(J)] = blk
            if J != I:
                chi[:, abf.sl(J), abf.sl(I)] = np.conj(np.transpose(blk, (0, 2, 1)))
        V = vq[iq]
        for iw in range(nw):
            Pi = V.real @ chi[iw].real if gamma_only else V @ chi[iw]
            val = np.trace(Pi) + ln_det_ref(np.eye(n) - Pi)
            contrib[iq] += complex(val) * grid['omega_w'][iw] * (1.0 / nq) / (2 * np.pi)
    return float(contrib.sum().real), contrib


def vq_upper_blocks(sys, ik):
    """Upper (I <= J) atom-pair blocks of V(k) as the API receives them."""
    abf = Basis(sys.nabf)
    V = sys.vq[ik]
    return {(I, J): V[abf.sl(I), abf.sl(J)] for (I, J) in upper_pairs(sys.natom)}


def rpa_from_system(sys, nbands_g=-1):
    """End-to-end oracle Ec for a System with the reference's minimax grid."""
    from . import tfgrids
    emin, emax = tfgrids.e_min_max(sys.eps, sys.occ, sys.nspin, sys.nk)
    grid = tfgrids.minimax(sys.nfreq, emin, emax)
    chi0s = chi0s_all_tau(sys, grid, nbands_g)
    chi0q = chi0_q_omega(sys, grid, chi0s)
    ec, contrib = rpa_ec(sys, grid, chi0q)
    return dict(grid=grid, chi0s=chi0s, chi0q=chi0q, ec=ec, contrib=contrib)
