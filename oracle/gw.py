"""FP64 oracle for the G0W0 correlation self-energy path.

Restates LibRPA's LibRI-routing G0W0 chain:

* ``wc_q_omega``: screened Coulomb interaction, ``compute_Wc_freq_q_blacs``
  (src/core/epsilon.cpp:1763-2321; plain twin compute_Wc_freq_q :1576-1761):
  Wc(q, w) = v_c^{1/2} [(1 - v^{1/2} chi0 v^{1/2})^{-1} - 1] v_c^{1/2}, square roots by
  eigen-decomposition with eigenvalues below ``sqrt_coulomb_threshold`` (default 0)
  dropped (power_hemat, src/math/matrix_m.h:916-970).  At q = 0 the square root of v is
  taken from Re v (power_hemat_blacs_real, src/math/utils_matrix_m_mpi.h:886-990,
  called at epsilon.cpp:2019-2024); v_c keeps its full complex form.
  With a macroscopic dielectric function (replace_w_head, option_dielect_func 0), the head
  of the dielectric matrix at q = 0 is overwritten in the eigenbasis of Re v
  (``eps_head_gamma``, epsilon.cpp:2106-2181).
* ``w_r_tau``: ``CT_FT_Wc_freq_q`` (epsilon.cpp:2490-2666) with ``FT_Wc_freq_q``
  (:2323-2489): W(R, w) = (1/N_k) sum_q e^{-i 2 pi q.R} Wc(q, w) (N_k = 1: plain copy),
  W(R, tau) = sum_w gamma_f2t(tau, w) W(R, w); the real part of each atom-pair block
  is what the real-tensor path hands to LibRI (src/core/gw.cpp:366-375).  With an IBZ
  mapping W is formed at the irreducible q only; a mapped q is taken as the time-reversal
  partner -q with W(-q) = conj W(q) (epsilon.cpp:2430-2452).
* ``sigma_labels``: ``RI::GW::cal_Sigmas`` (thirdparty/LibRI/include/RI/physics/GW.hpp:
  102-122), i.e. ``LRI::cal_loop3`` with labels a0b0_a1b1, a0b0_a1b2, a0b0_a2b1,
  a0b0_a2b2 (ri/LRI-cal_loop3.hpp:57-..): with C^I(K,Ra)(mu, i, k), W_IJ(R)(mu, nu),
  C^J(L,Rb)(nu, j, l) and G of one sign,
    a0b0_a2b2: Sigma_IJ(R)(i, j)         += C W G_KL(R+Rb-Ra)(k, l) C
    a0b0_a1b1: Sigma_KL(R+Rb-Ra)(k, l)   += C W G_IJ(R)(i, j) C
    a0b0_a1b2: Sigma_KJ(R-Ra)(k, j)      += C W G_IL(R+Rb)(i, l) C
    a0b0_a2b1: Sigma_IL(R+Rb)(i, l)      += C W G_KJ(R-Ra)(k, j) C
  all cells folded to the BvK supercell.
* ``sigma_supercell``: the same Sigma as one dense contraction over the BvK supercell,
  Sigma(a, b) = sum B[mu, a, c] W[mu, nu] B[nu, b, d] G(c, d) with the symmetric
  product tensor B[mu, a, c] = C(mu, a, c) + C(mu, c, a) (an independent restatement
  used only to pin ``sigma_labels``).
* ``sigma_factorized``: the factorisation the GPU path uses (U_J(S) = sum_L G C^J,
  T = W U, Y = W G C^J, mirror of the a0b0_a1b2 term for a0b0_a2b1).
* ``sigma_freq``: the cos/sin fold, gw.cpp:681-797:
  Sigma(R, i w) += 0.5 (S+ + S-) gamma_c(w, tau) + i 0.5 (S+ - S-) gamma_s(w, tau).
* ``sigma_ks``: build_sigc_matrix_KS_blacs (gw.cpp:884-1240): nearest-image cell
  R_bvk = argmin over R + period*(-1..1)^3 of |(f_I - f_J - R_bvk) A|, then
  Sigma(k) = sum_R e^{+i 2 pi k.R_bvk} Sigma_IJ(R) and Sigma_KS = conj(U) Sigma(k) U^T.
* ``Pade`` / ``qpe_solve``: AnalyContPade (src/core/analycont.cpp:17-80) and
  qpe_solver_pade_self_consistent (src/core/qpe_solver.cpp:10-55).

Pins (tests/test_oracle_gw.py): W against the closed form v (1 - chi0 v)^{-1} - v;
``sigma_labels`` against ``sigma_supercell`` (with a planted index mistake that
fails); ``sigma_factorized`` against both; the transpose symmetry
Sigma_XY(R) = Sigma_YX(-R)^T; Pade against -1/(x - x0) (src/test/test_analycont.cpp:
15-46); the QP solver against the fixed point of a pole model solved in closed form.
tests/test_oracle_libri.py: every block of ``sigma_labels`` equals the reference's own LibRI
``GW::cal_Sigmas`` (oracle/ref_libri.cpp) for both signs of tau on k-point systems, to 1e-12
relative, and the planted a1b1 cell mistake fails it.
tests/test_oracle_ref_qpe.py: ``Pade`` and ``qpe_solve`` equal the reference's own AnalyContPade and
qpe_solver_pade_self_consistent (oracle/ref_qpe.cpp, compiled from the reference's sources) to
1e-10 for every kind of parameter count, including the subsampled fit's last node 0.
"""
from __future__ import annotations

import numpy as np

from .rpa import Basis, fold_R, fold_cs, full_matrix, gf_R, r_grid


# ---------------------------------------------------------------------------------------
# W(q, omega) and W(R, tau)
# ---------------------------------------------------------------------------------------

def power_hemat(A, power, threshold=0.0):
    """Hermitian matrix power with eigenvalues below ``threshold`` dropped (matrix_m.h:916-970)."""
    w, v = np.linalg.eigh(A)
    wp = np.where(w < threshold, 0.0, w)
    wp = np.where(wp > 0, wp ** power, 0.0)
    return (v * wp[None, :]) @ v.conj().T


def is_gamma(q):
    return bool(np.all(np.abs(np.asarray(q)) < 1e-12))


def eps_head_gamma(V, chi, eps_mac, threshold=0.0):
    """Dielectric matrix at q = 0 with its head replaced by the macroscopic dielectric function
    (compute_Wc_freq_q_blacs, epsilon.cpp:2106-2181, option_dielect_func != 3): with the
    eigen-decomposition Re v = U diag(lam) U^T in descending order and the columns with
    lam >= threshold kept, M = (U lam^{1/2})^T chi (U lam^{1/2}); M[0, 0] is set to 1 - eps_mac,
    and eps = 1 - U M U^T."""
    w, U = np.linalg.eigh(V.real)
    w, U = w[::-1], U[:, ::-1]
    keep = int(np.count_nonzero(w >= threshold))
    Un = U[:, :keep].astype(np.complex128)
    S = Un * np.sqrt(w[:keep])[None, :]
    M = S.conj().T @ chi @ S
    M[0, 0] = 1.0 - eps_mac
    return np.eye(V.shape[0]) - Un @ M @ Un.conj().T


def ibz_partners(sys, ibz):
    """Irreducible q of every q of the full grid and whether q is its time-reversal partner, as
    FT_Wc_freq_q assumes (epsilon.cpp:2430-2452: a mapped q that is not the irreducible one is
    taken to be -q, and W(-q) = conj W(q))."""
    out = []
    for iq, q in enumerate(sys.kfrac):
        r = int(ibz[iq])
        d = np.asarray(q) - np.asarray(sys.kfrac[r])
        s = np.asarray(q) + np.asarray(sys.kfrac[r])
        if np.all(np.abs(d - np.round(d)) < 1e-8):
            out.append((r, False))
        elif np.all(np.abs(s - np.round(s)) < 1e-8):
            out.append((r, True))
        else:
            raise ValueError(f"q {iq} maps to {r}, which is neither q nor -q")
    return out


def wc_q_omega(sys, grid, chi0q, vq=None, vq_cut=None, threshold=0.0, head=None, ibz=None):
    """Wc[(iw, iq)] dense complex n_abf x n_abf; with ``ibz`` (irreducible index of each q) only
    the irreducible q are computed; ``head`` (eps_mac per omega) replaces the head of the
    dielectric matrix at q = 0."""
    abf = Basis(sys.nabf)
    vq = sys.vq if vq is None else vq
    vq_cut = vq if vq_cut is None else vq_cut
    n = abf.total
    out = {}
    for iq, q in enumerate(sys.kfrac):
        if ibz is not None and int(ibz[iq]) != iq:
            continue
        V = vq[iq]
        sv = power_hemat(V.real.astype(np.complex128) if is_gamma(q) else V, 0.5, threshold)
        svc = power_hemat(vq_cut[iq], 0.5, threshold)
        for iw in range(len(grid['omega'])):
            chi = full_matrix(chi0q, abf, sys.natom, lambda I, J: chi0q[(iw, iq, I, J)])
            if head is not None and is_gamma(q):
                eps = eps_head_gamma(V, chi, head[iw], threshold)
            else:
                eps = np.eye(n) - sv @ chi @ sv
            out[(iw, iq)] = svc @ (np.linalg.inv(eps) - np.eye(n)) @ svc
    return out


def w_r_tau(sys, grid, wq, ibz=None):
    """W[(it, R)] real dense n_abf x n_abf, R on the BvK grid; with ``ibz`` the q outside the
    irreducible set take conj W of their time-reversal partner."""
    nk = sys.nk
    Rs = r_grid(sys.period)
    nw = len(grid['omega'])
    part = ibz_partners(sys, ibz) if ibz is not None else [(iq, False) for iq in range(nk)]
    out = {}
    for R in Rs:
        wr = []
        for iw in range(nw):
            if nk == 1:
                wr.append(wq[(iw, 0)])
            else:
                acc = 0
                for iq, q in enumerate(sys.kfrac):
                    r, flip = part[iq]
                    wmat = wq[(iw, r)].conj() if flip else wq[(iw, r)]
                    acc = acc + np.exp(-2j * np.pi * float(np.dot(q, R))) * wmat
                wr.append(acc / nk)
        for it in range(len(grid['tau'])):
            acc = 0
            for iw in range(nw):
                acc = acc + grid['cos_f2t'][it, iw] * wr[iw]
            out[(it, R)] = np.real(acc).copy()
    return out


def w_closed_form(V, chi):
    """v (1 - chi v)^{-1} - v: equals wc for positive-definite v with v_c = v."""
    n = V.shape[0]
    return V @ np.linalg.inv(np.eye(n) - chi @ V) - V


# ---------------------------------------------------------------------------------------
# Sigma(R, tau) for one sign of tau
# ---------------------------------------------------------------------------------------

def _neighbours(cs):
    nb = {}
    for (I, K, R1) in cs:
        nb.setdefault(I, []).append((K, R1))
    return nb


def _wblk(Wt, abf, I, J, R, period):
    return Wt[fold_R(R, period)][abf.sl(I), abf.sl(J)]


def _gblk(G, ao, I, J, R, period):
    return G[fold_R(R, period)][ao.sl(I), ao.sl(J)]


def sigma_labels(sys, cs, G, Wt, mistake=None):
    """Sigma[R] dense nbasis x nbasis (rows atom-major AO of the home cell) from the four labels.

    ``G[R]`` dense nbasis^2, ``Wt[R]`` dense n_abf^2 (one sign / one tau).  ``mistake``
    plants a deliberate index error (for the pin test)."""
    ao, abf = Basis(sys.nao), Basis(sys.nabf)
    per = sys.period
    nb = _neighbours(cs)
    Rs = r_grid(per)
    n = ao.total
    out = {R: np.zeros((n, n)) for R in Rs}
    for I in range(sys.natom):
        for J in range(sys.natom):
            for R in Rs:
                Rv = np.array(R)
                w = _wblk(Wt, abf, I, J, Rv, per)
                for (K, Ra) in nb.get(I, []):
                    c1 = cs[(I, K, Ra)]
                    Rav = np.array(Ra)
                    for (L, Rb) in nb.get(J, []):
                        c2 = cs[(J, L, Rb)]
                        Rbv = np.array(Rb)
                        cw = np.einsum('mik,mn->ikn', c1, w)
                        # a0b0_a2b2
                        g = _gblk(G, ao, K, L, Rv + Rbv - Rav, per)
                        out[fold_R(Rv, per)][ao.sl(I), ao.sl(J)] += np.einsum('ikn,kl,njl->ij', cw, g, c2)
                        # a0b0_a1b1
                        g = _gblk(G, ao, I, J, Rv, per)
                        dst = fold_R(Rv + Rbv - Rav, per) if mistake != 'a1b1_cell' else fold_R(Rv + Rbv + Rav, per)
                        out[dst][ao.sl(K), ao.sl(L)] += np.einsum('ikn,ij,njl->kl', cw, g, c2)
                        # a0b0_a1b2
                        g = _gblk(G, ao, I, L, Rv + Rbv, per)
                        out[fold_R(Rv - Rav, per)][ao.sl(K), ao.sl(J)] += np.einsum('ikn,il,njl->kj', cw, g, c2)
                        # a0b0_a2b1
                        g = _gblk(G, ao, K, J, Rv - Rav, per)
                        out[fold_R(Rv + Rbv, per)][ao.sl(I), ao.sl(L)] += np.einsum('ikn,kj,njl->il', cw, g, c2)
    return out


def sigma_supercell(sys, cs, G, Wt):
    """Sigma[R] from the dense BvK-supercell contraction (independent restatement)."""
    ao, abf = Basis(sys.nao), Basis(sys.nabf)
    per = sys.period
    Rs = r_grid(per)
    nR = len(Rs)
    ridx = {R: i for i, R in enumerate(Rs)}
    n, m = ao.total, abf.total
    # supercell index: (cell, atom-major AO) and (cell, atom-major ABF)
    Gs = np.zeros((nR * n, nR * n))
    Ws = np.zeros((nR * m, nR * m))
    for a, Ra in enumerate(Rs):
        for b, Rb in enumerate(Rs):
            d = fold_R(np.array(Rb) - np.array(Ra), per)
            Gs[a * n:(a + 1) * n, b * n:(b + 1) * n] = G[d]
            Ws[a * m:(a + 1) * m, b * m:(b + 1) * m] = Wt[d]
    B = np.zeros((nR * m, nR * n, nR * n))
    for (P, X, Rx), c in cs.items():      # c: (mu on P, p on P, x on X)
        for a, Rp in enumerate(Rs):
            cx = ridx[fold_R(np.array(Rp) + np.array(Rx), per)]
            mu = slice(a * m + abf.off[P], a * m + abf.off[P + 1])
            p = slice(a * n + ao.off[P], a * n + ao.off[P + 1])
            x = slice(cx * n + ao.off[X], cx * n + ao.off[X + 1])
            B[mu, p, x] += c
            B[mu, x, p] += c.transpose(0, 2, 1)
    # Sigma(a, b) for a in the home cell only
    home = slice(ridx[(0, 0, 0)] * n, (ridx[(0, 0, 0)] + 1) * n)
    T = np.einsum('nbd,cd->nbc', B, Gs, optimize=True)          # (nu, b, c)
    Phi = np.tensordot(Ws, T, axes=([1], [0]))                   # (mu, b, c)
    S = np.einsum('mac,mbc->ab', B[:, home, :], Phi, optimize=True)
    return {R: S[:, ridx[R] * n:(ridx[R] + 1) * n].copy() for R in Rs}


def sigma_factorized(sys, cs, G, Wt):
    """Sigma[R] through the factorisation of the device path.

    U_J(S)[(X,x), (j,nu)] = sum_{(L,Rb) in N(J)} G_XL(S+Rb) C^J(L,Rb)(nu, j, l)
    a2b2: Sigma_IJ(R) = sum_{(K,Ra)} sum_{mu,k} C^I(mu,i,k) sum_nu W_IJ(R) U_J(R-Ra)[K](k, (j,nu))
    a1b2: P_IJR(mu,i,j) = sum_nu W_IJ(R)(mu,nu) U_J(R)[I](i,(j,nu));
          Sigma_KJ(R-Ra) += sum_{mu,i} C^I(K,Ra)(mu,i,k) P_IJR(mu,i,j)
    a1b1: Y_I(mu,i,(L,S)) = sum_{J,R,Rb: R+Rb=S} sum_{nu,j} W_IJ(R)(mu,nu) G_IJ(R)(i,j) C^J(L,Rb)(nu,j,l);
          Sigma_KL(S-Ra) += sum_{mu,i} C^I(K,Ra)(mu,i,k) Y_I(mu,i,(L,S))
    a2b1: the transpose of a1b2: Sigma_IL(S) += [a1b2]_LI(-S)^T."""
    ao, abf = Basis(sys.nao), Basis(sys.nabf)
    per = sys.period
    nb = _neighbours(cs)
    Rs = r_grid(per)
    n = ao.total
    U = {}
    for J in range(sys.natom):
        for S in Rs:
            acc = np.zeros((n, int(sys.nao[J]), int(sys.nabf[J])))
            for (L, Rb) in nb.get(J, []):
                g = G[fold_R(np.array(S) + np.array(Rb), per)][:, ao.sl(L)]
                acc += np.tensordot(g, cs[(J, L, Rb)], axes=([1], [2])).transpose(0, 2, 1)
            U[(J, S)] = acc
    main = {R: np.zeros((n, n)) for R in Rs}
    bterm = {R: np.zeros((n, n)) for R in Rs}
    for I in range(sys.natom):
        ni, nmu = int(sys.nao[I]), int(sys.nabf[I])
        Y = {}
        for J in range(sys.natom):
            nj, nnu = int(sys.nao[J]), int(sys.nabf[J])
            for R in Rs:
                Rv = np.array(R)
                w = _wblk(Wt, abf, I, J, Rv, per)
                # a2b2
                acc = np.zeros((ni, nj))
                for (K, Ra) in nb.get(I, []):
                    u = U[(J, fold_R(Rv - np.array(Ra), per))][ao.sl(K)]          # (k, j, nu)
                    t = np.tensordot(w, u, axes=([1], [2]))                      # (mu, k, j)
                    c = cs[(I, K, Ra)]                                           # (mu, i, k)
                    acc += np.tensordot(c, t, axes=([0, 2], [0, 1]))             # (i, j)
                main[fold_R(Rv, per)][ao.sl(I), ao.sl(J)] += acc
                # a1b2
                P = np.tensordot(w, U[(J, fold_R(Rv, per))][ao.sl(I)], axes=([1], [2]))   # (mu, i, j)
                for (K, Ra) in nb.get(I, []):
                    bterm[fold_R(Rv - np.array(Ra), per)][ao.sl(K), ao.sl(J)] += np.tensordot(
                        cs[(I, K, Ra)], P, axes=([0, 1], [0, 1]))
                # a1b1 intermediate: y(mu, i, l) = sum_{nu, j} w(mu, nu) g(i, j) c(nu, j, l)
                g = _gblk(G, ao, I, J, Rv, per)
                for (L, Rb) in nb.get(J, []):
                    S = fold_R(Rv + np.array(Rb), per)
                    wc = np.tensordot(w, cs[(J, L, Rb)], axes=([1], [0]))        # (mu, j, l)
                    y = np.einsum('ij,mjl->mil', g, wc, optimize=True)
                    Y[(L, S)] = Y.get((L, S), 0) + y
        for (K, Ra) in nb.get(I, []):
            c2 = cs[(I, K, Ra)].reshape(nmu * ni, -1)                            # ((mu, i), k)
            for (L, S), y in Y.items():
                main[fold_R(np.array(S) - np.array(Ra), per)][ao.sl(K), ao.sl(L)] += c2.T @ y.reshape(nmu * ni, -1)
    out = {}
    for R in Rs:
        mR = fold_R(-np.array(R), per)
        out[R] = main[R] + bterm[R] + bterm[mR].T
    return out


def sigma_freq(sys, grid, sig_pos, sig_neg):
    """Sigma[(iw, R)] complex from Sigma(R, +-tau) lists over tau (gw.cpp:681-797)."""
    Rs = r_grid(sys.period)
    out = {}
    for iw in range(len(grid['omega'])):
        for R in Rs:
            acc = 0
            for it in range(len(grid['tau'])):
                c = 0.5 * (sig_pos[it][R] + sig_neg[it][R])
                s = 0.5 * (sig_pos[it][R] - sig_neg[it][R])
                acc = acc + c * grid['cos_t2f'][iw, it] + 1j * s * grid['sin_t2f'][iw, it]
            out[(iw, R)] = acc
    return out


def sigma_all(sys, grid, wrt, spin=0, method='factorized'):
    """Sigma(R, i omega) for one spin from W(R, tau): dict (iw, R) -> complex nbasis^2."""
    cs = fold_cs(sys)
    Rs = r_grid(sys.period)
    f = {'factorized': sigma_factorized, 'labels': sigma_labels, 'supercell': sigma_supercell}[method]
    sp, sn = [], []
    for it, tau in enumerate(grid['tau']):
        Wt = {R: wrt[(it, R)] for R in Rs}
        sp.append(f(sys, cs, gf_R(sys, tau, Rs, ispin=spin), Wt))
        sn.append(f(sys, cs, gf_R(sys, -tau, Rs, ispin=spin), Wt))
    return sigma_freq(sys, grid, sp, sn)


# ---------------------------------------------------------------------------------------
# Kohn-Sham representation, analytic continuation, quasi-particle equation
# ---------------------------------------------------------------------------------------

def nearest_image(frac, latvec, I, J, R, period):
    best, Rb = np.inf, None
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                Rc = np.array([R[0] + i * period[0], R[1] + j * period[1], R[2] + k * period[2]])
                d = (frac[I] - frac[J] - Rc) @ latvec
                d2 = float(d @ d)
                if d2 < best:
                    best, Rb = d2, Rc
    return Rb


def sigma_ks_diag(sys, sig_w, iks, states, frac=None, spin=0):
    """Sigma_KS(k, i w)_nn of one spin: dict (ik, n) -> complex array over omega (ascending)."""
    ao = Basis(sys.nao)
    per = sys.period
    Rs = r_grid(per)
    nw = max(iw for (iw, _) in sig_w) + 1
    out = {}
    for ik in iks:
        kf = sys.kfrac[ik]
        U = sys.wfc[spin, ik]
        vals = np.zeros((len(states), nw), dtype=np.complex128)
        for iw in range(nw):
            Sk = np.zeros((ao.total, ao.total), dtype=np.complex128)
            for R in Rs:
                blk = sig_w[(iw, R)]
                for I in range(sys.natom):
                    for J in range(sys.natom):
                        Rb = nearest_image(frac, sys.latvec, I, J, R, per) if frac is not None else np.array(R)
                        ph = np.exp(2j * np.pi * float(np.dot(kf, Rb)))
                        Sk[ao.sl(I), ao.sl(J)] += ph * blk[ao.sl(I), ao.sl(J)]
            M = U.conj() @ Sk @ U.T
            for a, st in enumerate(states):
                vals[a, iw] = M[st, st]
        for a, st in enumerate(states):
            out[(ik, st)] = vals[a]
    return out


class Pade:
    """Thiele continued fraction (analycont.cpp:17-80)."""

    def __init__(self, n_pars, xs, data):
        xs = list(xs)
        data = list(data)
        nd = len(data)
        if n_pars < 0:
            n_pars = nd
        if nd <= n_pars:
            n_pars = nd
            px, pd = xs, data
        else:
            # n_pars - 1 evenly spaced nodes, then the last datum; LibRPA sizes its node list to
            # n_pars and fills only the first n_pars - 1, so that datum sits at the node 0
            # (analycont.cpp:35-45)
            step = nd // (n_pars - 1)
            px = [xs[i * step] for i in range(n_pars - 1)] + [0j]
            pd = [data[i * step] for i in range(n_pars - 1)] + [data[nd - 1]]
        self.n = n_pars
        self.x = np.array(px, dtype=np.complex128)
        g = np.zeros((n_pars, n_pars), dtype=np.complex128)
        g[:, 0] = pd
        for p in range(1, n_pars):
            for i in range(p, n_pars):
                g[i, p] = (g[p - 1, p - 1] - g[i, p - 1]) / ((self.x[i] - self.x[p - 1]) * g[i, p - 1])
        self.y = np.diag(g).copy()

    def __call__(self, z):
        t = 1.0 + 0j
        for p in range(self.n - 1, 0, -1):
            t = 1.0 + self.y[p] * (z - self.x[p - 1]) / t
        return self.y[0] / t


def qpe_solve(pade, e_mf, e_fermi, vxc, sigx, diff_init=1e-3, thres=1e-5, n_iter_max=200):
    """Damped fixed point of qpe_solver_pade_self_consistent (qpe_solver.cpp:10-55).

    Returns (info, e_qp, sigc)."""
    escale = 0.1
    n_iter = 0
    e_last = e_mf
    diff = diff_init
    sig_last = 0j
    while True:
        n_iter += 1
        if not n_iter - 1 < n_iter_max:
            break
        e_this = e_last + escale * diff
        s = pade(complex(e_this - e_fermi))
        diff = e_mf - vxc + sigx + s.real - e_this
        e_last = e_mf - vxc + sigx + s.real
        sig_last = s
        if abs(diff) < thres or n_iter == n_iter_max - 1:
            break
    info = 1 if n_iter > n_iter_max - 1 else 0
    return info, e_last, sig_last
