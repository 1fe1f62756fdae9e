// G0W0 on one GPU: W(q, i omega) -> W(R, i tau) and the correlation self-energy.
//
// Screened interaction (compute_Wc_freq_q_blacs, src/core/epsilon.cpp:1763-2321):
//   Wc = v^{1/2} [(1 - v^{1/2} chi0 v^{1/2})^{-1} - 1] v^{1/2}  =  L M^{-1} L^H - v,
// with v = L L^H (Cholesky) and M = 1 - L^H chi0 L, exact when the truncated Coulomb
// equals the bare one.  L M^{-1} L^H - v is read off as minus the Schur complement of
// the augmented matrix [[M, L^H], [L, v]] after eliminating its first block column with
// the signed blocked Cholesky; complex q use the real embedding [[Re, -Im], [Im, Re]].
// W(R, tau) = Re sum_w gamma_f2t(tau, w) (1/N_k) sum_q e^{-i 2 pi q.R} Wc(q, w)
// (CT_FT_Wc_freq_q, epsilon.cpp:2490-2666) is accumulated per q on the device.  With an IBZ
// mapping only the irreducible q are formed; the others must be their time-reversal partners
// (FT_Wc_freq_q, epsilon.cpp:2430-2452, takes W(-q) = conj W(q)), whose real parts of
// e^{-i 2 pi q.R} W(q) are equal, so each irreducible q is weighted by its star size.
// A Coulomb matrix with a negative pivot, a positive sqrt_coulomb_threshold or a head
// replacement at q = 0 takes the square roots from an eigen-decomposition instead
// (power_hemat, src/math/matrix_m.h:916-970, eigenvalues below the threshold dropped); then
// W = S_c eps^{-1} S_c - S_c S_c, the Schur complement of [[eps, S_c], [S_c, S_c S_c]].
//
// Self-energy (RI::GW::cal_Sigmas, thirdparty/LibRI/include/RI/physics/GW.hpp:102-122,
// called from G0W0::build_spacetime, src/core/gw.cpp:217-875), per tau, spin and sign:
//   U_J(S)[(X,x)][(j,nu)] = sum_{(L,Rb) in N(J)} G_XL(S+Rb) C^J(L,Rb)           (GEMM per (J,S))
//   T_I[(n,mu,k)][(R,J,j)] = sum_nu W_IJ(R)(mu,nu) U_J(R-Ra_n)[(K_n,k)][(j,nu)]    (rows gathered)
//   label a0b0_a2b2: Sigma_IJ(R)(i,j) = sum_(n,mu,k) C^I(n)(mu,i,k) T_I[(n,mu,k)][(R,J,j)]
//   label a0b0_a1b2: Sigma_KJ(R-Ra)(k,j) += sum_(mu,i) C^I(n)(mu,i,k) T_I[(on-site,mu,i)][(R,J,j)]
//   label a0b0_a1b1: Y_I[(mu,i)][(S,L,l)] = sum_{J,R} W_IJ(R) G_IJ(R) C^J(L,S-R)  (two GEMMs per (J,R))
//                    Sigma_KL(S-Ra)(k,l) += sum_(mu,i) C^I(n)(mu,i,k) Y_I[(mu,i)][(S,L,l)]
//   label a0b0_a2b1: the transpose of a0b0_a1b2, Sigma_XY(R) += [a1b2]_YX(-R)^T.
// Outputs indexed by a cell difference go through an extended (2p-1)^3 cell grid with
// atomic accumulation and are folded back to the BvK cell afterwards.  Sigma(R, i w) =
// sum_tau 0.5 (S+ + S-) gamma_c + i 0.5 (S+ - S-) gamma_s (gw.cpp:681-797).
#include "gw.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>

#include "dielec.h"
#include "gw_kernels.h"

namespace lra {

namespace {
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline int bvk_mod(int v, int n) { return (((v % n) + 3 * n / 2) % n) - n / 2; }

template <class T>
DevBuf upload_as(const std::vector<double> &v, hipStream_t s) {
    if constexpr (std::is_same<T, double>::value) {
        return upload(v, s);
    } else {
        std::vector<T> f(v.begin(), v.end());
        DevBuf b = upload(f, s);
        HIP_CHECK(hipStreamSynchronize(s));
        return b;
    }
}
}  // namespace

// ------------------------------------------------------------------------------------------
// plan
// ------------------------------------------------------------------------------------------
void GwEngine::plan(const Dataset &ds, std::shared_ptr<RpaEngine> r) {
    rpa = std::move(r);
    const RpaEngine &R0 = *rpa;
    prec = R0.prec;
    ts = R0.ts;
    natom = R0.natom;
    NAO = R0.NAO;
    NABF = R0.NABF;
    nR = R0.nR;
    nk = R0.nk;
    nspin = R0.nspin;
    ntau = R0.planned_ntau;
    LD = (long long)nR * NAO;
    const auto &nb = R0.nb;
    const auto &na = R0.na;
    const auto &ao = R0.ao;
    const auto &abf = R0.abf;
    const auto &cells = R0.cells;
    const hipStream_t st = R0.stream;
    const int cell0 = R0.cell_index(0, 0, 0);

    // ---- folded Cs, as in the chi0 plan ----
    std::map<std::tuple<int, int, int>, std::vector<double>> folded;
    for (const auto &kv : ds.cs) {
        const int I = std::get<0>(kv.first), K = std::get<1>(kv.first);
        const int rr = R0.fold_index(std::get<2>(kv.first), std::get<3>(kv.first), std::get<4>(kv.first));
        auto &dst = folded[{I, K, rr}];
        if (dst.empty()) dst = kv.second;
        else for (size_t i = 0; i < dst.size(); ++i) dst[i] += kv.second[i];
    }
    auto cval = [&](int I, int K, int rr, int mu, int i, int k) -> double {
        auto it = folded.find({I, K, rr});
        if (it == folded.end()) return 0.0;
        return it->second[((size_t)i * nb[K] + k) * na[I] + mu];
    };
    nbrs = R0.nbrs;
    onsite.assign(natom, -1);
    for (int I = 0; I < natom; ++I) {
        for (size_t n = 0; n < nbrs[I].size(); ++n)
            if (nbrs[I][n].K == I && nbrs[I][n].r == cell0) onsite[I] = (int)n;
        if (onsite[I] < 0) { onsite[I] = (int)nbrs[I].size(); nbrs[I].push_back({I, cell0}); }
    }
    loff.assign(natom, {});
    lsum.assign(natom, 0);
    for (int I = 0; I < natom; ++I) {
        long long l = 0;
        for (const auto &n : nbrs[I]) { loff[I].push_back(l); l += nb[n.K]; }
        lsum[I] = l;
    }

    // ---- packed Cs ----
    {
        cd_off.assign(natom + 1, 0);
        cb_off.assign(natom + 1, 0);
        ca_off.assign(natom + 1, 0);
        for (int I = 0; I < natom; ++I) {
            cd_off[I + 1] = cd_off[I] + (long long)nb[I] * na[I] * lsum[I];
            cb_off[I + 1] = cb_off[I] + lsum[I] * na[I] * nb[I];
            ca_off[I + 1] = ca_off[I] + (long long)nb[I] * na[I] * R0.lsum[I];
        }
        std::vector<double> cd(cd_off[natom]), cb(cb_off[natom]), ca(ca_off[natom]);
        for (int I = 0; I < natom; ++I) {
            const int ni = nb[I], nm = na[I];
            const long long kd = (long long)nm * lsum[I];
            for (size_t n = 0; n < nbrs[I].size(); ++n) {
                const int K = nbrs[I][n].K, rr = nbrs[I][n].r;
                for (int mu = 0; mu < nm; ++mu)
                    for (int i = 0; i < ni; ++i)
                        for (int k = 0; k < nb[K]; ++k) {
                            const double v = cval(I, K, rr, mu, i, k);
                            // CD_I[i][(n, mu, k)]
                            cd[cd_off[I] + (size_t)i * kd + (size_t)nm * loff[I][n] + (size_t)mu * nb[K] + k] = v;
                            // CB_I[(n, k)][(mu, i)]
                            cb[cb_off[I] + (size_t)(loff[I][n] + k) * nm * ni + (size_t)mu * ni + i] = v;
                        }
            }
            // CA_I[i][(mu, (n, k))] over the chi0 neighbour list of I (I in the role of J)
            const long long ls = R0.lsum[I];
            for (size_t n = 0; n < R0.nbrs[I].size(); ++n) {
                const int K = R0.nbrs[I][n].K, rr = R0.nbrs[I][n].r;
                for (int mu = 0; mu < nm; ++mu)
                    for (int i = 0; i < ni; ++i)
                        for (int k = 0; k < nb[K]; ++k)
                            ca[ca_off[I] + (size_t)i * nm * ls + (size_t)mu * ls + R0.loff[I][n] + k] =
                                cval(I, K, rr, mu, i, k);
            }
        }
        if (prec == Prec::F64) {
            d_CD = upload_as<double>(cd, st);
            d_CB = upload_as<double>(cb, st);
            d_CA = upload_as<double>(ca, st);
        } else {
            d_CD = upload_as<float>(cd, st);
            d_CB = upload_as<float>(cb, st);
            d_CA = upload_as<float>(ca, st);
        }
    }

    // ---- extended cell grid ----
    const int *p = R0.period;
    for (int a = 0; a < 3; ++a) E[a] = 2 * p[a] - 1;
    NE = E[0] * E[1] * E[2];
    c0 = linr(p[0] - 1, p[1] - 1, p[2] - 1);
    ext_fold.assign(NE, 0);
    ext_foldneg.assign(NE, 0);
    for (int x = 0; x < E[0]; ++x)
        for (int y = 0; y < E[1]; ++y)
            for (int z = 0; z < E[2]; ++z) {
                const int e = (x * E[1] + y) * E[2] + z;
                const int dx = x - (p[0] - 1), dy = y - (p[1] - 1), dz = z - (p[2] - 1);
                ext_fold[e] = R0.fold_index(dx, dy, dz);
                ext_foldneg[e] = R0.fold_index(-dx, -dy, -dz);
            }
    d_fold = upload(ext_fold, st);
    d_foldneg = upload(ext_foldneg, st);
    std::vector<long long> ext_cols(LD);
    for (int rr = 0; rr < nR; ++rr)
        for (int c = 0; c < NAO; ++c)
            ext_cols[(size_t)rr * NAO + c] =
                (linr(cells[rr][0], cells[rr][1], cells[rr][2]) + c0) * NAO * (long long)NAO + c;

    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));

    // ---- U_J(S) for all J, S (resident) and the gathered G panels by chunks of J ----
    u_off.assign((size_t)natom * nR + 1, 0);
    for (int J = 0; J < natom; ++J)
        for (int S = 0; S < nR; ++S)
            u_off[(size_t)J * nR + S + 1] = u_off[(size_t)J * nR + S] + (long long)NAO * nb[J] * na[J];
    u_elems = u_off[(size_t)natom * nR];
    const long long gpm_r = (long long)NAO * 2 * NAO;
    {
        const double budget = std::min(24e9, 0.08 * (double)free_b);
        int j = 0;
        while (j < natom) {
            UChunk c;
            c.j0 = j;
            long long g = 0;
            int jj = j;
            while (jj < natom) {
                const long long g1 = (long long)nR * NAO * R0.lsum[jj];
                if (jj > j && (double)(g + g1) * ts > budget) break;
                g += g1;
                ++jj;
            }
            c.j1 = jj;
            c.ghat_elems = g;
            long long go = 0;
            for (int J = c.j0; J < c.j1; ++J)
                for (int S = 0; S < nR; ++S) {
                    c.ghat_off.push_back(go);
                    go += (long long)NAO * R0.lsum[J];
                }
            for (int sg = 0; sg < 2; ++sg) {
                for (int J = c.j0; J < c.j1; ++J)
                    for (int S = 0; S < nR; ++S) {
                        GatherJob gj;
                        gj.dst_off = c.ghat_off[(size_t)(J - c.j0) * nR + S];
                        gj.ncol = (int)R0.lsum[J];
                        gj.col_off = (int)c.colsrc[sg].size();
                        for (const auto &n : R0.nbrs[J]) {
                            const auto &R2 = cells[n.r];
                            const auto &Sv = cells[S];
                            const int rr = R0.fold_index(Sv[0] + R2[0], Sv[1] + R2[1], Sv[2] + R2[2]);
                            for (int l = 0; l < nb[n.K]; ++l)
                                c.colsrc[sg].push_back(rr * gpm_r + 2 * ao[n.K] + (long long)sg * nb[n.K] + l);
                        }
                        c.max_gcol = std::max(c.max_gcol, gj.ncol);
                        c.jobs[sg].push_back(gj);
                    }
                c.d_jobs[sg] = upload(c.jobs[sg], st);
                c.d_colsrc[sg] = upload(c.colsrc[sg], st);
            }
            for (int J = c.j0; J < c.j1; ++J)
                for (int S = 0; S < nR; ++S) {
                    if (R0.lsum[J] == 0) continue;
                    c.ub.add(make_prob(NAO, nb[J] * na[J], (int)R0.lsum[J], c.ghat_off[(size_t)(J - c.j0) * nR + S],
                                       (int)R0.lsum[J], R0.cwb_off[J], nb[J] * na[J], u_off[(size_t)J * nR + S],
                                       nb[J] * na[J]));
                }
            // default tiles: the autotuner's choices on the 64-atom, 8-k cell (FP64)
            c.ub.finalize(prec == Prec::F64 ? T128x64w8 : big_tile(prec), st);
            flops_sigma += c.ub.flops;
            uchunks.push_back(std::move(c));
            j = jj;
        }
    }

    // ---- Y (partner-partner intermediate), GC, A1 / A2 / A3a ----
    y_off.assign(natom + 1, 0);
    for (int I = 0; I < natom; ++I) y_off[I + 1] = y_off[I] + (long long)na[I] * nb[I] * LD;
    y_elems = y_off[natom];
    gc_elems = 0;
    for (int J = 0; J < natom; ++J) gc_elems = std::max(gc_elems, (long long)na[J] * NAO * R0.lsum[J]);
    a1.clear();
    a2.clear();
    a1.resize((size_t)natom * nR);
    a2.resize((size_t)natom * nR);
    for (int J = 0; J < natom; ++J) {
        const long long ls = R0.lsum[J];
        if (ls == 0) continue;
        for (int rr = 0; rr < nR; ++rr) {
            GemmBatch &b1 = a1[(size_t)J * nR + rr];
            GemmBatch &b2 = a2[(size_t)J * nR + rr];
            for (int I = 0; I < natom; ++I) {
                // GC_I[nu][i][c] = sum_j G_IJ(R)(i, j) CA_J[j][(nu, c)]
                GemmProb q1 = make_prob(nb[I], na[J] * (int)ls, nb[J], rr * gpm_r + ao[I] * 2LL * NAO + 2 * ao[J],
                                        2 * NAO, ca_off[J], na[J] * (int)ls, (long long)na[J] * ls * ao[I], (int)ls);
                q1.ncb = (int)ls;
                q1.cbs = (long long)nb[I] * ls;
                b1.add(q1);
                // Y_I[(mu, i)][cols] += sum_nu W_IJ(R)(mu, nu) GC_I[nu][(i, c)]
                GemmProb q2 = make_prob(na[I], nb[I] * (int)ls, na[J], (long long)rr * NABF * NABF + abf[I] * NABF + abf[J],
                                        NABF, (long long)na[J] * ls * ao[I], nb[I] * (int)ls, y_off[I],
                                        (int)(nb[I] * LD));
                q2.ncb = (int)ls;
                q2.cbs = LD;
                q2.ct_off = 0;
                b2.add(q2);
            }
            // column map of the (L, Rb, l) index of J into the (S = R + Rb, L, l) index
            const auto &Rv = cells[rr];
            for (size_t n = 0; n < R0.nbrs[J].size(); ++n) {
                const auto &R2 = cells[R0.nbrs[J][n].r];
                const int S = R0.fold_index(Rv[0] + R2[0], Rv[1] + R2[1], Rv[2] + R2[2]);
                const int L = R0.nbrs[J][n].K;
                for (int l = 0; l < nb[L]; ++l) b2.coltab.push_back((long long)S * NAO + ao[L] + l);
            }
            b1.finalize(prec == Prec::F64 ? T16x256 : T32x128, st);
            b2.finalize(big_tile(prec), st);
            flops_sigma += b1.flops + b2.flops;
        }
    }
    auto add_a3 = [&](GemmBatch &b, int I, long long b_off) {
        GemmProb q = make_prob(0, (int)LD, na[I] * nb[I], 0, 0, b_off, (int)LD, 0, 0);
        q.rt_off = (int)(b.rowtab.size() / 2);
        q.ct_off = 0;
        int m = 0;
        for (size_t n = 0; n < nbrs[I].size(); ++n) {
            const int K = nbrs[I][n].K;
            const auto &Ra = cells[nbrs[I][n].r];
            const long long lr = linr(Ra[0], Ra[1], Ra[2]);
            for (int k = 0; k < nb[K]; ++k) {
                b.rowtab.push_back(cb_off[I] + (loff[I][n] + k) * na[I] * nb[I]);
                b.rowtab.push_back((ao[K] + k) * NAO - lr * NAO * (long long)NAO);
                ++m;
            }
        }
        q.M = m;
        b.add(q);
    };
    a3a = GemmBatch();
    a3a.atomic = true;
    a3a.coltab = ext_cols;
    for (int I = 0; I < natom; ++I) add_a3(a3a, I, y_off[I]);
    a3a.finalize(big_tile(prec), st);
    flops_sigma += a3a.flops;

    // ---- chunks of I for T^D ----
    std::vector<std::vector<std::pair<int, int>>> rev(natom);  // K -> (I, n)
    for (int I = 0; I < natom; ++I)
        for (size_t n = 0; n < nbrs[I].size(); ++n) rev[nbrs[I][n].K].push_back({I, (int)n});
    {
        const double fixed = (double)ts * (u_elems + y_elems + gc_elems + 4.0 * NE * NAO * (double)NAO) +
                             8.0 * ntau * nR * (double)NABF * NABF;
        const double budget = std::max(4e9, std::min(96e9, 0.6 * (double)free_b - fixed));
        int i = 0;
        while (i < natom) {
            IChunk c;
            c.i0 = i;
            long long t = 0;
            int ii = i;
            while (ii < natom) {
                const long long t1 = (long long)na[ii] * lsum[ii] * LD;
                if (ii > i && (double)(t + t1) * ts > budget) break;
                t += t1;
                ++ii;
            }
            c.i1 = ii;
            c.td_elems = t;
            long long o = 0;
            for (int I = c.i0; I < c.i1; ++I) { c.td_off.push_back(o); o += (long long)na[I] * lsum[I] * LD; }
            auto tdo = [&](int I) { return c.td_off[I - c.i0]; };
            // T^D build: per (J, S, K), rows (I, n, mu) of atoms I of the chunk with K_n = K
            c.tdb.trans_b = true;
            for (int J = 0; J < natom; ++J)
                for (int S = 0; S < nR; ++S)
                    for (int K = 0; K < natom; ++K) {
                        GemmProb q = make_prob(0, nb[K] * nb[J], na[J], 0, 0,
                                               u_off[(size_t)J * nR + S] + ao[K] * nb[J] * (long long)na[J], na[J], 0,
                                               0);
                        q.ncb = nb[J];
                        q.cbs = LD;
                        q.rt_off = (int)(c.tdb.rowtab.size() / 2);
                        int m = 0;
                        const auto &Sv = cells[S];
                        for (const auto &in : rev[K]) {
                            const int I = in.first, n = in.second;
                            if (I < c.i0 || I >= c.i1) continue;
                            const auto &Ra = cells[nbrs[I][n].r];
                            const int rr = R0.fold_index(Sv[0] + Ra[0], Sv[1] + Ra[1], Sv[2] + Ra[2]);
                            for (int mu = 0; mu < na[I]; ++mu) {
                                c.tdb.rowtab.push_back((long long)rr * NABF * NABF + (abf[I] + mu) * NABF + abf[J]);
                                c.tdb.rowtab.push_back(tdo(I) + ((long long)na[I] * loff[I][n] + (long long)mu * nb[K]) * LD +
                                                       (long long)rr * NAO + ao[J]);
                                ++m;
                            }
                        }
                        if (m == 0) { c.tdb.rowtab.resize((size_t)q.rt_off * 2); continue; }
                        q.M = m;
                        c.tdb.add(q);
                    }
            c.tdb.finalize(prec == Prec::F64 ? T128x64w8 : T128x64, st);
            // label a0b0_a2b2: Sigma_I.(R)[i][(R, J, j)] = CD_I * T_I
            for (int I = c.i0; I < c.i1; ++I) {
                GemmProb q = make_prob(nb[I], (int)LD, (int)(na[I] * lsum[I]), cd_off[I], (int)(na[I] * lsum[I]), tdo(I),
                                       (int)LD, ao[I] * NAO, NAO);
                q.ncb = NAO;
                q.cbs = (long long)NAO * NAO;
                c.dcon.add(q);
            }
            c.dcon.finalize(prec == Prec::F64 ? T32x64 : T32x128, st);
            // label a0b0_a1b2 from the on-site rows of T_I
            c.a3b.atomic = true;
            c.a3b.coltab = ext_cols;
            for (int I = c.i0; I < c.i1; ++I) add_a3(c.a3b, I, tdo(I) + (long long)na[I] * loff[I][onsite[I]] * LD);
            c.a3b.finalize(big_tile(prec), st);
            flops_sigma += c.tdb.flops + c.dcon.flops + c.a3b.flops;
            ichunks.push_back(std::move(c));
            i = ii;
        }
    }
    flops_sigma *= 2.0 * nspin * ntau;  // both signs, all spins and tau

    // ---- device buffers ----
    long long max_g = 0, max_td = 0;
    for (auto &c : uchunks) max_g = std::max(max_g, c.ghat_elems);
    for (auto &c : ichunks) max_td = std::max(max_td, c.td_elems);
    // work buffers of the Sigma phase are allocated when it starts (after W's temporaries are freed)
    max_ghat = max_g;
    max_tdc = max_td;
    HIP_CHECK(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------------------------------
// W(q, omega) -> W(R, tau)
// ------------------------------------------------------------------------------------------
namespace {
// Principal square root of a symmetric positive-definite n x n matrix A (in place) by the coupled
// Newton-Schulz iteration Y <- Y T, Z <- T Z, T = (3 I - Z Y) / 2 from Y = A / c, Z = I with
// c = ||A||_F >= lambda_max; Y -> (A / c)^{1/2}.  Equals the eigen-decomposition root of
// power_hemat (src/math/matrix_m.h:916-970) for positive-definite input.  Runs to convergence.
void ns_sqrt(double *A, int n, hipStream_t st) {
    const size_t nn = (size_t)n * n;
    DevBuf dY(nn * 8), dZ(nn * 8), dT(nn * 8), dP(nn * 8), dYn(nn * 8), dZn(nn * 8), dr(64);
    GemmBatch b;
    b.add(make_prob(n, n, n, 0, n, 0, n, 0, n));
    b.finalize(T64x64, st);
    launch_sumsq(A, (long long)nn, dr.as<double>(), st);
    double h[2];
    HIP_CHECK(hipMemcpyAsync(h, dr.get(), 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    const double c = std::sqrt(h[0]);
    if (!(c > 0) || !std::isfinite(c)) LRA_THROW("Coulomb matrix is zero or not finite");
    launch_scale_add_identity(A, dY.as<double>(), n, 1.0 / c, 0.0, st);
    launch_scale_add_identity(A, dZ.as<double>(), n, 0.0, 1.0, st);
    double prev = 1e300;
    int stall = 0;
    for (int k = 0; k < 200; ++k) {
        gemm_run(Prec::F64, b, dZ.get(), dY.get(), dP.get(), 1.0, 0.0, st);
        launch_scale_add_identity(dP.as<double>(), dT.as<double>(), n, -0.5, 1.5, st);
        gemm_run(Prec::F64, b, dY.get(), dT.get(), dYn.get(), 1.0, 0.0, st);
        gemm_run(Prec::F64, b, dT.get(), dZ.get(), dZn.get(), 1.0, 0.0, st);
        launch_absmax_diff(dYn.as<double>(), dY.as<double>(), (long long)nn, dr.as<double>(), st);
        HIP_CHECK(hipMemcpyAsync(h, dr.get(), 16, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        std::swap(dY, dYn);
        std::swap(dZ, dZn);
        const double rel = h[0] / std::max(h[1], 1e-300);
        if (!std::isfinite(rel)) LRA_THROW("square root of the Coulomb matrix did not converge (not positive definite?)");
        // quadratic phase ends at rounding level: stop once the update no longer shrinks
        if (rel < 1e-13) { if (rel >= prev || ++stall >= 3) break; }
        prev = rel;
        if (k == 199) LRA_THROW("square root of the Coulomb matrix did not converge");
    }
    launch_scale_add_identity(dY.as<double>(), A, n, std::sqrt(c), 0.0, st);
    HIP_CHECK(hipStreamSynchronize(st));
}

// Filtered square root of the symmetric n x n matrix A (in place): A <- sum_{lam_j >= thr, lam_j > 0}
// lam_j^{1/2} u_j u_j^T from its eigen-decomposition (power_hemat, src/math/matrix_m.h:916-970;
// the real embedding of a Hermitian matrix has every eigenvalue twice and the embedded root).
// The eigen-decomposition is rocSOLVER's syevd, once per q; the root is one GEMM B^T B with the
// eigenvectors scaled by lam^{1/4}.  top (optional, length n / halves): the eigenvector of the
// largest eigenvalue, taken from one diagonal half of an embedded matrix when halves = 2.
// Returns the largest and second largest eigenvalue.
std::pair<double, double> eig_sqrt(double *A, int n, double thr, double *top, int halves, hipStream_t st) {
    static rocblas_handle handle = nullptr;
    if (!handle && rocblas_create_handle(&handle) != rocblas_status_success)
        LRA_THROW("rocblas_create_handle failed");
    if (rocblas_set_stream(handle, st) != rocblas_status_success) LRA_THROW("rocblas_set_stream failed");
    const size_t nn = (size_t)n * n;
    DevBuf dU(nn * 8), dW((size_t)n * 8), dE((size_t)n * 8), dinfo(8), df((size_t)n * 8);
    HIP_CHECK(hipMemcpyAsync(dU.get(), A, nn * 8, hipMemcpyDeviceToDevice, st));
    // column-major eigenvectors of a symmetric matrix: row j of the row-major view is u_j
    if (rocsolver_dsyevd(handle, rocblas_evect_original, rocblas_fill_lower, n, dU.as<double>(), n, dW.as<double>(),
                         dE.as<double>(), reinterpret_cast<rocblas_int *>(dinfo.get())) != rocblas_status_success)
        LRA_THROW("rocsolver_dsyevd failed");
    std::vector<double> lam(n);
    int info = 0;
    HIP_CHECK(hipMemcpyAsync(lam.data(), dW.get(), (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(&info, dinfo.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (info != 0) LRA_THROW("eigen-decomposition of the Coulomb matrix did not converge");
    if (top) {
        std::vector<double> u(n);
        HIP_CHECK(hipMemcpyAsync(u.data(), dU.as<double>() + (size_t)(n - 1) * n, (size_t)n * 8, hipMemcpyDeviceToHost,
                                 st));
        HIP_CHECK(hipStreamSynchronize(st));
        const int h = n / halves;
        int best = 0;
        double bn = -1;
        for (int c = 0; c < halves; ++c) {
            double s2 = 0;
            for (int i = 0; i < h; ++i) s2 += u[(size_t)c * h + i] * u[(size_t)c * h + i];
            if (s2 > bn) { bn = s2; best = c; }
        }
        std::vector<double> t(u.begin() + (size_t)best * h, u.begin() + (size_t)(best + 1) * h);
        const double nr = std::sqrt(bn);
        for (double &x : t) x /= nr;
        HIP_CHECK(hipMemcpyAsync(top, t.data(), (size_t)h * 8, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
    std::vector<double> f(n);
    for (int j = 0; j < n; ++j) f[j] = (lam[j] >= thr && lam[j] > 0) ? std::pow(lam[j], 0.25) : 0.0;
    HIP_CHECK(hipMemcpyAsync(df.get(), f.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    launch_scale_rows(dU.as<double>(), n, df.as<double>(), st);
    GemmBatch b;
    b.trans_a = true;
    b.add(make_prob(n, n, n, 0, n, 0, n, 0, n));
    b.finalize(T64x64, st);
    gemm_run(Prec::F64, b, dU.get(), dU.get(), A, 1.0, 0.0, st);
    HIP_CHECK(hipStreamSynchronize(st));
    // largest distinct eigenvalues (an embedded matrix repeats each one)
    const double l0 = lam[n - 1], l1 = lam[std::max(0, n - 1 - halves)];
    return {l0, l1};
}

// The upper atom-pair blocks (I <= J) of a Coulomb input for every k, staged on the device for
// launch_v_scatter, with the largest |V| and |Im V| of each k.
struct CoulombStage {
    DevBuf blk, desc;
    std::vector<int> first;  // descriptors of k: [first[k], first[k + 1])
    std::vector<double> vmax, imax;
};
void stage_coulomb(const RpaEngine &R0, const std::map<std::tuple<int, int, int>, std::vector<cplx>> &store, int nk,
                   CoulombStage &out, hipStream_t st) {
    std::vector<VBlock> desc;
    std::vector<cplx> vals;
    out.first.assign(1, 0);
    out.vmax.assign(nk, 0.0);
    out.imax.assign(nk, 0.0);
    for (int k = 0; k < nk; ++k) {
        for (const auto &kv : store) {
            const int ik = std::get<0>(kv.first), I = std::get<1>(kv.first), J = std::get<2>(kv.first);
            if (ik != k || I > J) continue;
            if (kv.second.size() != (size_t)R0.na[I] * R0.na[J]) LRA_THROW("Coulomb block of the wrong size");
            for (const cplx &v : kv.second) {
                out.vmax[k] = std::max(out.vmax[k], std::abs(v));
                out.imax[k] = std::max(out.imax[k], std::fabs(v.imag()));
            }
            desc.push_back({(long long)vals.size(), (int)R0.abf[I], (int)R0.abf[J], R0.na[I], R0.na[J], I == J ? 1 : 0});
            vals.insert(vals.end(), kv.second.begin(), kv.second.end());
        }
        out.first.push_back((int)desc.size());
    }
    out.blk = upload(vals, st);
    out.desc = upload(desc, st);
    HIP_CHECK(hipStreamSynchronize(st));
}

// dense (embedded when complex) Hermitian matrix of k point kq into V (device)
void dense_coulomb(const CoulombStage &cs, int kq, int n, int n2, bool embed, bool real_part, double *V,
                   hipStream_t st) {
    HIP_CHECK(hipMemsetAsync(V, 0, sizeof(double) * (size_t)n2 * n2, st));
    launch_v_scatter(cs.blk.as<double>(), cs.desc.as<VBlock>() + cs.first[kq], cs.first[kq + 1] - cs.first[kq], V, n,
                     n2, embed ? 1 : 0, real_part ? 1 : 0, st);
}
}  // namespace

void GwEngine::build_w(const Dataset &ds, const TFGrid &g) {
    auto &prof = profiler();
    const double t0 = now_s();
    RpaEngine &R0 = *rpa;
    const hipStream_t st = R0.stream;
    const int n = NABF;
    const int nw = g.n;
    if (R0.ntau_done != g.n) LRA_THROW("chi0 must be built on the same grid before W");
    if (R0.host_staged) LRA_THROW("G0W0 needs the device-resident chi0 store (too large for this device)");
    d_Wrt.alloc(sizeof(double) * (size_t)g.n * nR * n * n);
    HIP_CHECK(hipMemsetAsync(d_Wrt.get(), 0, d_Wrt.bytes(), st));
    // no truncated Coulomb: v_c^{1/2} = 0 and W = 0 (epsilon.cpp:1862-1879)
    w_zero = ds.vq_cut.empty();
    if (w_zero) { HIP_CHECK(hipStreamSynchronize(st)); t_w = now_s() - t0; return; }
    // irreducible q and their star sizes; a reducible q must be the time-reversal partner of its
    // irreducible q (the assumption of FT_Wc_freq_q, epsilon.cpp:2430-2452)
    std::vector<int> mult(nk, 0);
    for (int k = 0; k < nk; ++k) {
        const int r = ds.ibz_map.empty() ? k : ds.ibz_map[k];
        if (r < 0 || r >= nk || (!ds.ibz_map.empty() && ds.ibz_map[r] != r))
            LRA_THROW("IBZ mapping: k point " + std::to_string(k) + " maps to a point that is not irreducible");
        if (r != k) {
            bool partner = true;
            for (int d = 0; d < 3; ++d) {
                const double sum = ds.kfrac[(size_t)k * 3 + d] + ds.kfrac[(size_t)r * 3 + d];
                partner = partner && std::fabs(sum - std::round(sum)) < 1e-8;
            }
            if (!partner)
                LRA_THROW("G0W0 with an IBZ mapping needs every reducible q to be -q of its irreducible q (k point " +
                          std::to_string(k) + ")");
        }
        mult[r]++;
    }
    // v_c = v lets W = L M^{-1} L^T - v use the Cholesky factor of v; otherwise the square roots
    // of v and v_c are formed (Newton-Schulz) and W = v_c^{1/2} eps^{-1} v_c^{1/2} - v_c
    bool same_cut = ds.vq_cut.size() == ds.vq.size();
    for (const auto &kv : ds.vq) {
        auto it = ds.vq_cut.find(kv.first);
        same_cut = same_cut && it != ds.vq_cut.end() && it->second.size() == kv.second.size();
        for (size_t i = 0; same_cut && i < kv.second.size(); ++i)
            same_cut = std::abs(it->second[i] - kv.second[i]) <= 1e-12 * (1.0 + std::abs(kv.second[i]));
    }

    const bool embed = nk > 1;
    const int n2 = embed ? 2 * n : n;
    const int N = 2 * n2;
    const size_t nn = (size_t)n2 * n2, NN = (size_t)N * N;
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    // chi0(R, omega) once for all q when it fits in a quarter of the free memory
    if (!R0.chiw_ok && R0.form_chiw(g, 0.25 * (double)free_b, st)) HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const double per_w = 8.0 * (NN + 2.0 * nn);
    const double budget = std::min(160e9, 0.7 * (double)free_b) - 8.0 * 10 * nn;
    const int WCH = std::max(1, std::min(nw, (int)(budget / per_w)));
    DevBuf dV0(nn * 8), dL(nn * 8), dD(nn * 8 * WCH), dX(nn * 8 * WCH), dA(NN * 8 * WCH), dacc(sizeof(double) * (WCH + 8));
    // pivot signs of v = L S L^T: a negative pivot means v^{1/2} must come from the eigen-decomposition
    // with the negative eigenvalues dropped (power_hemat, matrix_m.h:916-970)
    DevBuf dSg((size_t)n2 * 8), dProbe, dProbeFlag(16);
    std::vector<double> hsg;
    int *dflag = reinterpret_cast<int *>(dacc.as<double>() + WCH + 4);
    HIP_CHECK(hipMemsetAsync(dacc.get(), 0, dacc.bytes(), st));
    CholWork cw;
    DevBuf dSc;
    auto negative_pivot = [&](const double *M) {
        dProbe.reserve(nn * 8);
        HIP_CHECK(hipMemcpyAsync(dProbe.get(), M, nn * 8, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(hipMemsetAsync(dProbeFlag.get(), 0, 16, st));
        cholesky_logdet(dProbe.as<double>(), n2, n2, dProbeFlag.as<double>(), reinterpret_cast<int *>(dProbeFlag.as<double>() + 1),
                        true, cw, st, dSg.as<double>());
        hsg.resize(n2);
        HIP_CHECK(hipMemcpyAsync(hsg.data(), dSg.get(), (size_t)n2 * 8, hipMemcpyDeviceToHost, st));
        int zf = 0;
        HIP_CHECK(hipMemcpyAsync(&zf, dProbeFlag.as<double>() + 1, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        bool neg = zf != 0;  // a zero pivot also needs the filtered root
        for (double v : hsg) neg = neg || v < 0;
        return neg;
    };
    // head replacement at q = 0: eps_mac per frequency and the leading eigenvector of Re v(0)
    DevBuf dHead, dU0((size_t)n * 8), dQ(sizeof(double) * (WCH + 1));
    if (!head.empty()) {
        if ((int)head.size() != nw) LRA_THROW("dielectric function: one value per frequency of the grid is needed");
        dHead = upload(head, st);
    }
    CoulombStage vst, vcst;
    stage_coulomb(R0, ds.vq, nk, vst, st);
    if (!same_cut) stage_coulomb(R0, ds.vq_cut, nk, vcst, st);
    std::map<std::pair<int, bool>, std::pair<GemmBatch, GemmBatch>> batches;
    auto get_batches = [&](int cnt, bool tri) -> std::pair<GemmBatch, GemmBatch> & {
        auto it = batches.find({cnt, tri});
        if (it != batches.end()) return it->second;
        auto &pr = batches[{cnt, tri}];
        GemmBatch &bx = pr.first, &bm = pr.second;
        bm.trans_a = true;
        bm.lower_only = true;
        for (int w = 0; w < cnt; ++w) {
            GemmProb px = make_prob(n2, n2, n2, (long long)w * nn, n2, 0, n2, (long long)w * nn, n2);
            px.kmode = tri ? 1 : 0;  // triangular Cholesky factor, or the full square root
            bx.add(px);
            GemmProb pm = make_prob(n2, n2, n2, 0, n2, (long long)w * nn, n2, (long long)w * NN, N);
            pm.kmode = tri ? 2 : 0;
            bm.add(pm);
        }
        const int cfg = T64x64;
        bx.finalize(cfg, st);
        bm.finalize(cfg, st);
        bx.flops = 2.0 * cnt * (double)n2 * n2 * n2 / 2.0;
        bm.flops = 2.0 * cnt * (double)n2 * n2 * n2 / 6.0;
        return pr;
    };
    const auto &cells = R0.cells;
    for (int iq = 0; iq < nk; ++iq) {
        if (mult[iq] == 0) continue;
        const double *kf = &ds.kfrac[(size_t)iq * 3];
        const bool gamma = std::fabs(kf[0]) < 1e-12 && std::fabs(kf[1]) < 1e-12 && std::fabs(kf[2]) < 1e-12;
        const double vmax = vst.vmax[iq], imax = vst.imax[iq];
        const bool head_q = gamma && !head.empty();
        bool eig_v = sqrt_threshold > 0 || head_q, eig_vc = sqrt_threshold > 0;
        bool chol = false;  // v_c = v with v = L S L^T, S = I: the Cholesky identity
        // v^{1/2} from Re v at q = 0 (power_hemat_blacs_real, epsilon.cpp:2019-2024)
        if (same_cut && gamma && imax > 1e-12 * std::max(vmax, 1e-300))
            LRA_THROW("W at q = 0 with v_c = v needs a real Coulomb matrix (Im v(0) != 0)");
        if (same_cut && !eig_v) {
            dense_coulomb(vst, iq, n, n2, embed, gamma, dV0.as<double>(), st);
            HIP_CHECK(hipMemcpyAsync(dL.get(), dV0.get(), nn * 8, hipMemcpyDeviceToDevice, st));
            prof.begin("gw_w_chol_v", st);
            HIP_CHECK(hipMemsetAsync(dProbeFlag.get(), 0, 16, st));
            cholesky_logdet(dL.as<double>(), n2, n2, dProbeFlag.as<double>(),
                            reinterpret_cast<int *>(dProbeFlag.as<double>() + 1), true, cw, st, dSg.as<double>());
            hsg.resize(n2);
            int zf = 0;
            HIP_CHECK(hipMemcpyAsync(hsg.data(), dSg.get(), (size_t)n2 * 8, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipMemcpyAsync(&zf, dProbeFlag.as<double>() + 1, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            chol = zf == 0;  // a zero or negative pivot needs the filtered root
            for (double v : hsg) chol = chol && v > 0;
            eig_v = eig_vc = !chol;
            prof.end("gw_w_chol_v", st, (double)n2 * n2 * n2 / 3.0);
        }
        if (!chol) {
            // dL <- S = v^{1/2}; dSc <- S_c = v_c^{1/2} (complex v_c kept, as the reference does);
            // dV0 <- v_c, or S_c S_c when a root is filtered
            if (!same_cut && !embed && vcst.imax[iq] > 1e-12 * std::max(vcst.vmax[iq], 1e-300))
                LRA_THROW("a complex truncated Coulomb matrix at Gamma-only sampling is outside this build");
            prof.begin("gw_w_sqrt_v", st);
            dense_coulomb(vst, iq, n, n2, embed, gamma, dL.as<double>(), st);
            if (!eig_v) eig_v = negative_pivot(dL.as<double>());
            if (eig_v) {
                const auto top2 = eig_sqrt(dL.as<double>(), n2, sqrt_threshold, head_q ? dU0.as<double>() : nullptr,
                                           embed ? 2 : 1, st);
                if (head_q && !(top2.first - top2.second > 1e-8 * std::fabs(top2.first)))
                    LRA_THROW("head of the dielectric matrix: the largest eigenvalue of v(0) is degenerate");
            } else {
                ns_sqrt(dL.as<double>(), n2, st);
            }
            dSc.reserve(nn * 8);
            if (same_cut) {
                eig_vc = eig_v;
                HIP_CHECK(hipMemcpyAsync(dSc.get(), dL.get(), nn * 8, hipMemcpyDeviceToDevice, st));
            } else {
                dense_coulomb(vcst, iq, n, n2, embed, false, dV0.as<double>(), st);
                HIP_CHECK(hipMemcpyAsync(dSc.get(), dV0.get(), nn * 8, hipMemcpyDeviceToDevice, st));
                if (!eig_vc) eig_vc = negative_pivot(dSc.as<double>());
                if (eig_vc) eig_sqrt(dSc.as<double>(), n2, sqrt_threshold, nullptr, 1, st);
                else ns_sqrt(dSc.as<double>(), n2, st);
            }
            if (eig_v || eig_vc) {
                GemmBatch sq;
                sq.add(make_prob(n2, n2, n2, 0, n2, 0, n2, 0, n2));
                sq.finalize(T64x64, st);
                gemm_run(Prec::F64, sq, dSc.get(), dSc.get(), dV0.get(), 1.0, 0.0, st);
            }
            prof.end("gw_w_sqrt_v", st);
        }
        std::vector<double> csr(nR), snr(nR);
        for (int rr = 0; rr < nR; ++rr) {
            const double ang = 2 * M_PI * (kf[0] * cells[rr][0] + kf[1] * cells[rr][1] + kf[2] * cells[rr][2]);
            csr[rr] = std::cos(ang);
            snr[rr] = std::sin(ang);
        }
        DevBuf dcs = upload(csr, st), dsn = upload(snr, st);
        if (verbose) { fprintf(stderr, "[librpa-mi355x] W: q %d / %d (%.1f s)\n", iq + 1, nk, now_s() - t0); fflush(stderr); }
        for (int w0 = 0; w0 < nw; w0 += WCH) {
            const int cnt = std::min(WCH, nw - w0);
            prof.begin("gw_w_chi0q", st);
            R0.assemble_dense(g, kf, w0, cnt, dD.get(), true, (long long)nn, embed, st);
            prof.end("gw_w_chi0q", st);
            auto &bb = get_batches(cnt, chol);
            prof.begin("gw_w_gemm", st);
            gemm_run(Prec::F64, bb.first, dD.get(), dL.get(), dX.get(), 1.0, 0.0, st);
            HIP_CHECK(hipMemsetAsync(dA.get(), 0, (size_t)cnt * NN * 8, st));
            for (int w = 0; w < cnt; ++w) {
                double *A = dA.as<double>() + (size_t)w * NN;
                launch_set_identity<double>(A, n2, N, st);
                launch_copy_block(chol ? dL.as<double>() : dSc.as<double>(), n2, A + (size_t)n2 * N, N, n2, n2, st);
                launch_copy_block(dV0.as<double>(), n2, A + (size_t)n2 * N + n2, N, n2, n2, st);
            }
            gemm_run(Prec::F64, bb.second, dL.get(), dX.get(), dA.get(), -1.0, 1.0, st);
            if (head_q) {
                // eps += (eps_mac - u^T eps u) u u^T, u the leading eigenvector of Re v(0): the
                // head overwrite of epsilon.cpp:2106-2181 in the eigenbasis of v
                HIP_CHECK(hipMemsetAsync(dQ.get(), 0, sizeof(double) * cnt, st));
                launch_head_quad(dA.as<double>(), (long long)NN, N, n, cnt, dU0.as<double>(), dQ.as<double>(), st);
                launch_head_update(dA.as<double>(), (long long)NN, N, n, embed ? n : 0, cnt, dU0.as<double>(),
                                   dHead.as<double>() + w0, dQ.as<double>(), st);
            }
            prof.end("gw_w_gemm", st, bb.first.flops + bb.second.flops);
            prof.begin("gw_w_schur", st);
            cholesky_logdet_batched(dA.as<double>(), (long long)NN, cnt, N, N, dacc.as<double>(), 1, dflag, false, cw,
                                    st, n2);
            prof.end("gw_w_schur", st, (double)cnt * (7.0 / 3.0) * n2 * (double)n2 * n2);
            // W(R, tau) += sum_w f2t(tau, w) / N_k (cos Re W + sin Im W)
            prof.begin("gw_w_accum", st);
            for (int a0 = 0; a0 < cnt; a0 += 8) {
                const int cw_n = std::min(8, cnt - a0);
                std::vector<double> coef((size_t)cw_n * g.n);
                for (int w = 0; w < cw_n; ++w)
                    for (int t = 0; t < g.n; ++t)
                        coef[(size_t)w * g.n + t] = g.cos_f2t[(size_t)t * g.n + w0 + a0 + w] * mult[iq] / nk;
                DevBuf dcoef = upload(coef, st);
                launch_w_accum(dA.as<double>() + (size_t)a0 * NN, (long long)NN, cw_n, n, n2, N, embed,
                               dcoef.as<double>(), g.n, dcs.as<double>(), dsn.as<double>(), nR, d_Wrt.as<double>(), st);
                HIP_CHECK(hipStreamSynchronize(st));
            }
            prof.end("gw_w_accum", st, 0, 8.0 * cnt * (double)n * n + 16.0 * g.n * nR * (double)n * n);
        }
    }
    int hflag[2] = {0, 0};
    HIP_CHECK(hipMemcpyAsync(hflag, dflag, sizeof(hflag), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (hflag[0]) LRA_THROW("1 - v^1/2 chi0 v^1/2 is singular or not finite at some (q, omega)");
    const double nq_w = (double)std::count_if(mult.begin(), mult.end(), [](int m) { return m > 0; });
    flops_w = nq_w * ((double)n2 * n2 * n2 / 3.0 + nw * (2.0 * n2 * (double)n2 * n2 / 2.0 + 2.0 * n2 * (double)n2 * n2 / 6.0 +
                                                       (7.0 / 3.0) * n2 * (double)n2 * n2));
    t_w = now_s() - t0;
}

// ------------------------------------------------------------------------------------------
// Sigma(R, i omega)
// ------------------------------------------------------------------------------------------
void GwEngine::build_sigma(const Dataset &ds, const TFGrid &g) {
    auto &prof = profiler();
    const double t0 = now_s();
    RpaEngine &R0 = *rpa;
    const hipStream_t st = R0.stream;
    const int nw = g.n;
    const long long nel = (long long)nR * NAO * NAO;
    d_Sre.alloc(sizeof(double) * (size_t)nspin * nw * nel);
    d_Sim.alloc(sizeof(double) * (size_t)nspin * nw * nel);
    HIP_CHECK(hipMemsetAsync(d_Sre.get(), 0, d_Sre.bytes(), st));
    HIP_CHECK(hipMemsetAsync(d_Sim.get(), 0, d_Sim.bytes(), st));
    if (w_zero) { HIP_CHECK(hipStreamSynchronize(st)); t_sigma = now_s() - t0; return; }
    d_U.alloc(u_elems * ts);
    d_Ghat.alloc(max_ghat * ts);
    d_TD.alloc(max_tdc * ts);
    d_Y.alloc(y_elems * ts);
    d_GC.alloc(gc_elems * ts);
    for (int sg = 0; sg < 2; ++sg) d_sig[sg].alloc((size_t)nR * NAO * NAO * ts);
    d_extA.alloc((size_t)NE * NAO * NAO * ts);
    d_extB.alloc((size_t)NE * NAO * NAO * ts);
    if (prec != Prec::F64) d_Wf.alloc((size_t)nR * NABF * NABF * ts);
    R0.prepare_gf(ds.mf, -1);  // all bands (get_gf_cplx_imagtimes_Rs, meanfield.cpp:309-365)
    const long long gpm_r = (long long)NAO * 2 * NAO;
    (void)gpm_r;
    const size_t esz = ts;
    const long long nn_abf = (long long)NABF * NABF;
    // LIBRPA_AMD_TUNE_SIGMA=1 times the Sigma batches' candidate tiles on the first pass (the
    // trials cost about as much as they save in one G0W0 run, so the defaults are used otherwise)
    const char *tsg = std::getenv("LIBRPA_AMD_TUNE_SIGMA");
    const bool tune_sigma = R0.autotune && tsg && std::atoi(tsg) != 0;
    for (int it = 0; it < g.n; ++it) {
        const void *Wf;
        if (prec == Prec::F64) {
            Wf = d_Wrt.as<double>() + (size_t)it * nR * nn_abf;
        } else {
            launch_convert<double, float>(d_Wrt.as<double>() + (size_t)it * nR * nn_abf, d_Wf.as<float>(),
                                          (long long)nR * nn_abf, st);
            Wf = d_Wf.get();
        }
        std::vector<double> cw(nw), sw(nw);
        for (int w = 0; w < nw; ++w) {
            cw[w] = g.cos_t2f[(size_t)w * g.n + it];
            sw[w] = g.sin_t2f[(size_t)w * g.n + it];
        }
        DevBuf dcw = upload(cw, st), dsw = upload(sw, st);
        if (verbose) { fprintf(stderr, "[librpa-mi355x] Sigma_c: tau %d / %d (%.1f s)\n", it + 1, g.n, now_s() - t0); fflush(stderr); }
        for (int s = 0; s < nspin; ++s) {
            R0.build_gf(ds.mf, g.tau[it], s);
            for (int sg = 0; sg < 2; ++sg) {
                HIP_CHECK(hipMemsetAsync(d_sig[sg].get(), 0, d_sig[sg].bytes(), st));
                HIP_CHECK(hipMemsetAsync(d_extA.get(), 0, d_extA.bytes(), st));
                HIP_CHECK(hipMemsetAsync(d_extB.get(), 0, d_extB.bytes(), st));
                HIP_CHECK(hipMemsetAsync(d_Y.get(), 0, d_Y.bytes(), st));
                // U_J(S) for all J, S
                for (auto &c : uchunks) {
                    prof.begin("gw_gather", st);
                    if (prec == Prec::F64)
                        launch_gather_cols<double>(R0.d_Gpm.as<double>(), 2LL * NAO, d_Ghat.as<double>(),
                                                   c.d_jobs[sg].as<GatherJob>(), (int)c.jobs[sg].size(),
                                                   c.d_colsrc[sg].as<long long>(), NAO, c.max_gcol, st);
                    else
                        launch_gather_cols<float>(R0.d_Gpm.as<float>(), 2LL * NAO, d_Ghat.as<float>(),
                                                  c.d_jobs[sg].as<GatherJob>(), (int)c.jobs[sg].size(),
                                                  c.d_colsrc[sg].as<long long>(), NAO, c.max_gcol, st);
                    prof.end("gw_gather", st, 0, 2.0 * c.ghat_elems * esz);
                    prof.begin("gw_U", st);
                    if (tune_sigma && !sigma_tuned)
                        c.ub.tune(prec, d_Ghat.get(), R0.d_CWb.get(), d_U.get(), 1.0, 0.0, st);
                    gemm_run(prec, c.ub, d_Ghat.get(), R0.d_CWb.get(), d_U.get(), 1.0, 0.0, st);
                    prof.end("gw_U", st, c.ub.flops);
                }
                if (tune_sigma && !sigma_tuned) {
                    // time the candidate tiles once per batch shape; the trials leave garbage in
                    // GC, Y, T^D and Sigma, which are reset before the real pass
                    for (int J = 0; J < natom; ++J) {
                        const char *gbase = R0.d_Gpm.as<char>() + (size_t)sg * R0.nb[J] * esz;
                        for (int rr = 0; rr < nR; ++rr) {
                            GemmBatch &b1 = a1[(size_t)J * nR + rr];
                            GemmBatch &b2 = a2[(size_t)J * nR + rr];
                            if (b1.empty()) continue;
                            b1.tune(prec, gbase, d_CA.get(), d_GC.get(), 1.0, 0.0, st);
                            gemm_run(prec, b1, gbase, d_CA.get(), d_GC.get(), 1.0, 0.0, st);
                            b2.tune(prec, Wf, d_GC.get(), d_Y.get(), 1.0, 1.0, st);
                        }
                    }
                    for (auto &c : ichunks) {
                        c.tdb.tune(prec, Wf, d_U.get(), d_TD.get(), 1.0, 0.0, st);
                        gemm_run(prec, c.tdb, Wf, d_U.get(), d_TD.get(), 1.0, 0.0, st);
                        c.dcon.tune(prec, d_CD.get(), d_TD.get(), d_sig[sg].get(), 1.0, 1.0, st);
                    }
                    HIP_CHECK(hipMemsetAsync(d_Y.get(), 0, d_Y.bytes(), st));
                    HIP_CHECK(hipMemsetAsync(d_sig[sg].get(), 0, d_sig[sg].bytes(), st));
                    sigma_tuned = true;
                    if (verbose) {
                        auto nm = [](const GemmBatch &b) { return tile_name(b.cfg); };
                        for (auto &c : uchunks) fprintf(stderr, "[librpa-mi355x] tile U %s\n", nm(c.ub));
                        for (int J = 0; J < natom; ++J)
                            if (!a1[(size_t)J * nR].empty())
                                fprintf(stderr, "[librpa-mi355x] tile A1 J=%d %s A2 %s\n", J, nm(a1[(size_t)J * nR]),
                                        nm(a2[(size_t)J * nR]));
                        for (auto &c : ichunks)
                            fprintf(stderr, "[librpa-mi355x] tile TD %s D %s\n", nm(c.tdb), nm(c.dcon));
                        fflush(stderr);
                    }
                }
                // partner-partner term: Y_I over all (J, R), then one contraction
                prof.begin("gw_A12", st);
                double fa = 0;
                for (int J = 0; J < natom; ++J) {
                    const char *gbase = R0.d_Gpm.as<char>() + (size_t)sg * R0.nb[J] * esz;
                    for (int rr = 0; rr < nR; ++rr) {
                        const GemmBatch &b1 = a1[(size_t)J * nR + rr];
                        const GemmBatch &b2 = a2[(size_t)J * nR + rr];
                        if (b1.empty()) continue;
                        gemm_run(prec, b1, gbase, d_CA.get(), d_GC.get(), 1.0, 0.0, st);
                        gemm_run(prec, b2, Wf, d_GC.get(), d_Y.get(), 1.0, 1.0, st);
                        fa += b1.flops + b2.flops;
                    }
                }
                prof.end("gw_A12", st, fa);
                prof.begin("gw_A3", st);
                gemm_run(prec, a3a, d_CB.get(), d_Y.get(), d_extA.get(), 1.0, 1.0, st);
                prof.end("gw_A3", st, a3a.flops);
                // owner-row terms by chunks of I
                for (auto &c : ichunks) {
                    prof.begin("gw_TD", st);
                    gemm_run(prec, c.tdb, Wf, d_U.get(), d_TD.get(), 1.0, 0.0, st);
                    prof.end("gw_TD", st, c.tdb.flops);
                    prof.begin("gw_D", st);
                    gemm_run(prec, c.dcon, d_CD.get(), d_TD.get(), d_sig[sg].get(), 1.0, 1.0, st);
                    prof.end("gw_D", st, c.dcon.flops);
                    prof.begin("gw_A3", st);
                    gemm_run(prec, c.a3b, d_CB.get(), d_TD.get(), d_extB.get(), 1.0, 1.0, st);
                    prof.end("gw_A3", st, c.a3b.flops);
                }
                prof.begin("gw_fold", st);
                if (prec == Prec::F64)
                    launch_sigma_fold<double>(d_sig[sg].as<double>(), d_extA.as<double>(), d_extB.as<double>(),
                                              d_fold.as<int>(), d_foldneg.as<int>(), NE, nR, NAO, st);
                else
                    launch_sigma_fold<float>(d_sig[sg].as<float>(), d_extA.as<float>(), d_extB.as<float>(),
                                             d_fold.as<int>(), d_foldneg.as<int>(), NE, nR, NAO, st);
                prof.end("gw_fold", st);
            }
            prof.begin("gw_freq", st);
            double *sre = d_Sre.as<double>() + (size_t)s * nw * nel;
            double *sim = d_Sim.as<double>() + (size_t)s * nw * nel;
            if (prec == Prec::F64)
                launch_sigma_freq<double>(d_sig[0].as<double>(), d_sig[1].as<double>(), nel, dcw.as<double>(),
                                          dsw.as<double>(), nw, sre, sim, st);
            else
                launch_sigma_freq<float>(d_sig[0].as<float>(), d_sig[1].as<float>(), nel, dcw.as<double>(),
                                         dsw.as<double>(), nw, sre, sim, st);
            prof.end("gw_freq", st);
        }
        HIP_CHECK(hipStreamSynchronize(st));
    }
    for (DevBuf *b : {&d_U, &d_Ghat, &d_TD, &d_Y, &d_GC, &d_sig[0], &d_sig[1], &d_extA, &d_extB, &d_Wf}) b->release();
    t_sigma = now_s() - t0;
}

void GwEngine::w_block(int it, int I, int J, int r, double *out) const {
    const RpaEngine &R0 = *rpa;
    const size_t nn = (size_t)NABF * NABF;
    std::vector<double> row(R0.na[J]);
    for (int mu = 0; mu < R0.na[I]; ++mu) {
        const size_t off = ((size_t)it * nR + r) * nn + (size_t)(R0.abf[I] + mu) * NABF + R0.abf[J];
        HIP_CHECK(hipMemcpy(out + (size_t)mu * R0.na[J], d_Wrt.as<double>() + off, sizeof(double) * R0.na[J],
                            hipMemcpyDeviceToHost));
    }
}

void GwEngine::sigma_block(int spin, int iw, int I, int J, int r, double *re, double *im) const {
    const RpaEngine &R0 = *rpa;
    const long long nel = (long long)nR * NAO * NAO;
    for (int i = 0; i < R0.nb[I]; ++i) {
        const size_t off = ((size_t)spin * ntau + iw) * nel + (size_t)r * NAO * NAO + (size_t)(R0.ao[I] + i) * NAO +
                           R0.ao[J];
        HIP_CHECK(hipMemcpy(re + (size_t)i * R0.nb[J], d_Sre.as<double>() + off, sizeof(double) * R0.nb[J],
                            hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(im + (size_t)i * R0.nb[J], d_Sim.as<double>() + off, sizeof(double) * R0.nb[J],
                            hipMemcpyDeviceToHost));
    }
}

// Sigma_KS(k, i w)_nn = sum_ab conj(U_na) Sigma(k)_ab U_nb with Sigma(k) = sum_R e^{i 2 pi k.R'} Sigma(R)
// and R' the image of R nearest to the atom pair (build_sigc_matrix_KS_blacs, gw.cpp:992-1040).
std::vector<std::vector<cplx>> GwEngine::ks_diag(const Dataset &ds, int spin, int ik,
                                                  const std::vector<int> &states) {
    const RpaEngine &R0 = *rpa;
    const hipStream_t st = R0.stream;
    const int nw = ntau;
    const double *kf = &ds.kfrac[(size_t)ik * 3];
    const bool frac = ds.frac.size() == 3 * (size_t)natom && ds.lat_set;
    std::vector<double> ph((size_t)natom * natom * nR * 2);
    for (int I = 0; I < natom; ++I)
        for (int J = 0; J < natom; ++J)
            for (int rr = 0; rr < nR; ++rr) {
                int Rb[3] = {R0.cells[rr][0], R0.cells[rr][1], R0.cells[rr][2]};
                if (frac) {
                    double best = 1e300;
                    int bc[3] = {Rb[0], Rb[1], Rb[2]};
                    for (int a = -1; a < 2; ++a)
                        for (int b = -1; b < 2; ++b)
                            for (int c = -1; c < 2; ++c) {
                                const int Rc[3] = {a * R0.period[0] + Rb[0], b * R0.period[1] + Rb[1],
                                                   c * R0.period[2] + Rb[2]};
                                double d2 = 0;
                                for (int x = 0; x < 3; ++x) {
                                    double v = 0;
                                    for (int y = 0; y < 3; ++y)
                                        v += (ds.frac[3 * I + y] - ds.frac[3 * J + y] - Rc[y]) * ds.latvec[3 * y + x];
                                    d2 += v * v;
                                }
                                if (d2 < best) { best = d2; bc[0] = Rc[0]; bc[1] = Rc[1]; bc[2] = Rc[2]; }
                            }
                    Rb[0] = bc[0]; Rb[1] = bc[1]; Rb[2] = bc[2];
                }
                const double ang = 2 * M_PI * (kf[0] * Rb[0] + kf[1] * Rb[1] + kf[2] * Rb[2]);
                ph[(((size_t)I * natom + J) * nR + rr) * 2] = std::cos(ang);
                ph[(((size_t)I * natom + J) * nR + rr) * 2 + 1] = std::sin(ang);
            }
    std::vector<int> atom_of(NAO);
    for (int I = 0; I < natom; ++I)
        for (int i = 0; i < R0.nb[I]; ++i) atom_of[R0.ao[I] + i] = I;
    DevBuf dph = upload(ph, st), dat = upload(atom_of, st);
    const auto &Uk = ds.mf.wfc[(size_t)spin * nk + ik];
    std::vector<std::vector<cplx>> out(states.size(), std::vector<cplx>(nw));
    const long long nel = (long long)nR * NAO * NAO;
    for (size_t s0 = 0; s0 < states.size(); s0 += 16) {
        const int ns = (int)std::min<size_t>(16, states.size() - s0);
        std::vector<double> U((size_t)ns * NAO * 2);
        for (int a = 0; a < ns; ++a)
            for (int b = 0; b < NAO; ++b) {
                const cplx u = Uk[(size_t)states[s0 + a] * NAO + b];
                U[((size_t)a * NAO + b) * 2] = u.real();
                U[((size_t)a * NAO + b) * 2 + 1] = u.imag();
            }
        DevBuf dU = upload(U, st);
        DevBuf dY(sizeof(double) * (size_t)nw * NAO * ns * 2);
        launch_sigma_ks_rows(d_Sre.as<double>() + (size_t)spin * nw * nel, d_Sim.as<double>() + (size_t)spin * nw * nel,
                             nw, nR, NAO, dat.as<int>(), natom, dph.as<double>(), dU.as<double>(), ns, dY.as<double>(),
                             st);
        std::vector<double> Y((size_t)nw * NAO * ns * 2);
        HIP_CHECK(hipMemcpyAsync(Y.data(), dY.get(), Y.size() * 8, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (int a = 0; a < ns; ++a)
            for (int w = 0; w < nw; ++w) {
                cplx acc(0, 0);
                for (int x = 0; x < NAO; ++x) {
                    const cplx y(Y[(((size_t)w * NAO + x) * ns + a) * 2], Y[(((size_t)w * NAO + x) * ns + a) * 2 + 1]);
                    acc += std::conj(Uk[(size_t)states[s0 + a] * NAO + x]) * y;
                }
                out[s0 + a][w] = acc;
            }
    }
    return out;
}

// ------------------------------------------------------------------------------------------
// analytic continuation and QP equation (host)
// ------------------------------------------------------------------------------------------
PadeAC::PadeAC(int n_pars, const std::vector<cplx> &xs, const std::vector<cplx> &data) {
    const int nd = (int)data.size();
    std::vector<cplx> px, pd;
    if (n_pars < 0) n_pars = nd;
    if (nd <= n_pars) {
        n_pars = nd;
        px = xs;
        pd = data;
    } else {
        // n_pars - 1 evenly spaced nodes, then the last datum. LibRPA sizes its node list to
        // n_pars and fills only the first n_pars - 1, so that datum sits at the node 0
        // (src/core/analycont.cpp:35-45); the same node keeps the continuation identical.
        const int step = nd / (n_pars - 1);
        for (int i = 0; i < n_pars - 1; ++i) { px.push_back(xs[(size_t)i * step]); pd.push_back(data[(size_t)i * step]); }
        px.push_back(cplx(0.0, 0.0));
        pd.push_back(data[nd - 1]);
    }
    n = n_pars;
    x = px;
    // reciprocal differences, g(i, p) for i >= p
    std::vector<cplx> gm((size_t)n * n);
    for (int i = 0; i < n; ++i) gm[(size_t)i * n] = pd[i];
    for (int p = 1; p < n; ++p)
        for (int i = p; i < n; ++i)
            gm[(size_t)i * n + p] = (gm[(size_t)(p - 1) * n + p - 1] - gm[(size_t)i * n + p - 1]) /
                                    ((x[i] - x[p - 1]) * gm[(size_t)i * n + p - 1]);
    y.resize(n);
    for (int p = 0; p < n; ++p) y[p] = gm[(size_t)p * n + p];
}

cplx PadeAC::operator()(cplx z) const {
    cplx t(1.0, 0.0);
    for (int p = n - 1; p > 0; --p) t = 1.0 + y[p] * (z - x[p - 1]) / t;
    return y[0] / t;
}

int qpe_solve(const PadeAC &p, double e_mf, double e_fermi, double vxc, double sigx, double &e_qp, cplx &sigc,
              double diff_init, double thres, int n_iter_max) {
    const double escale = 0.1;
    int n_iter = 0;
    double e_last = e_mf, e_this;
    double diff = diff_init;
    cplx s_this, s_last;
    while (n_iter++ < n_iter_max) {
        e_this = e_last + escale * diff;
        s_this = p(cplx(e_this - e_fermi, 0.0));
        diff = e_mf - vxc + sigx + s_this.real() - e_this;
        e_last = e_mf - vxc + sigx + s_this.real();
        s_last = s_this;
        if (std::fabs(diff) < thres || n_iter == n_iter_max - 1) break;
    }
    sigc = s_last;
    e_qp = e_last;
    return n_iter > n_iter_max - 1 ? 1 : 0;
}

// ------------------------------------------------------------------------------------------
// API drivers
// ------------------------------------------------------------------------------------------
void gw_build_sigma(Dataset &d, const LibrpaOptions &o) {
    if (o.use_shrink_abfs == LIBRPA_SWITCH_ON) LRA_THROW("ABF shrinking is outside this build's path");
    if (o.sqrt_coulomb_threshold < 0) LRA_THROW("sqrt_coulomb_threshold must not be negative");
    lra_prepare_grid(d, o);
    // head of the dielectric matrix from the macroscopic dielectric function
    // (librpa_get_g0w0_sigc_kgrid path, src/api/compute_g0w0.cpp:102-118), on the grid's
    // frequencies as interpolate_dielec_func (src/core/dielecmodel.cpp:65-104) puts it:
    // 0 as given, 1 cubic spline, 2 Havriliak-Negami fit
    std::vector<double> head;
    if (o.replace_w_head == LIBRPA_SWITCH_ON && !d.dielect_omega.empty())
        head = interpolate_dielec(o.option_dielect_func, d.dielect_omega, d.dielect_val, d.grid.omega);
    auto eng = d.rpa;
    if (!eng || eng->prec != d.chi0_prec || eng->mixed != d.chi0_mixed || eng->planned_ntau != d.grid.n) {
        d.rpa.reset();
        eng = std::make_shared<RpaEngine>();
        eng->mixed = d.chi0_mixed;
        eng->plan(d, d.chi0_prec, d.ec_prec, d.grid.n);
        d.rpa = eng;
    }
    eng->build_chi0(d, d.grid, o.n_bands_chi0);
    eng->release_work();  // chi0(R, tau) stays; the chi0 work buffers are rebuilt on demand
    eng->release_ec();    // an earlier Ec evaluation's buffers
    auto gw = std::make_shared<GwEngine>();
    gw->verbose = o.output_level >= LIBRPA_VERBOSE_DEBUG;
    gw->head = head;
    gw->sqrt_threshold = o.sqrt_coulomb_threshold;
    gw->plan(d, eng);
    gw->build_w(d, d.grid);
    gw->build_sigma(d, d.grid);
    d.gw = gw;
}

void gw_sigc_kgrid(Dataset &d, const LibrpaOptions &o, int n_spins, int n_kpts_this, const int *iks_this,
                   int i_state_low, int i_state_high, const double *vxc, const double *vexx, double *sigc_re,
                   double *sigc_im) {
    const auto &mf = d.mf;
    i_state_low = std::max(0, i_state_low);
    i_state_high = std::min(mf.nbands, i_state_high);
    if (n_spins != mf.nspin) LRA_THROW("parsed nspins is not consistent with the SCF starting point");
    if (i_state_high <= i_state_low) return;
    const int nst = i_state_high - i_state_low;
    if (!d.gw) gw_build_sigma(d, o);
    GwEngine &g = *d.gw;
    std::vector<cplx> xs;
    for (int w = 0; w < d.grid.n; ++w) xs.push_back(cplx(0.0, d.grid.omega[w]));
    std::vector<int> states;
    for (int i = i_state_low; i < i_state_high; ++i) states.push_back(i);
    for (int isp = 0; isp < n_spins; ++isp) {
        const int start_isp = isp * n_kpts_this * nst;
        for (int kk = 0; kk < n_kpts_this; ++kk) {
            const int start_k = start_isp + kk * nst;
            const int ik = iks_this[kk];
            if (ik < 0 || ik >= mf.nk) LRA_THROW("k index out of range");
            auto vals = g.ks_diag(d, isp, ik, states);
            for (int i = 0; i < nst; ++i) {
                const double eks = mf.eps[((size_t)isp * mf.nk + ik) * mf.nbands + states[i]];
                sigc_re[start_k + i] = std::nan("");
                sigc_im[start_k + i] = std::nan("");
                PadeAC pade(o.n_params_anacon, xs, vals[i]);
                double eqp;
                cplx sc;
                if (qpe_solve(pade, eks, mf.efermi, vxc[start_k + i], vexx[start_k + i], eqp, sc) == 0) {
                    sigc_re[start_k + i] = sc.real();
                    sigc_im[start_k + i] = sc.imag();
                }
            }
        }
    }
}

}  // namespace lra
