// Space-time chi0 builder and RPA correlation energy on one GPU.
//
// Per imaginary time tau (LibRPA build_chi0_q_space_time_LibRI_routing,
// src/core/chi0.cpp:1014-1164) the device pipeline is
//   1. G(R, +-tau): per k, M_k = U^T diag(s) conj(U) as one real GEMM on the
//      stacked [Re; Im] eigenvectors, then the k -> R sum keeping the real part
//      (build_gf_Rt_libri_serial, chi0.cpp:303-426).  G(+tau) and G(-tau) are
//      stored side by side per atom column block: Gpm[R][row][2*ao(J) + s*nb(J) + j].
//   2. W(K, J, S)[k, (j, nu)] = sum_{(L,R2) in N(J)} G(S+R2)_KL C^J(L,R2): one GEMM
//      per (J, S, sign) on a gathered G panel, written as W[(s, j)][(K, k)][nu].
//   3. T(I,J,R)[(K,R1,k,i)][nu] = G_KJ(R-R1) W(I,J,R) + G_IJ(R) W(K,J,R-R1): two row-
//      indirect GEMM batches with K = 2 nb(J) (both signs at once).
//   4. chi0s(I,J,R) = Cfin_I * T: one GEMM per output block, K = sum_K nb(I) nb(K).
// The four terms of LibRI's labels a1b1_a2b2 and a1b2_a2b1 called with
// (G+, G-) and (G-, G+) (thirdparty/LibRI/include/RI/physics/RPA.hpp:104-138)
// are the four (sign, term) products of steps 3-4.
//
// chi0s(R, tau) of every tau stays resident; the tau -> omega cosine transform
// and the R -> q sum (chi_libri_ft_ct, chi0.cpp:507-596) are fused into the
// assembly of dense Hermitian chi0(q, omega) matrices, and Ec
// (compute_RPA_correlation_blacs_2d{,_gamma_only}, src/core/epsilon.cpp:43-558)
// uses ln |det(1 - V chi0)| = ln |det(S - L^T chi0 L)| with V = L S L^T (S = diag(+-1), the
// pivot signs; S = I for a positive-definite V).
#include "rpa.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>

namespace lra {

namespace {
inline int bvk_mod(int v, int n) { return (((v % n) + 3 * n / 2) % n) - n / 2; }
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

int RpaEngine::cell_index(int x, int y, int z) const {
    const int x0 = -(period[0] / 2), y0 = -(period[1] / 2), z0 = -(period[2] / 2);
    return ((x - x0) * period[1] + (y - y0)) * period[2] + (z - z0);
}

int RpaEngine::fold_index(int x, int y, int z) const {
    return cell_index(bvk_mod(x, period[0]), bvk_mod(y, period[1]), bvk_mod(z, period[2]));
}

void RpaEngine::plan(const Dataset &ds, Prec p, Prec pe, int ntau) {
    prec = p;
    ec_prec = pe;
    planned_ntau = ntau;
    ts = (prec == Prec::F64) ? 8 : 4;
    // K parts of each contraction problem (LIBRPA_AMD_SPLITK, default 4; 1 = one plain write)
    int nsk = 4;
    if (const char *e = std::getenv("LIBRPA_AMD_SPLITK")) nsk = std::max(1, std::atoi(e));
    const auto &mf = ds.mf;
    natom = (int)ds.nb_wfc.size();
    if (natom == 0 || ds.nb_aux.size() != (size_t)natom) LRA_THROW("atomic basis sizes not set");
    nb.assign(ds.nb_wfc.begin(), ds.nb_wfc.end());
    na.assign(ds.nb_aux.begin(), ds.nb_aux.end());
    ao.assign(natom + 1, 0);
    abf.assign(natom + 1, 0);
    for (int a = 0; a < natom; ++a) { ao[a + 1] = ao[a] + nb[a]; abf[a + 1] = abf[a] + na[a]; }
    NAO = (int)ao[natom];
    NABF = (int)abf[natom];
    if (NAO != mf.nbasis) LRA_THROW("sum of atomic basis sizes != nbasis of the mean-field input");
    nk = mf.nk;
    nspin = mf.nspin;
    nbands = mf.nbands;
    if (mf.nspinor != 1) LRA_THROW("spinor (SOC) input is outside this build's path");
    for (int a = 0; a < 3; ++a) period[a] = ds.period[a] > 0 ? ds.period[a] : 1;
    if (period[0] * period[1] * period[2] != nk) LRA_THROW("k mesh size does not match the number of k points");
    cells.clear();
    for (int x = -(period[0] / 2); x <= (period[0] - 1) / 2; ++x)
        for (int y = -(period[1] / 2); y <= (period[1] - 1) / 2; ++y)
            for (int z = -(period[2] / 2); z <= (period[2] - 1) / 2; ++z) cells.push_back({x, y, z});
    nR = (int)cells.size();

    // ---- fold Cs to the BvK cell (RI_Tools::cal_period) ----
    std::map<std::tuple<int, int, int>, std::vector<double>> folded;  // (I, K, r) -> API layout
    for (const auto &kv : ds.cs) {
        const int I = std::get<0>(kv.first), K = std::get<1>(kv.first);
        const int r = fold_index(std::get<2>(kv.first), std::get<3>(kv.first), std::get<4>(kv.first));
        auto &dst = folded[{I, K, r}];
        if (dst.empty()) dst = kv.second;
        else for (size_t i = 0; i < dst.size(); ++i) dst[i] += kv.second[i];
    }
    n_cs_blocks = (long long)folded.size();
    nbrs.assign(natom, {});
    for (const auto &kv : folded) nbrs[std::get<0>(kv.first)].push_back({std::get<1>(kv.first), std::get<2>(kv.first)});
    koff.assign(natom, {});
    loff.assign(natom, {});
    ksum.assign(natom, 0);
    lsum.assign(natom, 0);
    cfin_off.assign(natom + 1, 0);
    cwb_off.assign(natom + 1, 0);
    for (int I = 0; I < natom; ++I) {
        long long k = 0, l = 0;
        for (const auto &n : nbrs[I]) {
            koff[I].push_back(k);
            loff[I].push_back(l);
            k += (long long)nb[n.K] * nb[I];
            l += nb[n.K];
        }
        ksum[I] = k;
        lsum[I] = l;
        cfin_off[I + 1] = cfin_off[I] + (long long)na[I] * ksum[I];
        cwb_off[I + 1] = cwb_off[I] + lsum[I] * nb[I] * na[I];
    }
    // packed Cs in the two GEMM layouts
    {
        std::vector<double> cfin(cfin_off[natom]), cwb(cwb_off[natom]);
        cnorm.assign(natom, 0.0);
        for (int I = 0; I < natom; ++I) {
            for (size_t n = 0; n < nbrs[I].size(); ++n) {
                const int K = nbrs[I][n].K;
                const auto &blk = folded.at({I, K, nbrs[I][n].r});
                double ss = 0;
                for (double v : blk) ss += v * v;
                cnorm[I] += std::sqrt(ss);
                const int ni = nb[I], nkk = nb[K], nm = na[I];
                for (int i = 0; i < ni; ++i)
                    for (int k = 0; k < nkk; ++k)
                        for (int mu = 0; mu < nm; ++mu) {
                            const double v = blk[((size_t)i * nkk + k) * nm + mu];
                            // Cfin_I stored transposed: [koff + k * nb(I) + i][mu]
                            cfin[cfin_off[I] + (size_t)(koff[I][n] + (size_t)k * ni + i) * nm + mu] = v;
                            // CWb_I[(loff + k)][i * na(I) + mu]   (I plays the role of J, K of L)
                            cwb[cwb_off[I] + (size_t)(loff[I][n] + k) * ni * nm + (size_t)i * nm + mu] = v;
                        }
            }
        }
        if (prec == Prec::F64) {
            d_Cfin = upload(cfin, stream);
            d_CWb = upload(cwb, stream);
            if (mixed) {
                std::vector<float> f1(cfin.begin(), cfin.end()), f2(cwb.begin(), cwb.end());
                d_Cfin32 = upload(f1, stream);
                d_CWb32 = upload(f2, stream);
                HIP_CHECK(hipStreamSynchronize(stream));
            }
        } else {
            std::vector<float> f1(cfin.begin(), cfin.end()), f2(cwb.begin(), cwb.end());
            d_Cfin = upload(f1, stream);
            d_CWb = upload(f2, stream);
            HIP_CHECK(hipStreamSynchronize(stream));
        }
    }

    // ---- outputs (I <= J, all cells), ordered by J ----
    outs.clear();
    E_tot = 0;
    for (int J = 0; J < natom; ++J)
        for (int I = 0; I <= J; ++I)
            for (int r = 0; r < nR; ++r) {
                outs.push_back({I, J, r, E_tot, 0});
                E_tot += (long long)na[I] * na[J];
            }
    std::vector<int> out_index((size_t)natom * natom * nR, -1);
    for (size_t o = 0; o < outs.size(); ++o)
        out_index[((size_t)outs[o].I * natom + outs[o].J) * nR + outs[o].r] = (int)o;

    // ---- flop model (canonical count, per spin per tau) ----
    {
        double f = 0;
        for (int J = 0; J < natom; ++J) f += (double)nR * 2.0 * 2.0 * NAO * nb[J] * na[J] * lsum[J];
        for (const auto &o : outs)
            for (const auto &n : nbrs[o.I])
                f += 8.0 * nb[o.I] * nb[o.J] * nb[n.K] * na[o.J] + 2.0 * na[o.I] * nb[o.I] * nb[n.K] * na[o.J];
        flops_chi0_alg = f * ntau;
    }

    // ---- W(K, J, S) of every J stays resident: [(s', j)][(K, k)][nu] per (J, S) ----
    w_off.assign((size_t)natom * nR + 1, 0);
    for (int J = 0; J < natom; ++J)
        for (int S = 0; S < nR; ++S)
            w_off[(size_t)J * nR + S + 1] = w_off[(size_t)J * nR + S] + 2LL * nb[J] * NAO * na[J];
    // ---- chunking over J by memory budget (gathered G panels and T) ----
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    // chi0(R, tau) of every tau beside the W slabs, or staged through pinned host memory
    const double w_bytes = (double)w_off.back() * ts, slice_bytes = (double)ts * E_tot;
    free_host_store();
    host_staged = w_bytes + slice_bytes * ntau + 24e9 > (double)free_b;
    if (const char *hs = std::getenv("LIBRPA_AMD_HOST_CHI0")) host_staged = host_staged || std::atoi(hs) != 0;
    double budget = host_staged ? std::min(48e9, 0.5 * ((double)free_b - w_bytes - slice_bytes - 16e9))
                                : std::min(48e9, 0.35 * (double)free_b - w_bytes);
    // LIBRPA_AMD_CHUNK_BYTES=<bytes> caps the chunk size (smaller chunks, same results)
    if (const char *cb = std::getenv("LIBRPA_AMD_CHUNK_BYTES")) {
        const double v = std::atof(cb);
        if (v > 0) budget = std::min(budget, v);
    }
    chunks.clear();
    {
        int j = 0, o = 0;
        while (j < natom) {
            Chunk c;
            c.j_begin = j;
            c.o_begin = o;
            long long g = 0, t = 0;
            int jj = j, oo = o;
            while (jj < natom) {
                long long g1 = (long long)nR * 2 * NAO * lsum[jj];
                long long t1 = 0;
                int o2 = oo;
                while (o2 < (int)outs.size() && outs[o2].J == jj) { t1 += ksum[outs[o2].I] * na[jj]; ++o2; }
                const double bytes = (double)(g + g1 + t + t1) * ts;
                if (jj > j && bytes > budget) break;
                g += g1; t += t1; oo = o2; ++jj;
            }
            c.j_end = jj;
            c.o_end = oo;
            c.ghat_elems = g; c.t_elems = t;
            chunks.push_back(std::move(c));
            j = jj;
            o = oo;
        }
    }

    const long long gpm_r = (long long)NAO * 2 * NAO;  // elements per cell of Gpm
    flops_chi0_exec = 0;
    // reverse neighbour lists: for each K the (I, n) with nbrs[I][n].K == K
    std::vector<std::vector<std::pair<int, int>>> rev(natom);
    for (int I = 0; I < natom; ++I)
        for (size_t n = 0; n < nbrs[I].size(); ++n) rev[nbrs[I][n].K].push_back({I, (int)n});
    std::set<int> classes(nb.begin(), nb.end());
    std::map<std::tuple<int, int, int, int>, int> kt_pos;  // K-table start per species pair (per chunk)

    auto say = [&](const char *what, size_t a, size_t b) {
        if (!verbose) return;
        fprintf(stderr, "[librpa-mi355x] plan: %s %zu / %zu\n", what, a, b);
        fflush(stderr);
    };
    say("outputs (I <= J, R) and chunks", outs.size(), chunks.size());
    for (auto &c : chunks) {
        say("problem lists of chunk", (size_t)(&c - chunks.data()) + 1, chunks.size());
        kt_pos.clear();
        // offsets inside the chunk buffers
        long long go = 0, to = 0;
        for (int J = c.j_begin; J < c.j_end; ++J)
            for (int S = 0; S < nR; ++S)
                for (int sg = 0; sg < 2; ++sg) { c.ghat_off.push_back(go); go += (long long)NAO * lsum[J]; }
        for (int o = c.o_begin; o < c.o_end; ++o) { outs[o].toff = to; to += ksum[outs[o].I] * na[outs[o].J]; }
        auto ghat_at = [&](int J, int S, int sg) { return c.ghat_off[((size_t)(J - c.j_begin) * nR + S) * 2 + sg]; };
        auto w_at = [&](int J, int S) { return w_off[(size_t)J * nR + S]; };

        // gather jobs: Ghat(J,S,sg)[row][(n, l)] = G^sg(S + R2_n)[row][ao(L_n) + l]
        for (int J = c.j_begin; J < c.j_end; ++J)
            for (int S = 0; S < nR; ++S)
                for (int sg = 0; sg < 2; ++sg) {
                    GatherJob gj;
                    gj.dst_off = ghat_at(J, S, sg);
                    gj.ncol = (int)lsum[J];
                    gj.col_off = (int)c.colsrc.size();
                    for (const auto &n : nbrs[J]) {
                        const auto &R2 = cells[n.r];
                        const auto &Sv = cells[S];
                        const int rr = fold_index(Sv[0] + R2[0], Sv[1] + R2[1], Sv[2] + R2[2]);
                        for (int l = 0; l < nb[n.K]; ++l)
                            c.colsrc.push_back(rr * gpm_r + 2 * ao[n.K] + (long long)sg * nb[n.K] + l);
                    }
                    c.max_gcol = std::max(c.max_gcol, gj.ncol);
                    c.jobs.push_back(gj);
                }
        // W build: W[(s', j)][(K,k)][nu], s' = 1 holds W+ (from G+), s' = 0 holds W- (from G-)
        for (int J = c.j_begin; J < c.j_end; ++J)
            for (int S = 0; S < nR; ++S)
                for (int sg = 0; sg < 2; ++sg) {
                    const int sp = (sg == 0) ? 1 : 0;
                    GemmProb pr = make_prob(NAO, nb[J] * na[J], (int)lsum[J], ghat_at(J, S, sg), (int)lsum[J],
                                            cwb_off[J], nb[J] * na[J],
                                            w_at(J, S) + (long long)sp * nb[J] * NAO * na[J], na[J]);
                    pr.ncb = na[J];
                    pr.cbs = (long long)NAO * na[J];
                    if (pr.K > 0) c.wb.add(pr);
                }
        // G_KJ(R - R1) W(I, J, R) contracted with C^I directly: with
        // X(I, J, R)[mu][(j, i)] = sum_(K,R1) C^I(K,R1)(mu, i, k) G_KJ(R - R1)(k, j) = W(J, -R)-block of I,
        // chi0s(I, J, R) += sum_(s', j, i) W(I, -R)[(1 - s', i)][(J, j)][mu] W(J, R)[(s', j)][(I, i)][nu]
        for (int o = c.o_begin; o < c.o_end; ++o) {
            const auto &O = outs[o];
            const auto &Rv = cells[O.r];
            const int mr = fold_index(-Rv[0], -Rv[1], -Rv[2]);
            const int ni = nb[O.I], nj = nb[O.J];
            const auto key = std::make_tuple(ni, na[O.I], nj, na[O.J]);
            auto it = kt_pos.find(key);
            if (it == kt_pos.end()) {
                it = kt_pos.emplace(key, (int)(c.xb.ktab.size() / 2)).first;
                for (int sp = 0; sp < 2; ++sp)
                    for (int j = 0; j < nj; ++j)
                        for (int i = 0; i < ni; ++i) {
                            c.xb.ktab.push_back(((long long)(1 - sp) * ni + i) * NAO * na[O.I] + (long long)j * na[O.I]);
                            c.xb.ktab.push_back(((long long)sp * nj + j) * NAO * na[O.J] + (long long)i * na[O.J]);
                        }
            }
            GemmProb pr = make_prob(na[O.I], na[O.J], 2 * ni * nj, w_at(O.I, mr) + ao[O.J] * na[O.I], 0,
                                    w_at(O.J, O.r) + ao[O.I] * na[O.J], 0, O.eoff, na[O.J]);
            pr.kt_off = it->second;
            c.xb.add(pr);
            c.x_owner.push_back(o);
        }
        // T_A: per (J, S, K, class of I), rows (I, n, i) from G_IJ(S + R1), B = W(K, J, S)
        for (int J = c.j_begin; J < c.j_end; ++J)
            for (int S = 0; S < nR; ++S)
                for (int K = 0; K < natom; ++K)
                    for (int cls : classes) {
                        GemmProb pr = make_prob(0, nb[K] * na[J], 2 * nb[J], 0, 0, w_at(J, S) + ao[K] * na[J],
                                                NAO * na[J], 0, 0);
                        pr.ncb = na[J];
                        pr.cbs = (long long)cls * na[J];
                        pr.rt_off = (int)(c.ta.rowtab.size() / 2);
                        int m = 0;
                        const size_t owners0 = c.ta_outs.size();
                        const auto &Sv = cells[S];
                        for (const auto &in : rev[K]) {
                            const int I = in.first, n = in.second;
                            if (I > J || nb[I] != cls) continue;
                            const auto &R1 = cells[nbrs[I][n].r];
                            const int rr = fold_index(Sv[0] + R1[0], Sv[1] + R1[1], Sv[2] + R1[2]);
                            const int o = out_index[((size_t)I * natom + J) * nR + rr];
                            if (c.ta_outs.size() == owners0 || c.ta_outs.back() != o) c.ta_outs.push_back(o);
                            for (int i = 0; i < nb[I]; ++i) {
                                c.ta.rowtab.push_back(rr * gpm_r + (ao[I] + i) * 2LL * NAO + 2 * ao[J]);
                                c.ta.rowtab.push_back(outs[o].toff + (koff[I][n] + i) * na[J]);
                                ++m;
                            }
                        }
                        if (m == 0) {
                            c.ta.rowtab.resize((size_t)pr.rt_off * 2);
                            c.ta_outs.resize(owners0);
                            continue;
                        }
                        pr.M = m;
                        c.ta.add(pr);
                        if (c.ta_ptr.empty()) c.ta_ptr.push_back(0);
                        c.ta_ptr.push_back((int)c.ta_outs.size());
                    }
        // final: chi0s(o) = Cfin_I (na_I x ksum_I) * T_o (ksum_I x na_J), K split in nsk parts that
        // accumulate atomically: each output block has few, long tiles (K ~ 25k), and shorter ones
        // fill the last wave of a launch better
        for (int o = c.o_begin; o < c.o_end; ++o) {
            const auto &O = outs[o];
            if (ksum[O.I] == 0) continue;
            const int cls = (na[O.I] > 64 ? 2 : 0) + (na[O.J] > 64 ? 1 : 0);
            const int kt = (int)ksum[O.I];
            const int step = std::max(32, (kt / nsk + 31) / 32 * 32);
            for (int k0 = 0; k0 < kt; k0 += step) {
                const int kl = std::min(step, kt - k0);
                c.fin[cls].add(make_prob(na[O.I], na[O.J], kl, cfin_off[O.I] + (long long)k0 * na[O.I], na[O.I],
                                         O.toff + (long long)k0 * na[O.J], na[O.J], O.eoff, na[O.J]));
                c.fin_owner[cls].push_back(o);
            }
        }
        c.wb.finalize(big_tile(prec), stream);
        c.xb.trans_a = true;
        c.xb.finalize(T64x64, stream);
        c.ta.tag = 4;
        c.ta.finalize(T64x128, stream);
        c.fin_flops = 0;
        for (auto &f : c.fin) {
            f.tag = 3;
            f.atomic = nsk > 1;
            f.trans_a = true;  // Cfin_I is stored k-major, so A tiles load along mu
            f.finalize(T64x64, stream);
            c.fin_flops += f.flops;
        }
        c.d_jobs = upload(c.jobs, stream);
        c.d_colsrc = upload(c.colsrc, stream);
        flops_chi0_exec += c.wb.flops + c.xb.flops + c.ta.flops + c.fin_flops;
        // algorithmic HBM bytes of each contraction launch: Cfin_I and T read once, chi0s read and written
        for (int k = 0; k < 4; ++k) {
            c.fin_bytes[k] = 0;
            for (const auto &p : c.fin[k].probs)
                c.fin_bytes[k] += (double)ts * ((double)p.M * p.K + (double)p.K * p.N + 2.0 * p.M * p.N);
        }
        c.ta_bytes = 0;
        for (const auto &p : c.ta.probs)
            c.ta_bytes += (double)ts * ((double)p.M * p.K + (double)p.K * p.N + (double)p.M * p.N);
    }
    flops_chi0_exec *= ntau;
    size_t max_g = 0, max_t = 0;
    for (auto &c : chunks) {
        max_g = std::max(max_g, (size_t)c.ghat_elems);
        max_t = std::max(max_t, (size_t)c.t_elems);
    }
    ghat_max = max_g;
    t_max = max_t;
    d_Ghat.alloc(max_g * ts);
    d_W.alloc((size_t)w_off.back() * ts);
    d_T.alloc(max_t * ts);
    screen_plan(ds);

    // ---- G build data ----
    {
        std::vector<double> zz((size_t)nk * 2 * nbands * 2 * NAO), wre((size_t)nk * nbands * NAO),
            wim((size_t)nk * nbands * NAO);
        d_eps.alloc(sizeof(double) * (size_t)nspin * nk * nbands);
        d_wg.alloc(sizeof(double) * (size_t)nspin * nk * nbands);
        HIP_CHECK(hipMemcpyAsync(d_eps.get(), mf.eps.data(), d_eps.bytes(), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d_wg.get(), mf.wg.data(), d_wg.bytes(), hipMemcpyHostToDevice, stream));
        d_ZZ.alloc(sizeof(double) * zz.size() * nspin);
        d_wre.alloc(sizeof(double) * wre.size() * nspin);
        d_wim.alloc(sizeof(double) * wim.size() * nspin);
        for (int s = 0; s < nspin; ++s) {
            for (int k = 0; k < nk; ++k) {
                const auto &U = mf.wfc[(size_t)s * nk + k];
                if (U.size() != (size_t)nbands * NAO) LRA_THROW("eigenvectors missing for a k point");
                double *z = &zz[(size_t)k * 2 * nbands * 2 * NAO];
                for (int b = 0; b < nbands; ++b)
                    for (int j = 0; j < NAO; ++j) {
                        const double x = U[(size_t)b * NAO + j].real(), y = U[(size_t)b * NAO + j].imag();
                        z[(size_t)(2 * b) * 2 * NAO + j] = x;
                        z[(size_t)(2 * b) * 2 * NAO + NAO + j] = y;
                        z[(size_t)(2 * b + 1) * 2 * NAO + j] = y;
                        z[(size_t)(2 * b + 1) * 2 * NAO + NAO + j] = -x;
                        wre[((size_t)k * nbands + b) * NAO + j] = x;
                        wim[((size_t)k * nbands + b) * NAO + j] = y;
                    }
            }
            HIP_CHECK(hipMemcpy(d_ZZ.as<double>() + zz.size() * s, zz.data(), zz.size() * 8, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(d_wre.as<double>() + wre.size() * s, wre.data(), wre.size() * 8, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(d_wim.as<double>() + wim.size() * s, wim.data(), wim.size() * 8, hipMemcpyHostToDevice));
        }
        d_Zs.alloc(sizeof(double) * (size_t)nk * 2 * nbands * NAO);
        d_Mst.alloc(sizeof(double) * (size_t)nk * 2 * NAO * NAO);
        d_Gpm.alloc(ts * (size_t)nR * NAO * 2 * NAO);
        std::vector<double> cs((size_t)nR * nk), sn((size_t)nR * nk);
        for (int r = 0; r < nR; ++r)
            for (int k = 0; k < nk; ++k) {
                const double *kf = &ds.kfrac[(size_t)k * 3];
                const double ang = 2 * M_PI * (kf[0] * cells[r][0] + kf[1] * cells[r][1] + kf[2] * cells[r][2]);
                cs[(size_t)r * nk + k] = std::cos(ang);
                sn[(size_t)r * nk + k] = std::sin(ang);
            }
        d_cs = upload(cs, stream);
        d_sn = upload(sn, stream);
        std::vector<int> gcol(2 * (size_t)NAO);
        for (int J = 0; J < natom; ++J)
            for (int j = 0; j < nb[J]; ++j)
                for (int sg = 0; sg < 2; ++sg) gcol[(size_t)sg * NAO + ao[J] + j] = (int)(2 * ao[J] + sg * nb[J] + j);
        d_gcol = upload(gcol, stream);
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // ---- chi0s store and transform descriptors ----
    d_chi0s.alloc(ts * (size_t)E_tot * (host_staged ? 1 : ntau));
    if (host_staged) {
        // pageable memory: a store this size can exceed what the host lets a process pin
        h_chi0s.assign(ntau, {});
        for (int t = 0; t < ntau; ++t) h_chi0s[t].resize(ts * (size_t)E_tot);
    }
    pairs.clear();
    pair_eoff.clear();
    n_pair_elems = 0;
    for (int I = 0; I < natom; ++I)
        for (int J = I; J < natom; ++J) {
            PairDesc pd;
            pd.I = I; pd.J = J; pd.naI = na[I]; pd.naJ = na[J];
            pd.rowI = (int)abf[I]; pd.colJ = (int)abf[J];
            pd.e0 = n_pair_elems;
            n_pair_elems += (long long)na[I] * na[J];
            pairs.push_back(pd);
            for (int r = 0; r < nR; ++r) {
                const int o = out_index[((size_t)I * natom + J) * nR + r];
                pair_eoff.push_back(o >= 0 ? outs[o].eoff : -1);
            }
        }
    d_pairs = upload(pairs, stream);
    d_pair_eoff = upload(pair_eoff, stream);
    HIP_CHECK(hipStreamSynchronize(stream));
}

void RpaEngine::prepare_gf(const MeanField &mf, int nbands_g) {
    // band ranges with non-zero weight for each sign
    for (int sg = 0; sg < 2; ++sg) {
        int lo = nbands, hi = 0;
        for (int s = 0; s < nspin; ++s)
            for (int k = 0; k < nk; ++k)
                for (int b = 0; b < nbands; ++b) {
                    const double wgv = mf.wg[((size_t)s * nk + k) * nbands + b];
                    double w = (sg == 0) ? 1.0 / nk - wgv / 2 * nspin : wgv * 0.5 * nspin;
                    if (w < 0) w = 0;
                    if (nbands_g >= 0 && b >= nbands_g) w = 0;
                    if (w > 0) { lo = std::min(lo, b); hi = std::max(hi, b + 1); }
                }
        if (hi <= lo) { lo = 0; hi = 0; }
        band_lo[sg] = lo;
        band_hi[sg] = hi;
        GemmBatch &b = gb[sg];
        b = GemmBatch();
        b.trans_a = true;
        if (hi > lo)
            for (int k = 0; k < nk; ++k)
                b.add(make_prob(2 * NAO, NAO, 2 * (hi - lo), (long long)k * 2 * nbands * 2 * NAO + 2LL * lo * 2 * NAO,
                                2 * NAO, (long long)k * 2 * nbands * NAO + 2LL * lo * NAO, NAO,
                                (long long)k * 2 * NAO * NAO, NAO));
        b.finalize(T64x64, stream);
    }
    gf_nbands = nbands_g;
}

void RpaEngine::build_gf(const MeanField &mf, double tau, int s) {
    auto &prof = profiler();
    const size_t zz_s = (size_t)nk * 2 * nbands * 2 * NAO * s;
    const size_t w_s = (size_t)nk * nbands * NAO * s;
    const size_t e_s = (size_t)nk * nbands * s;
    const long long gpm_bytes = (long long)nR * NAO * 2 * NAO * ts;
    prof.begin("gf_build", stream);
    for (int sg = 0; sg < 2; ++sg) {
        const double t = sg == 0 ? tau : -tau;
        if (band_hi[sg] > band_lo[sg]) {
            launch_scale_bands(d_wre.as<double>() + w_s, d_wim.as<double>() + w_s, d_eps.as<double>() + e_s,
                               d_wg.as<double>() + e_s, d_Zs.as<double>(), nk, nbands, NAO, band_lo[sg], band_hi[sg],
                               t, mf.efermi, nspin, gf_nbands, stream);
            gemm_run(Prec::F64, gb[sg], d_ZZ.as<double>() + zz_s, d_Zs.get(), d_Mst.get(), 1.0, 0.0, stream);
        } else {
            HIP_CHECK(hipMemsetAsync(d_Mst.get(), 0, d_Mst.bytes(), stream));
        }
        const double sgn = t < 0 ? -1.0 : 1.0;
        if (cur == Prec::F64)
            launch_k2r<double>(d_Mst.as<double>(), d_cs.as<double>(), d_sn.as<double>(),
                               d_gcol.as<int>() + (size_t)sg * NAO, d_Gpm.as<double>(), nk, nR, NAO, sgn, stream);
        else
            launch_k2r<float>(d_Mst.as<double>(), d_cs.as<double>(), d_sn.as<double>(),
                              d_gcol.as<int>() + (size_t)sg * NAO, d_Gpm.as<float>(), nk, nR, NAO, sgn, stream);
    }
    prof.end("gf_build", stream, gb[0].flops + gb[1].flops, (double)gpm_bytes);
}

void RpaEngine::build_chi0(const Dataset &ds, const TFGrid &g, int nbands_g) {
    const auto &mf = ds.mf;
    auto &prof = profiler();
    const double t0 = now_s();
    const int ntau = g.n;
    chiw_ok = false;  // chi0(R, omega) of an earlier build is stale
    if (host_staged) d_chiw.release();  // its room goes back to W and T (reformed for Ec)
    prepare_gf(mf, nbands_g);
    d_Ghat.reserve(ghat_max * ts);
    d_W.reserve((size_t)w_off.back() * ts);
    d_T.reserve(t_max * ts);
    if (mixed) {
        if (prec != Prec::F64) LRA_THROW("mixed chi0 precision needs the FP64 plan");
        if (const char *e = std::getenv("LIBRPA_AMD_MIXED_AMAX")) mixed_amax = std::atof(e);
        // the low-weight tau points run in mixed_low, FP32 or BF16x3 (both read the FP32 copies)
        d_chi32.reserve(sizeof(float) * (size_t)E_tot);
    }
    cur = prec;
    if (const char *ov = std::getenv("LIBRPA_AMD_OVERLAP")) overlap_ta = std::atoi(ov) != 0;
    // LIBRPA_AMD_AUTOTUNE=0 keeps the default tile shapes (no timing runs on first use)
    if (const char *at = std::getenv("LIBRPA_AMD_AUTOTUNE")) autotune = std::atoi(at) != 0;
    if (overlap_ta && chunks.size() > 1) {
        if (!s2) {
            HIP_CHECK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            for (hipEvent_t *e : {&ev_ready, &ev_s2, &ev_ta[0], &ev_ta[1], &ev_c[0], &ev_c[1]})
                HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
        // a second T buffer when the device has room for it (else the chunks run in sequence)
        size_t free_b = 0, total_b = 0;
        HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
        if (d_T2.bytes() < t_max * ts && free_b > t_max * ts + (size_t)4e9) d_T2.reserve(t_max * ts);
    }
    HIP_CHECK(hipMemsetAsync(d_chi0s.get(), 0, d_chi0s.bytes(), stream));
    if (screen_mode != 0 && nspin > 1) LRA_THROW("chi0 block screening needs nspin = 1");
    if (host_staged && collect) LRA_THROW("screening samples need the device-resident chi0 store (too large here)");
    if (screen_mode == 2) {
        if (screen_model.size() != (size_t)MLP_NMODEL) LRA_THROW("learned screening without a model");
        d_model = upload(screen_model, stream);
    }
    scr_blocks = scr_skipped = scr_flops = scr_flops_skipped = 0;
    if (collect) features.assign((size_t)ntau * MLP_NFEAT * outs.size(), 0.0f);
    for (auto &c : chunks) {
        c.xb.use_skip = c.ta.use_skip = false;
        for (auto &f : c.fin) f.use_skip = false;
    }
    bool tuned_chi0[2] = {false, false};  // [plan precision, FP32 of the mixed mode]
    n_lowp_tau = 0;
    for (int it = 0; it < ntau; ++it) {
        const double tau = g.tau[it];
        // mixed precision: FP32 contractions for the tau points whose column of the cosine t->f
        // matrix is small (the transform amplifies chi0(R, tau) errors by that column's entries)
        cur = prec;
        if (mixed) {
            double amax = 0;
            for (int w = 0; w < g.n; ++w) amax = std::max(amax, std::fabs(g.cos_t2f[(size_t)w * g.n + it]));
            if (amax <= mixed_amax) cur = mixed_low;
        }
        const bool lowp = cur != prec;
        n_lowp_tau += lowp ? 1 : 0;
        // the small-K GEMMs (X: K = 2 n_b(I) n_b(J) in a K table; T_A: K = 2 n_b(J)) stay in FP32 when
        // the tau point runs in BF16x3, whose 32-deep K slices would pad them and cost more than FP32
        const Prec small = cur == Prec::BF16X3 ? Prec::F32 : cur;
        // low-precision tau points time under their own region names
        const char *lsuf = cur == Prec::BF16X3 ? "_bf16x3" : "_f32";
        auto rn = [lowp, lsuf](const char *r) { return lowp ? std::string(r) + lsuf : std::string(r); };
        auto rs = [&rn, small, cur_ = cur](const char *r) {  // regions of the small-K GEMMs
            return small == cur_ ? rn(r) : std::string(r) + "_f32";
        };
        const double bsc = lowp ? 4.0 / ts : 1.0;  // algorithmic bytes scale with the element size
        const void *Cfin_p = lowp ? d_Cfin32.get() : d_Cfin.get();
        const void *CWb_p = lowp ? d_CWb32.get() : d_CWb.get();
        if (lowp) HIP_CHECK(hipMemsetAsync(d_chi32.get(), 0, d_chi32.bytes(), stream));
        if (mixed)
            for (auto &c : chunks) {
                c.wb.use_tiles(cur, stream);
                c.xb.use_tiles(small, stream);
                c.ta.use_tiles(small, stream);
                for (auto &f : c.fin) f.use_tiles(cur, stream);
            }
        for (int s = 0; s < nspin; ++s) {
            build_gf(mf, tau, s);
            if (s == 0 && (screen_mode != 0 || collect)) screen_tau(it, tau);
            char *chi_base =
                lowp ? d_chi32.as<char>() : d_chi0s.as<char>() + (host_staged ? 0 : (size_t)it * E_tot * ts);
            // tiles are tuned per precision on the first tau point built at it
            const bool tune_now = autotune && s == 0 && !tuned_chi0[lowp];
            for (auto &c : chunks) {
                if (tune_now && verbose) {
                    fprintf(stderr, "[librpa-mi355x] chi0: tuning the W build of chunk %d / %zu\n",
                            (int)(&c - chunks.data()) + 1, chunks.size());
                    fflush(stderr);
                }
                prof.begin(rn("chi0_gather"), stream);
                if (cur == Prec::F64)
                    launch_gather_cols<double>(d_Gpm.as<double>(), 2LL * NAO, d_Ghat.as<double>(), c.d_jobs.as<GatherJob>(),
                                               (int)c.jobs.size(), c.d_colsrc.as<long long>(), NAO, c.max_gcol, stream);
                else
                    launch_gather_cols<float>(d_Gpm.as<float>(), 2LL * NAO, d_Ghat.as<float>(), c.d_jobs.as<GatherJob>(),
                                              (int)c.jobs.size(), c.d_colsrc.as<long long>(), NAO, c.max_gcol, stream);
                prof.end(rn("chi0_gather"), stream, 0, 2.0 * c.ghat_elems * (lowp ? 4 : ts));
                if (tune_now && c.wb.cfg_by_prec[(int)cur] < 0) c.wb.tune(cur, d_Ghat.get(), CWb_p, d_W.get(), 1.0, 0.0, stream);
                prof.begin(rn("chi0_W"), stream);
                gemm_run(cur, c.wb, d_Ghat.get(), CWb_p, d_W.get(), 1.0, 0.0, stream);
                prof.end(rn("chi0_W"), stream, c.wb.flops);
            }
            // T_A of chunk c + 1 (second stream, second T buffer) overlaps the contraction of chunk c
            const bool overlap = overlap_ta && !tune_now && chunks.size() > 1 && d_T2.bytes() >= t_max * ts;
            if (overlap) {
                HIP_CHECK(hipEventRecord(ev_ready, stream));  // G and W of this tau written
                HIP_CHECK(hipStreamWaitEvent(s2, ev_ready, 0));
                for (size_t ci = 0; ci < chunks.size(); ++ci) {
                    auto &c = chunks[ci];
                    void *Tb = (ci % 2) ? d_T2.get() : d_T.get();
                    // the contraction two chunks back has finished reading this buffer
                    if (ci >= 2) HIP_CHECK(hipStreamWaitEvent(s2, ev_c[ci % 2], 0));
                    prof.begin(rs("chi0_TA"), s2);
                    gemm_run(small, c.ta, d_Gpm.get(), d_W.get(), Tb, 1.0, 0.0, s2);
                    prof.end(rs("chi0_TA"), s2, c.ta.flops, c.ta_bytes * bsc);
                    HIP_CHECK(hipEventRecord(ev_ta[ci % 2], s2));
                    prof.begin(rs("chi0_X"), stream);
                    gemm_run(small, c.xb, d_W.get(), d_W.get(), chi_base, 1.0, s == 0 ? 0.0 : 1.0, stream);
                    prof.end(rs("chi0_X"), stream, c.xb.flops);
                    HIP_CHECK(hipStreamWaitEvent(stream, ev_ta[ci % 2], 0));
                    for (int k = 0; k < 4; ++k) {
                        if (c.fin[k].total_tiles == 0) continue;
                        prof.begin(rn("chi0_contract"), stream);
                        gemm_run(cur, c.fin[k], Cfin_p, Tb, chi_base, 1.0, 1.0, stream);
                        prof.end(rn("chi0_contract"), stream, c.fin[k].flops, c.fin_bytes[k] * bsc);
                    }
                    HIP_CHECK(hipEventRecord(ev_c[ci % 2], stream));
                }
                // the next G and W builds overwrite what T_A reads
                HIP_CHECK(hipEventRecord(ev_s2, s2));
                HIP_CHECK(hipStreamWaitEvent(stream, ev_s2, 0));
                continue;
            }
            for (auto &c : chunks) {
                if (tune_now && verbose) {
                    fprintf(stderr, "[librpa-mi355x] chi0: tuning T_A and the contraction of chunk %d / %zu\n",
                            (int)(&c - chunks.data()) + 1, chunks.size());
                    fflush(stderr);
                }
                if (tune_now && c.ta.cfg_by_prec[(int)cur] < 0) {
                    // TA overwrites T; the contraction trials accumulate into chi0s before the X
                    // term below resets it (beta = 0) for this tau
                    c.ta.tune(small, d_Gpm.get(), d_W.get(), d_T.get(), 1.0, 0.0, stream);
                    gemm_run(small, c.ta, d_Gpm.get(), d_W.get(), d_T.get(), 1.0, 0.0, stream);
                    for (auto &f : c.fin) f.tune(cur, Cfin_p, d_T.get(), chi_base, 1.0, 1.0, stream);
                }
                prof.begin(rs("chi0_X"), stream);
                gemm_run(small, c.xb, d_W.get(), d_W.get(), chi_base, 1.0, s == 0 ? 0.0 : 1.0, stream);
                prof.end(rs("chi0_X"), stream, c.xb.flops);
                prof.begin(rs("chi0_TA"), stream);
                gemm_run(small, c.ta, d_Gpm.get(), d_W.get(), d_T.get(), 1.0, 0.0, stream);
                prof.end(rs("chi0_TA"), stream, c.ta.flops, c.ta_bytes * bsc);
                // one timing region per launch (per block-shape class)
                for (int k = 0; k < 4; ++k) {
                    if (c.fin[k].total_tiles == 0) continue;
                    prof.begin(rn("chi0_contract"), stream);
                    gemm_run(cur, c.fin[k], Cfin_p, d_T.get(), chi_base, 1.0, 1.0, stream);
                    prof.end(rn("chi0_contract"), stream, c.fin[k].flops, c.fin_bytes[k] * bsc);
                }
            }
        }
        tuned_chi0[lowp] = tuned_chi0[lowp] || autotune;
        if (verbose) {
            HIP_CHECK(hipStreamSynchronize(stream));
            fprintf(stderr, "[librpa-mi355x] chi0: tau %d / %d (%.1f s)\n", it + 1, ntau, now_s() - t0);
            fflush(stderr);
        }
        // the FP32 result of a low-precision tau point goes into the FP64 chi0(R, tau) store
        if (lowp)
            launch_convert<float, double>(d_chi32.as<float>(), d_chi0s.as<double>() + (host_staged ? 0 : (size_t)it * E_tot),
                                          E_tot, stream);
        // a host-staged build keeps its slice in pageable host memory (stream order makes the next
        // tau's kernels wait for the copy)
        if (host_staged)
            HIP_CHECK(hipMemcpyAsync(h_chi0s[it].data(), d_chi0s.get(), ts * (size_t)E_tot, hipMemcpyDeviceToHost,
                                     stream));
    }
    cur = prec;  // later G builds (the self-energy's) run at the plan's precision
    HIP_CHECK(hipStreamSynchronize(stream));
    ntau_done = ntau;
    t_chi0 = now_s() - t0;
}

namespace {
// IBZ k list as in PeriodicBoundaryData::set_ibz_mapping (src/core/pbc.cpp:110-153)
void ibz_list(const Dataset &ds, int nk, std::vector<int> &ibz_k, std::vector<double> &w) {
    std::vector<int> map = ds.ibz_map;
    if ((int)map.size() != nk) { map.resize(nk); for (int k = 0; k < nk; ++k) map[k] = k; }
    std::map<int, int> pos;
    ibz_k.clear();
    w.clear();
    for (int k = 0; k < nk; ++k) {
        const int kb = map[k];
        auto it = pos.find(kb);
        if (it == pos.end()) { pos[kb] = (int)ibz_k.size(); ibz_k.push_back(kb); w.push_back(1.0 / nk); }
        else w[it->second] += 1.0 / nk;
    }
}
}  // namespace

RpaResult RpaEngine::ec(const Dataset &ds, const TFGrid &g) {
    auto &prof = profiler();
    const double t0 = now_s();
    RpaResult res;
    const bool embed = nk > 1;
    const int n = NABF, n2 = embed ? 2 * NABF : NABF;
    const size_t nn = (size_t)n2 * n2;
    std::vector<int> ibz;
    std::vector<double> wq;
    ibz_list(ds, nk, ibz, wq);
    res.qcontrib.assign(ibz.size(), cplx(0, 0));
    const bool f64 = ec_prec == Prec::F64;
    const size_t es = f64 ? 8 : 4;
    const int nw = g.n;
    const int nq = (int)ibz.size();
    // chi0(R, omega) once for all q when it fits in a third of the free memory
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    if (!chiw_ok) {
        prof.begin("chi0_tau_to_omega", stream);
        if (host_staged) {
            // the chi0 work buffers make room for chi0(R, omega); build_chi0 reserves them again
            release_work();
            HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
            if (!form_chiw(g, 0.8 * (double)free_b, stream))
                LRA_THROW("chi0(R, omega) does not fit on the device beside the Ec buffers");
        } else {
            form_chiw(g, 0.33 * (double)free_b, stream);
        }
        prof.end("chi0_tau_to_omega", stream, 0, (double)E_tot * g.n * (ts * ceil_div(g.n, 8) + 8.0));
    }
    // all omega of a q are factorised together, as far as memory allows; small matrices (k-point
    // workloads) also take several q per pass, so that the GEMM, Cholesky and log-det launches hold
    // enough matrices to fill the device
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    // the Ec buffers kept from an earlier call count as free here (they are reused or regrown)
    double held = 0;
    for (const DevBuf *b : {&ec_V, &ec_L32, &ec_D, &ec_X, &ec_M, &ec_M64, &ec_S, &ec_Vld, &ec_acc, &ec_cw.linv,
                            &ec_cw.panel})
        held += (double)b->bytes();
    const double panel_w = 16.0 * n2 * (n2 >= 4096 ? 2048 : 512);  // Cholesky panels per matrix
    const double per_w = (double)nn * (3 * es + (f64 ? 0 : 8)) + panel_w;
    const double per_q = (double)nn * 12 + panel_w;
    const double budget = std::min(96e9, 0.6 * ((double)free_b + held)) - (double)nn * 12;
    const int WCH = std::max(1, std::min(nw, (int)(budget / per_w)));
    int QG = 1;
    if (WCH == nw && nq > 1) {
        QG = std::max(1, std::min(nq, (int)(budget / (nw * per_w + per_q))));
        if (const char *qg = std::getenv("LIBRPA_AMD_EC_QGROUP")) QG = std::max(1, std::min(QG, std::atoi(qg)));
    }
    const int UCH = QG > 1 ? QG * nw : WCH;  // (q, omega) matrices per pass
    DevBuf &dV = ec_V, &dL32 = ec_L32, &dD = ec_D, &dX = ec_X, &dM = ec_M, &dM64 = ec_M64, &dacc = ec_acc,
           &dVld = ec_Vld, &dS = ec_S;
    dV.reserve(nn * 8 * QG);
    dD.reserve(nn * es * UCH);
    dX.reserve(nn * es * UCH);
    dM.reserve(nn * es * UCH);
    dVld.reserve(sizeof(double) * QG);
    // pivot signs S of V = L S L^T: a Coulomb matrix that is not positive definite keeps
    // det(1 - V chi0) = det(S) det(S - L^T chi0 L) and Tr(V chi0) = n - sum_i S_i M_ii with
    // M = S - L^T chi0 L, so Ec is exact for any V (as the LU of epsilon.cpp:600-693 is)
    dS.reserve((size_t)n2 * 8 * QG);
    if (!f64) { dL32.reserve(nn * 4 * QG); dM64.reserve(nn * 8 * UCH); }
    const size_t acc_bytes = sizeof(double) * 2 * nq * nw + 64;
    dacc.reserve(acc_bytes);
    HIP_CHECK(hipMemsetAsync(dacc.get(), 0, acc_bytes, stream));
    int *dflag = reinterpret_cast<int *>(dacc.as<char>() + sizeof(double) * 2 * nq * nw);
    CholWork &cw = ec_cw;
    if (v_ibz != ibz) {
        // stage the input's V(q) blocks once per plan (the engine is rebuilt when inputs change)
        std::vector<VBlock> desc;
        std::vector<cplx> store;
        vdesc_q.assign(1, 0);
        for (int iq = 0; iq < nq; ++iq) {
            for (const auto &kv : ds.vq) {
                const int ik = std::get<0>(kv.first), I = std::get<1>(kv.first), J = std::get<2>(kv.first);
                if (ik != ibz[iq] || I > J) continue;
                if (kv.second.size() != (size_t)na[I] * na[J]) LRA_THROW("Coulomb block of the wrong size");
                desc.push_back({(long long)store.size(), (int)abf[I], (int)abf[J], na[I], na[J], I == J ? 1 : 0});
                store.insert(store.end(), kv.second.begin(), kv.second.end());
            }
            vdesc_q.push_back((int)desc.size());
        }
        d_vblk = upload(store, stream);
        d_vdesc = upload(desc, stream);
        HIP_CHECK(hipStreamSynchronize(stream));
        v_ibz = ibz;
    }
    // (q per pass, omega per q, order, precision) -> (X = D L, M = S - L^T X) over the pass's
    // matrices u = qi * cnt + w, whose L is that of q qi; kept with their tuned tiles
    auto get_batches = [&](int qg, int cnt) -> std::pair<GemmBatch, GemmBatch> & {
        const auto key = std::make_tuple(qg, cnt, n2, (int)ec_prec);
        auto it = ec_batches.find(key);
        if (it != ec_batches.end()) return it->second;
        auto &pr = ec_batches[key];
        GemmBatch &bx = pr.first, &bm = pr.second;
        bm.trans_a = true;
        bm.lower_only = true;
        for (int qi = 0; qi < qg; ++qi)
            for (int w = 0; w < cnt; ++w) {
                const long long u = (long long)qi * cnt + w;
                GemmProb px = make_prob(n2, n2, n2, u * nn, n2, (long long)qi * nn, n2, u * nn, n2);
                px.kmode = 1;  // L is lower triangular: column tile n0 needs k >= n0 only
                bx.add(px);
                GemmProb pm = make_prob(n2, n2, n2, (long long)qi * nn, n2, u * nn, n2, u * nn, n2);
                pm.kmode = 2;  // A = L^T: row tile m0 needs k >= m0 only
                bm.add(pm);
            }
        const int cfg = n2 >= 1024 ? big_tile(ec_prec) : T64x64;
        bx.finalize(cfg, stream);
        bm.finalize(cfg, stream);
        // executed flops of the triangular products
        bx.flops = 2.0 * qg * cnt * (double)n2 * n2 * n2 / 2.0;
        bm.flops = 2.0 * qg * cnt * (double)n2 * n2 * n2 / 6.0;
        return pr;
    };
    auto set_m_identity = [&](int qg, int cnt) {
        for (int qi = 0; qi < qg; ++qi)
            for (int w = 0; w < cnt; ++w) {
                const size_t u = (size_t)qi * cnt + w;
                const double *sg = dS.as<double>() + (size_t)qi * n2;
                if (f64) launch_set_identity<double>(dM.as<double>() + u * nn, n2, n2, stream, sg);
                else launch_set_identity<float>(dM.as<float>() + u * nn, n2, n2, stream, sg);
            }
    };
    for (int iq0 = 0; iq0 < nq; iq0 += QG) {
        const int qg = std::min(QG, nq - iq0);
        if (verbose) {
            HIP_CHECK(hipStreamSynchronize(stream));
            fprintf(stderr, "[librpa-mi355x] Ec: q %d-%d / %d (%.1f s)\n", iq0 + 1, iq0 + qg, nq, now_s() - t0);
            fflush(stderr);
        }
        // dense V(q) from the staged upper atom-pair blocks (Hermitian fill, utils_matrix_m_mpi.h:430-441)
        prof.begin("ec_v_fill", stream);
        HIP_CHECK(hipMemsetAsync(dV.get(), 0, (size_t)qg * nn * 8, stream));
        for (int qi = 0; qi < qg; ++qi) {
            const int iq = iq0 + qi;
            launch_v_scatter(d_vblk.as<double>(), d_vdesc.as<VBlock>() + vdesc_q[iq], vdesc_q[iq + 1] - vdesc_q[iq],
                             dV.as<double>() + (size_t)qi * nn, n, n2, embed ? 1 : 0, 0, stream);
        }
        prof.end("ec_v_fill", stream, 0, (double)qg * nn * 8);
        prof.begin("ec_chol_v", stream);
        cholesky_logdet_batched(dV.as<double>(), (long long)nn, qg, n2, n2, dVld.as<double>(), 1, dflag + 1, true, cw,
                                stream, -1, dS.as<double>(), n2);
        prof.end("ec_chol_v", stream, (double)qg * n2 * n2 * n2 / 3.0);
        const void *Lp = dV.get();
        if (!f64) {
            launch_convert<double, float>(dV.as<double>(), dL32.as<float>(), (long long)qg * nn, stream);
            Lp = dL32.get();
        }
        const int cstep = qg > 1 ? nw : WCH;
        for (int w0 = 0; w0 < nw; w0 += cstep) {
            const int cnt = std::min(cstep, nw - w0);
            const int nu = qg * cnt;
            // tau -> omega (formed once above, or per chunk) and R -> q of chi0s into dense matrices
            prof.begin("chi0_transform", stream);
            std::vector<const double *> kfs(qg);
            for (int qi = 0; qi < qg; ++qi) kfs[qi] = &ds.kfrac[(size_t)ibz[iq0 + qi] * 3];
            assemble_dense_q(g, kfs, w0, cnt, dD.get(), f64, (long long)nn, embed, stream);
            const double tbytes = chiw_ok ? 8.0 * E_tot * cnt * ceil_div(qg, kChi0RqMaxQ) + (double)es * cnt * nn * qg
                                          : ((double)ts * E_tot * ntau_done * ceil_div(cnt, kChi0AssembleMaxW) +
                                             (double)es * cnt * nn) * qg;
            prof.end("chi0_transform", stream, 0, tbytes);
            auto &bb = get_batches(qg, cnt);
            if (autotune && !bb.first.tuned) {
                bb.first.tune(ec_prec, dD.get(), Lp, dX.get(), 1.0, 0.0, stream);
                gemm_run(ec_prec, bb.first, dD.get(), Lp, dX.get(), 1.0, 0.0, stream);
                bb.second.tune(ec_prec, Lp, dX.get(), dM.get(), -1.0, 1.0, stream, [&]() { set_m_identity(qg, cnt); });
            }
            prof.begin("ec_gemm_x", stream);
            gemm_run(ec_prec, bb.first, dD.get(), Lp, dX.get(), 1.0, 0.0, stream);
            prof.end("ec_gemm_x", stream, bb.first.flops);
            set_m_identity(qg, cnt);
            prof.begin("ec_gemm_m", stream);
            gemm_run(ec_prec, bb.second, Lp, dX.get(), dM.get(), -1.0, 1.0, stream);
            prof.end("ec_gemm_m", stream, bb.second.flops);
            double *M64 = dM.as<double>();
            if (!f64) {
                for (int u = 0; u < nu; ++u)
                    launch_copy_lower_f2d(dM.as<float>() + (size_t)u * nn, dM64.as<double>() + (size_t)u * nn, n2, n2,
                                          stream);
                M64 = dM64.as<double>();
            }
            // accumulators of (q, omega) at 2 (q nw + omega): the pass's matrices are consecutive
            // there (one q per pass, or all omega of each q)
            double *acc0 = dacc.as<double>() + 2 * ((size_t)iq0 * nw + w0);
            prof.begin("ec_logdet", stream);
            for (int u = 0; u < nu; ++u)
                launch_trace_acc(M64 + (size_t)u * nn, n2, n2, acc0 + 2 * u, stream,
                                 dS.as<double>() + (size_t)(u / cnt) * n2);
            cholesky_logdet_batched(M64, (long long)nn, nu, n2, n2, acc0 + 1, 2, dflag, false, cw, stream);
            prof.end("ec_logdet", stream, (double)nu * n2 * n2 * n2 / 3.0);
        }
    }
    std::vector<double> hacc(2 * (size_t)nq * nw);
    int hflag[2] = {0, 0};
    HIP_CHECK(hipMemcpyAsync(hacc.data(), dacc.get(), hacc.size() * 8, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(hflag, dflag, sizeof(hflag), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (hflag[1]) LRA_THROW("Coulomb matrix is singular or not finite (zero pivot in its factorisation)");
    if (hflag[0]) LRA_THROW("1 - V chi0 is singular or not finite at some (q, omega)");
    const double half = embed ? 0.5 : 1.0;
    for (int iq = 0; iq < nq; ++iq)
        for (int iw = 0; iw < nw; ++iw) {
            const double tr = hacc[2 * ((size_t)iq * nw + iw)];
            const double ld = hacc[2 * ((size_t)iq * nw + iw) + 1];
            const double trpi = (n2 - tr) * half;
            const double val = trpi + ld * half;
            res.qcontrib[iq] += cplx(val * g.omega_w[iw] * wq[iq] / (2 * M_PI), 0.0);
        }
    for (auto &c : res.qcontrib) res.ec_re += c.real();
    // a host-staged system needs the device memory back for the next chi0 build
    if (host_staged) release_ec();
    t_ec = now_s() - t0;
    return res;
}

bool RpaEngine::form_chiw(const TFGrid &g, double max_bytes, hipStream_t s) {
    chiw_ok = false;
    const double need = 8.0 * (double)E_tot * g.n;
    // a store kept from an earlier evaluation is reused (its bytes are not in the free memory)
    if (ntau_done != g.n || g.n > kTauOmegaMax || need > max_bytes + (double)d_chiw.bytes()) return false;
    d_chiw.reserve((size_t)need);
    std::vector<double> coef((size_t)g.n * g.n);
    for (int w = 0; w < g.n; ++w)
        for (int t = 0; t < g.n; ++t) coef[(size_t)w * g.n + t] = g.cos_t2f[(size_t)w * g.n + t] * 2.0 / nspin;
    DevBuf dc = upload(coef, s);
    if (host_staged) {
        // slices back through the one-slice device buffer, accumulated tau by tau
        for (int t = 0; t < ntau_done; ++t) {
            HIP_CHECK(hipMemcpyAsync(d_chi0s.get(), h_chi0s[t].data(), ts * (size_t)E_tot, hipMemcpyHostToDevice, s));
            if (prec == Prec::F64)
                launch_tau_accumulate(d_chi0s.as<double>(), E_tot, t, ntau_done, dc.as<double>(), g.n,
                                      d_chiw.as<double>(), s);
            else
                launch_tau_accumulate(d_chi0s.as<float>(), E_tot, t, ntau_done, dc.as<double>(), g.n,
                                      d_chiw.as<double>(), s);
        }
        HIP_CHECK(hipStreamSynchronize(s));
        chiw_ok = true;
        return true;
    }
    if (prec == Prec::F64)
        launch_tau_to_omega(d_chi0s.as<double>(), E_tot, ntau_done, dc.as<double>(), g.n, d_chiw.as<double>(), s);
    else
        launch_tau_to_omega(d_chi0s.as<float>(), E_tot, ntau_done, dc.as<double>(), g.n, d_chiw.as<double>(), s);
    HIP_CHECK(hipStreamSynchronize(s));  // dc is freed on return
    chiw_ok = true;
    return true;
}

void RpaEngine::assemble_dense_q(const TFGrid &g, const std::vector<const double *> &kfs, int w0, int cnt,
                                 void *dense, bool f64, long long nn, bool embed, hipStream_t s) {
    const size_t es = f64 ? 8 : 4;
    if (!chiw_ok) {
        for (size_t qi = 0; qi < kfs.size(); ++qi)
            assemble_dense(g, kfs[qi], w0, cnt, static_cast<char *>(dense) + qi * cnt * nn * es, f64, nn, embed, s);
        return;
    }
    const int n = NABF;
    for (size_t q0 = 0; q0 < kfs.size(); q0 += kChi0RqMaxQ) {
        const int nq = (int)std::min<size_t>(kChi0RqMaxQ, kfs.size() - q0);
        std::vector<double> ph((size_t)nq * nR * 2);
        for (int qi = 0; qi < nq; ++qi) {
            const double *kf = kfs[q0 + qi];
            for (int r = 0; r < nR; ++r) {
                const double ang = 2 * M_PI * (kf[0] * cells[r][0] + kf[1] * cells[r][1] + kf[2] * cells[r][2]);
                ph[((size_t)qi * nR + r) * 2] = std::cos(ang);
                ph[((size_t)qi * nR + r) * 2 + 1] = std::sin(ang);
            }
        }
        DevBuf dph = upload(ph, s);
        const double *cw = d_chiw.as<double>() + (size_t)w0 * E_tot;
        char *Dd = static_cast<char *>(dense) + q0 * cnt * nn * es;
        if (f64)
            launch_chi0_rq_multi<double>(cw, E_tot, d_pairs.as<PairDesc>(), (int)pairs.size(), n_pair_elems,
                                         d_pair_eoff.as<long long>(), nR, dph.as<double>(), nq, cnt, (double *)Dd, nn,
                                         (long long)cnt * nn, n, embed, s);
        else
            launch_chi0_rq_multi<float>(cw, E_tot, d_pairs.as<PairDesc>(), (int)pairs.size(), n_pair_elems,
                                        d_pair_eoff.as<long long>(), nR, dph.as<double>(), nq, cnt, (float *)Dd, nn,
                                        (long long)cnt * nn, n, embed, s);
        HIP_CHECK(hipStreamSynchronize(s));  // dph is freed on return
    }
}

void RpaEngine::assemble_dense(const TFGrid &g, const double *kf, int w0, int cnt, void *dense, bool f64, long long nn,
                               bool embed, hipStream_t s) {
    const int n = NABF;
    std::vector<double> ph((size_t)nR * 2);
    for (int r = 0; r < nR; ++r) {
        const double ang = 2 * M_PI * (kf[0] * cells[r][0] + kf[1] * cells[r][1] + kf[2] * cells[r][2]);
        ph[2 * r] = std::cos(ang);
        ph[2 * r + 1] = std::sin(ang);
    }
    if (chiw_ok) {
        DevBuf dph = upload(ph, s);
        const double *cw = d_chiw.as<double>() + (size_t)w0 * E_tot;
        if (f64)
            launch_chi0_rq<double>(cw, E_tot, d_pairs.as<PairDesc>(), (int)pairs.size(), n_pair_elems,
                                   d_pair_eoff.as<long long>(), nR, dph.as<double>(), cnt, (double *)dense, nn, n, embed, s);
        else
            launch_chi0_rq<float>(cw, E_tot, d_pairs.as<PairDesc>(), (int)pairs.size(), n_pair_elems,
                                  d_pair_eoff.as<long long>(), nR, dph.as<double>(), cnt, (float *)dense, nn, n, embed, s);
        HIP_CHECK(hipStreamSynchronize(s));  // dph is freed on return
        return;
    }
    const size_t es = f64 ? 8 : 4;
    for (int a0 = 0; a0 < cnt; a0 += kChi0AssembleMaxW) {
        const int cw_n = std::min(kChi0AssembleMaxW, cnt - a0);
        std::vector<double> gp((size_t)cw_n * nR * ntau_done * 2, 0.0);
        for (int w = 0; w < cw_n; ++w)
            for (int r = 0; r < nR; ++r)
                for (int t = 0; t < ntau_done; ++t) {
                    const double gam = g.cos_t2f[(size_t)(w0 + a0 + w) * g.n + t] * 2.0 / nspin;
                    gp[(((size_t)w * nR + r) * ntau_done + t) * 2] = gam * ph[2 * r];
                    gp[(((size_t)w * nR + r) * ntau_done + t) * 2 + 1] = gam * ph[2 * r + 1];
                }
        DevBuf dgp = upload(gp, s);
        char *Dd = static_cast<char *>(dense) + (size_t)a0 * nn * es;
        if (prec == Prec::F64 && f64)
            launch_chi0_assemble<double, double>(d_chi0s.as<double>(), E_tot, ntau_done, d_pairs.as<PairDesc>(),
                                                 (int)pairs.size(), n_pair_elems, d_pair_eoff.as<long long>(), nR,
                                                 dgp.as<double>(), cw_n, (double *)Dd, nn, n, embed, s);
        else if (prec == Prec::F64)
            launch_chi0_assemble<double, float>(d_chi0s.as<double>(), E_tot, ntau_done, d_pairs.as<PairDesc>(),
                                                (int)pairs.size(), n_pair_elems, d_pair_eoff.as<long long>(), nR,
                                                dgp.as<double>(), cw_n, (float *)Dd, nn, n, embed, s);
        else if (f64)
            launch_chi0_assemble<float, double>(d_chi0s.as<float>(), E_tot, ntau_done, d_pairs.as<PairDesc>(),
                                                (int)pairs.size(), n_pair_elems, d_pair_eoff.as<long long>(), nR,
                                                dgp.as<double>(), cw_n, (double *)Dd, nn, n, embed, s);
        else
            launch_chi0_assemble<float, float>(d_chi0s.as<float>(), E_tot, ntau_done, d_pairs.as<PairDesc>(),
                                               (int)pairs.size(), n_pair_elems, d_pair_eoff.as<long long>(), nR,
                                               dgp.as<double>(), cw_n, (float *)Dd, nn, n, embed, s);
        HIP_CHECK(hipStreamSynchronize(s));  // dgp is freed on return
    }
}

void RpaEngine::release_ec() {
    for (DevBuf *b : {&ec_V, &ec_L32, &ec_D, &ec_X, &ec_M, &ec_M64, &ec_S, &ec_Vld, &ec_acc, &ec_cw.linv, &ec_cw.panel})
        b->release();
}

void RpaEngine::release_work() {
    d_Ghat.release();
    d_W.release();
    d_T.release();
    d_T2.release();
}

void RpaEngine::free_host_store() {
    h_chi0s.clear();
    h_chi0s.shrink_to_fit();
}

RpaEngine::~RpaEngine() {
    free_host_store();
    if (s2) {
        (void)hipStreamSynchronize(s2);
        for (hipEvent_t e : {ev_ready, ev_s2, ev_ta[0], ev_ta[1], ev_c[0], ev_c[1]})
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(s2);
    }
}

void RpaEngine::chi0s_block(int it, int I, int J, int r, double *out) {
    for (const auto &o : outs)
        if (o.I == I && o.J == J && o.r == r) {
            const size_t n = (size_t)na[I] * na[J];
            const size_t off = (size_t)it * E_tot + o.eoff;
            if (host_staged) {
                if (it < 0 || it >= (int)h_chi0s.size()) LRA_THROW("tau index out of range");
                if (prec == Prec::F64) {
                    std::memcpy(out, reinterpret_cast<const double *>(h_chi0s[it].data()) + o.eoff, n * 8);
                } else {
                    const float *src = reinterpret_cast<const float *>(h_chi0s[it].data()) + o.eoff;
                    for (size_t i = 0; i < n; ++i) out[i] = src[i];
                }
                return;
            }
            if (prec == Prec::F64) {
                HIP_CHECK(hipMemcpy(out, d_chi0s.as<double>() + off, n * 8, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> tmp(n);
                HIP_CHECK(hipMemcpy(tmp.data(), d_chi0s.as<float>() + off, n * 4, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < n; ++i) out[i] = tmp[i];
            }
            return;
        }
    LRA_THROW("chi0s block not found");
}

void RpaEngine::chi0q_block(const Dataset &ds, const TFGrid &g, int iw, int iq, int I, int J, double *re,
                            double *im) {
    if (host_staged) LRA_THROW("chi0(q, omega) blocks are not kept with the host-staged chi0 store");
    const bool embed = nk > 1;
    const int n = NABF, n2 = embed ? 2 * NABF : NABF;
    DevBuf dD((size_t)n2 * n2 * 8);
    std::vector<double> gp((size_t)nR * ntau_done * 2);
    const double *kf = &ds.kfrac[(size_t)iq * 3];
    for (int r = 0; r < nR; ++r) {
        const double ang = 2 * M_PI * (kf[0] * cells[r][0] + kf[1] * cells[r][1] + kf[2] * cells[r][2]);
        for (int t = 0; t < ntau_done; ++t) {
            const double gam = g.cos_t2f[(size_t)iw * g.n + t] * 2.0 / nspin;
            gp[((size_t)r * ntau_done + t) * 2] = gam * std::cos(ang);
            gp[((size_t)r * ntau_done + t) * 2 + 1] = gam * std::sin(ang);
        }
    }
    DevBuf dgp = upload(gp, stream);
    if (prec == Prec::F64)
        launch_chi0_assemble<double, double>(d_chi0s.as<double>(), E_tot, ntau_done, d_pairs.as<PairDesc>(),
                                             (int)pairs.size(), n_pair_elems, d_pair_eoff.as<long long>(), nR,
                                             dgp.as<double>(), 1, dD.as<double>(), 0, n, embed, stream);
    else
        launch_chi0_assemble<float, double>(d_chi0s.as<float>(), E_tot, ntau_done, d_pairs.as<PairDesc>(),
                                            (int)pairs.size(), n_pair_elems, d_pair_eoff.as<long long>(), nR,
                                            dgp.as<double>(), 1, dD.as<double>(), 0, n, embed, stream);
    std::vector<double> h((size_t)n2 * n2);
    HIP_CHECK(hipMemcpyAsync(h.data(), dD.get(), h.size() * 8, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    for (int mu = 0; mu < na[I]; ++mu)
        for (int nu = 0; nu < na[J]; ++nu) {
            const size_t r = abf[I] + mu, c = abf[J] + nu;
            re[(size_t)mu * na[J] + nu] = h[r * n2 + c];
            im[(size_t)mu * na[J] + nu] = embed ? h[(n + r) * n2 + c] : 0.0;
        }
}

}  // namespace lra
