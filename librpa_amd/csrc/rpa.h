// Space-time chi0 builder and RPA correlation energy on one GPU.
#pragma once

#include <array>
#include <map>
#include <tuple>
#include <vector>

#include "dataset.h"
#include "gemm.h"
#include "kernels.h"
#include "screen.h"

namespace lra {

struct RpaResult {
    double ec_re = 0;
    std::vector<cplx> qcontrib;  // per IBZ k
};

struct RpaEngine {
    // ---- static plan (tau independent) ----
    Prec prec = Prec::F64, ec_prec = Prec::F64;
    int natom = 0, NAO = 0, NABF = 0, nR = 0, nk = 0, nspin = 1, nbands = 0;
    int period[3] = {1, 1, 1};
    std::vector<int> nb, na;
    std::vector<long long> ao, abf;
    std::vector<std::array<int, 3>> cells;

    struct Nbr { int K, r; };
    std::vector<std::vector<Nbr>> nbrs;           // folded Cs neighbours of each atom
    std::vector<std::vector<long long>> koff;     // row offset (in T / Cfin columns) per neighbour
    std::vector<std::vector<long long>> loff;     // row offset in CWb per neighbour
    std::vector<long long> ksum, lsum, cfin_off, cwb_off;
    long long n_cs_blocks = 0;

    struct Out { int I, J, r; long long eoff, toff; };
    std::vector<Out> outs;
    long long E_tot = 0;

    struct Chunk {
        int j_begin, j_end;          // atoms J in [j_begin, j_end)
        int o_begin, o_end;          // outputs
        long long ghat_elems = 0, t_elems = 0;
        double fin_bytes[4] = {0, 0, 0, 0}, ta_bytes = 0;  // algorithmic HBM bytes per launch
        std::vector<long long> ghat_off;  // per (J - j_begin, S, sigma)
        std::vector<GatherJob> jobs;
        std::vector<long long> colsrc;
        DevBuf d_jobs, d_colsrc;
        int max_gcol = 0;
        // contraction batches by (na(I) > 64, na(J) > 64), so that each shape class gets its own
        // tile (a tile covering all of na(I) reads T once)
        GemmBatch wb, xb, ta, fin[4];
        double fin_flops = 0;
        // screening maps: output of each X / contraction problem, outputs written by each T_A problem
        std::vector<int> x_owner, fin_owner[4], ta_ptr, ta_outs;
    };
    std::vector<Chunk> chunks;
    std::vector<long long> w_off;  // W(K, J, S) block of (J, S) in d_W

    // G build
    int band_lo[2] = {0, 0}, band_hi[2] = {0, 0};  // per sign (0: +tau, 1: -tau)
    int gf_nbands = -1;
    GemmBatch gb[2];

    // ---- device data ----
    size_t ts = 8;  // storage element size (8 fp64, 4 fp32)
    // mixed precision (FP64 plan): tau points whose cosine t->f column has no entry above
    // mixed_amax run their contractions in FP32 (operands converted once, result stored FP64)
    bool mixed = false;
    double mixed_amax = 1e3;
    int n_lowp_tau = 0;  // tau points of the last build that ran in FP32 (mixed mode)
    Prec mixed_low = Prec::F32;  // precision of those tau points
    Prec cur = Prec::F64;  // precision of the tau point being built
    DevBuf d_Cfin32, d_CWb32, d_chi32;
    DevBuf d_Cfin, d_CWb, d_ZZ, d_Zs, d_Mst, d_Gpm, d_Ghat, d_W, d_T, d_T2, d_chi0s;
    DevBuf d_wre, d_wim, d_eps, d_wg, d_cs, d_sn, d_gcol;
    hipStream_t stream = 0;
    // T_A of the next chunk on a second stream (default priority, non-blocking) into a second
    // T buffer, overlapping the current chunk's contraction.  Off unless LIBRPA_AMD_OVERLAP=1:
    // on the 128-atom cell it saves 1.5% of the wall time, but the two kernels then share the
    // memory system and each one's own timing no longer describes it.
    bool overlap_ta = false;
    hipStream_t s2 = nullptr;
    hipEvent_t ev_ready = nullptr, ev_s2 = nullptr, ev_ta[2] = {nullptr, nullptr}, ev_c[2] = {nullptr, nullptr};
    RpaEngine() = default;
    RpaEngine(const RpaEngine &) = delete;
    RpaEngine &operator=(const RpaEngine &) = delete;
    ~RpaEngine();

    // transform / Ec
    std::vector<PairDesc> pairs;
    std::vector<long long> pair_eoff;  // [pair][r]
    long long n_pair_elems = 0;
    DevBuf d_pairs, d_pair_eoff;
    // the input V(q) upper atom-pair blocks of the irreducible q points, staged on the device at
    // the first Ec call (block store, descriptors, first descriptor of each q)
    DevBuf d_vblk, d_vdesc;
    std::vector<int> vdesc_q, v_ibz;
    // Ec work buffers, kept between evaluations (allocating and freeing ~100 GB per call made
    // some calls 50% slower): V / L and pivot signs per q, chi0(q, omega), X, M per matrix, the
    // Cholesky workspace
    DevBuf ec_V, ec_L32, ec_D, ec_X, ec_M, ec_M64, ec_S, ec_Vld, ec_acc;
    CholWork ec_cw;

    // counters
    double flops_chi0_alg = 0;   // canonical count per spin, all tau
    double flops_chi0_exec = 0;
    double t_chi0 = 0, t_ec = 0;
    int ntau_done = 0;
    int planned_ntau = 0;
    bool autotune = true;  // pick each batch's tile shape by timing on first use
    std::map<std::tuple<int, int, int, int>, std::pair<GemmBatch, GemmBatch>> ec_batches;

    // ---- block-norm screening of chi0 outputs (nspin = 1) ----
    int screen_mode = 0;          // 0 off, 1 norm-product bound, 2 learned MLP and the bound
    double screen_thr = 1e-8;     // skip blocks whose bound (or prediction) is below this norm
    std::vector<float> screen_model;  // MLP_NMODEL values (parameters, feature and target scaling)
    bool collect = false;         // record the features of every (tau, output) of a build
    // chi0(R, tau) in host memory (one buffer per tau) with one tau slice on the device,
    // when the device store would not fit beside the resident W slabs (LIBRPA_AMD_HOST_CHI0=1
    // forces it); the Ec step then streams the slices back into chi0(R, omega)
    bool host_staged = false;
    bool verbose = false;  // progress lines on stderr (tau of the chi0 build, q of Ec)
    std::vector<std::vector<char>> h_chi0s;
    void free_host_store();
    std::vector<double> cnorm;    // per atom: sum of the Frobenius norms of its folded Cs blocks
    std::vector<float> feat_static;  // per output: d / 10, n_b(I) / 27, n_b(J) / 27, ln c_I, ln c_J
    std::vector<BlockRef> gblocks;   // per output and sign: the G_IJ(R) block inside Gpm
    DevBuf d_gblocks, d_gnorm, d_feat, d_pred, d_model, d_eoff, d_esize;
    std::vector<float> features;     // [tau][output][8] of the last build (collect)
    std::vector<double> gnorm_h;
    std::vector<unsigned char> skip_out;
    std::vector<float> pred_h, feat_h;
    double scr_blocks = 0, scr_skipped = 0, scr_flops = 0, scr_flops_skipped = 0;
    void screen_plan(const Dataset &ds);
    void screen_tau(int it, double tau);            // features and skip flags after build_gf
    void chi0_norms(std::vector<double> &out);      // ||chi0s(o, tau)||_F, [tau][output]

    void plan(const Dataset &ds, Prec p, Prec pe, int ntau);
    // G(R, +-tau) of spin s into d_Gpm (prepare_gf once per band cut first)
    void prepare_gf(const MeanField &mf, int nbands_g);
    void build_gf(const MeanField &mf, double tau, int s);
    void build_chi0(const Dataset &ds, const TFGrid &g, int nbands_g);
    void release_work();  // free the gathered-G, W and T work buffers (rebuilt by build_chi0)
    void release_ec();    // free the Ec work buffers (rebuilt by ec)
    // chi0(R, omega) = sum_tau (2 / nspin) gamma_c(omega, tau) chi0(R, tau) of the last build, FP64
    // [omega][E], formed once per Ec / W evaluation when it fits in max_bytes (returns whether it
    // was formed); the R -> q sums then read it once per q instead of chi0(R, tau) per chunk of 8
    // frequencies
    DevBuf d_chiw;
    bool chiw_ok = false;
    bool form_chiw(const TFGrid &g, double max_bytes, hipStream_t s);
    // dense chi0(q, omega) of the frequencies [w0, w0 + cnt) at fractional q into cnt matrices of nn
    // elements (FP64 or FP32), real at Gamma or the real embedding
    void assemble_dense(const TFGrid &g, const double *kf, int w0, int cnt, void *dense, bool f64, long long nn,
                        bool embed, hipStream_t s);
    // the same for several q points at once (chi0(R, omega) formed): q point qi of kfs writes its
    // cnt matrices at dense + qi * cnt * nn elements; chi0(R, omega) is read once per
    // kChi0RqMaxQ q points
    void assemble_dense_q(const TFGrid &g, const std::vector<const double *> &kfs, int w0, int cnt, void *dense,
                          bool f64, long long nn, bool embed, hipStream_t s);
    long long ghat_max = 0, t_max = 0;
    RpaResult ec(const Dataset &ds, const TFGrid &g);
    // dense chi0(q, omega) block for verification
    void chi0q_block(const Dataset &ds, const TFGrid &g, int iw, int iq, int I, int J, double *re, double *im);
    void chi0s_block(int it, int I, int J, int r, double *out);
    int cell_index(int x, int y, int z) const;
    int fold_index(int x, int y, int z) const;
};

}  // namespace lra
