// G0W0: screened interaction W(q, i omega) -> W(R, i tau) and the correlation
// self-energy Sigma(R, i tau) = C G W C contraction on top of the chi0 engine.
#pragma once

#include <cstdio>
#include <memory>
#include <vector>

#include "../../include/librpa_options.h"
#include "dataset.h"
#include "gemm.h"
#include "kernels.h"
#include "rpa.h"

namespace lra {

struct GwEngine {
    std::shared_ptr<RpaEngine> rpa;  // geometry, G(R, +-tau) builder, chi0(R, tau)
    Prec prec = Prec::F64;
    size_t ts = 8;
    int natom = 0, NAO = 0, NABF = 0, nR = 0, nk = 0, nspin = 1, ntau = 0;
    long long LD = 0;  // nR * NAO: column count of the (cell, atom, orbital) index

    // neighbour lists with the on-site block always present (zero if Cs has none)
    std::vector<std::vector<RpaEngine::Nbr>> nbrs;
    std::vector<std::vector<long long>> loff;
    std::vector<long long> lsum;
    std::vector<int> onsite;

    // cell differences R1 - R2 of two BvK cells on an extended (2p - 1)^3 grid
    int E[3] = {1, 1, 1};
    int NE = 1;
    long long c0 = 0;
    std::vector<int> ext_fold, ext_foldneg;
    long long linr(int x, int y, int z) const { return ((long long)x * E[1] + y) * E[2] + z; }

    // packed Cs (compute precision)
    DevBuf d_CD, d_CB, d_CA;
    std::vector<long long> cd_off, cb_off, ca_off;

    // U_J(S) build
    struct UChunk {
        int j0 = 0, j1 = 0;
        long long ghat_elems = 0;
        int max_gcol = 0;
        std::vector<long long> ghat_off;  // per (J - j0, S)
        std::vector<GatherJob> jobs[2];
        std::vector<long long> colsrc[2];
        DevBuf d_jobs[2], d_colsrc[2];
        GemmBatch ub;
    };
    std::vector<UChunk> uchunks;
    std::vector<long long> u_off;  // per (J, S)
    long long u_elems = 0;

    // T^D and the owner-row contractions per chunk of atoms I
    struct IChunk {
        int i0 = 0, i1 = 0;
        long long td_elems = 0;
        std::vector<long long> td_off;  // per I - i0
        GemmBatch tdb, dcon, a3b;
    };
    std::vector<IChunk> ichunks;

    // partner-partner term: per (J, R) GC = G C^J and Y += W GC, then one contraction
    std::vector<GemmBatch> a1, a2;
    GemmBatch a3a;
    std::vector<long long> y_off;
    long long y_elems = 0, gc_elems = 0;

    DevBuf d_U, d_Ghat, d_TD, d_Y, d_GC, d_Wf;
    DevBuf d_sig[2], d_extA, d_extB, d_fold, d_foldneg;
    DevBuf d_Wrt;         // W(R, tau) [tau][R][NABF][NABF], FP64
    DevBuf d_Sre, d_Sim;  // Sigma(R, i omega) [spin][omega][R][NAO][NAO], FP64
    bool w_zero = false;  // no truncated Coulomb given: W = 0

    double flops_sigma = 0, flops_w = 0, t_w = 0, t_sigma = 0;
    long long max_ghat = 0, max_tdc = 0;
    bool verbose = false;  // progress lines on stderr
    // eps_mac(i omega_w) per grid frequency: replaces the head of the dielectric matrix at q = 0
    // (replace_w_head with option_dielect_func 0); empty = no replacement
    std::vector<double> head;
    double sqrt_threshold = 0.0;  // sqrt_coulomb_threshold: eigenvalues of v, v_c below it are dropped
    bool sigma_tuned = false;  // Sigma batches' tiles timed (LIBRPA_AMD_TUNE_SIGMA=1) on the first pass

    void plan(const Dataset &ds, std::shared_ptr<RpaEngine> r);
    void build_w(const Dataset &ds, const TFGrid &g);
    void build_sigma(const Dataset &ds, const TFGrid &g);
    void w_block(int it, int I, int J, int r, double *out) const;
    void sigma_block(int spin, int iw, int I, int J, int r, double *re, double *im) const;
    // Sigma_KS(k, i omega)_nn for the given states: [state][omega]
    std::vector<std::vector<cplx>> ks_diag(const Dataset &ds, int spin, int ik, const std::vector<int> &states);
};

// the minimax / time-frequency grid of the current inputs (api.cpp)
void lra_prepare_grid(Dataset &d, const LibrpaOptions &o);

void gw_build_sigma(Dataset &d, const LibrpaOptions &o);
void gw_sigc_kgrid(Dataset &d, const LibrpaOptions &o, int n_spins, int n_kpts_this, const int *iks_this,
                   int i_state_low, int i_state_high, const double *vxc, const double *vexx, double *sigc_re,
                   double *sigc_im);

// Thiele continued-fraction analytic continuation (src/core/analycont.cpp:17-80)
struct PadeAC {
    int n = 0;
    std::vector<cplx> x, y;
    PadeAC(int n_pars, const std::vector<cplx> &xs, const std::vector<cplx> &data);
    cplx operator()(cplx z) const;
};
// Damped fixed-point QP solver (src/core/qpe_solver.cpp:10-55); returns 0 on success
int qpe_solve(const PadeAC &p, double e_mf, double e_fermi, double vxc, double sigx, double &e_qp, cplx &sigc,
              double diff_init = 1e-3, double thres = 1e-5, int n_iter_max = 200);

}  // namespace lra
