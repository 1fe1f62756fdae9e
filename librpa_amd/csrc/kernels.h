// Host wrappers of the streaming / small kernels of the chi0 -> Ec path.
#pragma once

#include "common.h"

namespace lra {

// G(R, +-tau) build ----------------------------------------------------------------
// Zs[(2b + c)][j] = s_b * (c == 0 ? Re U[b][j] : Im U[b][j]) for b in [b0, b1),
// s_b = exp(min(0, -tau (eps_b - mu))) * w_b with the occupation weights of
// build_gf_Rt_libri_serial (src/core/chi0.cpp:333-355).  U, eps, wg per k.
void launch_scale_bands(const double *wre, const double *wim, const double *eps, const double *wg,
                        double *Zs, int nk, int nbands, int nbasis, int b0, int b1, double tau, double efermi,
                        int nspin, int nbands_g, hipStream_t s);

// Gpm[R][row][gcol[col]] = sgn * sum_k (cs[R][k] Mre_k[row][col] + sn[R][k] Mim_k[row][col]),
// Mstack per k laid out [Re rows; Im rows] (2 nbasis x nbasis).  T = double or float.
template <class T>
void launch_k2r(const double *Mstack, const double *cs, const double *sn, const int *gcol, T *Gpm, int nk, int nR,
                int nbasis, double sgn, hipStream_t s);

// Column gather: dst[job][row][c] = src[colsrc[job_cols + c] + row * src_ld] for row < nrows.
struct GatherJob {
    long long dst_off;
    int ncol;
    int col_off;  // offset into colsrc
};
template <class T>
void launch_gather_cols(const T *src, long long src_ld, T *dst, const GatherJob *jobs, int njobs,
                        const long long *colsrc, int nrows, int maxcol, hipStream_t s);

// tau -> omega and R -> q of chi0s into dense Hermitian matrices ---------------------
// For each pair p = (I <= J) with element count na_I * na_J and each omega w of the
// chunk: val = (2 / nspin) sum_t sum_r gamma[w][t] e^{i 2 pi q.R_r} chi0s[t][eoff[p][r] + e].
// Output: real (Gamma) n x n, or the real embedding [[Re, -Im], [Im, Re]] (2n x 2n).
struct PairDesc {
    int I, J;
    int naI, naJ;
    int rowI, colJ;    // ABF offsets
    long long e0;      // first element in the flat pair-element index
};
constexpr int kChi0AssembleMaxW = 8;  // frequencies per launch (nw <= this)
template <class TS, class TD>
void launch_chi0_assemble(const TS *chi0s, long long E_tot, int ntau, const PairDesc *pairs, int npairs,
                          long long n_pair_elems, const long long *eoff, int nR, const double *gam_ph,
                          int nw, TD *dense, long long dense_stride, int n, bool embed, hipStream_t s);

// The same in two passes, when a chi0(R, omega) store fits: chiw[w][i] = sum_t coef[w][t]
// chi0s[t][i] once for all q (coef = (2 / nspin) gamma_c, nw x ntau, row-major), then per q and
// frequency chunk the R -> q phase sum ph[r] = (cos, sin)(2 pi q.R_r) into dense matrices.
constexpr int kTauOmegaMax = 34;  // the largest minimax grid
void launch_tau_to_omega(const double *chi0s, long long E_tot, int ntau, const double *coef, int nw, double *chiw,
                         hipStream_t s);
void launch_tau_to_omega(const float *chi0s, long long E_tot, int ntau, const double *coef, int nw, double *chiw,
                         hipStream_t s);
// chiw[w][i] = sum_{t' <= t} coef[w][t'] slice_t'[i], one slice per call (t = 0 overwrites)
template <class TS>
void launch_tau_accumulate(const TS *slice, long long E_tot, int t, int ntau, const double *coef, int nw, double *chiw,
                           hipStream_t s);
template <class TD>
void launch_chi0_rq(const double *chiw, long long E_tot, const PairDesc *pairs, int npairs, long long n_pair_elems,
                    const long long *eoff, int nR, const double *ph, int nw, TD *dense, long long dense_stride, int n,
                    bool embed, hipStream_t s);
// chi0_rq for nq <= kChi0RqMaxQ q points at once (phases ph[q][r] as (cos, sin)); q point qi writes
// its nw matrices at dense + qi * q_stride
constexpr int kChi0RqMaxQ = 8;
template <class TD>
void launch_chi0_rq_multi(const double *chiw, long long E_tot, const PairDesc *pairs, int npairs,
                          long long n_pair_elems, const long long *eoff, int nR, const double *ph, int nq, int nw,
                          TD *dense, long long dense_stride, long long q_stride, int n, bool embed, hipStream_t s);

// Dense helpers --------------------------------------------------------------------
// A = I, or diag(d) when d (n device doubles) is given
template <class T> void launch_set_identity(T *A, int n, int lda, hipStream_t s, const double *d = nullptr);
template <class T> void launch_zero_upper(T *A, int n, int lda, hipStream_t s);
template <class TS, class TD> void launch_convert(const TS *src, TD *dst, long long n, hipStream_t s);
void launch_copy_lower_f2d(const float *src, double *dst, int n, int lda, hipStream_t s);

// One upper atom-pair block V_IJ(q) (I <= J) of the Coulomb matrix, complex, na_I x na_J
// row-major at element offset off of the block store; r0 / c0 are the ABF offsets of I and J.
struct VBlock {
    long long off;
    int r0, c0, na, nb, diag;
};
// Dense real V(q) (n2 = n, Gamma) or its real embedding [[Re, -Im], [Im, Re]] (n2 = 2n) from the
// upper blocks, Hermitian fill; V must be zeroed first.  Every dense entry is written by at most
// one (block, element), so the blocks run in parallel.  drop_imag embeds Re V only.
void launch_v_scatter(const double *blk, const VBlock *d, int nblk, double *V, int n, int n2, int embed,
                      int drop_imag, hipStream_t s);
// *acc += sum_i A_ii, or sum_i signs_i A_ii when signs (n device doubles) is given
void launch_trace_acc(const double *A, int n, int lda, double *acc, hipStream_t s, const double *signs = nullptr);
void launch_copy_block(const double *src, int lds, double *dst, int ldd, int rows, int cols, hipStream_t s);

// Blocked Cholesky with signed pivots (lower, row-major, FP64): A = L S L^T in place,
// S = diag(+-1); equal to Cholesky for positive-definite A.  logdet (device scalar)
// += ln |det A|; *flag set to 1 on a zero / NaN pivot.  When keep_L is false only the
// log-determinant is produced.  signs (n device doubles, optional) receives S.
struct CholWork {
    DevBuf linv, panel;
};
void cholesky_logdet(double *A, int n, int lda, double *logdet, int *flag, bool keep_L, CholWork &w,
                     hipStream_t s, double *signs = nullptr);
// The same for nbatch matrices at A + b * strideA, all factorised in the same launches;
// matrix b adds its ln|det| to logdet[b * ld_stride].  With ncols in (0, n) only the
// first ncols columns are eliminated: the lower triangle of the trailing block then
// holds its Schur complement A22 - A21 A11^{-1} A21^T and logdet gets ln|det A11|.
// Matrix b writes its pivot signs to signs[b * sstride + j] when signs is given.
void cholesky_logdet_batched(double *A, long long strideA, int nbatch, int n, int lda, double *logdet,
                             long long ld_stride, int *flag, bool keep_L, CholWork &w, hipStream_t s,
                             int ncols = -1, double *signs = nullptr, long long sstride = 0);

}  // namespace lra
