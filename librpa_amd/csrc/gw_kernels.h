// Streaming kernels of the G0W0 path (gfx950): W(q, omega) -> W(R, tau) accumulation,
// folding of the self-energy cell images, the tau -> omega fold and the Kohn-Sham
// projection of Sigma(R, i omega).
#pragma once

#include "common.h"

namespace lra {

// W(R, tau)[a][b] += sum_w coef[w][tau] (cos_R Re W_w(a, b) + sin_R Im W_w(a, b)) for all R, tau,
// with W_w = -(trailing block of the partially factorised augmented matrix w), that block
// starting at (n2, n2) of an N x N row-major matrix (lower triangle valid).  For complex q
// the trailing block is the real embedding [[Re W, -Im W], [Im W, Re W]] (embed = true).
// Wrt layout: [tau][R][n][n].  At most 8 omega per launch.
void launch_w_accum(const double *aug, long long aug_stride, int cnt, int n, int n2, int N, bool embed,
                    const double *coef, int ntau, const double *cs_r, const double *sn_r, int nR, double *Wrt,
                    hipStream_t s);

// sig[R][a][b] += sum_{e : fold[e] == R} (extA[e][a][b] + extB[e][a][b])
//              + sum_{e : foldneg[e] == R} extB[e][b][a]
template <class T>
void launch_sigma_fold(T *sig, const T *extA, const T *extB, const int *fold, const int *foldneg, int NE, int nR,
                       int n, hipStream_t s);

// Sigma(R, i w) += 0.5 (S+ + S-) cos_t2f(w) + i 0.5 (S+ - S-) sin_t2f(w) for all w
// (sre / sim layout [w][R][n][n]).
template <class T>
void launch_sigma_freq(const T *sp, const T *sn, long long nelem, const double *cw, const double *sw, int nw,
                       double *sre, double *sim, hipStream_t s);

// Y[w][a][st] = sum_b sum_R ph[(I(a), J(b), R)] Sigma(R, w)[a][b] U[st][b] for one k point
// (complex, st < nst <= 16); ph layout [I][J][R] (re, im interleaved), U [st][b] (re, im).
void launch_sigma_ks_rows(const double *sre, const double *sim, int nw, int nR, int n, const int *atom_of,
                          int natom, const double *ph, const double *U, int nst, double *Y, hipStream_t s);

// Dense helpers of the Newton-Schulz square root (FP64, row-major n x n, lda = n):
// B = alpha A + beta I
void launch_scale_add_identity(const double *A, double *B, int n, double alpha, double beta, hipStream_t s);
// out[0] = max |A - B|, out[1] = max |B| (single workgroup, overwritten)
void launch_absmax_diff(const double *A, const double *B, long long n, double *out, hipStream_t s);
// A[i][:] *= f[i] (row-major n x n)
void launch_scale_rows(double *A, int n, const double *f, hipStream_t s);

// Head of the dielectric matrix at q = 0 (replace_w_head): for each of cnt matrices eps_w (lower
// triangle valid, rows of length ld, stride apart), q[w] = u^T eps_w u over the n x n block at the
// origin (q is accumulated, clear it first) ...
void launch_head_quad(const double *A, long long stride, int ld, int n, int cnt, const double *u, double *q,
                      hipStream_t s);
// ... then eps_w += (eps_mac[w] - q[w]) u u^T on the lower triangle of that block and, when
// off2 > 0, of the copy at (off2, off2) (the real embedding of a q = 0 matrix).
void launch_head_update(double *A, long long stride, int ld, int n, int off2, int cnt, const double *u,
                        const double *eps_mac, const double *q, hipStream_t s);
// out[0] = sum A^2 (overwritten)
void launch_sumsq(const double *A, long long n, double *out, hipStream_t s);

}  // namespace lra
