/* Extensions of the MI355X build.  None of these has a counterpart in
 * LibRPA; they select precision, expose intermediates for verification and
 * report device timings. */
#pragma once

#include "librpa_handler.h"
#include "librpa_options.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LIBRPA_AMD_FP64 0
#define LIBRPA_AMD_FP32 1
#define LIBRPA_AMD_BF16X3 2
/* chi0 only: FP64 at the tau points whose column of the cosine tau->omega
 * matrix has an entry above 1e3 (the transform multiplies their chi0 errors
 * by up to that much), FP32 at the others; chi0(R, tau) is stored in FP64. */
#define LIBRPA_AMD_MIXED 3
/* chi0 only: as LIBRPA_AMD_MIXED with BF16x3 instead of FP32 at the low-weight tau points (their
 * small-K X and T_A GEMMs stay FP32). */
#define LIBRPA_AMD_MIXED_BF16 4

/* Contraction precision of the chi0 build (G, W, T and the final Cs.T
 * contraction) and of the Ec transforms (chi0 L and L^T X).  The Cholesky
 * log-determinant, the tau->omega quadrature and all accumulation of Ec stay
 * in FP64.  Default FP64 / FP64. */
void librpa_amd_set_precision(LibrpaHandler *h, int chi0_prec, int ec_prec);

/* 0 (default): errors are C++ exceptions thrown through the C ABI as in
 * LibRPA.  1: errors are caught at the API boundary and recorded; the call
 * returns NaN / leaves outputs untouched.  Process-wide. */
void librpa_amd_set_error_mode(int record);
/* Copy the recorded error message (if any) into buf, clear it, and return 1
 * if there was one. */
int librpa_amd_get_last_error(char *buf, int len);

/* Keep chi0(R, tau) and chi0(q, omega) on the device after the Ec call. */
void librpa_amd_set_keep_chi0(LibrpaHandler *h, int keep);
/* chi0s block (naux_I x naux_J, row-major) of minimax time point itau for
 * the BvK cell R; I <= J.  Returns 0 on success. */
int librpa_amd_get_chi0s_block(LibrpaHandler *h, int itau, int I, int J, const int R[3], double *out);
/* chi0(q, omega) block, I <= J, row-major re / im.  Returns 0 on success. */
int librpa_amd_get_chi0q_block(LibrpaHandler *h, int ifreq, int iq, int I, int J, double *re, double *im);

/* Minimax grid of the current mean-field data: tau, tau weights, omega,
 * omega weights (nfreq each) and the cosine t->f (nfreq x nfreq, omega rows),
 * cosine f->t (tau rows) and sine t->f matrices.  Returns nfreq. */
int librpa_amd_get_tfgrids(LibrpaHandler *h, const LibrpaOptions *p_opts, double *tau, double *tau_w,
                           double *omega, double *omega_w, double *cos_t2f, double *cos_f2t, double *sin_t2f);

/* The macroscopic dielectric function eps_M(i omega) given at n_in imaginary frequencies, on the
 * n_out frequencies omega_out, as the head replacement of G0W0 uses it (option_dielect_func):
 * 0 as given (n_in == n_out), 1 LibRPA's cubic spline, 2 its double Havriliak-Negami fit.
 * LibRPA keeps this step internal (interpolate_dielec_func, src/core/dielecmodel.cpp:65-104);
 * it is exported here so that callers can see the values W is built with.  Host only.
 * Returns n_out, or -1 on error. */
int librpa_amd_interpolate_dielec_func(int option, int n_in, const double *omega_in, const double *df_in,
                                       int n_out, const double *omega_out, double *df_out);

/* Device timing table as JSON {"region": [ms, calls, flops, bytes], ...}.
 * Returns the length written (excluding the terminator). */
int librpa_amd_profile_json(char *buf, int len);
void librpa_amd_profile_reset(void);

/* Counters of the last chi0 build: [0] algorithmic chi0 flops (canonical count, one spin,
 * all tau), [1] executed GEMM flops, [2] chi0 wall seconds, [3] Ec wall seconds,
 * [4] output blocks (I<=J, R), [5] Cs blocks after folding, [6] n_abf, [7] n_basis,
 * [8] the mixed mode's cut: tau points whose cosine-transform column stays at or below it run in
 * FP32 (1e3, or LIBRPA_AMD_MIXED_AMAX), [9] tau points of the last build that ran in FP32.
 * Returns the number of entries written. */
int librpa_amd_get_stats(LibrpaHandler *h, double *out, int n);

/* G0W0 intermediates after librpa_build_g0w0_sigma.  W(R, tau) block
 * (naux_I x naux_J, row-major) of minimax time itau for the BvK cell R, all
 * ordered atom pairs.  Returns 0 on success. */
int librpa_amd_get_wc_block(LibrpaHandler *h, int itau, int I, int J, const int R[3], double *out);
/* Sigma_c(R, i omega) block (nbasis_I x nbasis_J, row-major re / im) of spin
 * ispin, frequency ifreq and BvK cell R.  Returns 0 on success. */
int librpa_amd_get_sigma_block(LibrpaHandler *h, int ispin, int ifreq, int I, int J, const int R[3], double *re,
                               double *im);
/* Diagonal Kohn-Sham matrix element Sigma_c(k, i omega)_nn of state i_state at
 * k point ik for all nfreq frequencies (ascending).  Returns nfreq, < 0 on error. */
int librpa_amd_get_sigc_ks(LibrpaHandler *h, int ispin, int ik, int i_state, double *re, double *im);

/* Block-norm screening of the chi0 build (nspin = 1).  mode 0: off (default);
 * 1: skip output blocks (I, J, R) at tau whose norm-product bound
 * c_I c_J (g+_IJ gm- + g-_IJ gm+ + 2 gm+ gm-) is below threshold; 2: also skip
 * blocks whose learned prediction of ||chi0s(I,J,R; tau)||_F is below threshold
 * (model: LIBRPA_AMD_SCREEN_MODEL floats from librpa_amd_screen_train).  Skipped
 * blocks are zero in chi0(R, tau). */
#define LIBRPA_AMD_SCREEN_NFEAT 8
#define LIBRPA_AMD_SCREEN_NPARAM 4801
#define LIBRPA_AMD_SCREEN_MODEL (LIBRPA_AMD_SCREEN_NPARAM + 2 * LIBRPA_AMD_SCREEN_NFEAT + 2)
void librpa_amd_set_screening(LibrpaHandler *h, int mode, double threshold, const float *model);
/* Screening counters of the last chi0 build: [0] blocks x tau considered, [1] skipped,
 * [2] GEMM flops of the screened batches, [3] flops skipped.  Returns 4. */
int librpa_amd_get_screen_stats(LibrpaHandler *h, double *out);
/* Training samples of the last chi0 build made with librpa_amd_set_keep_chi0(h, 1) and
 * screening off: per (tau, output block) the 8 features (distance / 10 bohr, n_b(I) / 27,
 * n_b(J) / 27, ln tau, ln c_I, ln c_J, ln(g+_IJ g-_IJ), ln(gm+ gm-)) into X [n][8] and
 * ln ||chi0s||_F into t [n]; blocks with zero norm are left out.  NULL pointers return the
 * count only.  Returns n. */
int librpa_amd_get_screen_samples(LibrpaHandler *h, float *X, float *t);
/* Trains the screening MLP (8 -> 64 -> 64 -> 1, SiLU, FP32) on the device: standardises
 * X and t, runs iters full-batch Adam steps (learning rate lr decaying to lr / 20) from a
 * seeded initialisation, and writes the model.  Returns the final mean squared error in
 * standardised units (NaN on error). */
double librpa_amd_screen_train(const float *X, const float *t, int n, int iters, double lr, int seed, float *model);
/* MLP on standardised inputs, for verification: forward pass y [n], and the mean squared
 * error against t with its gradient in the parameter layout W1[8][64], b1[64], W2[64][64],
 * b2[64], w3[64], b3.  Return 0 on success. */
int librpa_amd_mlp_predict(const float *params, const float *X, int n, float *y);
int librpa_amd_mlp_loss_grad(const float *params, const float *X, const float *t, int n, float *grad, double *loss);
/* No counterpart in LibRPA.  librpa_amd_mlp_loss_grad for callers that cannot allocate device
 * memory (code under HIP graph capture, tight retraining loops): the same kernels on
 * device-resident arrays d_params [4801], d_X [n][8], d_t [n] (in) and d_grad [4801], d_loss
 * [1 double] (out), enqueued on stream (a hipStream_t; NULL = the null stream) without
 * synchronising.  The per-sample activations and chunk partial sums live in a static device
 * array of the library (35 MB, held by each GPU it runs on until the process ends), so
 * n <= LIBRPA_AMD_MLP_NOALLOC_MAX; a larger n is an error.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads) race.  Returns 0 when
 * the work is enqueued, nonzero on error. */
#define LIBRPA_AMD_MLP_NOALLOC_MAX 32768
int librpa_amd_mlp_loss_grad_noalloc(const float *d_params, const float *d_X, const float *d_t, int n, float *d_grad,
                                     double *d_loss, void *stream);

/* One dense GEMM C = alpha op(A) op(B) + beta C on host arrays through the
 * device engine (prec as above; tile_cfg 0..8 selects the workgroup tile).
 * For engine verification.  Returns 0 on success. */
int librpa_amd_gemm(int prec, int trans_a, int trans_b, int M, int N, int K, double alpha, const double *A, int lda,
                    const double *B, int ldb, double beta, double *C, int ldc, int tile_cfg);

#ifdef __cplusplus
}
#endif
