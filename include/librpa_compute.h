/* Compute entry points (LibRPA include/librpa_compute.h:26-149).
 * Errors are reported as C++ exceptions thrown through the C ABI, as in
 * LibRPA (src/api/compute_rpa.cpp:62,157); librpa_amd_set_error_mode()
 * switches to recorded errors for callers that cannot catch them. */
#pragma once

#include "librpa_handler.h"
#include "librpa_options.h"

#ifdef __cplusplus
extern "C" {
#endif

/* librpa_compute.h:26 / compute_rpa.cpp:15-34: nfreq frequency nodes and weights. */
void librpa_get_imaginary_frequency_grids(LibrpaHandler *h, const LibrpaOptions *p_opts, double *omegas,
                                          double *weights);

/* librpa_compute.h:42 / compute_rpa.cpp:36-168: returns Re Ec and fills the
 * per-IBZ-k contributions (indexed like LibRPA's get_k_index_ibz). */
double librpa_get_rpa_correlation_energy(LibrpaHandler *h, const LibrpaOptions *p_opts, int n_ibz_kpoints,
                                         double *rpa_corr_ibzk_contrib_re, double *rpa_corr_ibzk_contrib_im);

/* librpa_compute.h:51-87: exact exchange is outside this library's path; these
 * raise an error naming the missing feature. */
void librpa_build_exx(LibrpaHandler *h, const LibrpaOptions *p_opts);
void librpa_get_exx_pot_kgrid(LibrpaHandler *h, const LibrpaOptions *p_opts, const int n_spins,
                              const int n_kpts_this, const int *iks_this, int i_state_low, int i_state_high,
                              double *vexx);
void librpa_get_exx_pot_band_k(LibrpaHandler *h, const LibrpaOptions *p_opts, const int n_spins,
                               const int n_kpts_band_this, const int *iks_band_this, int i_state_low,
                               int i_state_high, double *vexx_band);

/* librpa_compute.h:94 / compute_g0w0.cpp:23-160: chi0 -> W -> Sigma(R, i omega). */
void librpa_build_g0w0_sigma(LibrpaHandler *h, const LibrpaOptions *p_opts);

/* librpa_compute.h:117 / compute_g0w0.cpp:162-260: KS-projected Sigma_c and QP solve. */
void librpa_get_g0w0_sigc_kgrid(LibrpaHandler *h, const LibrpaOptions *p_opts, const int n_spins,
                                const int n_kpts_this, const int *iks_this, int i_state_low, int i_state_high,
                                const double *vxc, const double *vexx, double *sigc_re, double *sigc_im);

/* librpa_compute.h:142 */
void librpa_get_g0w0_sigc_band_k(LibrpaHandler *h, const LibrpaOptions *p_opts, const int n_spins,
                                 const int n_kpts_band_this, const int *iks_band_this, int i_state_low,
                                 int i_state_high, const double *vxc_band, const double *vexx_band,
                                 double *sigc_band_re, double *sigc_band_im);

#ifdef __cplusplus
}
#endif
