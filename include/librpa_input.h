/* Input setters (LibRPA include/librpa_input.h:25-325).
 * Every setter deep-copies the caller's arrays into the instance
 * (src/api/input.cpp); device upload happens at the first compute call. */
#pragma once

#include <stddef.h>

#include "librpa_enums.h"
#include "librpa_handler.h"

#ifdef __cplusplus
#define LIBRPA_DEFAULT(x) = x
extern "C" {
#else
#define LIBRPA_DEFAULT(x)
#endif

/* input.h:25 / input.cpp:27-60 */
void librpa_set_scf_dimension(LibrpaHandler *h, int nspins, int nkpts, int nstates, int nbasis,
                              int nspinor LIBRPA_DEFAULT(1));
/* input.h:37 / input.cpp:62-99: occupations are divided by nkpts on entry. */
void librpa_set_wg_ekb_efermi(LibrpaHandler *h, int nspins, int nkpts, int nstates, const double *wg,
                              const double *ekb, double efermi);
/* input.h:50 / input.cpp:101-126: row-major nstates x nbasis. */
void librpa_set_wfc(LibrpaHandler *h, int ispin, int ik, int nstates_local, int nbasis_local,
                    const double *wfc_real, const double *wfc_imag);
/* input.h:65 */
void librpa_set_wfc_spinor(LibrpaHandler *h, int ik, int nstates_local, int nbasis_local,
                           const double *wfc_up_real, const double *wfc_up_imag, const double *wfc_dn_real,
                           const double *wfc_dn_imag);
/* input.h:78: interleaved (re, im) pairs. */
void librpa_set_wfc_packed(LibrpaHandler *h, int ispin, int ik, int nstates_local, int nbasis_local,
                           const double *wfc_ri);
/* input.h:90 */
void librpa_set_wfc_spinor_packed(LibrpaHandler *h, int ik, int nstates_local, int nbasis_local,
                                  const double *wfc_up_ri, const double *wfc_dn_ri);
/* input.h:99, 107 */
void librpa_set_ao_basis_wfc(LibrpaHandler *h, int natoms, const size_t *nbs_wfc);
void librpa_set_ao_basis_aux(LibrpaHandler *h, int natoms, const size_t *nbs_aux);
/* input.h:115: lattice vectors a_i = lat_mat[3i..3i+2] (bohr), G likewise. */
void librpa_set_latvec_and_G(LibrpaHandler *h, const double lat_mat[9], const double G_mat[9]);
/* input.h:124 */
void librpa_set_atoms(LibrpaHandler *h, int natoms, const int *types, const double *posi_cart);
/* input.h:135: Cartesian k (bohr^-1); the BvK supercell is nk1 x nk2 x nk3. */
void librpa_set_kgrids_kvec(LibrpaHandler *h, int nk1, int nk2, int nk3, const double *kvecs);
/* input.h:147 */
void librpa_set_ibz_mapping(LibrpaHandler *h, int nkpts, const int *map_ibzk);
/* input.h:163 / input.cpp:480-554: Cs_in[(i * nbasis_j + j) * naux_mu + mu], ABF on atom I. */
void librpa_set_lri_coeff(LibrpaHandler *h, LibrpaParallelRouting routing, int I, int J, int nbasis_i,
                          int nbasis_j, int naux_mu, const int R[3], const double *Cs_in);
/* input.h:179 / input.cpp:556-597: row-major naux_mu x naux_nu, stored if max|Re| >= vq_threshold. */
void librpa_set_aux_bare_coulomb_k_atom_pair(LibrpaHandler *h, int ik, int I, int J, int naux_mu, int naux_nu,
                                             const double *Vq_real_in, const double *Vq_imag_in,
                                             double vq_threshold);
/* input.h:195 */
void librpa_set_aux_cut_coulomb_k_atom_pair(LibrpaHandler *h, int ik, int I, int J, int naux_mu, int naux_nu,
                                            const double *Vq_real_in, const double *Vq_imag_in,
                                            double vq_threshold);
/* input.h:211, 226: [begin, end) index ranges of a dense row-major block. */
void librpa_set_aux_bare_coulomb_k_2d_block(LibrpaHandler *h, int ik, int mu_begin, int mu_end, int nu_begin,
                                            int nu_end, const double *Vq_real_in, const double *Vq_imag_in);
void librpa_set_aux_cut_coulomb_k_2d_block(LibrpaHandler *h, int ik, int mu_begin, int mu_end, int nu_begin,
                                           int nu_end, const double *Vq_real_in, const double *Vq_imag_in);
/* input.h:238 */
void librpa_set_dielect_func_imagfreq(LibrpaHandler *h, int nfreq, const double *omegas_imag,
                                      const double *dielect_func);
/* input.h:247-325: band k-points for G0W0 band structures. */
void librpa_set_band_kvec(LibrpaHandler *h, int n_kpts_band, const double *kfrac_band);
void librpa_set_band_occ_eigval(LibrpaHandler *h, int n_spins, int n_kpts_band, int n_states, const double *occ,
                                const double *eig);
void librpa_set_wfc_band(LibrpaHandler *h, int ispin, int ik_band, int nstates_local, int nbasis_local,
                         const double *wfc_real, const double *wfc_imag);
void librpa_set_wfc_band_spinor(LibrpaHandler *h, int ik_band, int nstates_local, int nbasis_local,
                                const double *wfc_up_real, const double *wfc_up_imag, const double *wfc_dn_real,
                                const double *wfc_dn_imag);
void librpa_set_wfc_band_packed(LibrpaHandler *h, int ispin, int ik_band, int nstates_local, int nbasis_local,
                                const double *wfc_ri);
void librpa_set_wfc_band_spinor_packed(LibrpaHandler *h, int ik_band, int nstates_local, int nbasis_local,
                                       const double *wfc_up_ri, const double *wfc_dn_ri);
void librpa_reset_band_data(LibrpaHandler *h);

#ifdef __cplusplus
}
#endif
#undef LIBRPA_DEFAULT
