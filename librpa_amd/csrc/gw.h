// G0W0: screened interaction W(q, i omega) -> W(R, i tau) and the self-energy
// Sigma(R, i tau) = C G W C contraction that shares the chi0 engine.
#pragma once

#include "../../include/librpa_options.h"
#include "dataset.h"

namespace lra {

void gw_build_sigma(Dataset &d, const LibrpaOptions &o);
void gw_sigc_kgrid(Dataset &d, const LibrpaOptions &o, int n_spins, int n_kpts_this, const int *iks_this,
                   int i_state_low, int i_state_high, const double *vxc, const double *vexx, double *sigc_re,
                   double *sigc_im);

}  // namespace lra
