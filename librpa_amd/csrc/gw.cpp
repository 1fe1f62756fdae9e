#include "gw.h"

namespace lra {

struct GwEngine {};

void gw_build_sigma(Dataset &, const LibrpaOptions &) { LRA_THROW("G0W0 self-energy is not available yet"); }

void gw_sigc_kgrid(Dataset &, const LibrpaOptions &, int, int, const int *, int, int, const double *, const double *,
                   double *, double *) {
    LRA_THROW("G0W0 self-energy is not available yet");
}

}  // namespace lra
