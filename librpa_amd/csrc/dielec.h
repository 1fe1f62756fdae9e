// Macroscopic dielectric function on the grid frequencies (host).
#pragma once

#include <vector>

namespace lra {

// interpolate_dielec_func (src/core/dielecmodel.cpp:65-104): the values eps_M(i omega) given on
// the caller's imaginary frequencies om_in, on the frequencies om_target.
//   0: as given (one value per target frequency);
//   1: LibRPA's cubic spline (src/math/interpolate.cpp);
//   2: the double Havriliak-Negami model fitted by Levenberg-Marquardt (src/math/fitting.cpp).
std::vector<double> interpolate_dielec(int option, const std::vector<double> &om_in,
                                       const std::vector<double> &df_in,
                                       const std::vector<double> &om_target);

}  // namespace lra
