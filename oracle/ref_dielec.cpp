// Checker driver, test infrastructure only: runs the reference's own treatment of
// a macroscopic dielectric function given on the caller's imaginary frequencies
// (interpolate_dielec_func, src/core/dielecmodel.cpp:65-104), with the
// reference's cubic spline (src/math/interpolate.cpp) and Levenberg-Marquardt
// fit (src/math/fitting.cpp) compiled from its sources:
//   option 1: interp_cubic_spline(frequencies_in, df_in, frequencies_target);
//   option 2: LevMarqFitting::fit_eval of the double Havriliak-Negami model, all
//             parameters 1 at the start but pars[0] = pars[4] = df_in[0].
// dielecmodel.cpp itself needs LibRI and BLACS; the two lambdas of the model
// (DoubleHavriliakNegami::func_imfreq / grad_imfreq, dielecmodel.cpp:33-63) are
// therefore written out here, formula for formula.
//
// Input and output are little-endian binary streams (int32, float64):
//   in : count, then per case {option, n_in, n_out, xs[n_in], ys[n_in], xs_new[n_out]}
//   out: per case {ys_new[n_out], then for option 2 the fitted pars[8]}
//
// Usage: ref_dielec <in> <out>

#include "math/fitting.h"
#include "math/interpolate.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

double dhn(double u, const std::vector<double> &p) {
  return 1.0 + (p[0] - 1.0) / std::pow(1.0 + std::pow(u * p[3], p[1]), p[2]) +
         (p[4] - 1.0) / std::pow(1.0 + std::pow(u * p[7], p[5]), p[6]);
}

void dhn_grad(std::vector<double> &g, double u, const std::vector<double> &p) {
  using std::log;
  using std::pow;
  g[0] = 1.0 / pow(1.0 + pow(u * p[3], p[1]), p[2]);
  g[1] = (p[0] - 1.0) * (-p[2]) / pow(1.0 + pow(u * p[3], p[1]), p[2] + 1) * log(u * p[3]) *
         pow(u * p[3], p[1]);
  g[2] = (1.0 - p[0]) * log(1.0 + pow(u * p[3], p[1])) / pow(1.0 + pow(u * p[3], p[1]), p[2]);
  g[3] = (p[0] - 1.0) * (-p[2]) / pow(1.0 + pow(u * p[3], p[1]), p[2] + 1) * p[1] / p[3] *
         pow(u * p[3], p[1]);
  g[4] = 1.0 / pow(1.0 + pow(u * p[7], p[5]), p[6]);
  g[5] = (p[4] - 1.0) * (-p[6]) / pow(1.0 + pow(u * p[7], p[5]), p[6] + 1) * log(u * p[7]) *
         pow(u * p[7], p[5]);
  g[6] = (1.0 - p[4]) * log(1.0 + pow(u * p[7], p[5])) / pow(1.0 + pow(u * p[7], p[5]), p[6]);
  g[7] = (p[4] - 1.0) * (-p[6]) / pow(1.0 + pow(u * p[7], p[5]), p[6] + 1) * p[5] / p[7] *
         pow(u * p[7], p[5]);
}

struct Reader {
  FILE *f;
  template <class T> T get() {
    T v;
    if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("short input");
    return v;
  }
  std::vector<double> get_doubles(size_t n) {
    std::vector<double> v(n);
    if (n && std::fread(v.data(), sizeof(double), n, f) != n) throw std::runtime_error("short input");
    return v;
  }
};

void run(const char *in_path, const char *out_path) {
  FILE *fin = std::fopen(in_path, "rb");
  if (!fin) throw std::runtime_error(std::string("cannot open ") + in_path);
  FILE *fout = std::fopen(out_path, "wb");
  if (!fout) throw std::runtime_error(std::string("cannot open ") + out_path);
  Reader r{fin};
  const int32_t count = r.get<int32_t>();
  for (int32_t c = 0; c < count; ++c) {
    const int32_t option = r.get<int32_t>(), n_in = r.get<int32_t>(), n_out = r.get<int32_t>();
    const std::vector<double> xs = r.get_doubles(n_in), ys = r.get_doubles(n_in);
    const std::vector<double> xs_new = r.get_doubles(n_out);
    if (option == 1) {
      const std::vector<double> out = librpa_int::interp_cubic_spline(xs, ys, xs_new);
      std::fwrite(out.data(), sizeof(double), out.size(), fout);
    } else if (option == 2) {
      librpa_int::LevMarqFitting lm;
      std::vector<double> pars(8, 1.0);
      pars[0] = pars[4] = ys[0];
      const std::vector<double> out = lm.fit_eval(pars, xs, ys, dhn, dhn_grad, xs_new);
      std::fwrite(out.data(), sizeof(double), out.size(), fout);
      std::fwrite(pars.data(), sizeof(double), pars.size(), fout);
    } else {
      throw std::runtime_error("option must be 1 or 2");
    }
  }
  std::fclose(fin);
  std::fclose(fout);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <in> <out>\n", argv[0]);
    return 2;
  }
  try {
    run(argv[1], argv[2]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "ref_dielec: %s\n", e.what());
    return 1;
  }
  return 0;
}
