// Checker driver, test infrastructure only: runs the reference's own Pade
// continuation (AnalyContPade, src/core/analycont.cpp:17-80) and QP solver
// (qpe_solver_pade_self_consistent, src/core/qpe_solver.cpp:10-55), compiled
// from the reference's sources, so that tests can hold the oracle's and the
// library's restatements against them. LibRPA calls them per state as
// src/api/compute_g0w0.cpp:241-246 does: the imaginary frequencies i*omega as
// the nodes, Sigma_c(i*omega) of the state as the data.
//
// Input and output are little-endian binary streams (int32 and float64,
// complex numbers as (re, im) pairs):
//   in : count, then per case {n_pars, n_data, n_eval, n_iter_max,
//        xs[n_data], data[n_data], z[n_eval], e_mf, e_fermi, vxc, sigma_x,
//        diff_init, thres}
//   out: per case {pade(z)[n_eval], info (int32), e_qp, sigc}
//
// Usage: ref_qpe <in> <out>

#include "core/analycont.h"
#include "core/qpe_solver.h"

#include <complex>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using cplx = std::complex<double>;

struct Reader {
  FILE *f;
  template <class T> T get() {
    T v;
    if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("short input");
    return v;
  }
  std::vector<cplx> get_cplx(size_t n) {
    std::vector<cplx> v(n);
    for (auto &c : v) {
      const double re = get<double>(), im = get<double>();
      c = cplx(re, im);
    }
    return v;
  }
};

void put_cplx(FILE *f, cplx c) {
  const double v[2] = {c.real(), c.imag()};
  std::fwrite(v, sizeof(double), 2, f);
}

void run(const char *in_path, const char *out_path) {
  FILE *fin = std::fopen(in_path, "rb");
  if (!fin) throw std::runtime_error(std::string("cannot open ") + in_path);
  FILE *fout = std::fopen(out_path, "wb");
  if (!fout) throw std::runtime_error(std::string("cannot open ") + out_path);
  Reader r{fin};
  const int32_t count = r.get<int32_t>();
  for (int32_t c = 0; c < count; ++c) {
    const int32_t n_pars = r.get<int32_t>(), n_data = r.get<int32_t>();
    const int32_t n_eval = r.get<int32_t>(), n_iter_max = r.get<int32_t>();
    const std::vector<cplx> xs = r.get_cplx(n_data), data = r.get_cplx(n_data);
    const std::vector<cplx> zs = r.get_cplx(n_eval);
    const double e_mf = r.get<double>(), e_fermi = r.get<double>();
    const double vxc = r.get<double>(), sigma_x = r.get<double>();
    const double diff_init = r.get<double>(), thres = r.get<double>();

    const librpa_int::AnalyContPade pade(n_pars, xs, data);
    for (const cplx &z : zs) put_cplx(fout, pade.get(z));
    double e_qp = 0.0;
    cplx sigc;
    const int32_t info = librpa_int::qpe_solver_pade_self_consistent(
        pade, e_mf, e_fermi, vxc, sigma_x, e_qp, sigc, diff_init, thres, n_iter_max);
    std::fwrite(&info, sizeof info, 1, fout);
    std::fwrite(&e_qp, sizeof e_qp, 1, fout);
    put_cplx(fout, sigc);
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
    std::fprintf(stderr, "ref_qpe: %s\n", e.what());
    return 1;
  }
  return 0;
}
