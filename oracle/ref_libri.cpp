// Checker driver, test infrastructure only: runs the reference's own LibRI
// label contractions on one tau point of a small system and writes every block
// they return, so that tests can hold the oracle against the reference's code
// itself. Two modes:
//
//   chi0s : RPA::cal_chi0s (thirdparty/LibRI/include/RI/physics/RPA.hpp:104-138),
//           fed as src/core/chi0.cpp:956-1066 feeds it: set_parallel over the
//           atoms and the BvK period, set_Cs with the raw (unfolded) Cs
//           tensors (mu, i, j) of src/api/input.cpp:525-540, set_Gs_pos /
//           set_Gs_neg with G(R, +tau) / G(R, -tau) as (i, j) blocks
//           (src/core/chi0.cpp:395-419), then cal_chi0s.
//   sigmas: GW::cal_Sigmas (thirdparty/LibRI/include/RI/physics/GW.hpp:99-122),
//           fed as src/core/gw.cpp:425-470 and :616-618 feed it: set_Cs, set_Ws
//           with W(R, tau) as (mu, nu) blocks, set_Gs with G(R, tau) of one
//           sign, then cal_Sigmas.
//
// Input and output are little-endian binary streams (int32 and float64):
//   in : natom, period[3], latvec[3][3], pos[natom][3], nao[natom],
//        nabf[natom], then three block lists, each a count followed by blocks
//        {I, J, R[3], data}:
//          chi0s : Cs (nabf_I*nao_I*nao_J values), G+ and G- (nao_I*nao_J)
//          sigmas: Cs, W (nabf_I*nabf_J), G (nao_I*nao_J)
//   out: count, then blocks {I, J, R[3], rows, cols, data[rows*cols]}
//
// Usage: ref_libri chi0s|sigmas <in> <out>; one process, no launcher needed.

#include <RI/physics/GW.h>
#include <RI/physics/RPA.h>

#include <mpi.h>

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <valarray>
#include <vector>

namespace {

using Tensor = RI::Tensor<double>;
using Key = std::pair<int, std::array<int, 3>>;
using BlockMap = std::map<int, std::map<Key, Tensor>>;

struct Reader {
  FILE *f;
  template <class T> T get() {
    T v;
    if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("short input");
    return v;
  }
  void get_doubles(double *p, size_t n) {
    if (std::fread(p, sizeof(double), n, f) != n) throw std::runtime_error("short input");
  }
};

// Row and column sizes of one block list, per atom.
enum class Shape { cs, ao_ao, abf_abf };

BlockMap read_blocks(Reader &r, const std::vector<int> &nao, const std::vector<int> &nabf,
                     Shape shape) {
  BlockMap out;
  const int32_t count = r.get<int32_t>();
  for (int32_t b = 0; b < count; ++b) {
    const int32_t I = r.get<int32_t>(), J = r.get<int32_t>();
    std::array<int, 3> R;
    for (int &x : R) x = r.get<int32_t>();
    std::vector<size_t> dims;
    switch (shape) {
      case Shape::cs: dims = {size_t(nabf.at(I)), size_t(nao.at(I)), size_t(nao.at(J))}; break;
      case Shape::ao_ao: dims = {size_t(nao.at(I)), size_t(nao.at(J))}; break;
      case Shape::abf_abf: dims = {size_t(nabf.at(I)), size_t(nabf.at(J))}; break;
    }
    size_t n = 1;
    for (size_t d : dims) n *= d;
    auto data = std::make_shared<std::valarray<double>>(n);
    r.get_doubles(&(*data)[0], n);
    out[I][{J, R}] = dims.size() == 3 ? Tensor({dims[0], dims[1], dims[2]}, data)
                                      : Tensor({dims[0], dims[1]}, data);
  }
  return out;
}

void write_blocks(const char *out_path, const BlockMap &blocks) {
  FILE *fout = std::fopen(out_path, "wb");
  if (!fout) throw std::runtime_error(std::string("cannot open ") + out_path);
  int32_t count = 0;
  for (const auto &I_blocks : blocks) count += static_cast<int32_t>(I_blocks.second.size());
  std::fwrite(&count, sizeof count, 1, fout);
  for (const auto &I_blocks : blocks) {
    for (const auto &JR_t : I_blocks.second) {
      const Tensor &t = JR_t.second;
      const int32_t head[7] = {I_blocks.first, JR_t.first.first, JR_t.first.second[0],
                               JR_t.first.second[1], JR_t.first.second[2],
                               static_cast<int32_t>(t.shape[0]),
                               static_cast<int32_t>(t.shape[1])};
      std::fwrite(head, sizeof(int32_t), 7, fout);
      std::fwrite(t.ptr(), sizeof(double), t.get_shape_all(), fout);
    }
  }
  std::fclose(fout);
}

void run(const std::string &mode, const char *in_path, const char *out_path) {
  if (mode != "chi0s" && mode != "sigmas") throw std::runtime_error("unknown mode " + mode);
  const bool sigmas = mode == "sigmas";
  FILE *fin = std::fopen(in_path, "rb");
  if (!fin) throw std::runtime_error(std::string("cannot open ") + in_path);
  Reader r{fin};
  const int natom = r.get<int32_t>();
  std::array<int, 3> period;
  for (int &p : period) p = r.get<int32_t>();
  std::array<std::array<double, 3>, 3> latvec;
  for (auto &row : latvec)
    for (double &x : row) x = r.get<double>();
  std::map<int, std::array<double, 3>> pos;
  for (int a = 0; a < natom; ++a)
    for (double &x : pos[a]) x = r.get<double>();
  std::vector<int> nao(natom), nabf(natom);
  for (int &x : nao) x = r.get<int32_t>();
  for (int &x : nabf) x = r.get<int32_t>();
  const BlockMap cs = read_blocks(r, nao, nabf, Shape::cs);
  const BlockMap second = read_blocks(r, nao, nabf, sigmas ? Shape::abf_abf : Shape::ao_ao);
  const BlockMap third = read_blocks(r, nao, nabf, Shape::ao_ao);
  std::fclose(fin);

  if (sigmas) {
    RI::GW<int, int, 3, double> gw;
    gw.set_parallel(MPI_COMM_WORLD, pos, latvec, period);
    gw.set_Cs(cs, 0.0);
    gw.set_Ws(second, 0.0);
    gw.set_Gs(third, 0.0);
    gw.cal_Sigmas();
    write_blocks(out_path, gw.Sigmas);
  } else {
    RI::RPA<int, int, 3, double> rpa;
    rpa.set_parallel(MPI_COMM_WORLD, pos, latvec, period);
    rpa.set_Cs(cs, 0.0);
    rpa.set_Gs_pos(second, 0.0);
    rpa.set_Gs_neg(third, 0.0);
    rpa.cal_chi0s();
    write_blocks(out_path, rpa.chi0s);
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s chi0s|sigmas <in> <out>\n", argv[0]);
    return 2;
  }
  MPI_Init(&argc, &argv);
  int rc = 0;
  try {
    run(argv[1], argv[2], argv[3]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "ref_libri: %s\n", e.what());
    rc = 1;
  }
  MPI_Finalize();
  return rc;
}
