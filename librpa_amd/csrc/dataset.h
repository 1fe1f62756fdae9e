// Host-side instance state behind a LibrpaHandler: deep copies of every input
// array (LibRPA src/api/dataset.h:27-130 holds the same items) plus the
// device-side engines and retained results.
#pragma once

#include <array>
#include <complex>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "common.h"
#include "gemm.h"

namespace lra {

using cplx = std::complex<double>;

struct MeanField {
    int nspin = 0, nk = 0, nbands = 0, nbasis = 0, nspinor = 1;
    std::vector<double> eps;  // [s][k][b]
    std::vector<double> wg;   // occupations / nk, [s][k][b] (input.cpp:82)
    double efermi = 0;
    std::vector<std::vector<cplx>> wfc;  // [s * nk + k] -> nbands x nbasis
    bool dims_set = false, eig_set = false;
};

struct TFGrid {
    int n = 0;
    std::vector<double> tau, tau_w, omega, omega_w;
    std::vector<double> cos_t2f, cos_f2t, sin_t2f, sin_f2t;  // row-major n x n
    double e_min = 0, e_max = 0;
};

using CsKey = std::tuple<int, int, int, int, int>;  // I, J, R1, R2, R3

struct BandData {
    int nk = 0, nspin = 0, nstates = 0;
    std::vector<double> kfrac;   // [k][3]
    std::vector<double> occ, eig;
    std::vector<std::vector<cplx>> wfc;
};

struct RpaEngine;  // rpa.cpp
struct GwEngine;   // gw.cpp

struct Dataset {
    MeanField mf;
    std::vector<int> nb_wfc, nb_aux;
    bool lat_set = false;
    double latvec[9] = {0}, Gvec[9] = {0};
    std::vector<int> types;
    std::vector<double> pos;
    std::vector<double> frac;  // fractional coordinates, empty unless the lattice was set first
    int period[3] = {0, 0, 0};
    std::vector<double> kcart, kfrac;  // [k][3]
    std::vector<int> ibz_map;
    std::map<CsKey, std::vector<double>> cs;  // API layout, last set wins (input.cpp:539)
    // Coulomb atom-pair blocks per (ik, I, J), row-major naux_I x naux_J
    std::map<std::tuple<int, int, int>, std::vector<cplx>> vq, vq_cut;
    std::vector<double> dielect_omega, dielect_val;
    BandData band;

    // precision / behaviour knobs of this build
    Prec chi0_prec = Prec::F64;
    Prec ec_prec = Prec::F64;
    bool chi0_mixed = false;  // FP64 chi0 plan with FP32 contractions at the tau points of small weight
    Prec chi0_mixed_low = Prec::F32;  // the precision of those tau points (FP32 or BF16x3)
    bool keep_chi0 = false;
    int screen_mode = 0;
    double screen_thr = 1e-8;
    std::vector<float> screen_model;

    // results of the last compute calls
    TFGrid grid;
    // the parameters grid was made from (type, n, e_min, e_max, regulation, frequency and time
    // minima and intervals): the least-squares fits of the transform matrices take up to a second
    // for 24 points, so an unchanged grid is not made again
    std::vector<double> grid_key;
    std::shared_ptr<RpaEngine> rpa;
    std::shared_ptr<GwEngine> gw;
    double stats[10] = {0};
};

// minimax / other grids (tfgrids.cpp)
void e_min_max(const MeanField &mf, double &emin, double &emax);
TFGrid make_grid(int type, int n, double emin, double emax, double reg, double fmin, double fintv, double tmin,
                 double tintv);

}  // namespace lra
