// Imaginary time / frequency grids.
//
// Minimax grids follow LibRPA TFGrids::generate_minimax (src/core/timefreq.cpp:193-236)
// which delegates to greenX gx_minimax_grid (GX-TimeFrequency/src/minimax_grids.F90:53-237):
// tabulated optimal coefficients are selected by the energy-range bin
// (minimax_utils.F90:38-86), scaled by e_min, and the cosine / sine transform
// weights are least-squares fits on log-spaced transition energies
// (minimax_grids.F90:286-406), solved here with a one-sided Jacobi SVD and the
// same Tikhonov filter sigma / (sigma^2 + reg^2).
#include <algorithm>
#include <cmath>
#include <sstream>
#include <string>

#include "dataset.h"

extern const char kMinimaxTable[];  // generated from librpa_amd/data/minimax_coeffs.txt

namespace lra {

namespace {

struct TableEntry {
    std::vector<double> edges;
    std::vector<std::vector<double>> cols;
};

const std::map<std::pair<std::string, int>, TableEntry> &table() {
    static std::map<std::pair<std::string, int>, TableEntry> tab;
    static bool done = false;
    if (done) return tab;
    std::istringstream is(kMinimaxTable);
    std::string line;
    std::vector<std::string> lines;
    while (std::getline(is, line))
        if (!line.empty() && line[0] != '#') lines.push_back(line);
    size_t i = 0;
    auto nums = [](const std::string &s) {
        std::istringstream ls(s);
        std::vector<double> v;
        double x;
        while (ls >> x) v.push_back(x);
        return v;
    };
    while (i < lines.size()) {
        std::istringstream hs(lines[i]);
        std::string kind;
        int n, nb;
        hs >> kind >> n >> nb;
        TableEntry e;
        e.edges = nums(lines[i + 1]);
        for (int c = 0; c <= nb; ++c) e.cols.push_back(nums(lines[i + 2 + c]));
        tab[{kind, n}] = std::move(e);
        i += 3 + nb;
    }
    done = true;
    return tab;
}

std::vector<double> coeffs(const std::string &kind, int n, double e_range, double &e_ratio) {
    auto it = table().find({kind, n});
    if (it == table().end()) LRA_THROW("unsupported minimax grids size: " + std::to_string(n));
    const auto &e = it->second;
    size_t idx = e.edges.size();
    double best = HUGE_VAL;
    for (size_t j = 0; j < e.edges.size(); ++j)
        if (e_range < e.edges[j] && e.edges[j] < best) { idx = j; best = e.edges[j]; }
    e_ratio = 1.0;
    if (idx == 0 && n > 20) {
        e_ratio = e.edges[0] / e_range;
        if (e_ratio > 1.5) e_ratio /= 1.5;
    }
    return e.cols[idx];
}

// Least-squares weights w minimising |A w - psi| through the SVD of A (m x n, column-major).
std::vector<double> svd_solve(std::vector<double> A, int m, int n, const std::vector<double> &psi, double reg) {
    std::vector<double> V(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 80; ++sweep) {
        double off = 0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double a = 0, b = 0, c = 0;
                const double *cp = &A[(size_t)p * m], *cq = &A[(size_t)q * m];
                for (int i = 0; i < m; ++i) { a += cp[i] * cp[i]; b += cq[i] * cq[i]; c += cp[i] * cq[i]; }
                if (std::fabs(c) <= 1e-300) continue;
                off = std::max(off, std::fabs(c) / std::sqrt(a * b));
                const double zeta = (b - a) / (2 * c);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
                double *xp = &A[(size_t)p * m], *xq = &A[(size_t)q * m];
                for (int i = 0; i < m; ++i) {
                    const double u = xp[i], v = xq[i];
                    xp[i] = cs * u - sn * v;
                    xq[i] = sn * u + cs * v;
                }
                for (int i = 0; i < n; ++i) {
                    const double u = V[p * n + i], v = V[q * n + i];
                    V[p * n + i] = cs * u - sn * v;
                    V[q * n + i] = sn * u + cs * v;
                }
            }
        if (off < 1e-15) break;
    }
    std::vector<double> w(n, 0.0), sv(n);
    double smax = 0;
    for (int j = 0; j < n; ++j) {
        double s2 = 0;
        for (int i = 0; i < m; ++i) s2 += A[(size_t)j * m + i] * A[(size_t)j * m + i];
        sv[j] = std::sqrt(s2);
        smax = std::max(smax, sv[j]);
    }
    for (int j = 0; j < n; ++j) {
        const double s = sv[j];
        // singular values at roundoff level carry no information; drop them
        if (s <= 1e-15 * smax) continue;
        double ut = 0;
        for (int i = 0; i < m; ++i) ut += A[(size_t)j * m + i] / s * psi[i];
        const double f = s / (reg * reg + s * s) * ut;
        for (int i = 0; i < n; ++i) w[i] += V[j * n + i] * f;
    }
    return w;
}

// transformation types of minimax_grids.F90:420-500
std::vector<double> fit(int n, const std::vector<double> &tau, const std::vector<double> &om, double emin,
                        double emax, int type, double reg) {
    int nx = ((int)std::log10(emax / emin) + 1) * 200;
    nx = std::max(nx, n);
    const double fac = std::pow(emax / emin, 1.0 / (nx - 1.0));
    std::vector<double> x(nx);
    for (int i = 0; i < nx; ++i) x[i] = emin * std::pow(fac, (double)i);
    std::vector<double> W((size_t)n * n);
    std::vector<double> A((size_t)nx * n), psi(nx);
    for (int ip = 0; ip < n; ++ip) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < nx; ++i) {
                double a;
                if (type == 1) a = std::cos(om[ip] * tau[j]) * std::exp(-x[i] * tau[j]);
                else if (type == 2) a = std::cos(tau[ip] * om[j]) * 2.0 * x[i] / (x[i] * x[i] + om[j] * om[j]);
                else if (type == 3) a = std::sin(om[ip] * tau[j]) * std::exp(-x[i] * tau[j]);
                else a = std::sin(tau[ip] * om[j]) * 2.0 * om[j] / (x[i] * x[i] + om[j] * om[j]);
                A[(size_t)j * nx + i] = a;
            }
        for (int i = 0; i < nx; ++i) {
            if (type == 1) psi[i] = 2.0 * x[i] / (x[i] * x[i] + om[ip] * om[ip]);
            else if (type == 3) psi[i] = 2.0 * om[ip] / (x[i] * x[i] + om[ip] * om[ip]);
            else psi[i] = std::exp(-x[i] * tau[ip]);
        }
        auto w = svd_solve(A, nx, n, psi, reg);
        for (int j = 0; j < n; ++j) W[(size_t)ip * n + j] = w[j];
    }
    return W;
}

}  // namespace

void e_min_max(const MeanField &mf, double &emin, double &emax) {
    // MeanField::get_E_min_max / get_band_gap (src/core/meanfield.cpp:151-198)
    double lb = HUGE_VAL, ub = 2.2250738585072014e-308;
    double homo = -1e6, lumo = 1e6;
    const double midpoint = 1.0 / (mf.nspin * mf.nk * mf.nspinor);
    for (int s = 0; s < mf.nspin; ++s)
        for (int k = 0; k < mf.nk; ++k) {
            const double *e = &mf.eps[((size_t)s * mf.nk + k) * mf.nbands];
            const double *w = &mf.wg[((size_t)s * mf.nk + k) * mf.nbands];
            lb = std::min(lb, e[0]);
            ub = std::max(ub, e[mf.nbands - 1]);
            int hl = -1;
            for (int b = 0; b < mf.nbands; ++b)
                if (w[b] >= midpoint) hl = b;
            if (hl + 1 < mf.nbands) lumo = std::min(lumo, e[hl + 1]);
            if (hl != -1) homo = std::max(homo, e[hl]);
        }
    emax = ub - lb;
    emin = lumo - homo;
}

TFGrid make_grid(int type, int n, double emin, double emax, double reg, double fmin, double fintv, double tmin,
                 double tintv) {
    TFGrid g;
    g.n = n;
    g.e_min = emin;
    g.e_max = emax;
    const size_t nn = (size_t)n * n;
    g.cos_t2f.assign(nn, 0);
    g.cos_f2t.assign(nn, 0);
    g.sin_t2f.assign(nn, 0);
    g.sin_f2t.assign(nn, 0);
    // Gauss-type requests fall through to the minimax grid in LibRPA (timefreq.cpp:121-139).
    if (type == 3 || type == 0 || type == 1 || type == 2 || type == -101) {
        double r;
        auto xo = coeffs("omega", n, emax / emin, r);
        for (auto &v : xo) v /= r;
        auto xt = coeffs("tau", n, emax / emin, r);
        for (auto &v : xt) v *= r;
        g.omega.assign(xo.begin(), xo.begin() + n);
        g.omega_w.assign(xo.begin() + n, xo.end());
        for (auto &v : g.omega) v *= emin;
        for (auto &v : g.omega_w) v *= emin;
        g.tau.assign(xt.begin(), xt.begin() + n);
        g.tau_w.assign(xt.begin() + n, xt.end());
        for (auto &v : g.tau) v /= 2 * emin;
        for (auto &v : g.tau_w) v /= 2 * emin;
        auto ct = fit(n, g.tau, g.omega, emin, emax, 1, reg);
        auto cf = fit(n, g.tau, g.omega, emin, emax, 2, reg);
        auto st = fit(n, g.tau, g.omega, emin, emax, 3, reg);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                g.cos_t2f[j * n + i] = ct[j * n + i] * std::cos(g.tau[i] * g.omega[j]);
                g.cos_f2t[i * n + j] = cf[i * n + j] * std::cos(g.tau[i] * g.omega[j]);
                g.sin_t2f[j * n + i] = st[j * n + i] * std::sin(g.tau[i] * g.omega[j]);
            }
    } else if (type == 5) {
        // EvenSpaced_TF (timefreq.cpp:171-191): diagonal stand-in transforms
        const double w = 1.0 / n;
        for (int i = 0; i < n; ++i) {
            g.omega.push_back(fmin + fintv * i);
            g.omega_w.push_back(w);
            g.tau.push_back(tmin + tintv * i);
            g.tau_w.push_back(w);
            g.cos_t2f[i * n + i] = w;
            g.sin_t2f[i * n + i] = w;
            g.cos_f2t[i * n + i] = 1 / w;
            g.sin_f2t[i * n + i] = 1 / w;
        }
    } else {
        LRA_THROW("grid type " + std::to_string(type) +
                  " has no imaginary-time grid; the space-time path needs Minimax or EvenSpaced_TF");
    }
    return g;
}

}  // namespace lra
