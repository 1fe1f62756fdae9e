// Macroscopic dielectric function on the grid frequencies (host).
//
// The results must be those of LibRPA's own routines, so both follow their steps and their order
// of operations:
// * the cubic spline of src/math/interpolate.cpp:9-153. It differs from a textbook natural
//   spline: the right-hand side holds the slopes 3 (y[i+1] - y[i]) / h[i] themselves, and the
//   back substitution runs inside the forward sweep, reading entries of z that the sweep has not
//   written yet (zero). Outside the data it is the cubic Hermite extrapolation of the end
//   interval, with slope 0 at the last point;
// * the Levenberg-Marquardt fit of src/math/fitting.cpp:80-196 (damping 0.1, factors 10,
//   Cholesky of the damped normal matrix with a 1e-30 pivot floor, at most 10000 iterations,
//   stop when the error falls by less than 1e-8) of the double Havriliak-Negami model of
//   src/core/dielecmodel.cpp:33-63, started from all parameters 1 and both static constants at
//   the first datum (dielecmodel.cpp:86-96).

#include "dielec.h"

#include <cmath>
#include <string>

#include "common.h"

namespace lra {
namespace {

std::vector<double> cubic_spline(const std::vector<double> &xs, const std::vector<double> &ys,
                                 const std::vector<double> &xs_new) {
    const int n = (int)xs.size();
    if (n < 2) LRA_THROW("cubic spline of the dielectric function needs at least 2 frequencies");
    std::vector<double> h(n - 1), alpha(n - 1), l(n, 0.0), mu(n - 1, 0.0), z(n, 0.0), c(n, 0.0);
    for (int i = 0; i < n - 1; ++i) {
        h[i] = xs[i + 1] - xs[i];
        alpha[i] = 3 * (ys[i + 1] - ys[i]) / h[i];
    }
    l[0] = 1;
    for (int i = 1; i < n - 1; ++i) {
        l[i] = 2 * (xs[i + 1] - xs[i - 1]) - h[i - 1] * mu[i - 1];
        mu[i] = h[i] / l[i];
    }
    l[n - 1] = 1;
    for (int i = 1; i < n; ++i) {
        z[i] = (alpha[i - 1] - h[i - 1] * z[i - 1]) / l[i];
        c[n - i - 1] = z[n - i] - mu[n - i - 1] * c[n - i];
    }
    std::vector<double> a(n, 0.0), b(n, 0.0), d(n, 0.0);
    for (int i = 0; i < n - 1; ++i) {
        a[i] = ys[i];
        b[i] = (ys[i + 1] - ys[i]) / h[i] - h[i] * (c[i + 1] + 2 * c[i]) / 3;
        d[i] = (c[i + 1] - c[i]) / (3 * h[i]);
    }
    a[n - 1] = ys[n - 1];

    std::vector<double> out;
    out.reserve(xs_new.size());
    for (const double x : xs_new) {
        double v = 0.0;
        if (x < xs[0] || x > xs[n - 1]) {
            const int k = x < xs[0] ? 0 : n - 2;
            const double x0 = xs[k], x1 = xs[k + 1];
            const double y0 = a[k], y1 = a[k + 1], m0 = b[k], m1 = b[k + 1];
            const double hh = x1 - x0;
            const double dx = x - x0, dy = y1 - y0, dm = m1 - m0;
            v = y0 + m0 * dx +
                ((3 * dy / std::pow(hh, 2)) - (2 * m0 / hh) - (dm / hh)) * std::pow(dx, 2) +
                ((-2 * dy / std::pow(hh, 3)) + (m0 / std::pow(hh, 2)) + (dm / std::pow(hh, 2))) *
                    std::pow(dx, 3);
        } else {
            for (int i = 0; i < n - 1; ++i)
                if (x >= xs[i] && x <= xs[i + 1]) {
                    const double dx = x - xs[i];
                    v = a[i] + b[i] * dx + c[i] * std::pow(dx, 2) + d[i] * std::pow(dx, 3);
                    break;
                }
        }
        out.push_back(v);
    }
    return out;
}

// eps(i u) of two Havriliak-Negami terms: p = (eps0, alpha, beta, tau) of each
double dhn(double u, const std::vector<double> &p) {
    return 1.0 + (p[0] - 1.0) / std::pow(1.0 + std::pow(u * p[3], p[1]), p[2]) +
           (p[4] - 1.0) / std::pow(1.0 + std::pow(u * p[7], p[5]), p[6]);
}

void dhn_grad(std::vector<double> &g, double u, const std::vector<double> &p) {
    for (int t = 0; t < 2; ++t) {
        const double e0 = p[4 * t], al = p[4 * t + 1], be = p[4 * t + 2], ta = p[4 * t + 3];
        const double ut = u * ta, pw = std::pow(ut, al), base = 1.0 + pw;
        g[4 * t] = 1.0 / std::pow(base, be);
        g[4 * t + 1] = (e0 - 1.0) * (-be) / std::pow(base, be + 1) * std::log(ut) * pw;
        g[4 * t + 2] = (1.0 - e0) * std::log(base) / std::pow(base, be);
        g[4 * t + 3] = (e0 - 1.0) * (-be) / std::pow(base, be + 1) * al / ta * pw;
    }
}

double sq_error(const std::vector<double> &p, const std::vector<double> &xs, const std::vector<double> &ys) {
    double e = 0;
    for (size_t i = 0; i < ys.size(); ++i) {
        const double r = dhn(xs[i], p) - ys[i];
        e += r * r;
    }
    return e;
}

// lower Cholesky factor of the lower triangle of a; false below the pivot floor
bool cholesky(int n, std::vector<double> &lo, const std::vector<double> &a) {
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < i; ++j) {
            double s = 0;
            for (int k = 0; k < j; ++k) s += lo[i * n + k] * lo[j * n + k];
            lo[i * n + j] = (a[i * n + j] - s) / lo[j * n + j];
        }
        double s = 0;
        for (int k = 0; k < i; ++k) s += lo[i * n + k] * lo[i * n + k];
        s = a[i * n + i] - s;
        if (s < 1e-30) return false;
        lo[i * n + i] = std::sqrt(s);
    }
    return true;
}

void chol_solve(int n, const std::vector<double> &lo, std::vector<double> &x, const std::vector<double> &b) {
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < i; ++j) s += lo[i * n + j] * x[j];
        x[i] = (b[i] - s) / lo[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = 0;
        for (int j = i + 1; j < n; ++j) s += lo[j * n + i] * x[j];
        x[i] = (x[i] - s) / lo[i * n + i];
    }
}

std::vector<double> levmarq_dhn(std::vector<double> p, const std::vector<double> &xs,
                                const std::vector<double> &ys) {
    const int np = (int)p.size(), maxiter = 10000;
    const double up = 10, down = 1.0 / 10, target = 1e-8;
    double lambda = 0.1, err = sq_error(p, xs, ys), newerr = 0, derr = 0;
    std::vector<double> h((size_t)np * np, 0.0), lo((size_t)np * np, 0.0), g(np), d(np), delta(np),
        newpar(np, 0.0);
    for (int it = 0; it < maxiter; ++it) {
        for (int i = 0; i < np; ++i) {
            d[i] = 0;
            for (int j = 0; j <= i; ++j) h[i * np + j] = 0;
        }
        for (size_t ix = 0; ix < xs.size(); ++ix) {
            dhn_grad(g, xs[ix], p);
            for (int i = 0; i < np; ++i) {
                d[i] += (ys[ix] - dhn(xs[ix], p)) * g[i] * 1.0;
                for (int j = 0; j <= i; ++j) h[i * np + j] += g[i] * g[j] * 1.0;
            }
        }
        double mult = 1 + lambda;
        bool ill = true;
        while (ill && it < maxiter) {
            for (int i = 0; i < np; ++i) h[i * np + i] = h[i * np + i] * mult;
            ill = !cholesky(np, lo, h);
            if (!ill) {
                chol_solve(np, lo, delta, d);
                for (int i = 0; i < np; ++i) newpar[i] = p[i] + delta[i];
                newerr = sq_error(newpar, xs, ys);
                derr = newerr - err;
                ill = derr > 0;
            }
            if (ill) {
                mult = (1 + lambda * up) / (1 + lambda);
                lambda *= up;
                ++it;
            }
        }
        p = newpar;
        err = newerr;
        lambda *= down;
        if (!ill && -derr < target) break;
    }
    return p;
}

}  // namespace

std::vector<double> interpolate_dielec(int option, const std::vector<double> &om_in,
                                       const std::vector<double> &df_in,
                                       const std::vector<double> &om_target) {
    if (om_in.size() != df_in.size())
        LRA_THROW("dielectric function: as many values as frequencies are needed");
    switch (option) {
        case 0:
            if (df_in.size() != om_target.size())
                LRA_THROW("option_dielect_func 0 needs the dielectric function at the nfreq grid frequencies");
            return df_in;
        case 1:
            return cubic_spline(om_in, df_in, om_target);
        case 2: {
            if (df_in.empty()) LRA_THROW("dielectric function: no values to fit");
            std::vector<double> p(8, 1.0);
            p[0] = p[4] = df_in[0];
            p = levmarq_dhn(p, om_in, df_in);
            std::vector<double> out;
            out.reserve(om_target.size());
            for (const double u : om_target) out.push_back(dhn(u, p));
            return out;
        }
        default:
            LRA_THROW("option_dielect_func " + std::to_string(option) +
                      ": 0 (values on the grid), 1 (cubic spline) and 2 (Havriliak-Negami fit) are in this "
                      "build; 3 (the head-and-wing model) is not");
    }
}

}  // namespace lra
