// Streaming and small kernels of the chi0 -> Ec path (gfx950).
#include "kernels.h"

#include <cmath>
#include <map>
#include <tuple>

#include "gemm.h"

namespace lra {

// ------------------------------------------------------------------------------------
// G build helpers
// ------------------------------------------------------------------------------------
__global__ void scale_bands_kernel(const double *__restrict__ wre, const double *__restrict__ wim,
                                   const double *__restrict__ eps, const double *__restrict__ wg,
                                   double *__restrict__ Zs, int nk, int nbands, int nbasis, int b0, int b1,
                                   double tau, double efermi, int nspin, int nbands_g) {
    const long long per_k = (long long)(b1 - b0) * nbasis;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (idx >= per_k || k >= nk) return;
    const int b = b0 + (int)(idx / nbasis);
    const int j = (int)(idx % nbasis);
    const double e = eps[(long long)k * nbands + b];
    const double g = wg[(long long)k * nbands + b];
    double w;
    if (tau > 0) {
        w = 1.0 / nk - g / 2 * nspin;
        if (w < 0) w = 0;
    } else {
        w = g * 0.5 * nspin;
    }
    double sc = -tau * (e - efermi);
    if (sc > 0) sc = 0;
    double sb = exp(sc) * w;
    if (nbands_g >= 0 && b >= nbands_g) sb = 0;
    const long long src = ((long long)k * nbands + b) * nbasis + j;
    double *z = Zs + (long long)k * 2 * nbands * nbasis;
    z[(2ll * b) * nbasis + j] = sb * wre[src];
    z[(2ll * b + 1) * nbasis + j] = sb * wim[src];
}

void launch_scale_bands(const double *wre, const double *wim, const double *eps, const double *wg, double *Zs,
                        int nk, int nbands, int nbasis, int b0, int b1, double tau, double efermi, int nspin,
                        int nbands_g, hipStream_t s) {
    if (b1 <= b0) return;
    const long long per_k = (long long)(b1 - b0) * nbasis;
    dim3 grid(ceil_div(per_k, 256), nk);
    hipLaunchKernelGGL(scale_bands_kernel, grid, dim3(256), 0, s, wre, wim, eps, wg, Zs, nk, nbands, nbasis, b0,
                       b1, tau, efermi, nspin, nbands_g);
    HIP_CHECK(hipGetLastError());
}

template <class T>
__global__ void k2r_kernel(const double *__restrict__ M, const double *__restrict__ cs, const double *__restrict__ sn,
                           const int *__restrict__ gcol, T *__restrict__ G, int nk, int nR, int nbasis, double sgn) {
    const long long nn = (long long)nbasis * nbasis;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nn) return;
    const int row = (int)(idx / nbasis), col = (int)(idx % nbasis);
    const long long kstride = 2 * nn;
    const long long dst = (long long)row * 2 * nbasis + gcol[col];
    for (int r = 0; r < nR; ++r) {
        double acc = 0;
        for (int k = 0; k < nk; ++k) {
            const double re = M[k * kstride + idx];
            const double im = M[k * kstride + nn + idx];
            acc += cs[r * nk + k] * re + sn[r * nk + k] * im;
        }
        G[(long long)r * 2 * nn + dst] = static_cast<T>(sgn * acc);
    }
}

template <class T>
void launch_k2r(const double *Mstack, const double *cs, const double *sn, const int *gcol, T *Gpm, int nk, int nR,
                int nbasis, double sgn, hipStream_t s) {
    const long long nn = (long long)nbasis * nbasis;
    hipLaunchKernelGGL((k2r_kernel<T>), dim3(ceil_div(nn, 256)), dim3(256), 0, s, Mstack, cs, sn, gcol, Gpm, nk,
                       nR, nbasis, sgn);
    HIP_CHECK(hipGetLastError());
}
template void launch_k2r<double>(const double *, const double *, const double *, const int *, double *, int, int,
                                 int, double, hipStream_t);
template void launch_k2r<float>(const double *, const double *, const double *, const int *, float *, int, int,
                                int, double, hipStream_t);

template <class T>
__global__ void gather_cols_kernel(const T *__restrict__ src, long long src_ld, T *__restrict__ dst,
                                   const GatherJob *__restrict__ jobs, const long long *__restrict__ colsrc,
                                   int nrows) {
    const GatherJob J = jobs[blockIdx.y];
    const long long n = (long long)nrows * J.ncol;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const int row = (int)(idx / J.ncol), c = (int)(idx % J.ncol);
        dst[J.dst_off + idx] = src[colsrc[J.col_off + c] + row * src_ld];
    }
}

template <class T>
void launch_gather_cols(const T *src, long long src_ld, T *dst, const GatherJob *jobs, int njobs,
                        const long long *colsrc, int nrows, int maxcol, hipStream_t s) {
    if (njobs == 0) return;
    const long long n = (long long)nrows * maxcol;
    const int bx = std::min(ceil_div(n, 256), 2048);
    hipLaunchKernelGGL((gather_cols_kernel<T>), dim3(bx, njobs), dim3(256), 0, s, src, src_ld, dst, jobs, colsrc,
                       nrows);
    HIP_CHECK(hipGetLastError());
}
template void launch_gather_cols<double>(const double *, long long, double *, const GatherJob *, int,
                                         const long long *, int, int, hipStream_t);
template void launch_gather_cols<float>(const float *, long long, float *, const GatherJob *, int,
                                        const long long *, int, int, hipStream_t);

// ------------------------------------------------------------------------------------
// chi0 (R, tau) -> dense chi0 (q, omega)
// ------------------------------------------------------------------------------------
// NW accumulators per thread (a multiple of 8 covering the chunk): one pass over chi0(R, tau) per
// chunk of up to kChi0AssembleMaxW frequencies
template <class TS, class TD, int NW>
__global__ void chi0_assemble_kernel(const TS *__restrict__ chi0s, long long E_tot, int ntau,
                                     const PairDesc *__restrict__ pairs, int npairs, long long nelem,
                                     const long long *__restrict__ eoff, int nR, const double *__restrict__ gp,
                                     int nw, TD *__restrict__ dense, long long dstride, int n, int embed) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nelem) return;
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pairs[mid].e0 <= idx) lo = mid; else hi = mid - 1;
    }
    const PairDesc P = pairs[lo];
    const long long e = idx - P.e0;
    const int mu = (int)(e / P.naJ), nu = (int)(e % P.naJ);
    double re[NW], im[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) { re[w] = 0; im[w] = 0; }
    for (int r = 0; r < nR; ++r) {
        const long long off = eoff[(long long)lo * nR + r];
        if (off < 0) continue;
        for (int t = 0; t < ntau; ++t) {
            const double v = static_cast<double>(chi0s[(long long)t * E_tot + off + e]);
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                if (w < nw) {
                    const double *g = gp + (((long long)w * nR + r) * ntau + t) * 2;
                    re[w] += g[0] * v;
                    im[w] += g[1] * v;
                }
            }
        }
    }
    const long long row = P.rowI + mu, col = P.colJ + nu;
    const bool mirror = P.I != P.J;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w >= nw) break;
        TD *D = dense + w * dstride;
        if (!embed) {
            const long long ld = n;
            D[row * ld + col] = static_cast<TD>(re[w]);
            if (mirror) D[col * ld + row] = static_cast<TD>(re[w]);
        } else {
            const long long ld = 2ll * n;
            D[row * ld + col] = static_cast<TD>(re[w]);
            D[(n + row) * ld + n + col] = static_cast<TD>(re[w]);
            D[(n + row) * ld + col] = static_cast<TD>(im[w]);
            D[row * ld + n + col] = static_cast<TD>(-im[w]);
            if (mirror) {
                D[col * ld + row] = static_cast<TD>(re[w]);
                D[(n + col) * ld + n + row] = static_cast<TD>(re[w]);
                D[(n + col) * ld + row] = static_cast<TD>(-im[w]);
                D[col * ld + n + row] = static_cast<TD>(im[w]);
            }
        }
    }
}

template <class TS, class TD>
void launch_chi0_assemble(const TS *chi0s, long long E_tot, int ntau, const PairDesc *pairs, int npairs,
                          long long n_pair_elems, const long long *eoff, int nR, const double *gam_ph, int nw,
                          TD *dense, long long dense_stride, int n, bool embed, hipStream_t s) {
    if (nw < 1 || nw > kChi0AssembleMaxW) LRA_THROW("omega chunk too large");
    hipLaunchKernelGGL((chi0_assemble_kernel<TS, TD, kChi0AssembleMaxW>), dim3(ceil_div(n_pair_elems, 256)), dim3(256),
                       0, s, chi0s, E_tot, ntau, pairs, npairs, n_pair_elems, eoff, nR, gam_ph, nw, dense, dense_stride,
                       n, (int)embed);
    HIP_CHECK(hipGetLastError());
}
template void launch_chi0_assemble<double, double>(const double *, long long, int, const PairDesc *, int, long long,
                                                   const long long *, int, const double *, int, double *, long long,
                                                   int, bool, hipStream_t);
template void launch_chi0_assemble<float, float>(const float *, long long, int, const PairDesc *, int, long long,
                                                 const long long *, int, const double *, int, float *, long long,
                                                 int, bool, hipStream_t);
template void launch_chi0_assemble<float, double>(const float *, long long, int, const PairDesc *, int, long long,
                                                  const long long *, int, const double *, int, double *, long long,
                                                  int, bool, hipStream_t);
template void launch_chi0_assemble<double, float>(const double *, long long, int, const PairDesc *, int, long long,
                                                  const long long *, int, const double *, int, float *, long long,
                                                  int, bool, hipStream_t);

// chiw[w][i] = sum_t coef[w][t] chi0s[t][i] for every element i of the chi0(R, tau) store: the
// cosine tau -> omega transform, done once for all q.  Coefficients in LDS; each thread owns one
// element and re-reads its ntau values once per group of 8 frequencies.
template <class TS>
__global__ __launch_bounds__(256) void tau_to_omega_kernel(const TS *__restrict__ chi0s, long long E_tot, int ntau,
                                                           const double *__restrict__ coef, int nw,
                                                           double *__restrict__ chiw) {
    __shared__ double c[kTauOmegaMax * kTauOmegaMax];
    for (int k = threadIdx.x; k < nw * ntau; k += blockDim.x) c[k] = coef[k];
    __syncthreads();
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < E_tot;
         i += (long long)gridDim.x * blockDim.x) {
        for (int w0 = 0; w0 < nw; w0 += 8) {
            double acc[8];
#pragma unroll
            for (int w = 0; w < 8; ++w) acc[w] = 0.0;
            for (int t = 0; t < ntau; ++t) {
                const double v = static_cast<double>(chi0s[(long long)t * E_tot + i]);
#pragma unroll
                for (int w = 0; w < 8; ++w)
                    if (w0 + w < nw) acc[w] += c[(w0 + w) * ntau + t] * v;
            }
#pragma unroll
            for (int w = 0; w < 8; ++w)
                if (w0 + w < nw) chiw[(long long)(w0 + w) * E_tot + i] = acc[w];
        }
    }
}

void launch_tau_to_omega(const double *chi0s, long long E_tot, int ntau, const double *coef, int nw, double *chiw,
                         hipStream_t s) {
    if (ntau > kTauOmegaMax || nw > kTauOmegaMax) LRA_THROW("tau -> omega: grid too large");
    const int bx = (int)std::min<long long>(ceil_div(E_tot, 256), 8192);
    hipLaunchKernelGGL(tau_to_omega_kernel<double>, dim3(bx), dim3(256), 0, s, chi0s, E_tot, ntau, coef, nw, chiw);
    HIP_CHECK(hipGetLastError());
}
void launch_tau_to_omega(const float *chi0s, long long E_tot, int ntau, const double *coef, int nw, double *chiw,
                         hipStream_t s) {
    if (ntau > kTauOmegaMax || nw > kTauOmegaMax) LRA_THROW("tau -> omega: grid too large");
    const int bx = (int)std::min<long long>(ceil_div(E_tot, 256), 8192);
    hipLaunchKernelGGL(tau_to_omega_kernel<float>, dim3(bx), dim3(256), 0, s, chi0s, E_tot, ntau, coef, nw, chiw);
    HIP_CHECK(hipGetLastError());
}

// chiw[w][i] (+)= coef[w][t] slice[i] for one tau slice t (the first slice overwrites): the same
// sums as tau_to_omega_kernel, accumulated slice by slice in the same order of t
template <class TS>
__global__ __launch_bounds__(256) void tau_accumulate_kernel(const TS *__restrict__ slice, long long E_tot, int t,
                                                             int ntau, const double *__restrict__ coef, int nw,
                                                             double *__restrict__ chiw) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < E_tot;
         i += (long long)gridDim.x * blockDim.x) {
        const double v = static_cast<double>(slice[i]);
        for (int w = 0; w < nw; ++w) {
            double *d = chiw + (long long)w * E_tot + i;
            const double base = t == 0 ? 0.0 : *d;
            *d = base + coef[w * ntau + t] * v;
        }
    }
}

template <class TS>
void launch_tau_accumulate(const TS *slice, long long E_tot, int t, int ntau, const double *coef, int nw, double *chiw,
                           hipStream_t s) {
    const int bx = (int)std::min<long long>(ceil_div(E_tot, 256), 16384);
    hipLaunchKernelGGL(tau_accumulate_kernel<TS>, dim3(bx), dim3(256), 0, s, slice, E_tot, t, ntau, coef, nw, chiw);
    HIP_CHECK(hipGetLastError());
}
template void launch_tau_accumulate<double>(const double *, long long, int, int, const double *, int, double *,
                                            hipStream_t);
template void launch_tau_accumulate<float>(const float *, long long, int, int, const double *, int, double *,
                                           hipStream_t);

// dense chi0(q, omega) of nw frequencies from chi0(R, omega): per pair element and frequency,
// sum_r (cos_r + i sin_r) chiw[w][eoff[p][r] + e], written like chi0_assemble (Hermitian fill,
// real or embedded).  Reads chiw once per (element, R, frequency).
template <class TD>
__global__ __launch_bounds__(256) void chi0_rq_kernel(const double *__restrict__ chiw, long long E_tot,
                                                      const PairDesc *__restrict__ pairs, int npairs, long long nelem,
                                                      const long long *__restrict__ eoff, int nR,
                                                      const double *__restrict__ ph, int nw, TD *__restrict__ dense,
                                                      long long dstride, int n, int embed) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nelem) return;
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pairs[mid].e0 <= idx) lo = mid; else hi = mid - 1;
    }
    const PairDesc P = pairs[lo];
    const long long e = idx - P.e0;
    const int mu = (int)(e / P.naJ), nu = (int)(e % P.naJ);
    const long long row = P.rowI + mu, col = P.colJ + nu;
    const bool mirror = P.I != P.J;
    for (int w = 0; w < nw; ++w) {
        double re = 0, im = 0;
        const double *cw = chiw + (long long)w * E_tot + e;
        for (int r = 0; r < nR; ++r) {
            const long long off = eoff[(long long)lo * nR + r];
            if (off < 0) continue;
            const double v = cw[off];
            re += ph[2 * r] * v;
            im += ph[2 * r + 1] * v;
        }
        TD *D = dense + w * dstride;
        if (!embed) {
            D[row * n + col] = static_cast<TD>(re);
            if (mirror) D[col * n + row] = static_cast<TD>(re);
        } else {
            const long long ld = 2ll * n;
            D[row * ld + col] = static_cast<TD>(re);
            D[(n + row) * ld + n + col] = static_cast<TD>(re);
            D[(n + row) * ld + col] = static_cast<TD>(im);
            D[row * ld + n + col] = static_cast<TD>(-im);
            if (mirror) {
                D[col * ld + row] = static_cast<TD>(re);
                D[(n + col) * ld + n + row] = static_cast<TD>(re);
                D[(n + col) * ld + row] = static_cast<TD>(-im);
                D[col * ld + n + row] = static_cast<TD>(im);
            }
        }
    }
}

// The same for up to kChi0RqMaxQ q points at once: each chiw value read feeds every q of the launch
// (phases ph[q][r] as (cos, sin)); q point qi writes its nw matrices at dense + qi * qstride.
template <class TD>
__global__ __launch_bounds__(256) void chi0_rq_multi_kernel(const double *__restrict__ chiw, long long E_tot,
                                                            const PairDesc *__restrict__ pairs, int npairs,
                                                            long long nelem, const long long *__restrict__ eoff,
                                                            int nR, const double *__restrict__ ph, int nq, int nw,
                                                            TD *__restrict__ dense, long long dstride,
                                                            long long qstride, int n, int embed) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nelem) return;
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pairs[mid].e0 <= idx) lo = mid; else hi = mid - 1;
    }
    const PairDesc P = pairs[lo];
    const long long e = idx - P.e0;
    const int mu = (int)(e / P.naJ), nu = (int)(e % P.naJ);
    const long long row = P.rowI + mu, col = P.colJ + nu;
    const bool mirror = P.I != P.J;
    for (int w = 0; w < nw; ++w) {
        double re[kChi0RqMaxQ], im[kChi0RqMaxQ];
#pragma unroll
        for (int q = 0; q < kChi0RqMaxQ; ++q) re[q] = im[q] = 0.0;
        const double *cw = chiw + (long long)w * E_tot + e;
        for (int r = 0; r < nR; ++r) {
            const long long off = eoff[(long long)lo * nR + r];
            if (off < 0) continue;
            const double v = cw[off];
#pragma unroll
            for (int q = 0; q < kChi0RqMaxQ; ++q) {
                if (q < nq) {
                    re[q] += ph[2 * ((long long)q * nR + r)] * v;
                    im[q] += ph[2 * ((long long)q * nR + r) + 1] * v;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kChi0RqMaxQ; ++q) {
            if (q >= nq) break;
            TD *D = dense + q * qstride + w * dstride;
            if (!embed) {
                D[row * n + col] = static_cast<TD>(re[q]);
                if (mirror) D[col * n + row] = static_cast<TD>(re[q]);
            } else {
                const long long ld = 2ll * n;
                D[row * ld + col] = static_cast<TD>(re[q]);
                D[(n + row) * ld + n + col] = static_cast<TD>(re[q]);
                D[(n + row) * ld + col] = static_cast<TD>(im[q]);
                D[row * ld + n + col] = static_cast<TD>(-im[q]);
                if (mirror) {
                    D[col * ld + row] = static_cast<TD>(re[q]);
                    D[(n + col) * ld + n + row] = static_cast<TD>(re[q]);
                    D[(n + col) * ld + row] = static_cast<TD>(-im[q]);
                    D[col * ld + n + row] = static_cast<TD>(im[q]);
                }
            }
        }
    }
}

template <class TD>
void launch_chi0_rq_multi(const double *chiw, long long E_tot, const PairDesc *pairs, int npairs,
                          long long n_pair_elems, const long long *eoff, int nR, const double *ph, int nq, int nw,
                          TD *dense, long long dense_stride, long long q_stride, int n, bool embed, hipStream_t s) {
    if (nq < 1 || nq > kChi0RqMaxQ) LRA_THROW("chi0_rq_multi: 1 to kChi0RqMaxQ q points per launch");
    hipLaunchKernelGGL((chi0_rq_multi_kernel<TD>), dim3(ceil_div(n_pair_elems, 256)), dim3(256), 0, s, chiw, E_tot,
                       pairs, npairs, n_pair_elems, eoff, nR, ph, nq, nw, dense, dense_stride, q_stride, n, (int)embed);
    HIP_CHECK(hipGetLastError());
}
template void launch_chi0_rq_multi<double>(const double *, long long, const PairDesc *, int, long long,
                                           const long long *, int, const double *, int, int, double *, long long,
                                           long long, int, bool, hipStream_t);
template void launch_chi0_rq_multi<float>(const double *, long long, const PairDesc *, int, long long,
                                          const long long *, int, const double *, int, int, float *, long long,
                                          long long, int, bool, hipStream_t);

template <class TD>
void launch_chi0_rq(const double *chiw, long long E_tot, const PairDesc *pairs, int npairs, long long n_pair_elems,
                    const long long *eoff, int nR, const double *ph, int nw, TD *dense, long long dense_stride, int n,
                    bool embed, hipStream_t s) {
    hipLaunchKernelGGL((chi0_rq_kernel<TD>), dim3(ceil_div(n_pair_elems, 256)), dim3(256), 0, s, chiw, E_tot, pairs,
                       npairs, n_pair_elems, eoff, nR, ph, nw, dense, dense_stride, n, (int)embed);
    HIP_CHECK(hipGetLastError());
}
template void launch_chi0_rq<double>(const double *, long long, const PairDesc *, int, long long, const long long *,
                                     int, const double *, int, double *, long long, int, bool, hipStream_t);
template void launch_chi0_rq<float>(const double *, long long, const PairDesc *, int, long long, const long long *, int,
                                    const double *, int, float *, long long, int, bool, hipStream_t);

// ------------------------------------------------------------------------------------
// dense helpers
// ------------------------------------------------------------------------------------
template <class T>
__global__ void set_identity_kernel(T *A, int n, int lda, const double *__restrict__ diag) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    A[(long long)i * lda + j] = (i == j) ? (diag ? static_cast<T>(diag[i]) : T(1)) : T(0);
}
template <class T> void launch_set_identity(T *A, int n, int lda, hipStream_t s, const double *diag) {
    hipLaunchKernelGGL((set_identity_kernel<T>), dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, s, A, n, lda,
                       diag);
    HIP_CHECK(hipGetLastError());
}
template void launch_set_identity<double>(double *, int, int, hipStream_t, const double *);
template void launch_set_identity<float>(float *, int, int, hipStream_t, const double *);

template <class T>
__global__ void zero_upper_kernel(T *A, int n, int lda) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    if (j > i) A[(long long)i * lda + j] = T(0);
}
template <class T> void launch_zero_upper(T *A, int n, int lda, hipStream_t s) {
    hipLaunchKernelGGL((zero_upper_kernel<T>), dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, s, A, n, lda);
    HIP_CHECK(hipGetLastError());
}
template void launch_zero_upper<double>(double *, int, int, hipStream_t);
template void launch_zero_upper<float>(float *, int, int, hipStream_t);

template <class TS, class TD>
__global__ void convert_kernel(const TS *__restrict__ src, TD *__restrict__ dst, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = static_cast<TD>(src[i]);
}
template <class TS, class TD> void launch_convert(const TS *src, TD *dst, long long n, hipStream_t s) {
    const int bx = std::min(ceil_div(n, 256), 65536);
    hipLaunchKernelGGL((convert_kernel<TS, TD>), dim3(bx), dim3(256), 0, s, src, dst, n);
    HIP_CHECK(hipGetLastError());
}
template void launch_convert<double, float>(const double *, float *, long long, hipStream_t);
template void launch_convert<float, double>(const float *, double *, long long, hipStream_t);
template void launch_convert<double, double>(const double *, double *, long long, hipStream_t);

__global__ void copy_lower_f2d_kernel(const float *__restrict__ src, double *__restrict__ dst, int n, int lda) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    if (j <= i) dst[(long long)i * lda + j] = static_cast<double>(src[(long long)i * lda + j]);
}
void launch_copy_lower_f2d(const float *src, double *dst, int n, int lda, hipStream_t s) {
    hipLaunchKernelGGL(copy_lower_f2d_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, s, src, dst, n,
                       lda);
    HIP_CHECK(hipGetLastError());
}

__global__ void v_scatter_kernel(const double2 *__restrict__ blk, const VBlock *__restrict__ desc, double *V,
                                 long long n, long long n2, int embed, int drop_imag) {
    const VBlock b = desc[blockIdx.x];
    const int tot = b.na * b.nb;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int mu = e / b.nb, nu = e % b.nb;
        double2 v = blk[b.off + e];
        if (drop_imag) v.y = 0.0;
        const long long r = b.r0 + mu, c = b.c0 + nu;
        V[r * n2 + c] = v.x;
        if (!b.diag) V[c * n2 + r] = v.x;
        if (embed) {
            V[(n + r) * n2 + n + c] = v.x;
            V[(n + r) * n2 + c] = v.y;
            V[r * n2 + n + c] = -v.y;
            if (!b.diag) {
                V[(n + c) * n2 + n + r] = v.x;
                V[(n + c) * n2 + r] = -v.y;
                V[c * n2 + n + r] = v.y;
            }
        }
    }
}

void launch_v_scatter(const double *blk, const VBlock *d, int nblk, double *V, int n, int n2, int embed,
                      int drop_imag, hipStream_t s) {
    if (nblk == 0) return;
    hipLaunchKernelGGL(v_scatter_kernel, dim3(nblk), dim3(256), 0, s, reinterpret_cast<const double2 *>(blk), d, V,
                       (long long)n, (long long)n2, embed, drop_imag);
    HIP_CHECK(hipGetLastError());
}

__global__ void trace_kernel(const double *__restrict__ A, int n, int lda, double *acc,
                             const double *__restrict__ signs) {
    __shared__ double red[256];
    double v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += (signs ? signs[i] : 1.0) * A[(long long)i * lda + i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += red[0];
}
void launch_trace_acc(const double *A, int n, int lda, double *acc, hipStream_t s, const double *signs) {
    hipLaunchKernelGGL(trace_kernel, dim3(1), dim3(256), 0, s, A, n, lda, acc, signs);
    HIP_CHECK(hipGetLastError());
}

__global__ void copy_block_kernel(const double *__restrict__ src, int lds, double *__restrict__ dst, int ldd,
                                  int rows, int cols) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)rows * cols) return;
    const int i = (int)(idx / cols), j = (int)(idx % cols);
    dst[(long long)i * ldd + j] = src[(long long)i * lds + j];
}
void launch_copy_block(const double *src, int lds, double *dst, int ldd, int rows, int cols, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL(copy_block_kernel, dim3(ceil_div((long long)rows * cols, 256)), dim3(256), 0, s, src, lds,
                       dst, ldd, rows, cols);
    HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------------------------
// Blocked Cholesky with log-determinant
// ------------------------------------------------------------------------------------
constexpr int CHOL_NB = 64;

// A double of another lane (lane is wave-uniform).
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u & 0xffffffffu), lane));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), lane));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Factor the nb x nb diagonal block at (j0, j0) of matrix blockIdx.x as L S L^T with
// S = diag(+-1) (a Cholesky factorisation that tolerates negative pivots), write L11 back, write
// Linv = L11^{-1} (row-major, CHOL_NB pitch) and the pivot signs, and accumulate
// sum log|d_j| = ln |det| of the block.  A zero or NaN pivot sets *flag.
//
// One wave per matrix; lane t keeps row t of the unit-padded 64 x 64 block in registers (the
// loops are unrolled so every index is static).  At step j every lane takes the pivot from lane j
// with readlane and forms s_j and sqrt|d_j| itself, lane t > j scales its entry, and each lane
// updates its row with the scaled column read lane by lane.  The inverse is then formed with lane c
// holding column c, again from readlane of the rows.  (Rows in LDS with lane-local loops took
// three times as long: every step then waits on LDS round trips.)
__global__ __launch_bounds__(64) void potrf_diag_kernel(double *A0, long long strideA, int lda, int j0, int nb,
                                                        double *work, double *logdet0, long long ld_stride,
                                                        int *flag, double *signs, long long sstride) {
    __shared__ double L[CHOL_NB][CHOL_NB + 1];
    const int t = threadIdx.x;
    double *A = A0 + strideA * blockIdx.x;
    double *Linv = work + (long long)blockIdx.x * (CHOL_NB * CHOL_NB + CHOL_NB);
    double *sgn = Linv + CHOL_NB * CHOL_NB;
    // coalesced row loads through LDS into row t of lane t
    for (int r = 0; r < CHOL_NB; ++r)
        L[r][t] = (r < nb && t < nb) ? A[(long long)(j0 + r) * lda + j0 + t] : (r == t ? 1.0 : 0.0);
    __syncthreads();
    double a[CHOL_NB];
#pragma unroll
    for (int k = 0; k < CHOL_NB; ++k) a[k] = L[t][k];
    bool bad = false;
    double my_s = 1.0, my_r = 1.0;
#pragma unroll
    for (int j = 0; j < CHOL_NB; ++j) {
        double d = readlane_f64(a[j], j);
        if (!(fabs(d) > 0.0)) { bad = true; d = 1.0; }
        const double s = d > 0 ? 1.0 : -1.0;
        const double r = sqrt(fabs(d));
        // L_tj = a_tj / (s_j sqrt|d_j|), so that A = L S L^T also for a negative pivot; the update
        // a_tk -= s_j L_tj L_kj then equals a_tk - a_tj a_kj / d_j
        const double l = a[j] * (s / r);
        if (t == j) { a[j] = r; my_s = s; my_r = r; }
        if (t > j) a[j] = l;
        const double f = t > j ? s * l : 0.0;
        // entries k > t (upper triangle) are updated too and never used: no lane mask needed
#pragma unroll
        for (int k = j + 1; k < CHOL_NB; ++k) a[k] -= f * readlane_f64(a[j], k);
    }
    // L back through LDS as coalesced rows (lower triangle)
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CHOL_NB; ++k) L[t][k] = a[k];
    __syncthreads();
    for (int r = 0; r < nb; ++r)
        if (t <= r) A[(long long)(j0 + r) * lda + j0 + t] = L[r][t];
    // column t of L^{-1}: x_r = (delta_rt - sum_{k < r} L_rk x_k) / L_rr (x_k = 0 for k < t)
    double x[CHOL_NB];
    const double dinv = 1.0 / my_r;
#pragma unroll
    for (int r = 0; r < CHOL_NB; ++r) {
        double acc = r == t ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < r; ++k) acc -= readlane_f64(a[k], r) * x[k];
        x[r] = acc * readlane_f64(dinv, r);
    }
#pragma unroll
    for (int r = 0; r < CHOL_NB; ++r) Linv[r * CHOL_NB + t] = (r < nb && t < nb) ? x[r] : 0.0;
    sgn[t] = t < nb ? my_s : 1.0;
    if (signs != nullptr && t < nb) signs[sstride * blockIdx.x + j0 + t] = my_s;
    double v = t < nb ? 2.0 * log(my_r) : 0.0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (t == 0) {
        logdet0[ld_stride * blockIdx.x] += v;
        if (bad) *flag = 1;
    }
}

// Qs[r][c] = Q[r][c] * sgn[c] for matrix blockIdx.y of the batch; Q and Qs are m x nb regions of
// the given pitch inside that matrix's pair of wide panels
constexpr int CHOL_PW = 8 * CHOL_NB;   // middle panel: its update of the outer panel runs with K = 512
constexpr int CHOL_PO = 4 * CHOL_PW;   // outer panel of large matrices: the trailing update has K = 2048

__global__ void scale_cols_kernel(double *__restrict__ Q, double *__restrict__ Qs, long long pstride, int pitch,
                                  const double *__restrict__ work, int m, int nb) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)m * nb) return;
    const long long r = idx / nb, c = idx % nb;
    const double *sgn = work + (long long)blockIdx.y * (CHOL_NB * CHOL_NB + CHOL_NB) + CHOL_NB * CHOL_NB;
    const long long o = (long long)blockIdx.y * pstride + r * pitch + c;
    Qs[o] = Q[o] * sgn[c];
}

struct CholPlan {
    std::vector<int> j0s, nbs;
    std::vector<GemmBatch> trsm, inpanel;  // per 64-column step
    std::vector<GemmBatch> mid;            // per step: the K = CHOL_PW update of the outer panel's other columns
    std::vector<GemmBatch> trail;          // per step: the trailing update after an outer panel
};

static std::map<std::tuple<int, int, int, long long, hipStream_t, int>, CholPlan> &chol_plans() {
    static std::map<std::tuple<int, int, int, long long, hipStream_t, int>, CholPlan> p;
    return p;
}

// Right-looking blocked L S L^T factorisation in three levels.  64-column steps (diagonal block
// by potrf_diag_kernel, Q = A21 L11^-T by GEMM, signs applied) update only the rest of their
// 512-column middle panel (K = 64); after each middle panel one GEMM with K = 512 updates the
// remaining columns of its outer panel; after each outer panel one lower-triangular GEMM updates
// the trailing matrix with K = the outer panel width (2048 for n >= 4096, else 512, when the
// middle level is the outer one).  Few, deep trailing updates read and write the trailing matrix
// four times less often than 512-wide ones.  Q and Q S of the outer panel's columns are kept in
// two wide buffers per matrix (pitch = outer panel width, indexed by global row).
void cholesky_logdet_batched(double *A, long long strideA, int nbatch, int n, int lda, double *logdet,
                             long long ld_stride, int *flag, bool keep_L, CholWork &w, hipStream_t s, int ncols,
                             double *signs, long long sstride) {
    const int nf = (ncols > 0 && ncols < n) ? ncols : n;
    const int PO = n >= 4096 ? CHOL_PO : CHOL_PW;
    const long long wstride = CHOL_NB * CHOL_NB + CHOL_NB;
    const long long pstride = 2LL * n * PO;  // Q then Q S, each n x PO
    w.linv.reserve(sizeof(double) * wstride * nbatch);
    w.panel.reserve(sizeof(double) * pstride * nbatch);
    const long long qs_off = (long long)n * PO;
    auto key = std::make_tuple(n, lda, nbatch, strideA, s, nf);
    auto &plans = chol_plans();
    auto it = plans.find(key);
    if (it == plans.end()) {
        CholPlan cp;
        for (int o0 = 0; o0 < nf; o0 += PO) {
            const int oend = std::min(o0 + PO, nf);
            for (int p0 = o0; p0 < oend; p0 += CHOL_PW) {
                const int pend = std::min(p0 + CHOL_PW, oend);
                for (int j0 = p0; j0 < pend; j0 += CHOL_NB) {
                    const int nb = std::min(CHOL_NB, pend - j0);
                    const int m = n - j0 - nb;
                    const int wcols = pend - j0 - nb;  // middle panel columns right of this step
                    cp.j0s.push_back(j0);
                    cp.nbs.push_back(nb);
                    GemmBatch tr, ip, md, tl;
                    if (m > 0) {
                        tr.trans_b = true;
                        ip.trans_b = true;
                        for (int b = 0; b < nbatch; ++b) {
                            const long long wq = pstride * b + (long long)(j0 + nb) * PO + (j0 - o0);
                            // Q = A21 * Linv^T  (A21: m x nb at (j0+nb, j0)) into the wide buffer
                            tr.add(make_prob(m, nb, nb, strideA * b + (long long)(j0 + nb) * lda + j0, lda,
                                             wstride * b, CHOL_NB, wq, PO));
                            // A[j0+nb:, j0+nb:pend] -= (Q S) Q[0:wcols]^T
                            if (wcols > 0)
                                ip.add(make_prob(m, wcols, nb, wq + qs_off, PO, wq, PO,
                                                 strideA * b + (long long)(j0 + nb) * lda + (j0 + nb), lda));
                        }
                        tr.finalize(T64x64, s);
                        if (wcols > 0) ip.finalize(T64x64, s);
                    }
                    if (j0 + nb == pend && pend < oend) {
                        // the middle panel's update of the outer panel: A[pend:, pend:oend] -= (Q S) Q^T
                        // over its pend - p0 columns (rectangular; the upper part of the diagonal
                        // block is not read)
                        md.trans_b = true;
                        for (int b = 0; b < nbatch; ++b) {
                            const long long wq = pstride * b + (long long)pend * PO + (p0 - o0);
                            md.add(make_prob(n - pend, oend - pend, pend - p0, wq + qs_off, PO, wq, PO,
                                             strideA * b + (long long)pend * lda + pend, lda));
                        }
                        md.finalize(T128x128w8, s);
                    }
                    const int mt = n - oend;
                    if (j0 + nb == oend && mt > 0) {
                        // trailing update of the outer panel: A22 -= (Q S) Q^T over its oend - o0 columns
                        tl.trans_b = true;
                        tl.lower_only = true;
                        for (int b = 0; b < nbatch; ++b) {
                            const long long wq = pstride * b + (long long)oend * PO;
                            tl.add(make_prob(mt, mt, oend - o0, wq + qs_off, PO, wq, PO,
                                             strideA * b + (long long)oend * lda + oend, lda));
                        }
                        tl.finalize(T128x128w8, s);
                    }
                    cp.trsm.push_back(std::move(tr));
                    cp.inpanel.push_back(std::move(ip));
                    cp.mid.push_back(std::move(md));
                    cp.trail.push_back(std::move(tl));
                }
            }
        }
        it = plans.emplace(key, std::move(cp)).first;
    }
    CholPlan &cp = it->second;
    double *work = w.linv.as<double>();
    double *panel = w.panel.as<double>();
    for (size_t b = 0; b < cp.j0s.size(); ++b) {
        const int j0 = cp.j0s[b], nb = cp.nbs[b];
        const int o0 = (j0 / PO) * PO;
        hipLaunchKernelGGL(potrf_diag_kernel, dim3(nbatch), dim3(CHOL_NB), 0, s, A, strideA, lda, j0, nb, work,
                           logdet, ld_stride, flag, signs, sstride);
        HIP_CHECK(hipGetLastError());
        const int m = n - j0 - nb;
        if (m > 0) {
            double *wq = panel + (long long)(j0 + nb) * PO + (j0 - o0);
            gemm_run(Prec::F64, cp.trsm[b], A, work, panel, 1.0, 0.0, s);
            hipLaunchKernelGGL(scale_cols_kernel, dim3(ceil_div((long long)m * nb, 256), nbatch), dim3(256), 0, s, wq,
                               wq + qs_off, pstride, PO, work, m, nb);
            HIP_CHECK(hipGetLastError());
            if (keep_L)
                for (int bb = 0; bb < nbatch; ++bb)
                    launch_copy_block(wq + pstride * bb + qs_off, PO, A + strideA * bb + (long long)(j0 + nb) * lda + j0,
                                      lda, m, nb, s);
            gemm_run(Prec::F64, cp.inpanel[b], panel, panel, A, -1.0, 1.0, s);
        }
        gemm_run(Prec::F64, cp.mid[b], panel, panel, A, -1.0, 1.0, s);
        gemm_run(Prec::F64, cp.trail[b], panel, panel, A, -1.0, 1.0, s);
    }
    if (keep_L)
        for (int bb = 0; bb < nbatch; ++bb) launch_zero_upper<double>(A + strideA * bb, n, lda, s);
}

void cholesky_logdet(double *A, int n, int lda, double *logdet, int *flag, bool keep_L, CholWork &w, hipStream_t s,
                     double *signs) {
    cholesky_logdet_batched(A, 0, 1, n, lda, logdet, 0, flag, keep_L, w, s, -1, signs, 0);
}

}  // namespace lra
