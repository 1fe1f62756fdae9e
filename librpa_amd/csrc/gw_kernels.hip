// Streaming kernels of the G0W0 path (gfx950).
#include "gw_kernels.h"

namespace lra {

constexpr int kWMax = 8;

__global__ __launch_bounds__(256) void w_accum_kernel(const double *__restrict__ aug, long long aug_stride, int cnt,
                                                      int n, int n2, int N, int embed, const double *__restrict__ coef,
                                                      int ntau, const double *__restrict__ cs_r,
                                                      const double *__restrict__ sn_r, int nR,
                                                      double *__restrict__ Wrt) {
    const long long nn = (long long)n * n;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nn) return;
    const int a = (int)(idx / n), b = (int)(idx % n);
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    double re[kWMax], im[kWMax];
#pragma unroll
    for (int w = 0; w < kWMax; ++w) {
        re[w] = 0;
        im[w] = 0;
        if (w < cnt) {
            const double *A = aug + (long long)w * aug_stride;
            re[w] = -A[(long long)(n2 + hi) * N + n2 + lo];
            if (embed) im[w] = -A[(long long)(n2 + n + a) * N + n2 + b];
        }
    }
    for (int t = 0; t < ntau; ++t) {
        double x = 0, y = 0;
#pragma unroll
        for (int w = 0; w < kWMax; ++w) {
            if (w < cnt) {
                const double c = coef[w * ntau + t];
                x += c * re[w];
                y += c * im[w];
            }
        }
        double *dst = Wrt + (long long)t * nR * nn + idx;
        for (int R = 0; R < nR; ++R) dst[(long long)R * nn] += cs_r[R] * x + sn_r[R] * y;
    }
}

void launch_w_accum(const double *aug, long long aug_stride, int cnt, int n, int n2, int N, bool embed,
                    const double *coef, int ntau, const double *cs_r, const double *sn_r, int nR, double *Wrt,
                    hipStream_t s) {
    if (cnt < 1 || cnt > kWMax) LRA_THROW("w_accum: 1..8 frequencies per launch");
    hipLaunchKernelGGL(w_accum_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, s, aug, aug_stride, cnt,
                       n, n2, N, (int)embed, coef, ntau, cs_r, sn_r, nR, Wrt);
    HIP_CHECK(hipGetLastError());
}

template <class T>
__global__ void sigma_fold_kernel(T *__restrict__ sig, const T *__restrict__ extA, const T *__restrict__ extB,
                                  const int *__restrict__ fold, const int *__restrict__ foldneg, int NE, int nR,
                                  int n) {
    const long long nn = (long long)n * n;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nn * nR) return;
    const int R = (int)(idx / nn);
    const long long ab = idx % nn;
    const int a = (int)(ab / n), b = (int)(ab % n);
    T acc = 0;
    for (int e = 0; e < NE; ++e) {
        if (fold[e] == R) acc += extA[(long long)e * nn + ab] + extB[(long long)e * nn + ab];
        if (foldneg[e] == R) acc += extB[(long long)e * nn + (long long)b * n + a];
    }
    sig[idx] += acc;
}

template <class T>
void launch_sigma_fold(T *sig, const T *extA, const T *extB, const int *fold, const int *foldneg, int NE, int nR,
                       int n, hipStream_t s) {
    hipLaunchKernelGGL(sigma_fold_kernel<T>, dim3(ceil_div((long long)n * n * nR, 256)), dim3(256), 0, s, sig, extA,
                       extB, fold, foldneg, NE, nR, n);
    HIP_CHECK(hipGetLastError());
}
template void launch_sigma_fold<double>(double *, const double *, const double *, const int *, const int *, int, int,
                                        int, hipStream_t);
template void launch_sigma_fold<float>(float *, const float *, const float *, const int *, const int *, int, int,
                                       int, hipStream_t);

template <class T>
__global__ void sigma_freq_kernel(const T *__restrict__ sp, const T *__restrict__ sn, long long nelem,
                                  const double *__restrict__ cw, const double *__restrict__ sw, int nw,
                                  double *__restrict__ sre, double *__restrict__ sim) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nelem) return;
    const double p = static_cast<double>(sp[i]), m = static_cast<double>(sn[i]);
    const double c = 0.5 * (p + m), d = 0.5 * (p - m);
    for (int w = 0; w < nw; ++w) {
        sre[(long long)w * nelem + i] += c * cw[w];
        sim[(long long)w * nelem + i] += d * sw[w];
    }
}

template <class T>
void launch_sigma_freq(const T *sp, const T *sn, long long nelem, const double *cw, const double *sw, int nw,
                       double *sre, double *sim, hipStream_t s) {
    hipLaunchKernelGGL(sigma_freq_kernel<T>, dim3(ceil_div(nelem, 256)), dim3(256), 0, s, sp, sn, nelem, cw, sw, nw,
                       sre, sim);
    HIP_CHECK(hipGetLastError());
}
template void launch_sigma_freq<double>(const double *, const double *, long long, const double *, const double *,
                                        int, double *, double *, hipStream_t);
template void launch_sigma_freq<float>(const float *, const float *, long long, const double *, const double *, int,
                                       double *, double *, hipStream_t);

constexpr int kStMax = 16;

// one workgroup per (row a, omega); threads stride over the columns b
__global__ __launch_bounds__(256) void sigma_ks_rows_kernel(const double *__restrict__ sre,
                                                            const double *__restrict__ sim, int nR, int n,
                                                            const int *__restrict__ atom_of, int natom,
                                                            const double *__restrict__ ph,
                                                            const double *__restrict__ U, int nst,
                                                            double *__restrict__ Y) {
    const int a = blockIdx.x, w = blockIdx.y;
    const int t = threadIdx.x;
    const int I = atom_of[a];
    double ar[kStMax], ai[kStMax];
#pragma unroll
    for (int st = 0; st < kStMax; ++st) { ar[st] = 0; ai[st] = 0; }
    const long long nn = (long long)n * n;
    for (int b = t; b < n; b += blockDim.x) {
        const int J = atom_of[b];
        const double *p = ph + ((long long)I * natom + J) * nR * 2;
        double vr = 0, vi = 0;
        for (int R = 0; R < nR; ++R) {
            const long long e = ((long long)w * nR + R) * nn + (long long)a * n + b;
            const double sr = sre[e], si = sim[e];
            const double cr = p[2 * R], ci = p[2 * R + 1];
            vr += cr * sr - ci * si;
            vi += cr * si + ci * sr;
        }
#pragma unroll
        for (int st = 0; st < kStMax; ++st) {
            if (st < nst) {
                const double ur = U[((long long)st * n + b) * 2], ui = U[((long long)st * n + b) * 2 + 1];
                ar[st] += vr * ur - vi * ui;
                ai[st] += vr * ui + vi * ur;
            }
        }
    }
    __shared__ double red[4][2 * kStMax];
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int st = 0; st < kStMax; ++st) {
        double x = ar[st], y = ai[st];
        for (int o = 32; o > 0; o >>= 1) {
            x += __shfl_xor(x, o, 64);
            y += __shfl_xor(y, o, 64);
        }
        if (lane == 0) { red[wave][2 * st] = x; red[wave][2 * st + 1] = y; }
    }
    __syncthreads();
    if (t < 2 * nst) {
        const double v = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        Y[((long long)w * n + a) * nst * 2 + t] = v;
    }
}

void launch_sigma_ks_rows(const double *sre, const double *sim, int nw, int nR, int n, const int *atom_of,
                          int natom, const double *ph, const double *U, int nst, double *Y, hipStream_t s) {
    if (nst < 1 || nst > kStMax) LRA_THROW("sigma_ks_rows: 1..16 states per launch");
    hipLaunchKernelGGL(sigma_ks_rows_kernel, dim3(n, nw), dim3(256), 0, s, sre, sim, nR, n, atom_of, natom, ph, U,
                       nst, Y);
    HIP_CHECK(hipGetLastError());
}

__global__ void scale_add_identity_kernel(const double *__restrict__ A, double *__restrict__ B, int n, double alpha,
                                          double beta) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    B[idx] = alpha * A[idx] + (i == j ? beta : 0.0);
}
void launch_scale_add_identity(const double *A, double *B, int n, double alpha, double beta, hipStream_t s) {
    hipLaunchKernelGGL(scale_add_identity_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, s, A, B, n, alpha,
                       beta);
    HIP_CHECK(hipGetLastError());
}

__global__ void scale_rows_kernel(double *__restrict__ A, int n, const double *__restrict__ f) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * n) return;
    A[idx] *= f[idx / n];
}
void launch_scale_rows(double *A, int n, const double *f, hipStream_t s) {
    hipLaunchKernelGGL(scale_rows_kernel, dim3(ceil_div((long long)n * n, 256)), dim3(256), 0, s, A, n, f);
    HIP_CHECK(hipGetLastError());
}

// one workgroup per (row i, matrix w): u_i (eps_ii u_i + 2 sum_{j<i} eps_ij u_j), summed into q[w]
__global__ __launch_bounds__(256) void head_quad_kernel(const double *__restrict__ A, long long stride, int ld,
                                                        const double *__restrict__ u, double *q) {
    __shared__ double red[256];
    const int i = blockIdx.x, w = blockIdx.y;
    const double *row = A + (long long)w * stride + (long long)i * ld;
    double acc = 0.0;
    for (int j = threadIdx.x; j < i; j += blockDim.x) acc += row[j] * u[j];
    acc *= 2.0;
    if (threadIdx.x == 0) acc += row[i] * u[i];
    red[threadIdx.x] = acc * u[i];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(q + w, red[0]);
}
void launch_head_quad(const double *A, long long stride, int ld, int n, int cnt, const double *u, double *q,
                      hipStream_t s) {
    hipLaunchKernelGGL(head_quad_kernel, dim3(n, cnt), dim3(256), 0, s, A, stride, ld, u, q);
    HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(256) void head_update_kernel(double *__restrict__ A, long long stride, int ld, int off2,
                                                          const double *__restrict__ u,
                                                          const double *__restrict__ eps_mac,
                                                          const double *__restrict__ q) {
    const int i = blockIdx.x, w = blockIdx.y;
    const double d = (eps_mac[w] - q[w]) * u[i];
    double *row = A + (long long)w * stride + (long long)i * ld;
    for (int j = threadIdx.x; j <= i; j += blockDim.x) {
        const double add = d * u[j];
        row[j] += add;
        if (off2 > 0) row[(long long)off2 * ld + off2 + j] += add;
    }
}
void launch_head_update(double *A, long long stride, int ld, int n, int off2, int cnt, const double *u,
                        const double *eps_mac, const double *q, hipStream_t s) {
    hipLaunchKernelGGL(head_update_kernel, dim3(n, cnt), dim3(256), 0, s, A, stride, ld, off2, u, eps_mac, q);
    HIP_CHECK(hipGetLastError());
}

// single workgroup of 1024 threads: deterministic reductions without atomics
__global__ __launch_bounds__(1024) void absmax_diff_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                                           long long n, double *out) {
    __shared__ double r0[1024], r1[1024];
    double d = 0, m = 0;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
        d = fmax(d, fabs(A[i] - B[i]));
        m = fmax(m, fabs(B[i]));
    }
    r0[threadIdx.x] = d;
    r1[threadIdx.x] = m;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r0[threadIdx.x] = fmax(r0[threadIdx.x], r0[threadIdx.x + o]);
            r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = r0[0]; out[1] = r1[0]; }
}
void launch_absmax_diff(const double *A, const double *B, long long n, double *out, hipStream_t s) {
    hipLaunchKernelGGL(absmax_diff_kernel, dim3(1), dim3(1024), 0, s, A, B, n, out);
    HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(1024) void sumsq_kernel(const double *__restrict__ A, long long n, double *out) {
    __shared__ double r[1024];
    double v = 0;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) v += A[i] * A[i];
    r[threadIdx.x] = v;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) r[threadIdx.x] += r[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = r[0];
}
void launch_sumsq(const double *A, long long n, double *out, hipStream_t s) {
    hipLaunchKernelGGL(sumsq_kernel, dim3(1), dim3(1024), 0, s, A, n, out);
    HIP_CHECK(hipGetLastError());
}

}  // namespace lra
