// Block norms and the screener MLP (forward, backward, Adam) on gfx950.
//
// MLP, FP32: z1 = x W1 + b1, h1 = silu(z1), z2 = h1 W2 + b2, h2 = silu(z2), y = h2 w3 + b3.
// Parameter vector: W1[8][64], b1[64], W2[64][64], b2[64], w3[64], b3.
// Training: full-batch mean squared error; one thread per sample computes the forward pass and
// the per-sample deltas (dz1, dz2, dy); a second kernel reduces them to the parameter gradient
// over sample chunks, a third sums the chunks; Adam updates the parameters in place.
#include "screen.h"

#include <map>
#include <mutex>
#include <string>

namespace lra {

namespace {
constexpr int OW1 = 0, OB1 = MLP_NFEAT * MLP_H, OW2 = OB1 + MLP_H, OB2 = OW2 + MLP_H * MLP_H, OW3 = OB2 + MLP_H,
              OB3 = OW3 + MLP_H;
static_assert(OB3 + 1 == MLP_NPARAM, "parameter layout");
constexpr int ACT = 4 * MLP_H;  // per-sample activations: h1, dz1, h2, dz2
constexpr int NROW = MLP_NFEAT + 1 + MLP_H + 1 + 1 + 1;  // gradient rows: W1, b1, W2, b2, w3, b3
constexpr int GCHUNKS = 64;

__device__ __forceinline__ float sig(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    return s;
}
}  // namespace

template <class T>
__global__ __launch_bounds__(256) void block_norms_kernel(const T *__restrict__ base, long long ld,
                                                          const BlockRef *__restrict__ blocks, double *out) {
    __shared__ double red[4];
    const BlockRef b = blocks[blockIdx.x];
    const int n = b.rows * b.cols;
    double v = 0;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const double x = static_cast<double>(base[b.off + (long long)(e / b.cols) * ld + e % b.cols]);
        v += x * x;
    }
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) out[blockIdx.x] = sqrt(s);
}

template <class T>
void launch_block_norms(const T *base, long long ld, const BlockRef *blocks, int nblocks, double *out, hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(block_norms_kernel<T>, dim3(nblocks), dim3(256), 0, s, base, ld, blocks, out);
    HIP_CHECK(hipGetLastError());
}
template void launch_block_norms<double>(const double *, long long, const BlockRef *, int, double *, hipStream_t);
template void launch_block_norms<float>(const float *, long long, const BlockRef *, int, double *, hipStream_t);

template <class T>
__global__ __launch_bounds__(256) void chi0_block_norms_kernel(const T *__restrict__ chi0s, long long E_tot,
                                                               const long long *__restrict__ eoff,
                                                               const int *__restrict__ esize, int nout,
                                                               double *out) {
    __shared__ double red[4];
    const int o = blockIdx.x, t = blockIdx.y;
    const T *p = chi0s + (long long)t * E_tot + eoff[o];
    double v = 0;
    for (int e = threadIdx.x; e < esize[o]; e += blockDim.x) {
        const double x = static_cast<double>(p[e]);
        v += x * x;
    }
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) out[(long long)t * nout + o] = sqrt(s);
}

template <class T>
void launch_chi0_block_norms(const T *chi0s, long long E_tot, int ntau, const long long *eoff, const int *esize,
                             int nout, double *out, hipStream_t s) {
    if (nout <= 0 || ntau <= 0) return;
    hipLaunchKernelGGL(chi0_block_norms_kernel<T>, dim3(nout, ntau), dim3(256), 0, s, chi0s, E_tot, eoff, esize,
                       nout, out);
    HIP_CHECK(hipGetLastError());
}
template void launch_chi0_block_norms<double>(const double *, long long, int, const long long *, const int *, int,
                                              double *, hipStream_t);
template void launch_chi0_block_norms<float>(const float *, long long, int, const long long *, const int *, int,
                                             double *, hipStream_t);

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *__restrict__ params,
                                                          const float *__restrict__ X, int n, float *__restrict__ y) {
    __shared__ float P[MLP_NPARAM];
    for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[MLP_NFEAT];
#pragma unroll
    for (int i = 0; i < MLP_NFEAT; ++i) x[i] = X[(long long)s * MLP_NFEAT + i];
    float h1[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        float z = P[OB1 + j];
#pragma unroll
        for (int i = 0; i < MLP_NFEAT; ++i) z += x[i] * P[OW1 + i * MLP_H + j];
        h1[j] = z * sig(z);
    }
    float out = P[OB3];
    for (int j = 0; j < MLP_H; ++j) {
        float z = P[OB2 + j];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) z += h1[k] * P[OW2 + k * MLP_H + j];
        out += z * sig(z) * P[OW3 + j];
    }
    y[s] = out;
}

void launch_mlp_forward(const float *params, const float *X, int n, float *y, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, params, X, n, y);
    HIP_CHECK(hipGetLastError());
}

// per sample: forward pass, r = y - t, dy = 2 r / n, dz2 = dy w3 silu'(z2), dz1 = (W2 dz2) silu'(z1);
// act[s] = [h1 | dz1 | h2 | dz2], dy[s], r2[s] = r^2.  The silu' factors are parked in the dz
// slots of act (each thread re-reads only what it wrote), so that at most one 64-vector
// (h1, then dz2) lives in registers.
__global__ __launch_bounds__(256) void mlp_fwdbwd_kernel(const float *__restrict__ params,
                                                         const float *__restrict__ X, const float *__restrict__ t,
                                                         int n, float *__restrict__ act, float *__restrict__ dyo,
                                                         double *__restrict__ r2) {
    __shared__ float P[MLP_NPARAM];
    for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float *a = act + (long long)s * ACT;
    float x[MLP_NFEAT];
#pragma unroll
    for (int i = 0; i < MLP_NFEAT; ++i) x[i] = X[(long long)s * MLP_NFEAT + i];
    float h1[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        float z = P[OB1 + j];
#pragma unroll
        for (int i = 0; i < MLP_NFEAT; ++i) z += x[i] * P[OW1 + i * MLP_H + j];
        const float g = sig(z);
        h1[j] = z * g;
        a[j] = h1[j];
        a[MLP_H + j] = g * (1.0f + z * (1.0f - g));  // silu'(z1), scaled below
    }
    float y = P[OB3];
    for (int j = 0; j < MLP_H; ++j) {
        float z = P[OB2 + j];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) z += h1[k] * P[OW2 + k * MLP_H + j];
        const float g = sig(z);
        const float h2 = z * g;
        a[2 * MLP_H + j] = h2;
        a[3 * MLP_H + j] = g * (1.0f + z * (1.0f - g));  // silu'(z2), scaled below
        y += h2 * P[OW3 + j];
    }
    const float r = y - t[s];
    const float dy = 2.0f * r / (float)n;
    float d2[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        d2[j] = a[3 * MLP_H + j] * dy * P[OW3 + j];
        a[3 * MLP_H + j] = d2[j];
    }
    for (int k = 0; k < MLP_H; ++k) {
        float dh = 0;
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) dh += d2[j] * P[OW2 + k * MLP_H + j];
        a[MLP_H + k] *= dh;
    }
    dyo[s] = dy;
    r2[s] = (double)r * (double)r;
}

// part[chunk][param] = sum over the chunk's samples of the per-sample gradient, one gradient row
// (64 or 1 columns) per blockIdx.x
__global__ __launch_bounds__(256) void mlp_grad_rows_kernel(const float *__restrict__ X,
                                                            const float *__restrict__ act,
                                                            const float *__restrict__ dy, int n, int chunk,
                                                            float *__restrict__ part) {
    __shared__ float red[4][MLP_H];
    const int row = blockIdx.x, ch = blockIdx.y;
    const int g = threadIdx.x >> 6, j = threadIdx.x & 63;
    // row -> (A source, B source, output offset, columns)
    int asrc, bsrc, out, ncol = MLP_H;
    if (row < MLP_NFEAT) { asrc = row; bsrc = 1; out = OW1 + row * MLP_H; }
    else if (row == MLP_NFEAT) { asrc = -1; bsrc = 1; out = OB1; }
    else if (row < MLP_NFEAT + 1 + MLP_H) { asrc = 100 + (row - MLP_NFEAT - 1); bsrc = 3; out = OW2 + (row - MLP_NFEAT - 1) * MLP_H; }
    else if (row == MLP_NFEAT + 1 + MLP_H) { asrc = -1; bsrc = 3; out = OB2; }
    else if (row == MLP_NFEAT + 2 + MLP_H) { asrc = -2; bsrc = 2; out = OW3; }
    else { asrc = -1; bsrc = -2; out = OB3; ncol = 1; }
    const int s0 = ch * chunk, s1 = min(n, s0 + chunk);
    float acc = 0;
    if (j < ncol)
        for (int s = s0 + g; s < s1; s += 4) {
            const float *as = act + (long long)s * ACT;
            const float av = asrc >= 100 ? as[asrc - 100] : asrc >= 0 ? X[(long long)s * MLP_NFEAT + asrc]
                                                           : asrc == -2 ? dy[s] : 1.0f;
            const float bv = bsrc == -2 ? dy[s] : as[bsrc * MLP_H + j];
            acc += av * bv;
        }
    red[g][j] = acc;
    __syncthreads();
    if (g == 0 && j < ncol) part[(long long)ch * MLP_NPARAM + out + j] = red[0][j] + red[1][j] + red[2][j] + red[3][j];
}

__global__ void mlp_grad_sum_kernel(const float *__restrict__ part, int nch, float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    float s = 0;
    for (int c = 0; c < nch; ++c) s += part[(long long)c * MLP_NPARAM + p];
    grad[p] = s;
}

__global__ __launch_bounds__(1024) void mean_kernel(const double *__restrict__ v, int n, double *out) {
    __shared__ double red[16];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
    const double tot = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = tot / (double)n;
}

namespace {
// The loss and gradient launches on caller-provided work memory: act [n][ACT] followed by dy [n],
// r2 [n], part [GCHUNKS][MLP_NPARAM].  Every entry the later launches read is written by an earlier
// one in the same call, so the work memory needs no clearing.
void loss_grad_launches(const float *params, const float *X, const float *t, int n, float *grad, double *loss,
                        float *act, double *r2, float *part, hipStream_t s) {
    float *dy = act + (size_t)n * ACT;
    hipLaunchKernelGGL(mlp_fwdbwd_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, params, X, t, n, act, dy, r2);
    HIP_CHECK(hipGetLastError());
    const int chunk = ceil_div(n, GCHUNKS);
    const int nch = ceil_div(n, chunk);
    hipLaunchKernelGGL(mlp_grad_rows_kernel, dim3(NROW, nch), dim3(256), 0, s, X, act, dy, n, chunk, part);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(mlp_grad_sum_kernel, dim3(ceil_div(MLP_NPARAM, 256)), dim3(256), 0, s, part, nch, grad);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(mean_kernel, dim3(1), dim3(1024), 0, s, r2, n, loss);
    HIP_CHECK(hipGetLastError());
}

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kActBytes = align256(sizeof(float) * (size_t)kMlpNoallocMax * (ACT + 1));
constexpr size_t kR2Bytes = align256(sizeof(double) * (size_t)kMlpNoallocMax);
constexpr size_t kPartBytes = align256(sizeof(float) * (size_t)GCHUNKS * MLP_NPARAM);
constexpr size_t kScratchBytes = kActBytes + kR2Bytes + kPartBytes;
static_assert(kScratchBytes == kMlpNoallocScratchBytes, "scratch size");
}  // namespace

// Work memory of the allocation-free loss and gradient (about 35 MB): each GPU the library runs
// on holds its own copy from the code object's load until the process ends.  Host code reaches it
// only through the per-device address from hipGetSymbolAddress below, never through this name.
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_dev;  // device ordinal -> device address of scratch
}  // namespace

char *mlp_noalloc_setup() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_dev.find(dev);
    if (it != g_scratch_dev.end()) return it->second;
    void *p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || p == nullptr)
        LRA_THROW(std::string("hipGetSymbolAddress of the MLP scratch array failed: ") + hipGetErrorString(rc));
    g_scratch_dev[dev] = static_cast<char *>(p);
    return static_cast<char *>(p);
}

void mlp_loss_grad_noalloc(const float *params, const float *X, const float *t, int n, float *grad, double *loss,
                           hipStream_t s) {
    if (n <= 0) LRA_THROW("mlp: no samples");
    if (n > kMlpNoallocMax)
        LRA_THROW("allocation-free MLP loss/gradient: " + std::to_string(n) + " samples exceed the maximum of " +
                  std::to_string(kMlpNoallocMax));
    if (!params || !X || !t || !grad || !loss) LRA_THROW("allocation-free MLP loss/gradient: null device pointer");
    char *base = mlp_noalloc_setup();
    loss_grad_launches(params, X, t, n, grad, loss, reinterpret_cast<float *>(base),
                       reinterpret_cast<double *>(base + kActBytes), reinterpret_cast<float *>(base + kActBytes + kR2Bytes),
                       s);
}

void mlp_loss_grad(const float *params, const float *X, const float *t, int n, float *grad, double *loss,
                   MlpWork &w, hipStream_t s) {
    if (n <= 0) LRA_THROW("mlp: no samples");
    w.acts.reserve(sizeof(float) * ((size_t)n * ACT + n));
    w.r2.reserve(sizeof(double) * n);
    w.part.reserve(sizeof(float) * (size_t)GCHUNKS * MLP_NPARAM);
    loss_grad_launches(params, X, t, n, grad, loss, w.acts.as<float>(), w.r2.as<double>(), w.part.as<float>(), s);
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

void launch_adam(float *params, const float *grad, float *m, float *v, int n, float lr, int step, hipStream_t s) {
    const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, params, grad, m, v, n, lr, c1, c2);
    HIP_CHECK(hipGetLastError());
}

}  // namespace lra
