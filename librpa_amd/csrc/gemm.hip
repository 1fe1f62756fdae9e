// Grouped MFMA GEMM kernels for gfx950.
//
// One workgroup (4 or 8 waves) computes a BM x BN tile of one problem.  Tiles of all problems of a batch form one flat grid; a binary
// search over the per-problem tile prefix finds the problem.  The flat tile
// index is remapped so that each XCD (8 per MI355X, workgroups dispatched
// round-robin) receives a contiguous range of tiles, which keeps tiles that
// share A or B panels in the same L2.
//
// Operands are staged through LDS as As[m][k] / Bs[n][k] (k contiguous, row
// pitch padded) with a register prefetch of the next K slice.  Inner products
// use the 16x16 MFMA families:
//   F64    v_mfma_f64_16x16x4_f64   (A, B, C in double)
//   F32    v_mfma_f32_16x16x4_f32   (A, B, C in float)
//   BF16X3 v_mfma_f32_16x16x32_bf16 three times per K slice with operands split
//          x = hi + lo (hi = bf16(x), lo = bf16(x - hi)); acc += lo*hi + hi*lo + hi*hi,
//          operands read as float and split once when a tile is staged (LDS holds bf16 hi and
//          lo planes), accumulation in float.
#include "gemm.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

namespace lra {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));

#ifndef LRA_GEMM_BK64
#define LRA_GEMM_BK64 16
#endif
#ifndef LRA_GEMM_BK32
#define LRA_GEMM_BK32 16
#endif
template <int MODE> struct ModeTraits;
template <> struct ModeTraits<0> { using T = double; using V = v4d; static constexpr int BK = LRA_GEMM_BK64; static constexpr int PAD = 1; };
template <> struct ModeTraits<1> { using T = float; using V = v4f; static constexpr int BK = LRA_GEMM_BK32; static constexpr int PAD = 1; };
template <> struct ModeTraits<2> { using T = float; using V = v4f; static constexpr int BK = 32; static constexpr int PAD = 4; };

__device__ __forceinline__ unsigned short f2bf_rne(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(static_cast<unsigned int>(h) << 16); }

__device__ __forceinline__ int find_prob(const GemmProb *__restrict__ probs, int nprob, int t) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (probs[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// XCD-aware remap of the dispatch index: XCD x (= bid % 8) gets a contiguous
// chunk of logical tiles.
__device__ __forceinline__ int xcd_remap(int bid, int G) {
    const int per = G / 8, rem = G % 8;
    const int x = bid & 7, idx = bid >> 3;
    return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

// VAR selects the addressing extras compiled in: 0 plain, 1 K table, 2 column table and
// atomic epilogue (kept out of the plain kernels, where they cost registers), 3 and 4 plain
// code under separate symbols (the chi0 contraction and T_A, so that kernel traces can
// isolate them).
template <int MODE, int BM, int BN, int WM, int WN, bool TA, bool TB, int VAR>
__global__ __launch_bounds__(WM * WN * 64) void ggemm_kernel(const GemmProb *__restrict__ probs, int nprob, int total,
                                                    const typename ModeTraits<MODE>::T *__restrict__ A,
                                                    const typename ModeTraits<MODE>::T *__restrict__ B,
                                                    typename ModeTraits<MODE>::T *__restrict__ C,
                                                    const long long *__restrict__ rowtab,
                                                    const long long *__restrict__ coltab,
                                                    const long long *__restrict__ ktab,
                                                    typename ModeTraits<MODE>::T alpha,
                                                    typename ModeTraits<MODE>::T beta, int lower_only, int atomic,
                                                    const unsigned char *__restrict__ skip, int skip_zero) {
    using T = typename ModeTraits<MODE>::T;
    using V = typename ModeTraits<MODE>::V;
    constexpr int NT = WM * WN * 64;
    // the chi0 contraction (VAR 3, K ~ 25k) runs with a 32-deep K slice: half the barriers per
    // flop; T_A (VAR 4, K = 2 n_basis(J), e.g. 40) with an 8-deep one, which pads K by at most 7
    constexpr bool bk8 = VAR == 4 && MODE != 2 && (BM * 8) % NT == 0 && (BN * 8) % NT == 0;
    constexpr int BK = (VAR == 3 && MODE != 2) ? 32 : bk8 ? 8 : ModeTraits<MODE>::BK;
    constexpr int PAD = ModeTraits<MODE>::PAD;
    constexpr int LDK = BK + PAD;
    static_assert(WM * WN == 4 || WM * WN == 8, "4 or 8 waves");
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    static_assert(FM >= 1 && FN >= 1, "wave tile");
    constexpr int LA = BM * BK / NT, LB = BN * BK / NT;
    static_assert(LA * NT == BM * BK && LB * NT == BN * BK, "tile load split");

    // MODE 2 stages bf16 hi / lo planes instead (LDK2 keeps 16-byte rows for the fragment loads)
    constexpr int LDK2 = BK + 8;
    __shared__ T As[MODE == 2 ? 1 : BM * LDK];
    __shared__ T Bs[MODE == 2 ? 1 : BN * LDK];
    __shared__ __align__(16) unsigned short Ahl[MODE == 2 ? 2 * BM * LDK2 : 8];
    __shared__ __align__(16) unsigned short Bhl[MODE == 2 ? 2 * BN * LDK2 : 8];

    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= total) return;
    const int pi = find_prob(probs, nprob, tile);
    const GemmProb P = probs[pi];
    // a screened-out problem: return, or (skip_zero) write alpha * 0 + beta * C without the K loop
    const bool dead = skip != nullptr && skip[pi] != 0;
    if (dead && !skip_zero) return;
    const int lt = tile - P.tile0;
    const int tm = lt / P.tiles_n, tn = lt % P.tiles_n;
    const int m0 = tm * BM, n0 = tn * BN;
    if (lower_only && n0 > m0 + BM - 1) return;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const bool rt = P.rt_off >= 0;

    // per-thread A row bases (non-transposed A)
    long long arow[LA];
    int am[LA], ak[LA];
#pragma unroll
    for (int i = 0; i < LA; ++i) {
        const int e = t + i * NT;
        if (TA) { ak[i] = e / BM; am[i] = e % BM; }
        else { am[i] = e / BK; ak[i] = e % BK; }
        const int gm = m0 + am[i];
        if (!TA) {
            if (gm < P.M) arow[i] = rt ? rowtab[2ll * (P.rt_off + gm)] : P.a_off + (long long)gm * P.lda;
            else arow[i] = -1;
        } else {
            arow[i] = gm < P.M ? P.a_off + gm : -1;
        }
    }
    int bk[LB], bn[LB];
#pragma unroll
    for (int i = 0; i < LB; ++i) {
        const int e = t + i * NT;
        if (TB) { bn[i] = e / BK; bk[i] = e % BK; }
        else { bk[i] = e / BN; bn[i] = e % BN; }
    }

    V acc[FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[a][b][r] = 0;

    T ra[LA], rb[LB];
    constexpr bool kt = VAR == 1;
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int gk = k0 + ak[i];
            T v = 0;
            if (arow[i] >= 0 && gk < P.K) {
                const long long ko = kt ? ktab[2ll * (P.kt_off + gk)] : (TA ? (long long)gk * P.lda : (long long)gk);
                v = A[arow[i] + ko];
            }
            ra[i] = v;
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            const int gk = k0 + bk[i], gn = n0 + bn[i];
            T v = 0;
            if (gk < P.K && gn < P.N) {
                const long long ko =
                    kt ? ktab[2ll * (P.kt_off + gk) + 1] : (TB ? (long long)gk : (long long)gk * P.ldb);
                v = TB ? B[P.b_off + (long long)gn * P.ldb + ko] : B[P.b_off + ko + gn];
            }
            rb[i] = v;
        }
    };
    auto sstore = [&]() {
        if constexpr (MODE == 2) {
#pragma unroll
            for (int i = 0; i < LA; ++i) {
                const unsigned short h = f2bf_rne(ra[i]);
                Ahl[am[i] * LDK2 + ak[i]] = h;
                Ahl[BM * LDK2 + am[i] * LDK2 + ak[i]] = f2bf_rne(ra[i] - bf2f(h));
            }
#pragma unroll
            for (int i = 0; i < LB; ++i) {
                const unsigned short h = f2bf_rne(rb[i]);
                Bhl[bn[i] * LDK2 + bk[i]] = h;
                Bhl[BN * LDK2 + bn[i] * LDK2 + bk[i]] = f2bf_rne(rb[i] - bf2f(h));
            }
        } else {
#pragma unroll
            for (int i = 0; i < LA; ++i) As[am[i] * LDK + ak[i]] = ra[i];
#pragma unroll
            for (int i = 0; i < LB; ++i) Bs[bn[i] * LDK + bk[i]] = rb[i];
        }
    };

    const int rowA = wm * WTM + (lane & 15);
    const int rowB = wn * WTN + (lane & 15);
    const int kq = lane >> 4;

    // a triangular operand lets the K loop start at the tile's diagonal
    int kbeg = 0;
    if (P.kmode == 1) kbeg = (n0 / BK) * BK;
    else if (P.kmode == 2) kbeg = (m0 / BK) * BK;
    if (dead) kbeg = P.K;
    gload(kbeg);
    for (int k0 = kbeg; k0 < P.K; k0 += BK) {
        __syncthreads();
        sstore();
        __syncthreads();
        if (k0 + BK < P.K) gload(k0 + BK);
        if constexpr (MODE == 0 || MODE == 1) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                T fa[FM], fb[FN];
#pragma unroll
                for (int a = 0; a < FM; ++a) fa[a] = As[(rowA + a * 16) * LDK + kk + kq];
#pragma unroll
                for (int b = 0; b < FN; ++b) fb[b] = Bs[(rowB + b * 16) * LDK + kk + kq];
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) {
                        if constexpr (MODE == 0)
                            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
                        else
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a], fb[b], acc[a][b], 0, 0, 0);
                    }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 32) {
                v8s ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
                for (int a = 0; a < FM; ++a) {
                    const int o = (rowA + a * 16) * LDK2 + kk + kq * 8;
                    ah[a] = *reinterpret_cast<const v8s *>(&Ahl[o]);
                    al[a] = *reinterpret_cast<const v8s *>(&Ahl[BM * LDK2 + o]);
                }
#pragma unroll
                for (int b = 0; b < FN; ++b) {
                    const int o = (rowB + b * 16) * LDK2 + kk + kq * 8;
                    bh[b] = *reinterpret_cast<const v8s *>(&Bhl[o]);
                    bl[b] = *reinterpret_cast<const v8s *>(&Bhl[BN * LDK2 + o]);
                }
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) {
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(v8bf, al[a]), __builtin_bit_cast(v8bf, bh[b]), acc[a][b], 0, 0, 0);
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(v8bf, ah[a]), __builtin_bit_cast(v8bf, bl[b]), acc[a][b], 0, 0, 0);
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(v8bf, ah[a]), __builtin_bit_cast(v8bf, bh[b]), acc[a][b], 0, 0, 0);
                    }
            }
        }
    }

    // epilogue: acc[a][b][r] holds C[row][n0 + wn*WTN + b*16 + lane%16].  The output column of
    // each b (column split and table included) depends on the lane only: worked out once here,
    // not for every row (a split's division by ncb is some thirty VALU instructions)
    const bool split = P.ncb < P.N;
    const bool ct = VAR == 2 && P.ct_off >= 0;
    long long ccol[FN];
#pragma unroll
    for (int b = 0; b < FN; ++b) {
        const int gn = n0 + wn * WTN + b * 16 + (lane & 15);
        ccol[b] = -1;
        if (gn < P.N) {
            const int inner = split ? gn % P.ncb : gn;
            long long col = ct ? coltab[P.ct_off + inner] : inner;
            if (split) col += (long long)(gn / P.ncb) * P.cbs;
            ccol[b] = col;
        }
    }
    // a full tile with alpha = 1, beta = 0 and plain stores (the chi0 build's W, X and T_A GEMMs):
    // no per-element checks, scaling or reads of C
    // T_A (tag 4) writes the contraction's T operand, gigabytes per launch that no cache keeps
    // until the contraction reads them: in FP64 streamed past the caches with non-temporal stores
    // (6% faster on the 128-atom cell). 16 lanes then write whole 128-byte lines; with 4-byte
    // elements they write half lines, and non-temporal stores made the FP32 T_A 17% slower.
    const bool plain = !((VAR == 2 || VAR == 3) && atomic) && alpha == T(1) && beta == T(0);
    if (plain && m0 + BM <= P.M && n0 + BN <= P.N) {
#pragma unroll
        for (int a = 0; a < FM; ++a) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * WTM + a * 16 + (MODE == 0 ? 4 * r + kq : 4 * kq + r);
                const long long crow = rt ? rowtab[2ll * (P.rt_off + gm) + 1] : P.c_off + (long long)gm * P.ldc;
#pragma unroll
                for (int b = 0; b < FN; ++b) {
                    if constexpr (VAR == 4 && sizeof(T) == 8) __builtin_nontemporal_store(acc[a][b][r], C + crow + ccol[b]);
                    else C[crow + ccol[b]] = acc[a][b][r];
                }
            }
        }
        return;
    }
#pragma unroll
    for (int a = 0; a < FM; ++a) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // f64 16x16x4 places register r of lane group kq at row 4r + kq, the f32 / bf16
            // variants at row 4kq + r
            const int gm = m0 + wm * WTM + a * 16 + (MODE == 0 ? 4 * r + kq : 4 * kq + r);
            if (gm >= P.M) continue;
            const long long crow = rt ? rowtab[2ll * (P.rt_off + gm) + 1] : P.c_off + (long long)gm * P.ldc;
#pragma unroll
            for (int b = 0; b < FN; ++b) {
                if (ccol[b] < 0) continue;
                T *dst = C + crow + ccol[b];
                const T v0 = alpha * acc[a][b][r];
                if ((VAR == 2 || VAR == 3) && atomic) {
                    atomicAdd(dst, v0);
                } else {
                    T v = v0;
                    if (beta != T(0)) v += beta * *dst;
                    if constexpr (VAR == 4 && sizeof(T) == 8) __builtin_nontemporal_store(v, dst);
                    else *dst = v;
                }
            }
        }
    }
}

struct TileShape { int bm, bn, wm, wn; const char *name; };
static const TileShape kTiles[NUM_TILE_CFG] = {
    {64, 64, 2, 2, "64x64"},   {128, 128, 2, 2, "128x128"}, {32, 128, 1, 4, "32x128"},
    {128, 64, 2, 2, "128x64"}, {64, 128, 2, 2, "64x128"},   {16, 256, 1, 4, "16x256"},
    {32, 64, 2, 2, "32x64"},     {128, 128, 2, 4, "128x128w8"}, {128, 64, 4, 2, "128x64w8"},
};
int tile_bm(int cfg) { return kTiles[cfg].bm; }
int tile_bn(int cfg) { return kTiles[cfg].bn; }
const char *tile_name(int cfg) { return kTiles[cfg].name; }

GemmProb make_prob(int M, int N, int K, long long a_off, int lda, long long b_off, int ldb, long long c_off,
                   int ldc) {
    GemmProb p;
    std::memset(&p, 0, sizeof(p));
    p.M = M; p.N = N; p.K = K;
    p.ncb = 1 << 30; p.cbs = 0;
    p.a_off = a_off; p.lda = lda; p.b_off = b_off; p.ldb = ldb; p.c_off = c_off; p.ldc = ldc;
    p.rt_off = -1;
    p.ct_off = -1;
    p.kt_off = -1;
    return p;
}

void GemmBatch::finalize(int cfg_, hipStream_t s) {
    cfg = cfg_;
    const int bm = tile_bm(cfg), bn = tile_bn(cfg);
    int t0 = 0;
    flops = 0;
    for (auto &p : probs) {
        p.tiles_n = ceil_div(p.N, bn);
        p.tile0 = t0;
        t0 += ceil_div(p.M, bm) * p.tiles_n;
        if (p.kmode == 0 && !lower_only) {
            flops += 2.0 * p.M * p.N * (double)p.K;
        } else {
            // executed work: tiles above the diagonal skipped, K started at the tile's diagonal
            for (int m0 = 0; m0 < p.M; m0 += bm)
                for (int n0 = 0; n0 < p.N; n0 += bn) {
                    if (lower_only && n0 > m0 + bm - 1) continue;
                    const int kb = p.kmode == 1 ? n0 : (p.kmode == 2 ? m0 : 0);
                    flops += 2.0 * std::min(bm, p.M - m0) * std::min(bn, p.N - n0) * (double)std::max(0, p.K - kb);
                }
        }
    }
    total_tiles = t0;
    d_probs = upload(probs, s);
    // the tables do not depend on the tile shape: upload once
    if (!rowtab.empty() && d_rowtab.bytes() != rowtab.size() * sizeof(long long)) d_rowtab = upload(rowtab, s);
    if (!coltab.empty() && d_coltab.bytes() != coltab.size() * sizeof(long long)) d_coltab = upload(coltab, s);
    if (!ktab.empty() && d_ktab.bytes() != ktab.size() * sizeof(long long)) d_ktab = upload(ktab, s);
}

template <int MODE, int CFG, bool TA, bool TB, int VAR>
static void launch_one(const GemmBatch &b, const void *A, const void *B, void *C, double alpha, double beta,
                       hipStream_t s) {
    using T = typename ModeTraits<MODE>::T;
    constexpr int BM = (CFG == 0) ? 64 : (CFG == 1) ? 128 : (CFG == 2) ? 32 : (CFG == 3) ? 128 : (CFG == 4) ? 64
                     : (CFG == 5) ? 16 : (CFG == 6) ? 32 : 128;
    constexpr int BN = (CFG == 0) ? 64 : (CFG == 1) ? 128 : (CFG == 2) ? 128 : (CFG == 3) ? 64 : (CFG == 4) ? 128
                     : (CFG == 5) ? 256 : (CFG == 6) ? 64 : (CFG == 7) ? 128 : 64;
    // 8-wave tiles: 128x128 as 2 x 4 waves of 64x32, 128x64 as 4 x 2 waves of 32x32
    constexpr int WM = (CFG == 2 || CFG == 5) ? 1 : (CFG == 8) ? 4 : 2;
    constexpr int WN = (CFG == 7) ? 4 : (CFG == 8) ? 2 : 4 / WM;
    hipLaunchKernelGGL((ggemm_kernel<MODE, BM, BN, WM, WN, TA, TB, VAR>), dim3(b.total_tiles), dim3(WM * WN * 64), 0, s,
                       b.d_probs.as<GemmProb>(), (int)b.probs.size(), b.total_tiles, static_cast<const T *>(A),
                       static_cast<const T *>(B), static_cast<T *>(C), b.d_rowtab.as<long long>(),
                       b.d_coltab.as<long long>(), b.d_ktab.as<long long>(), (T)alpha, (T)beta,
                       (int)b.lower_only, (int)b.atomic, b.use_skip ? b.d_skip.as<unsigned char>() : nullptr,
                       (int)b.skip_zero);
    HIP_CHECK(hipGetLastError());
}

template <int MODE, bool TA, bool TB, int VAR = 0>
static void launch_cfg(const GemmBatch &b, const void *A, const void *B, void *C, double alpha, double beta,
                       hipStream_t s) {
    switch (b.cfg) {
        case 0: launch_one<MODE, 0, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        case 1: launch_one<MODE, 1, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        case 2: launch_one<MODE, 2, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        case 3: launch_one<MODE, 3, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        case 4: launch_one<MODE, 4, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        case 5: launch_one<MODE, 5, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        case 6: launch_one<MODE, 6, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        case 7: launch_one<MODE, 7, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        case 8: launch_one<MODE, 8, TA, TB, VAR>(b, A, B, C, alpha, beta, s); break;
        default: LRA_THROW("bad tile cfg");
    }
}

template <int MODE>
static void launch_mode(const GemmBatch &b, const void *A, const void *B, void *C, double alpha, double beta,
                        hipStream_t s) {
    if (!b.ktab.empty()) {
        // K-table batches: transposed A, plain B, 64x64 tiles
        if (!b.trans_a || b.trans_b || b.cfg != T64x64 || b.atomic || !b.coltab.empty())
            LRA_THROW("K-table GEMM: transposed A, plain B, 64x64 tiles, plain epilogue");
        launch_one<MODE, 0, true, false, 1>(b, A, B, C, alpha, beta, s);
        return;
    }
    if (b.tag == 3 || b.tag == 4) {
        // same code as the plain kernel, instantiated separately so that traces can tell it apart
        if (b.trans_b || (b.trans_a && b.tag == 4) || !b.coltab.empty() || (b.atomic && b.tag == 4))
            LRA_THROW("tagged GEMM: plain B, no column table (and plain A, plain epilogue for tag 4)");
        if (b.tag == 3 && b.trans_a) launch_cfg<MODE, true, false, 3>(b, A, B, C, alpha, beta, s);
        else if (b.tag == 3) launch_cfg<MODE, false, false, 3>(b, A, B, C, alpha, beta, s);
        else launch_cfg<MODE, false, false, 4>(b, A, B, C, alpha, beta, s);
        return;
    }
    if (b.atomic || !b.coltab.empty()) {
        // column-table / atomic batches: plain A and B, 64x64 or 128x128 tiles
        if (b.trans_a || b.trans_b || (b.cfg != T128x128 && b.cfg != T64x64))
            LRA_THROW("column-table GEMM: plain A and B, 64x64 or 128x128 tiles");
        if (b.cfg == T64x64) launch_one<MODE, 0, false, false, 2>(b, A, B, C, alpha, beta, s);
        else launch_one<MODE, 1, false, false, 2>(b, A, B, C, alpha, beta, s);
        return;
    }
    if (b.trans_a && b.trans_b) launch_cfg<MODE, true, true>(b, A, B, C, alpha, beta, s);
    else if (b.trans_a) launch_cfg<MODE, true, false>(b, A, B, C, alpha, beta, s);
    else if (b.trans_b) launch_cfg<MODE, false, true>(b, A, B, C, alpha, beta, s);
    else launch_cfg<MODE, false, false>(b, A, B, C, alpha, beta, s);
}

namespace {
// tile choices of batches already tuned in this process, keyed by the batch's shape signature
// (precision, variant, transposes, and every problem's M, N, K, strides and addressing mode), so
// that a second handler with the same inputs does not time the candidates again
std::mutex g_tune_mu;
std::map<std::vector<long long>, std::pair<int, float>> &tune_cache() {
    static std::map<std::vector<long long>, std::pair<int, float>> c;
    return c;
}
std::vector<long long> tune_key(Prec prec, const GemmBatch &b) {
    std::vector<long long> k = {(long long)prec, b.tag, b.trans_a, b.trans_b, b.lower_only,
                                (long long)b.rowtab.size(), (long long)b.coltab.size(), (long long)b.ktab.size()};
    k.reserve(k.size() + 8 * b.probs.size());
    for (const auto &p : b.probs) {
        k.push_back(p.M);
        k.push_back(p.N);
        k.push_back(p.K);
        k.push_back(p.lda);
        k.push_back(p.ldb);
        k.push_back(p.ldc);
        k.push_back(p.kmode + 4 * (p.rt_off >= 0) + 8 * (p.ct_off >= 0) + 16 * (p.kt_off >= 0));
        k.push_back(p.ncb < p.N ? p.ncb : 0);
    }
    return k;
}
}  // namespace

int GemmBatch::tune(Prec prec, const void *A, const void *B, void *C, double alpha, double beta, hipStream_t s,
                    const std::function<void()> &before) {
    if (probs.empty() || (atomic && tag != 3)) return cfg;
    const auto key = tune_key(prec, *this);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = tune_cache().find(key);
        if (it != tune_cache().end()) {
            finalize(it->second.first, s);
            tuned = true;
            tuned_ms = it->second.second;
            cfg_by_prec[(int)prec] = cfg;
            return cfg;
        }
    }
    const bool was_skip = use_skip;
    use_skip = false;
    std::vector<int> cands;
    if (!ktab.empty()) cands = {T64x64};
    else if (!coltab.empty()) cands = {T64x64, T128x128};
    else for (int c = 0; c < NUM_TILE_CFG; ++c) cands.push_back(c);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    auto timed = [&](void) {
        if (before) before();
        HIP_CHECK(hipEventRecord(e0, s));
        gemm_run(prec, *this, A, B, C, alpha, beta, s);
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        return ms;
    };
    float best = 1e30f;
    int bc = cfg;
    for (int c : cands) {
        finalize(c, s);
        // first run warms the code object; a candidate already 1.5x slower than the best
        // is not timed again
        const float first = timed();
        if (first > 1.5f * best) continue;
        const float ms = std::min(first, timed());
        if (ms < best) { best = ms; bc = c; }
    }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    finalize(bc, s);
    use_skip = was_skip;
    tuned = true;
    tuned_ms = best;
    cfg_by_prec[(int)prec] = bc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_cache()[key] = {bc, best};
    }
    return bc;
}

void gemm_run(Prec prec, const GemmBatch &b, const void *A, const void *B, void *C, double alpha, double beta,
              hipStream_t s) {
    if (b.total_tiles == 0) return;
    if (b.trans_a && !b.rowtab.empty()) LRA_THROW("row indirection needs non-transposed A");
    if (b.atomic && beta != 1.0) LRA_THROW("atomic GEMM epilogue accumulates (beta must be 1)");
    switch (prec) {
        case Prec::F64: launch_mode<0>(b, A, B, C, alpha, beta, s); break;
        case Prec::F32: launch_mode<1>(b, A, B, C, alpha, beta, s); break;
        case Prec::BF16X3: launch_mode<2>(b, A, B, C, alpha, beta, s); break;
    }
}

}  // namespace lra
