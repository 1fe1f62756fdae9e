// Grouped (persistent-style) GEMM engine for CDNA4 MFMA.
//
// Every dense contraction of the chi0 / Ec / Sigma path is expressed as a
// list of problems C[M x N] = alpha * A[M x K] * B[K x N] + beta * C that
// share one launch.  Operands live in three base buffers and each problem
// addresses them through element offsets and leading dimensions, plus two
// generalisations needed by the atom-pair layouts:
//   * row indirection: row m of A and C may come from a table of absolute
//     element offsets (gathered rows, scattered outputs);
//   * column split on C: output column c goes to (c / ncb) * cbs + c % ncb,
//     which writes a [j][rows][nu] layout directly from a [rows][(j, nu)] GEMM;
//   * column table on C: the within-block part c % ncb may be replaced by an
//     entry of a table of element offsets (scattered output columns);
//   * atomic epilogue: C += alpha A B with atomic adds, for outputs that several
//     problems of one launch accumulate into;
//   * K table: the inner index k may address A and B through a table of element
//     offsets (a_k, b_k), for contractions whose summed index spans several strides.
#pragma once

#include <functional>
#include <vector>

#include "common.h"

namespace lra {

enum class Prec : int { F64 = 0, F32 = 1, BF16X3 = 2 };

inline const char *prec_name(Prec p) {
    return p == Prec::F64 ? "fp64" : (p == Prec::F32 ? "fp32" : "bf16x3");
}

struct GemmProb {
    int M, N, K;
    int tiles_n;   // filled by GemmBatch::finalize
    int tile0;     // first global tile of this problem (filled by finalize)
    int ncb;       // C column-split block (>= N means none)
    long long cbs; // C column-split stride
    long long a_off, b_off, c_off;
    int lda, ldb, ldc;
    int rt_off;    // offset into the row table, -1 if not row-indirect
    int kmode;     // triangular operand: 0 none, 1 B lower (k < n is zero), 2 A = L^T (k < m is zero)
    int ct_off;    // offset into the column table, -1 if none
    int kt_off;    // offset (in pairs) into the K table, -1 if none
};

// Tile configurations (BM x BN; 256 threads, the w8 shapes 512).
enum TileCfg : int {
    T64x64 = 0,
    T128x128 = 1,
    T32x128 = 2,
    T128x64 = 3,
    T64x128 = 4,
    T16x256 = 5,
    T32x64 = 6,
    T128x128w8 = 7,
    T128x64w8 = 8,
    NUM_TILE_CFG = 9
};
// Largest square tile that runs well for a precision: 64x64 for FP64 (two waves per SIMD;
// 128x128 FP64 tiles leave one wave per SIMD and measure ~20% slower), 128x128 otherwise.
inline int big_tile(Prec p) { return p == Prec::F64 ? T64x64 : T128x128; }
int tile_bm(int cfg);
int tile_bn(int cfg);
const char *tile_name(int cfg);

// A device-resident list of problems with its row table.
struct GemmBatch {
    std::vector<GemmProb> probs;
    std::vector<long long> rowtab;  // pairs (a_row_off, c_row_off)
    std::vector<long long> coltab;  // C column offsets (within a column-split block)
    std::vector<long long> ktab;    // pairs (a_k, b_k): offsets of inner index k in A and B
    int cfg = T64x64;
    bool trans_a = false;  // A stored as A^T (element (m,k) at a_off + k*lda + m)
    bool trans_b = false;  // B stored as B^T (element (k,n) at b_off + n*ldb + k)
    bool lower_only = false;
    bool atomic = false;   // C += alpha A B by atomic adds (beta must be 1)
    // per-problem skip flags (screening): skipped problems return, or with skip_zero write
    // alpha * 0 + beta * C; d_skip is filled by the owner before a launch with use_skip set
    bool use_skip = false, skip_zero = false;
    std::vector<unsigned char> skipv;
    DevBuf d_skip;
    int tag = 0;           // 3 / 4: launch under a separately named (traceable) plain instantiation
    int total_tiles = 0;
    double flops = 0;      // flops of one launch: 2 M N K summed, less the skipped triangular work
    DevBuf d_probs, d_rowtab, d_coltab, d_ktab;

    bool tuned = false;
    float tuned_ms = 0;    // time of the chosen tile shape in the tuning run
    // tuned tile shape per precision (index Prec), -1 where not tuned: a caller that launches one
    // batch at two precisions re-finalizes with use_tiles before switching
    int cfg_by_prec[3] = {-1, -1, -1};

    void add(const GemmProb &p) { probs.push_back(p); }
    void finalize(int cfg_, hipStream_t s = 0);
    // switch to the shape tuned for prec (if any and different from the current one)
    void use_tiles(Prec prec, hipStream_t s) {
        const int c = cfg_by_prec[(int)prec];
        if (c >= 0 && c != cfg) finalize(c, s);
    }
    bool empty() const { return probs.empty(); }
    // Autotuner: time one launch of every tile shape this batch can use (after a warm-up
    // launch) on the real operands and keep the fastest.  The launches must be harmless
    // to repeat (overwriting outputs); `before` runs before every launch (e.g. to reset an
    // accumulated output).  Returns the chosen tile configuration.
    int tune(Prec prec, const void *A, const void *B, void *C, double alpha, double beta, hipStream_t s,
             const std::function<void()> &before = nullptr);
};

GemmProb make_prob(int M, int N, int K, long long a_off, int lda, long long b_off, int ldb,
                   long long c_off, int ldc);

// Launch all problems of a batch.  A, B, C are the base pointers the offsets refer to.
void gemm_run(Prec prec, const GemmBatch &b, const void *A, const void *B, void *C, double alpha,
              double beta, hipStream_t s);

}  // namespace lra
