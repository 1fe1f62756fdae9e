// Block-norm screener of the chi0 build: block norms, the norm-product bound's inputs and a
// small FP32 MLP (8 -> 64 -> 64 -> 1, SiLU) with its own forward, backward and Adam kernels.
#pragma once

#include <vector>

#include "common.h"

namespace lra {

constexpr int MLP_NFEAT = 8;
constexpr int MLP_H = 64;
constexpr int MLP_NPARAM = MLP_NFEAT * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;  // 4801
// model = parameters, then feature means [8], feature scales [8], target mean, target scale
constexpr int MLP_NMODEL = MLP_NPARAM + 2 * MLP_NFEAT + 2;

// A dense block inside a row-major matrix: first element, rows, columns (leading dimension shared)
struct BlockRef {
    long long off;
    int rows, cols;
};

// out[b] = ||block b||_F (one workgroup per block)
template <class T>
void launch_block_norms(const T *base, long long ld, const BlockRef *blocks, int nblocks, double *out, hipStream_t s);
// out[t][o] = ||chi0s(o, t)||_F for contiguous blocks of esize[o] elements at t * E_tot + eoff[o]
template <class T>
void launch_chi0_block_norms(const T *chi0s, long long E_tot, int ntau, const long long *eoff, const int *esize,
                             int nout, double *out, hipStream_t s);

// y[s] = MLP(X[s]) for standardised features X [n][8]
void launch_mlp_forward(const float *params, const float *X, int n, float *y, hipStream_t s);

struct MlpWork {
    DevBuf acts, part, r2, grad, m, v;
};
// mean squared error of the MLP on (X, t) and its gradient (layout of the parameter vector);
// grad and loss are device pointers (loss: one double)
void mlp_loss_grad(const float *params, const float *X, const float *t, int n, float *grad, double *loss,
                   MlpWork &w, hipStream_t s);
// The same on device arrays without allocating: the work memory is a static device array of the
// library (kMlpNoallocScratchBytes per GPU), for n <= kMlpNoallocMax; larger n throws.  Calls that
// overlap on one GPU share that array and race.
constexpr int kMlpNoallocMax = 32768;
constexpr size_t kMlpNoallocScratchBytes = 35176704;  // activations + dy, r2, 64 chunk partials
void mlp_loss_grad_noalloc(const float *params, const float *X, const float *t, int n, float *grad, double *loss,
                           hipStream_t s);
// Device address of the static work array on the current device (resolved once per device).
char *mlp_noalloc_setup();
// one Adam step on params with gradient grad (step >= 1)
void launch_adam(float *params, const float *grad, float *m, float *v, int n, float lr, int step, hipStream_t s);

// Trains the MLP on raw features X [n][8] and targets t [n] (ln block norms): standardises both,
// runs `iters` full-batch Adam steps on the device and writes the model (MLP_NMODEL floats:
// parameters, feature means, feature scales, target mean, target scale).  Returns the final
// mean squared error in standardised units.
double mlp_train(const float *X, const float *t, int n, int iters, double lr, int seed, float *model);

}  // namespace lra
