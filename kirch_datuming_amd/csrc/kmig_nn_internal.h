// Internal declarations of the traveltime surrogate (kmig_nn.hip), shared
// with the migration's tuple preparation (kmig_api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "kmig_nn.h"

namespace kmig {

typedef __bf16 bf16;

constexpr int NN_FEAT = KMIG_NN_FEATURES;  // Fourier features (3 x 16)
constexpr int NN_KIN = 64;                 // features padded to the MFMA K step
constexpr int NN_W = KMIG_NN_WIDTH;
constexpr int NN_OUT = KMIG_NN_OUTPUTS;
constexpr int NN_PARAMS = KMIG_NN_PARAMS;
static_assert(NN_PARAMS == NN_W * NN_FEAT + NN_W + 3 * (NN_W * NN_W + NN_W) + NN_OUT * NN_W + NN_OUT,
              "parameter count");

// input scaling: u = (c - c0) * c_inv per coordinate; source depth and
// reference slowness of the output transform
struct NnScale {
    float s0, z0, x0, s_inv, z_inv, x_inv, zsrc, S0;
};

// FP32 master parameters and their BF16 operand copies per layer:
// Wb[l] = W_l as [M][K padded], WT[l] = W_l^T as [K][M padded]
struct NnNet {
    float *w;
    bf16 *Wb[5];
    bf16 *WT[5];
};

void nn_cast_weights(const NnNet &net, hipStream_t st);

struct SlotSpec;

// Brings the network's inference weights up to date (synchronous on the
// network's stream); call before enqueueing nn_replace_tuples.
int nn_prepare(kmig_nn *nn);

// Surrogate step of the tuple preparation: for every (slot, grid point) of
// T/A ([nslots][nzx], filled by the table interpolation), evaluates the
// network at (surf_x of the slot, z0 + iz dz, x0 + ix dx) and, where
// |T_nn - T| <= max_dt, replaces T by T_nn and A by |dT/ds_nn * dside|.
// used[slot] counts the replaced points, maxres[slot] keeps the largest
// |T_nn - T| (float bits); both must be cleared by the caller.
// With Tn/An given (per-slot fallback), the network's values go there and a
// slot takes them at every point only if its largest residual is within
// max_dt (used = its points or 0); without them each point is decided alone.
// With tab given (per-slot fallback only), the tables' T is read at each
// slot's node of tab ([npos][nzx]) instead of from T, which the table
// interpolation then need not fill: T/A receive the network's values for the
// slots that take it, and only those.
int nn_replace_tuples(const kmig_nn *nn, const SlotSpec *spec, int nslots, int nz, size_t nzx,
                      float z0, float dz, float x0, float dx, float dside, float max_dt, float *T,
                      float *A, int *used, unsigned *maxres, float *Tn, float *An,
                      const float *tab, hipStream_t st);

}  // namespace kmig
