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

}  // namespace kmig
