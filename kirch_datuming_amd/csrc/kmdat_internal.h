// Internal declarations shared by the datuming host code and its kernels.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace kmdat {

// Largest filter length in samples (length / dt + 2) the table kernel takes.
constexpr int KMDAT_MAX_NSAM = 4096;

// Banded operator: out[b][o] = sum_i w(o, i) ktab[key(o, i)] in[b][i] over
// |i - o| <= aper; pkey / pw are [n_o][nb] (nb = 2 aper + 1, key -1 where
// the pair adds nothing); strides in float2 elements.
struct BandArgs {
    const float2 *in;
    float2 *out;
    const float2 *ktab;
    const int *pkey;
    const float *pw;
    int n_o, n_i, n_b, nb, aper, nw;
    size_t in_b, in_i, out_b, out_o;
    int super_o = 16, super_b = 8;  // tile grouping of the VALU kernel (1 x 1: output-fastest)
};

struct PairRec {
    int tr;   // input trace
    int key;  // table key
    float w;  // geometric weight (taper, obliquity, spreading)
    int pad;
};

// Pair-list operator: out[o] = sum_q w ktab[key] in[tr] over rowptr[o] ..
// rowptr[o + 1].
struct PairArgs {
    const float2 *in;
    float2 *out;
    const float2 *ktab;
    const int *rowptr;
    const PairRec *pairs;
    int nw;
};

int launch_pad(const float *data, const float *wtr, float *pad, int tr0, int ntr, int nt, int nfft,
               void *stream);
int launch_post(const float *pad, float *out, int tr0, int ntr, int nt, int nfft, int reverse,
                void *stream);
int launch_zero2(float2 *p, size_t n, void *stream);
int launch_ktab(float2 *ktab, const float *tau, const float *aarg, int nkey, int nw, int nfft,
                float dt, int nsam, float antialias, void *stream);
int launch_band(const BandArgs &a, void *stream);
// the same operator on the matrix cores (v_mfma_f32_32x32x2f32)
int launch_band_mfma(const BandArgs &a, void *stream);
int launch_pairs(const PairArgs &a, int n_out, void *stream);

}  // namespace kmdat
