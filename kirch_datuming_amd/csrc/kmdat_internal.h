// Internal declarations shared by the datuming host code and its kernels.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "kmdat_bf.h"

namespace kmdat {

// Largest filter length in samples (length / dt + 2) the table kernel takes.
constexpr int KMDAT_MAX_NSAM = 4096;

#if defined(__HIPCC__)
// antialias_weight_iw (src/kirchdat2d_frequency_cs.cpp:203-233) at bin iw for
// a key whose traveltime slope times the summation step is aarg.
__device__ __forceinline__ float antialias_weight(float antialias, float aarg, int iw, int nfft,
                                                  float dt)
{
    if (!(antialias > 0.0f) || iw <= 0) return 1.0f;
    const float dtau = fabsf(aarg) * antialias;
    if (!(dtau > 1e-12f) || !isfinite(dtau)) return 1.0f;
    const float freq = (float)iw / ((float)nfft * dt);
    const float fnyq = 0.5f / dt;
    const float fmax = 0.5f / dtau;
    if (fmax >= fnyq) return 1.0f;
    if (fmax <= 0.0f) return 0.0f;
    const float fpass = 0.8f * fmax;
    if (freq <= fpass) return 1.0f;
    if (freq >= fmax) return 0.0f;
    const float x = (freq - fpass) / (fmax - fpass);
    return 0.5f * (1.0f + cosf(3.14159265358979323846f * x));
}
#endif

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

// ------------------------------------------------ butterfly (use_bf=1) --

// One operator as the butterfly planner reads it: banded (pkey / pw, one
// matrix) or pair lists (pairs, ninst matrices whose row r of instance k is
// output trace row_o[k n + r] (-1: none) with pairs [row_q0, row_q1) in
// increasing column; a pair's column is tr / col_base or tr % col_base).
struct BfPlanIn {
    int n;                   // matrix size (outputs = summation indices)
    int nb, aper;            // band of pkey / pw: [n][nb], column c = r + j - aper
    const int *pkey;
    const float *pw;
    const PairRec *pairs = nullptr;
    const int *row_o = nullptr, *row_q0 = nullptr, *row_q1 = nullptr;
    int ninst = 1, col_base = 1, col_by_div = 0;
    const float *tau_key, *aarg;  // per table key
    int nkey;
    int nw, nfft;
    float dt, antialias;
    int p, leaf, levels;     // bf_p, bf_n_leaf, bf_panel_levels
    float phase_tol;         // bf_phase_tol after the reference's default
};

struct BfPlan {
    std::vector<bf::Node> nodes;
    std::vector<double> tau;
    std::vector<bf::Item> items;
    std::vector<int> grp_ptr, grp_row0, grp_rows, grp_chunk, grp;
    std::vector<bf::Quad> quads;
    std::vector<float> lag;
    std::vector<bf::SwSample> sws;
    std::vector<float> omega;
    std::vector<int> node_pair0, node_out;   // pair-list operators
    int lag_off[bf::MAX_L + 1] = {}, sws_off[bf::MAX_L + 1] = {};
    long long d_elems = 0, sw_elems = 0;  // float2 elements of the work buffers
    int maxL = 0;
};

int bf_plan(const BfPlanIn &in, BfPlan &out, std::string &err);

// the four butterfly steps of one operator (kmdat_bf.hip), after its banded
// sums: masks, switch kernels, difference blocks, correction of `out`
int launch_bf(const bf::Args &a, void *stream);
// the same for a pair-list operator after k_pairs: one fused pass per group
int launch_bf_pairs(const bf::Args &a, void *stream);

}  // namespace kmdat
