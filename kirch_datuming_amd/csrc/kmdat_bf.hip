// Segmented butterfly of kirchdat2d_freq (use_bf=1), after the exact sums
// of an operator have been written: the banded (cmp=0) operators through
// difference blocks shared by all batch vectors, the pair-list (cmp=1) ones
// through k_bf_pairs at the end of this file.
//
// The reference evaluates every frequency's operator as exact leaf panels
// plus butterfly panels (SERECKIRCH/src/bf1d.cpp:1002-1033).  The exact
// panels hold the operator's own entries, which the banded kernel has
// already summed; on a butterfly panel P the reference applies the factor
// B_P instead of the entries K_P.  So the butterfly's output is the banded
// output plus (B_P - K_P) u_P over the butterfly panels, which is what the
// four kernels below add:
//   k_bf_mask    per (panel, 64-frequency chunk): which frequencies make the
//                panel a butterfly (fully active entries, residual passed,
//                no containing panel a butterfly itself) -- one bit a lane;
//   k_bf_switch  the switch kernel of each (panel, chunk) (bf1d.cpp:405-418):
//                amplitude and traveltime interpolated bilinearly;
//   k_bf_dense   B_P - K_P column by column: the factor's stages applied to
//                the unit vector of the column (bf1d.cpp:522-645), the
//                coefficients of a level in LDS;
//   k_bf_corr    out += (B_P - K_P) u_P for every batch vector, the panels of
//                overlapping rows in one pass so no two waves update a row.
// Lanes are the 64 frequencies of a chunk throughout, as in the banded
// kernel: spectrum, table and difference loads are 512-byte rows.  Built
// without contraction: the phases (float)(alpha t) come from double
// products as the reference forms them, the stage sums in float.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "kmdat_internal.h"

namespace kmdat {
namespace bf {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

__device__ __forceinline__ void cadd(float2 &a, float2 b)
{
    a.x += b.x;
    a.y += b.y;
}

// exp(i (float)(alpha t))
__device__ __forceinline__ float2 phase(double alpha, double t)
{
    const float f = (float)(alpha * t);
    float s, c;
    sincosf(f, &s, &c);
    return make_float2(c, s);
}

// Table key and weight of entry (r, c) of a node: from the band (cmp=0) or
// from the node's rows of the pair list (cmp=1).
template <bool Pairs>
__device__ __forceinline__ int entry_key(const Args &a, const Node &nd, int r, int c, float &w)
{
    if (Pairs) {
        const PairRec q = a.pairs[a.node_pair0[nd.rows_off + r] + c];
        w = q.w;
        return q.key;
    }
    const int row = nd.row0 + r, col = nd.col0 + c;
    const size_t e = (size_t)row * a.nb + (col - row + a.aper);
    w = a.pw[e];
    return a.pkey[e];
}

// Every entry of the panel active at this lane's frequency (bf1d.cpp:703-727
// and kirchdat2d_frequency_cs.cpp:731-736: |w aa| >= 1e-20, |amp|^2 >
// amp_eps^2).  The planner has checked that every entry has a pair.
template <bool Pairs>
__device__ bool panel_full(const Args &a, const Node &nd, int iw, bool live)
{
    const int nloc = a.leaf << nd.L;
    bool act = live;
    for (int r = 0; r < nloc; r++) {
        if (!__any(act)) break;
        for (int c = 0; c < nloc; c++) {
            float w;
            const int key = entry_key<Pairs>(a, nd, r, c, w);
            const float2 kt = a.ktab[(size_t)key * a.nw + iw];
            const float aa = antialias_weight(a.antialias, a.aarg[key], iw, a.nfft, a.dt);
            const float ar = w * kt.x, ai = w * kt.y;
            act = act && !(fabsf(w * aa) < 1e-20f) && (ar * ar + ai * ai > a.eps2);
        }
    }
    return act;
}

__global__ __launch_bounds__(64) void k_bf_mask(const Args a)
{
    const int item = blockIdx.x, lane = threadIdx.x;
    const Item it = a.items[item];
    const Node nd = a.nodes[it.node];
    const int iw = it.chunk * 64 + lane;
    const bool ok = iw < a.nw;
    const int iwc = ok ? iw : a.nw - 1;
    bool bfly = panel_full<false>(a, nd, iwc, ok && iw <= nd.iw_res);
    for (int q = 0; q < nd.nanc; q++) {
        const Node an = a.nodes[nd.anc[q]];
        bfly = bfly && !panel_full<false>(a, an, iwc, bfly && iw <= an.iw_res);
    }
    const unsigned long long m = __ballot(bfly);
    if (lane == 0) a.mask[item] = m;
}

__global__ __launch_bounds__(64) void k_bf_switch(const Args a)
{
    const int item = blockIdx.x, lane = threadIdx.x;
    if (a.mask[item] == 0) return;
    const Item it = a.items[item];
    const Node nd = a.nodes[it.node];
    const int L = nd.L, p = a.p, nloc = a.leaf << L;
    const int iw = min(it.chunk * 64 + lane, a.nw - 1);
    const float om = a.omega[iw];
    const double alpha = -(double)om;
    const TauLayout lay = tau_layout(L, p, a.leaf);
    const double *T = a.tau + nd.tau_off;
    const SwSample *S = a.sws + a.sws_off[L];
    const int ns = (1 << L) * p * p;
    for (int s = 0; s < ns; s++) {
        double re = 0.0, im = 0.0;
        for (int k = 0; k < 4; k++) {
            const int e = S[s].e[k];
            float w;
            const int key = entry_key<false>(a, nd, e / nloc, e % nloc, w);
            const float2 kt = a.ktab[(size_t)key * a.nw + iw];
            // amp = w aa H exp(i omega tau) (kirchdat2d_frequency_cs.cpp:737-745)
            const float ph = om * a.tau_key[key];
            float sp, cp;
            sincosf(ph, &sp, &cp);
            const float ar = w * (kt.x * cp - kt.y * sp), ai = w * (kt.x * sp + kt.y * cp);
            re += S[s].w[k] * (double)ar;
            im += S[s].w[k] * (double)ai;
        }
        const float2 amp = make_float2((float)re, (float)im);
        a.sw[it.sw_off + (size_t)s * 64 + lane] = cmul(amp, phase(alpha, T[lay.sw + s]));
    }
}

// LDS per wave: two levels of coefficients and one row of products, each
// [index][64 lanes].
__global__ __launch_bounds__(64) void k_bf_dense(const Args a, int ncol)
{
    extern __shared__ float2 lds[];
    const int item = blockIdx.x / ncol, k = blockIdx.x % ncol, lane = threadIdx.x;
    const Item it = a.items[item];
    const Node nd = a.nodes[it.node];
    const int L = nd.L, p = a.p, leaf = a.leaf, nloc = leaf << L, mid = L / 2;
    const unsigned long long m = a.mask[item];
    if (k >= nloc || m == 0) return;  // the correction skips items without butterfly lanes
    float2 *D = a.d + it.d_off;
    const int iw = it.chunk * 64 + lane;
    const bool live = iw < a.nw && ((m >> lane) & 1ull);
    const int iwc = min(iw, a.nw - 1);
    const int S = (1 << L) * p;
    float2 *A = lds, *B = lds + (size_t)S * 64, *tmp = lds + (size_t)2 * S * 64;
#define SG(arr, i) arr[(size_t)(i) * 64 + lane]
    const double alpha = -(double)a.omega[iwc];
    const TauLayout t = tau_layout(L, p, leaf);
    const LagLayout g = lag_layout(L, p, leaf);
    const double *T = a.tau + nd.tau_off;
    const float *Lg = a.lag + a.lag_off[L];
    const int bsL = k / leaf, kk = k % leaf;

    // leaf initialisation of the unit vector e_k: only its source box is
    // nonzero, and the single term of the sum is exact
    {
        const float2 pp = phase(alpha, T[t.inpt + k]);
        for (int i = 0; i < p; i++)
            SG(A, i) = cmul(cscale(pp, Lg[g.in + (bsL * leaf + kk) * p + i]),
                            phase(-alpha, T[t.innode + bsL * p + i]));
    }
    // source-side levels: the column's box at each level, all target boxes
    for (int lv = 1; lv <= mid; lv++) {
        const int nr = 1 << lv, cs = bsL >> (lv - 1), bs = cs >> 1;
        const int nsc = 1 << (L - lv + 1), nsp = nsc >> 1;
        for (int br = 0; br < nr; br++) {
            for (int j = 0; j < p; j++)
                SG(tmp, j) = cmul(SG(A, (br >> 1) * p + j),
                                  phase(alpha, T[t.r1in[lv] + (br * nsc + cs) * p + j]));
            for (int i = 0; i < p; i++) {
                float2 s = make_float2(0.0f, 0.0f);
                for (int j = 0; j < p; j++)
                    cadd(s, cscale(SG(tmp, j), Lg[g.r1[lv] + (cs * p + j) * p + i]));
                SG(B, br * p + i) = cmul(s, phase(-alpha, T[t.r1out[lv] + (br * nsp + bs) * p + i]));
            }
        }
        float2 *x = A;
        A = B;
        B = x;
    }
    // switch
    {
        const int nr = 1 << mid, bs = bsL >> mid;
        const float2 *W = a.sw + it.sw_off;
        for (int br = 0; br < nr; br++)
            for (int i = 0; i < p; i++) {
                float2 s = make_float2(0.0f, 0.0f);
                const size_t w0 = ((size_t)((bs * nr + br) * p + i) * p) * 64 + lane;
                for (int j = 0; j < p; j++) cadd(s, cmul(W[w0 + (size_t)j * 64], SG(A, br * p + j)));
                SG(B, br * p + i) = s;
            }
        float2 *x = A;
        A = B;
        B = x;
    }
    // target-side levels: only the column's source child is nonzero
    for (int lv = mid + 1; lv <= L; lv++) {
        const int nr = 1 << lv, cs = bsL >> (lv - 1), nsc = 1 << (L - lv + 1);
        for (int br = 0; br < nr; br++) {
            for (int j = 0; j < p; j++)
                SG(tmp, j) = cmul(SG(A, (br >> 1) * p + j),
                                  phase(-alpha, T[t.r2in[lv] + (br * nsc + cs) * p + j]));
            for (int i = 0; i < p; i++) {
                float2 y = make_float2(0.0f, 0.0f);
                for (int j = 0; j < p; j++)
                    cadd(y, cscale(SG(tmp, j), Lg[g.r2[lv] + (br * p + i) * p + j]));
                SG(B, br * p + i) = cmul(y, phase(alpha, T[t.r2out[lv] + (br * nsc + cs) * p + i]));
            }
        }
        float2 *x = A;
        A = B;
        B = x;
    }
    // termination, minus the exact entry the banded kernel has summed
    for (int br = 0; br < (1 << L); br++) {
        for (int i = 0; i < p; i++)
            SG(tmp, i) = cmul(SG(A, br * p + i), phase(-alpha, T[t.outnode + br * p + i]));
        for (int q = 0; q < leaf; q++) {
            const int r = br * leaf + q;
            float2 s = make_float2(0.0f, 0.0f);
            for (int i = 0; i < p; i++) cadd(s, cscale(SG(tmp, i), Lg[g.out + (br * leaf + q) * p + i]));
            const float2 b = cmul(s, phase(alpha, T[t.outpt + r]));
            float w;
            const int key = entry_key<false>(a, nd, r, k, w);
            const float2 kt = a.ktab[(size_t)key * a.nw + iwc];
            const float2 d = make_float2(b.x - w * kt.x, b.y - w * kt.y);
            D[((size_t)r * nloc + k) * 64 + lane] = live ? d : make_float2(0.0f, 0.0f);
        }
    }
#undef SG
}

// out[b][row] += sum_c D[row][c] u[b][col0 + c] over the items of a group.
// A block: four waves of TB batch vectors, one quad of rows, one chunk.
template <int TB>
__global__ __launch_bounds__(256) void k_bf_corr(const Args a)
{
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int nbb = (a.n_b + 4 * TB - 1) / (4 * TB);
    const int qi = (int)blockIdx.x / nbb, bb = (int)blockIdx.x % nbb;
    const Quad Q = a.quads[qi];
    const int g = Q.g, rq = 4 * Q.q, r0 = a.grp_row0[g] + rq, rows = a.grp_rows[g];
    const int iw = a.grp_chunk[g] * 64 + lane;
    const bool ok = iw < a.nw;
    const int iwc = ok ? iw : a.nw - 1;
    const int b0 = (bb * 4 + wave) * TB;
    float2 acc[TB][4];
#pragma unroll
    for (int tb = 0; tb < TB; tb++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[tb][t] = make_float2(0.0f, 0.0f);
    for (int e = a.grp_ptr[g]; e < a.grp_ptr[g + 1]; e++) {
        const int item = a.grp[e];
        if (a.mask[item] == 0) continue;
        const Item it = a.items[item];
        const Node nd = a.nodes[it.node];
        const int nloc = a.leaf << nd.L, rl0 = r0 - nd.row0;
        if (rl0 + 3 < 0 || rl0 >= nloc) continue;
        const float2 *D = a.d + it.d_off;
        for (int c = 0; c < nloc; c++) {
            float2 u[TB];
#pragma unroll
            for (int tb = 0; tb < TB; tb++) {
                const int b = min(b0 + tb, a.n_b - 1);
                u[tb] = a.in[(size_t)b * a.in_b + (size_t)(nd.col0 + c) * a.in_i + iwc];
            }
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int rl = rl0 + t;
                if (rl < 0 || rl >= nloc) continue;
                const float2 dv = D[((size_t)rl * nloc + c) * 64 + lane];
#pragma unroll
                for (int tb = 0; tb < TB; tb++) {
                    acc[tb][t].x += dv.x * u[tb].x - dv.y * u[tb].y;
                    acc[tb][t].y += dv.x * u[tb].y + dv.y * u[tb].x;
                }
            }
        }
    }
    if (!ok) return;
#pragma unroll
    for (int tb = 0; tb < TB; tb++) {
        const int b = b0 + tb;
        if (b >= a.n_b) continue;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (rq + t >= rows) continue;
            float2 *o = a.out + (size_t)b * a.out_b + (size_t)(r0 + t) * a.out_o + iw;
            float2 v = *o;
            v.x += acc[tb][t].x;
            v.y += acc[tb][t].y;
            *o = v;
        }
    }
}


// Switch-kernel entry (bs, br, it, jt) of a node at this lane's frequency:
// the bilinear amplitude sample times exp(i alpha t~) (bf1d.cpp:405-418).
template <bool Pairs>
__device__ float2 switch_entry(const Args &a, const Node &nd, int s, float om, double alpha,
                               const double *T, const TauLayout &lay, int iw)
{
    const int nloc = a.leaf << nd.L;
    const SwSample *S = a.sws + a.sws_off[nd.L] + s;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < 4; k++) {
        const int e = S->e[k];
        float w;
        const int key = entry_key<Pairs>(a, nd, e / nloc, e % nloc, w);
        const float2 kt = a.ktab[(size_t)key * a.nw + iw];
        const float ph = om * a.tau_key[key];
        float sp, cp;
        sincosf(ph, &sp, &cp);
        const float ar = w * (kt.x * cp - kt.y * sp), ai = w * (kt.x * sp + kt.y * cp);
        re += S->w[k] * (double)ar;
        im += S->w[k] * (double)ai;
    }
    return cmul(make_float2((float)re, (float)im), phase(alpha, T[lay.sw + s]));
}

// Pair-list operators (cmp=1): each node's matrix is applied to one vector,
// so the factor runs on that vector (bf1d.cpp:522-645) -- every source box
// of every level live -- and (B_P - K_P) u_P is added to the node's output
// traces.  One wave per group; the group's items run one after the other, so
// an output trace is only ever updated by one wave.
__global__ __launch_bounds__(64) void k_bf_pairs(const Args a)
{
    extern __shared__ float2 lds[];
    const int g = blockIdx.x, lane = threadIdx.x;
#define SG(arr, i) arr[(size_t)(i) * 64 + lane]
    for (int e = a.grp_ptr[g]; e < a.grp_ptr[g + 1]; e++) {
        const Item it = a.items[a.grp[e]];
        const Node nd = a.nodes[it.node];
        const int iw = it.chunk * 64 + lane;
        const bool ok = iw < a.nw;
        const int iwc = ok ? iw : a.nw - 1;
        bool bfly = panel_full<true>(a, nd, iwc, ok && iw <= nd.iw_res);
        for (int q = 0; q < nd.nanc; q++) {
            const Node an = a.nodes[nd.anc[q]];
            bfly = bfly && !panel_full<true>(a, an, iwc, bfly && iw <= an.iw_res);
        }
        if (!__any(bfly)) continue;
        const int L = nd.L, p = a.p, leaf = a.leaf, nloc = leaf << L, mid = L / 2,
                  nbx = 1 << L;
        const int S = nbx * p;
        float2 *A = lds, *B = lds + (size_t)S * 64, *tmp = lds + (size_t)2 * S * 64;
        const float om = a.omega[iwc];
        const double alpha = -(double)om;
        const TauLayout t = tau_layout(L, p, leaf);
        const LagLayout lg = lag_layout(L, p, leaf);
        const double *T = a.tau + nd.tau_off;
        const float *Lg = a.lag + a.lag_off[L];
        const int q0 = a.node_pair0[nd.rows_off];  // column c's input trace: pairs[q0 + c].tr
        auto U = [&](int c) { return a.in[(size_t)a.pairs[q0 + c].tr * a.nw + iwc]; };

        // leaf initialisation, every source leaf box
        for (int bs = 0; bs < nbx; bs++) {
            for (int i = 0; i < p; i++) SG(A, bs * p + i) = make_float2(0.0f, 0.0f);
            for (int k = 0; k < leaf; k++) {
                const int c = bs * leaf + k;
                const float2 v = cmul(U(c), phase(alpha, T[t.inpt + c]));
                for (int i = 0; i < p; i++) {
                    float2 s = SG(A, bs * p + i);
                    cadd(s, cscale(v, Lg[lg.in + (bs * leaf + k) * p + i]));
                    SG(A, bs * p + i) = s;
                }
            }
            for (int i = 0; i < p; i++)
                SG(A, bs * p + i) = cmul(SG(A, bs * p + i), phase(-alpha, T[t.innode + bs * p + i]));
        }
        // source-side levels: coefficients [source box][target box][node]
        for (int lv = 1; lv <= mid; lv++) {
            const int nr = 1 << lv, nrp = nr >> 1, nsp = 1 << (L - lv), nsc = 2 * nsp;
            for (int br = 0; br < nr; br++)
                for (int bs = 0; bs < nsp; bs++) {
                    const int o = (bs * nr + br) * p;
                    for (int i = 0; i < p; i++) SG(B, o + i) = make_float2(0.0f, 0.0f);
                    for (int ch = 0; ch < 2; ch++) {
                        const int cs = 2 * bs + ch;
                        for (int j = 0; j < p; j++) {
                            const float2 w = cmul(SG(A, (cs * nrp + (br >> 1)) * p + j),
                                                  phase(alpha, T[t.r1in[lv] + (br * nsc + cs) * p + j]));
                            for (int i = 0; i < p; i++) {
                                float2 s = SG(B, o + i);
                                cadd(s, cscale(w, Lg[lg.r1[lv] + (cs * p + j) * p + i]));
                                SG(B, o + i) = s;
                            }
                        }
                    }
                    for (int i = 0; i < p; i++)
                        SG(B, o + i) = cmul(SG(B, o + i),
                                            phase(-alpha, T[t.r1out[lv] + (br * nsp + bs) * p + i]));
                }
            float2 *x = A;
            A = B;
            B = x;
        }
        // switch, every (source box, target box) of the middle level
        {
            const int nr = 1 << mid, nsp = 1 << (L - mid);
            for (int bs = 0; bs < nsp; bs++)
                for (int br = 0; br < nr; br++) {
                    const int o = (bs * nr + br) * p;
                    for (int i = 0; i < p; i++) {
                        float2 s = make_float2(0.0f, 0.0f);
                        for (int j = 0; j < p; j++)
                            cadd(s, cmul(switch_entry<true>(a, nd, ((bs * nr + br) * p + i) * p + j,
                                                            om, alpha, T, t, iwc),
                                         SG(A, o + j)));
                        SG(B, o + i) = s;
                    }
                }
            float2 *x = A;
            A = B;
            B = x;
        }
        // target-side levels
        for (int lv = mid + 1; lv <= L; lv++) {
            const int nr = 1 << lv, nrp = nr >> 1, nsp = 1 << (L - lv), nsc = 2 * nsp;
            for (int br = 0; br < nr; br++) {
                const float *Lt = Lg + lg.r2[lv] + br * p * p;
                for (int bs = 0; bs < nsp; bs++) {
                    const int o = (bs * nr + br) * p;
                    for (int i = 0; i < p; i++) SG(B, o + i) = make_float2(0.0f, 0.0f);
                    for (int ch = 0; ch < 2; ch++) {
                        const int cs = 2 * bs + ch;
                        for (int j = 0; j < p; j++)
                            SG(tmp, j) = cmul(SG(A, (cs * nrp + (br >> 1)) * p + j),
                                              phase(-alpha, T[t.r2in[lv] + (br * nsc + cs) * p + j]));
                        for (int i = 0; i < p; i++) {
                            float2 y = make_float2(0.0f, 0.0f);
                            for (int j = 0; j < p; j++) cadd(y, cscale(SG(tmp, j), Lt[i * p + j]));
                            float2 s = SG(B, o + i);
                            cadd(s, cmul(y, phase(alpha, T[t.r2out[lv] + (br * nsc + cs) * p + i])));
                            SG(B, o + i) = s;
                        }
                    }
                }
            }
            float2 *x = A;
            A = B;
            B = x;
        }
        // termination, minus the exact entries k_pairs has summed
        for (int br = 0; br < nbx; br++) {
            for (int i = 0; i < p; i++)
                SG(tmp, i) = cmul(SG(A, br * p + i), phase(-alpha, T[t.outnode + br * p + i]));
            for (int k = 0; k < leaf; k++) {
                const int r = br * leaf + k;
                float2 s = make_float2(0.0f, 0.0f);
                for (int i = 0; i < p; i++) cadd(s, cscale(SG(tmp, i), Lg[lg.out + (br * leaf + k) * p + i]));
                const float2 bu = cmul(s, phase(alpha, T[t.outpt + r]));
                float2 ku = make_float2(0.0f, 0.0f);
                for (int c = 0; c < nloc; c++) {
                    float w;
                    const int key = entry_key<true>(a, nd, r, c, w);
                    const float2 kt = a.ktab[(size_t)key * a.nw + iwc];
                    const float kr = w * kt.x, ki = w * kt.y;
                    const float2 u = U(c);
                    ku.x += kr * u.x - ki * u.y;
                    ku.y += kr * u.y + ki * u.x;
                }
                if (bfly) {
                    float2 *y = a.out + (size_t)a.node_out[nd.rows_off + r] * a.nw + iw;
                    float2 v = *y;
                    v.x += bu.x - ku.x;
                    v.y += bu.y - ku.y;
                    *y = v;
                }
            }
        }
    }
#undef SG
}

}  // namespace bf

int launch_bf(const bf::Args &a, void *stream)
{
    using namespace bf;
    hipStream_t st = (hipStream_t)stream;
    if (a.nitems <= 0) return 0;
    hipLaunchKernelGGL(k_bf_mask, dim3((unsigned)a.nitems), dim3(64), 0, st, a);
    hipLaunchKernelGGL(k_bf_switch, dim3((unsigned)a.nitems), dim3(64), 0, st, a);
    if (a.maxL < 1 || a.maxL > MAX_L || ((1 << a.maxL) * a.p) > MAX_SIGMA)
        return (int)hipErrorInvalidValue;
    const int ncol = a.leaf << a.maxL;
    const size_t lds = (size_t)(2 * (1 << a.maxL) * a.p + a.p) * 64 * sizeof(float2);
    const size_t nblk = (size_t)a.nitems * ncol;
    if (nblk > 0x7fffffffu || lds > 64 * 1024) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bf_dense, dim3((unsigned)nblk), dim3(64), lds, st, a, ncol);
    constexpr int TB = 8;
    const size_t nbb = (size_t)(a.n_b + 4 * TB - 1) / (4 * TB);
    const size_t ncorr = (size_t)a.nquads * nbb;
    if (ncorr > 0x7fffffffu) return (int)hipErrorInvalidValue;
    if (ncorr) hipLaunchKernelGGL((k_bf_corr<TB>), dim3((unsigned)ncorr), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

int launch_bf_pairs(const bf::Args &a, void *stream)
{
    using namespace bf;
    if (a.nitems <= 0 || a.ngroup <= 0) return 0;
    if (a.maxL < 1 || a.maxL > MAX_L || ((1 << a.maxL) * a.p) > MAX_SIGMA)
        return (int)hipErrorInvalidValue;
    const size_t lds = (size_t)(2 * (1 << a.maxL) * a.p + a.p) * 64 * sizeof(float2);
    if (lds > 64 * 1024) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bf_pairs, dim3((unsigned)a.ngroup), dim3(64), lds, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

}  // namespace kmdat
