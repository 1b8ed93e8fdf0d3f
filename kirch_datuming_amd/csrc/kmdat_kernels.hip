// Device kernels of the frequency-domain Kirchhoff datuming (kirchdat2d_freq).
//
// Spectra are stored trace-major, frequency fastest: U[trace][iw], float2
// (re, im), trace = is * nh + ih.  Every kernel puts 64 consecutive
// frequencies on the 64 lanes of a wave, so each load and store of a
// spectrum row is one coalesced 512-byte access; everything that depends
// only on the summation pair (indices, weights, table keys) is wave-uniform.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "kmdat_internal.h"

namespace kmdat {

#define KPI_F 3.14159265358979323846f

// Consecutive block ids go to the 8 XCDs in turn; remap so that each XCD
// walks a contiguous range of tiles (neighbouring tiles share inputs in L2).
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    constexpr int NX = 8;
    if (nb < NX) return b;
    const int x = b % NX, k = b / NX;
    const int per = nb / NX, extra = nb % NX;
    return x * per + min(x, extra) + k;
}

// ---------------------------------------------------------------- traces --

// read_traces_to_freq (src/kirchdat2d_frequency_cs.cpp:281-313) for traces
// [tr0, tr0 + ntr): time reversal, the data-edge taper weight of the trace
// and zero padding to nfft samples.
__global__ void k_pad(const float *__restrict__ data, const float *__restrict__ wtr,
                      float *__restrict__ pad, int tr0, int ntr, int nt, int nfft)
{
    const size_t n = (size_t)ntr * nfft;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / nfft), k = (int)(e % nfft);
        float v = 0.0f;
        if (k < nt) v = wtr[tr0 + t] * data[(size_t)(tr0 + t) * nt + (nt - 1 - k)];
        pad[e] = v;
    }
}

// write_traces_from_freq (:470-491): the first nt samples of the inverse
// transform divided by nfft, reversed when asked.
__global__ void k_post(const float *__restrict__ pad, float *__restrict__ out, int tr0, int ntr,
                       int nt, int nfft, int reverse)
{
    const size_t n = (size_t)ntr * nt;
    const float fn = (float)nfft;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / nt), it = (int)(e % nt);
        const float v = pad[(size_t)t * nfft + it] / fn;
        out[(size_t)(tr0 + t) * nt + (reverse ? nt - 1 - it : it)] = v;
    }
}

__global__ void k_zero2(float2 *__restrict__ p, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        p[e] = make_float2(0.0f, 0.0f);
}

// -------------------------------------------------------- operator table --

// One row per table key (a (summation, output) pair of table positions):
// ktab[key][iw] = antialias(iw, slope) * H(tau, iw), with H the spectrum of
// the datuming filter delayed by tau (build_filter_F, build_H_iw, :100-184)
// and the antialias weight (:203-233) of the key's traveltime slope.  Keys
// with tau <= 0 or non-finite tau have H = 0.  The filter taps of the key
// are formed once per block in LDS.
__global__ __launch_bounds__(256) void k_ktab(float2 *__restrict__ ktab,
                                              const float *__restrict__ tau,
                                              const float *__restrict__ aarg, int nkey, int nw,
                                              int nchunk, int nfft, float dt, int nsam,
                                              float antialias)
{
    __shared__ float Fk[KMDAT_MAX_NSAM];
    const int key = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    const float t = tau[key];
    const bool live = t > 0.0f && isfinite(t);
    const int ntap = nsam - 1;
    // F[k] = f(k+1) - 2 f(k) + f(k-1), f(0) = 0, f(m) = sqrt(((tau + m dt)/tau)^2 - 1)
    if (live) {
        for (int k = threadIdx.x; k < ntap; k += blockDim.x) {
            auto f = [&](int m) -> float {
                if (m == 0) return 0.0f;
                const float q = (t + (float)m * dt) / t;
                return sqrtf(q * q - 1.0f);
            };
            Fk[k] = k == 0 ? f(1) - 0.0f : f(k + 1) - 2.0f * f(k) + f(k - 1);
        }
    }
    __syncthreads();
    const int iw = chunk * blockDim.x + threadIdx.x;
    if (iw >= nw) return;
    float2 K = make_float2(0.0f, 0.0f);
    if (live) {
        const float omega = 2.0f * KPI_F * (float)iw / (float)nfft;
        float step_i, step_r;
        sincosf(-omega, &step_i, &step_r);
        float z_r = 1.0f, z_i = 0.0f, sum_r = 0.0f, sum_i = 0.0f;
        for (int k = 0; k < ntap; k++) {
            const float fk = Fk[k] / dt;
            sum_r += fk * z_r;
            sum_i += fk * z_i;
            const float nr = z_r * step_r - z_i * step_i;
            const float ni = z_r * step_i + z_i * step_r;
            z_r = nr;
            z_i = ni;
        }
        const float q = t / dt;
        const int m = (int)ceilf(q);
        const float delta = (float)m - q;
        float sin_m, cos_m, sin_m1, cos_m1;
        sincosf(-omega * (float)m, &sin_m, &cos_m);
        sincosf(-omega * (float)(m - 1), &sin_m1, &cos_m1);
        const float lin_r = (1.0f - delta) * cos_m + delta * cos_m1;
        const float lin_i = (1.0f - delta) * sin_m + delta * sin_m1;
        K.x = lin_r * sum_r - lin_i * sum_i;
        K.y = lin_r * sum_i + lin_i * sum_r;
        const float aa = antialias_weight(antialias, aarg[key], iw, nfft, dt);
        K.x *= aa;
        K.y *= aa;
    }
    ktab[(size_t)key * nw + iw] = K;
}

// ------------------------------------------------------- banded operator --

// out[b][o] = sum_{|i - o| <= aper} w(o, i) ktab[key(o, i)] in[b][i], for
// operators whose pairs do not depend on the batch index b (cmp=0: the
// receiver side sums receivers ic of one shot, b = shot; the source side sums
// shots of one receiver, b = receiver).  A wave holds a TB x TO register tile
// of (batch, output) at its 64 frequencies: per summation index it loads TB
// input rows and, per output whose band holds the index, one table row.  The
// block's four waves split WB ways along the batch tile (same 64
// frequencies, so their table-row loads are the same lines and hit L1) and
// 4 / WB ways along frequency.  Tiles run output-fastest, so consecutive
// tiles on one XCD slide along the same input rows.
template <int TB, int TO, int WB>
__global__ __launch_bounds__(256) void k_band(const BandArgs a)
{
    constexpr int WF = 4 / WB;
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int wb = wave % WB, wf = wave / WB;
    const int nbt = (a.n_b + TB * WB - 1) / (TB * WB), not_ = (a.n_o + TO - 1) / TO;
    const int tile = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    // super-tiles of so output tiles x sb batch tiles (output tiles fastest
    // inside): the blocks an XCD runs together share table rows and input rows
    const int so = a.super_o, sb = a.super_b;
    const int nog = (not_ + so - 1) / so, nbg = (nbt + sb - 1) / sb;
    const int per_chunk = nog * nbg * so * sb;
    const int chunk = tile / per_chunk, r = tile % per_chunk;
    const int grp = r / (so * sb), w_ = r % (so * sb);
    const int ot = (grp % nog) * so + w_ % so, bt = (grp / nog) * sb + w_ / so;
    if (ot >= not_ || bt >= nbt) return;
    const int iw = (chunk * WF + wf) * 64 + lane;
    const bool ok = iw < a.nw;
    const int iwc = ok ? iw : a.nw - 1;  // keeps the tail lanes' loads in bounds
    const int o0 = ot * TO, b0 = (bt * WB + wb) * TB;
    float2 acc[TB][TO];
#pragma unroll
    for (int tb = 0; tb < TB; tb++)
#pragma unroll
        for (int to = 0; to < TO; to++) acc[tb][to] = make_float2(0.0f, 0.0f);
    const int i_lo = max(0, o0 - a.aper), i_hi = min(a.n_i - 1, o0 + TO - 1 + a.aper);
    for (int i = i_lo; i <= i_hi; i++) {
        float2 u[TB];
#pragma unroll
        for (int tb = 0; tb < TB; tb++) {
            const int b = min(b0 + tb, a.n_b - 1);
            u[tb] = a.in[(size_t)b * a.in_b + (size_t)i * a.in_i + iwc];
        }
#pragma unroll
        for (int to = 0; to < TO; to++) {
            const int o = o0 + to, j = i - o + a.aper;
            if (o >= a.n_o || j < 0 || j >= a.nb) continue;
            const int key = a.pkey[(size_t)o * a.nb + j];
            if (key < 0) continue;
            const float w = a.pw[(size_t)o * a.nb + j];
            const float2 k = a.ktab[(size_t)key * a.nw + iwc];
            const float kr = w * k.x, ki = w * k.y;
#pragma unroll
            for (int tb = 0; tb < TB; tb++) {
                acc[tb][to].x += kr * u[tb].x - ki * u[tb].y;
                acc[tb][to].y += kr * u[tb].y + ki * u[tb].x;
            }
        }
    }
    if (!ok) return;
#pragma unroll
    for (int tb = 0; tb < TB; tb++) {
        const int b = b0 + tb;
        if (b >= a.n_b) continue;
#pragma unroll
        for (int to = 0; to < TO; to++) {
            const int o = o0 + to;
            if (o < a.n_o) a.out[(size_t)b * a.out_b + (size_t)o * a.out_o + iw] = acc[tb][to];
        }
    }
}

// The banded operator on the matrix cores.  Per frequency the operator is a
// complex product C[b][o] = sum_i A[b][i] K[o][i]; a block owns a 32 (batch)
// x 32 (output) tile at MF = 16 consecutive frequencies and walks the
// summation index in stages of MK = 8.  Each stage brings the 32 input rows
// and the 32 pre-weighted table rows (w(o, i) ktab[key(o, i)], zero outside
// the band) of its 8 indices into LDS as [index][frequency][row] (rows
// fastest, padded to 33, so the operand reads of a wave and the staging
// writes are free of bank conflicts); two blocks per CU overlap one's
// staging with the other's products.  Wave w takes frequencies 4w .. 4w+3: per pair of indices and
// frequency four v_mfma_f32_32x32x2f32 form the real and imaginary parts
// (Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br), 32 accumulators per frequency.
constexpr int MF = 16, MB = 32, MO = 32, MK = 8, MPAD = 33;
typedef float floatx16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256, 2) void k_band_mfma(const BandArgs a)
{
    __shared__ float2 su[MK][MF][MPAD];
    __shared__ float2 sk[MK][MF][MPAD];
    const int t = (int)threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int sf = t % MF, sg = t / MF;  // staging: frequency, row group of 16
    const int nbt = (a.n_b + MB - 1) / MB, not_ = (a.n_o + MO - 1) / MO;
    const int tile = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int ot = tile % not_, rest = tile / not_;
    const int bt = rest % nbt, chunk = rest / nbt;
    const int o0 = ot * MO, b0 = bt * MB, f0 = chunk * MF;
    const int siw = min(f0 + sf, a.nw - 1);  // staging frequency, kept in bounds
    const int i_lo = max(0, o0 - a.aper);
    const int i_hi = min(a.n_i - 1, min(o0 + MO, a.n_o) - 1 + a.aper);

    floatx16 cr[4], ci[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            cr[q][j] = 0.0f;
            ci[q][j] = 0.0f;
        }
    const int row = lane % 32, kh = lane / 32;
    for (int is = i_lo; is <= i_hi; is += MK) {
        // stage: rows r = 16 p + sg of the 8 x 32 input and table rows
#pragma unroll 4
        for (int p = 0; p < (MK * MB) / 16; p++) {
            const int r = p * 16 + sg;
            const int ii = r / MB, rr = r % MB;
            const int i = is + ii;
            const int b = b0 + rr;
            float2 v = make_float2(0.0f, 0.0f);
            if (i <= i_hi && b < a.n_b) v = a.in[(size_t)b * a.in_b + (size_t)i * a.in_i + siw];
            su[ii][sf][rr] = v;
            const int o = o0 + rr, j = i - o + a.aper;
            float2 k = make_float2(0.0f, 0.0f);
            if (i <= i_hi && o < a.n_o && j >= 0 && j < a.nb) {
                const int key = a.pkey[(size_t)o * a.nb + j];
                if (key >= 0) {
                    const float w = a.pw[(size_t)o * a.nb + j];
                    const float2 kv = a.ktab[(size_t)key * a.nw + siw];
                    k = make_float2(w * kv.x, w * kv.y);
                }
            }
            sk[ii][sf][rr] = k;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < MK / 2; ks++) {
            const int ii = 2 * ks + kh;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int f = wave * 4 + q;
                const float2 av = su[ii][f][row];
                const float2 bv = sk[ii][f][row];
                cr[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x, bv.x, cr[q], 0, 0, 0);
                cr[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(-av.y, bv.y, cr[q], 0, 0, 0);
                ci[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x, bv.y, ci[q], 0, 0, 0);
                ci[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(av.y, bv.x, ci[q], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // accumulator j of lane l holds C[8 (j / 4) + 4 (l / 32) + j % 4][l % 32]
    const int o = o0 + row;
    if (o >= a.n_o) return;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int iw = f0 + wave * 4 + q;
        if (iw >= a.nw) continue;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int b = b0 + 8 * (j / 4) + 4 * kh + (j % 4);
            if (b < a.n_b)
                a.out[(size_t)b * a.out_b + (size_t)o * a.out_o + iw] = make_float2(cr[q][j], ci[q][j]);
        }
    }
}

// ---------------------------------------------------- pair-list operator --

// out[o] = sum_q w[q] ktab[key[q]] in[tr[q]], q over the pairs of output
// trace o (rowptr[o] .. rowptr[o+1]): the cmp=1 operators, whose pairs
// change with the shot (receiver side) or follow the common-receiver
// regrouping (source side).
__global__ __launch_bounds__(256) void k_pairs(const PairArgs a)
{
    const int nchunk = (a.nw + 255) / 256;
    const int o = (int)blockIdx.x / nchunk, chunk = (int)blockIdx.x % nchunk;
    const int iw = chunk * 256 + (int)threadIdx.x;
    if (iw >= a.nw) return;
    float2 acc = make_float2(0.0f, 0.0f);
    const int q0 = a.rowptr[o], q1 = a.rowptr[o + 1];
    for (int q = q0; q < q1; q++) {
        const PairRec p = a.pairs[q];
        const float2 k = a.ktab[(size_t)p.key * a.nw + iw];
        const float2 u = a.in[(size_t)p.tr * a.nw + iw];
        const float kr = p.w * k.x, ki = p.w * k.y;
        acc.x += kr * u.x - ki * u.y;
        acc.y += kr * u.y + ki * u.x;
    }
    a.out[(size_t)o * a.nw + iw] = acc;
}

// ------------------------------------------------------------- launchers --

static unsigned grid_for(size_t n, int block)
{
    size_t g = (n + block - 1) / block;
    if (g > 65536 * 4) g = 65536 * 4;
    return (unsigned)(g ? g : 1);
}

int launch_pad(const float *data, const float *wtr, float *pad, int tr0, int ntr, int nt, int nfft,
               void *stream)
{
    hipLaunchKernelGGL(k_pad, dim3(grid_for((size_t)ntr * nfft, 256)), dim3(256), 0,
                       (hipStream_t)stream, data, wtr, pad, tr0, ntr, nt, nfft);
    return (int)hipGetLastError();
}

int launch_post(const float *pad, float *out, int tr0, int ntr, int nt, int nfft, int reverse,
                void *stream)
{
    hipLaunchKernelGGL(k_post, dim3(grid_for((size_t)ntr * nt, 256)), dim3(256), 0,
                       (hipStream_t)stream, pad, out, tr0, ntr, nt, nfft, reverse);
    return (int)hipGetLastError();
}

int launch_zero2(float2 *p, size_t n, void *stream)
{
    hipLaunchKernelGGL(k_zero2, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, p, n);
    return (int)hipGetLastError();
}

int launch_ktab(float2 *ktab, const float *tau, const float *aarg, int nkey, int nw, int nfft,
                float dt, int nsam, float antialias, void *stream)
{
    if (nkey <= 0) return 0;
    if (nsam - 1 > KMDAT_MAX_NSAM) return (int)hipErrorInvalidValue;
    const int nchunk = (nw + 255) / 256;
    const size_t nblk = (size_t)nkey * nchunk;
    if (nblk > 0x7fffffffu) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ktab, dim3((unsigned)nblk), dim3(256), 0, (hipStream_t)stream, ktab,
                       tau, aarg, nkey, nw, nchunk, nfft, dt, nsam, antialias);
    return (int)hipGetLastError();
}

template <int TB, int TO, int WB>
static int launch_band_t(const BandArgs &a, hipStream_t st)
{
    constexpr int WF = 4 / WB;
    const int not_ = (a.n_o + TO - 1) / TO, nbt = (a.n_b + TB * WB - 1) / (TB * WB);
    const int so = a.super_o, sb = a.super_b;
    const size_t nblk = (size_t)((not_ + so - 1) / so) * ((nbt + sb - 1) / sb) * so * sb *
                        ((a.nw + 64 * WF - 1) / (64 * WF));
    if (nblk == 0) return 0;
    if (nblk > 0x7fffffffu) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL((k_band<TB, TO, WB>), dim3((unsigned)nblk), dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

int launch_band_mfma(const BandArgs &a, void *stream)
{
    const size_t nblk = (size_t)((a.n_o + MO - 1) / MO) * ((a.n_b + MB - 1) / MB) *
                        ((a.nw + MF - 1) / MF);
    if (nblk == 0) return 0;
    if (nblk > 0x7fffffffu) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_band_mfma, dim3((unsigned)nblk), dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int launch_band(const BandArgs &a, void *stream)
{
    // 8 x 4 tile per wave, the block's four waves along the batch (measured on
    // the 768 x 768 datuming workload: 148 ms per side, against 175 ms with the
    // waves along frequency, 265 ms with 8 x 8, 174 ms with 16 x 4, 207 ms for
    // an LDS-staged 32 x 32 block tile at two waves per SIMD, and 187 ms with
    // the summation loop unrolled twice, three waves per SIMD); super-tiles of
    // 16 x 8 tiles 145 ms (8 x 8: 146, 32 x 4: 149, output-fastest: 148)
    return launch_band_t<8, 4, 4>(a, (hipStream_t)stream);
}

int launch_pairs(const PairArgs &a, int n_out, void *stream)
{
    const size_t nblk = (size_t)n_out * ((a.nw + 255) / 256);
    if (nblk == 0) return 0;
    if (nblk > 0x7fffffffu) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pairs, dim3((unsigned)nblk), dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

}  // namespace kmdat
