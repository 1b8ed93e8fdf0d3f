// Learned traveltime surrogate for gfx950: a coordinate MLP trained and
// evaluated with BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
// Network, parameter layout and loss: include/kmig_nn.h.
//
// Every matrix product of training and inference is one kernel, k_gemm_nt:
//     C[i][j] = sum_k A[i][k] * B[j][k]      (both operands K-contiguous)
// with a per-element epilogue.  Activations are kept point-major ([n][f],
// the B operand of the next layer's forward product) and, where a weight
// gradient needs them as an A or B operand with the points as the reduction
// axis, also feature-major ([f][n]).
//
//   forward  layer l:  Z[n][m] = sum_k W[m][k] X[n][k] + b[m]   (i = m, j = n)
//   backward data:     dX[n][k] = sum_m WT[k][m] dZ[n][m]       (i = k, j = n)
//   weight gradient:   dW[m][k] = sum_n dZt[m][n] Xt[k][n]      (i = m, j = k,
//                      split over point chunks, FP32 atomics)
//
// A training step (batch sampling, forward, loss, backward, Adam, BF16 weight
// copies) is captured once as a HIP graph and replayed; its step counter,
// schedule and loss history live in device memory, so replays need no host
// work.  Memory is cleared with fill kernels only (no memset nodes).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kmig.h"
#include "kmig_internal.h"
#include "kmig_nn.h"
#include "kmig_nn_internal.h"

namespace kmig {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ GEMM --

// One block: 64 (i) x 64 (j) outputs, four waves of 32 x 32, each as 2 x 2
// MFMA tiles of 16 x 16.  Operand fragments (16 x 32 of A, 32 x 16 of B) are
// loaded straight from global memory, 16 bytes per lane: lane l holds row
// (l % 16) and k = 8 (l / 16) .. + 7.  The accumulator of a tile holds
// C[4 (l / 16) + e][l % 16], e = 0..3.  Rows past I (or J) read zeros and
// are not stored.  K ranges are multiples of 32; lda, ldb multiples of 8.
template <class Epi>
__global__ __launch_bounds__(256) void k_gemm_nt(const bf16 *__restrict__ A, int lda,
                                                 const bf16 *__restrict__ B, int ldb, int I,
                                                 int J, int K, int kchunk, Epi epi)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = blockIdx.y * 64 + (w & 1) * 32;
    const int j0 = blockIdx.x * 64 + (w >> 1) * 32;
    const int kb = blockIdx.z * kchunk;
    const int ke = min(K, kb + kchunk);
    const int r = lane & 15, q = lane >> 4;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const bf16x8 zero = {};
    for (int k = kb; k < ke; k += 32) {
        bf16x8 fa[2], fb[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int ii = i0 + 16 * t + r, jj = j0 + 16 * t + r;
            fa[t] = ii < I ? *(const bf16x8 *)(A + (size_t)ii * lda + k + 8 * q) : zero;
            fb[t] = jj < J ? *(const bf16x8 *)(B + (size_t)jj * ldb + k + 8 * q) : zero;
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0,
                                                                    0);
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int i = i0 + 16 * a + 4 * q + e, j = j0 + 16 * b + r;
                if (i < I && j < J) epi(i, j, acc[a][b][e]);
            }
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

__device__ __forceinline__ float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// forward hidden layer: Z (nullable, FP32 [n][M]), H = silu(Z) as BF16 [n][M]
// and (nullable) [M][n]
struct EpiFwd {
    const float *bias;
    float *Z;
    bf16 *H, *Ht;
    int M;
    long N;
    __device__ void operator()(int i, int j, float v) const
    {
        const float z = v + bias[i];
        if (Z) Z[(size_t)j * M + i] = z;
        const bf16 h = (bf16)silu(z);
        H[(size_t)j * M + i] = h;
        if (Ht) Ht[(size_t)i * N + j] = h;
    }
};

// output layer: Y [n][4]; in training also the loss terms and dL/dy as BF16
// [n][32] (columns 4..31 stay zero) and [4][n]
struct EpiOut {
    const float *bias;
    float *Y;
    const float *tgt;   // [n][4] T, dT/dx, dT/dz, dT/ds (training only)
    const float *dist;  // [n] distance factor
    const float *wt;    // [n] record weight (1/N, 0 for padding)
    float *loss;        // [n][4]
    bf16 *dY, *dYt;
    long N;
    float S0, inv_st2, inv_sg2;
    __device__ void operator()(int i, int j, float v) const
    {
        const float y = v + bias[i];
        Y[(size_t)j * 4 + i] = y;
        if (!tgt) return;
        float r, dpdy, inv2;
        if (i == 0) {
            dpdy = dist[j] * S0;
            r = dpdy * (1.0f + y) - tgt[(size_t)j * 4];
            inv2 = inv_st2;
        } else {
            dpdy = S0;
            r = S0 * y - tgt[(size_t)j * 4 + i];
            inv2 = inv_sg2;
        }
        const float w = wt[j];
        loss[(size_t)j * 4 + i] = w * r * r * inv2;
        const bf16 g = (bf16)(2.0f * w * r * inv2 * dpdy);
        dY[(size_t)j * 32 + i] = g;
        dYt[(size_t)i * N + j] = g;
    }
};

// backward through SiLU: dZ = dH * silu'(Zprev) as BF16 [n][K] and [K][n]
struct EpiBwd {
    const float *Zp;
    bf16 *dZ, *dZt;
    int K;
    long N;
    __device__ void operator()(int i, int j, float v) const
    {
        const bf16 g = (bf16)(v * silu_grad(Zp[(size_t)j * K + i]));
        dZ[(size_t)j * K + i] = g;
        dZt[(size_t)i * N + j] = g;
    }
};

struct EpiWgrad {
    float *dW;
    int ldw;
    __device__ void operator()(int i, int j, float v) const
    {
        atomicAdd(dW + (size_t)i * ldw + j, v);
    }
};

template <class Epi>
static void gemm(hipStream_t st, const bf16 *A, int lda, const bf16 *B, int ldb, int I, int J,
                 int K, int ksplit, const Epi &e)
{
    const int kchunk = ((K + ksplit - 1) / ksplit + 31) / 32 * 32;
    const int nz = (K + kchunk - 1) / kchunk;
    hipLaunchKernelGGL(k_gemm_nt<Epi>, dim3((unsigned)((J + 63) / 64), (unsigned)((I + 63) / 64),
                                            (unsigned)nz),
                       dim3(256), 0, st, A, lda, B, ldb, I, J, K, kchunk, e);
}

// -------------------------------------------------------- small kernels --

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// sin and cos of pi 2^f u for f = f0 .. f0+3, as (sin, cos) pairs: one
// sincosf, then double-angle steps (the error of the highest frequency is a
// few 1e-6, far below the BF16 rounding of the features)
__device__ __forceinline__ void fourier4(float u, int f0, float *o)
{
    float sv, cv;
    sincosf(3.14159265358979f * (float)(1 << f0) * u, &sv, &cv);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        o[2 * k] = sv;
        o[2 * k + 1] = cv;
        const float s2 = 2.0f * sv * cv, c2 = 1.0f - 2.0f * sv * sv;
        sv = s2;
        cv = c2;
    }
}

// Fourier features of one point, BF16, into row n of X [n][64] (48 used) and
// column n of Xt [64][n] (nullable); column 16 c + 2 f + {0, 1} holds
// {sin, cos}(pi 2^f u_c)
__device__ void features(const NnScale &sc, float s, float z, float x, bf16 *X, bf16 *Xt,
                         long n, long N)
{
    const float u[3] = {(s - sc.s0) * sc.s_inv, (z - sc.z0) * sc.z_inv, (x - sc.x0) * sc.x_inv};
    bf16 *row = X + (size_t)n * NN_KIN;
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int half = 0; half < 2; half++) {
            float f[8];
            fourier4(u[c], 4 * half, f);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int col = c * 16 + half * 8 + k;
                row[col] = (bf16)f[k];
                if (Xt) Xt[(size_t)col * N + n] = (bf16)f[k];
            }
        }
    }
#pragma unroll
    for (int col = NN_FEAT; col < NN_KIN; col++) {
        row[col] = (bf16)0.0f;
        if (Xt) Xt[(size_t)col * N + n] = (bf16)0.0f;
    }
}

__device__ __forceinline__ float src_dist(const NnScale &sc, float s, float z, float x)
{
    const float dx = x - s, dz = z - sc.zsrc;
    return sqrtf(dx * dx + dz * dz);
}

// one training batch: N records drawn uniformly (counter-based hash of seed,
// step and slot) from the device records
__global__ void k_sample(const float *rec, long nrec, const int *ctr, uint64_t seed, NnScale sc,
                         bf16 *X, bf16 *Xt, float *dist, float *tgt, float *wt, long N)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const uint64_t h = mix64(seed ^ mix64((uint64_t)ctr[0] * 0x100000001B3ull + (uint64_t)n));
    const long idx = (long)((h >> 11) % (uint64_t)nrec);
    const float *r = rec + (size_t)idx * 7;
    features(sc, r[0], r[1], r[2], X, Xt, n, N);
    dist[n] = src_dist(sc, r[0], r[1], r[2]);
    for (int k = 0; k < 4; k++) tgt[(size_t)n * 4 + k] = r[3 + k];
    wt[n] = 1.0f / (float)N;
}

// given records (n_real of them, padded with zero weights up to N)
__global__ void k_records(const float *rec, long nreal, NnScale sc, bf16 *X, bf16 *Xt,
                          float *dist, float *tgt, float *wt, long N)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const bool real = n < nreal;
    const float *r = rec + (size_t)(real ? n : 0) * 7;
    const float s = real ? r[0] : sc.s0, z = real ? r[1] : sc.z0, x = real ? r[2] : sc.x0;
    features(sc, s, z, x, X, Xt, n, N);
    dist[n] = src_dist(sc, s, z, x);
    for (int k = 0; k < 4; k++) tgt[(size_t)n * 4 + k] = real ? r[3 + k] : 0.0f;
    wt[n] = real ? 1.0f / (float)nreal : 0.0f;
}

// db[m] += sum_n G[m][n] (one block per m)
__global__ __launch_bounds__(256) void k_rowsum(const bf16 *G, long N, float *db)
{
    __shared__ float red[256];
    const int m = blockIdx.x;
    float s = 0.0f;
    for (long n = threadIdx.x; n < N; n += 256) s += (float)G[(size_t)m * N + n];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[m] += red[0];
}

// sum of the loss terms into out[slot] (slot = ctr[1] when `slot` < 0)
__global__ __launch_bounds__(256) void k_losssum(const float *loss, long n, float *out,
                                                 const int *ctr, int slot)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (long k = threadIdx.x; k < n; k += 256) s += loss[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot >= 0 ? slot : ctr[1]] = red[0];
}

__global__ void k_fillf(float *p, size_t n, float v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fillb(bf16 *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (bf16)0.0f;
}

// Adam on the FP32 master parameters; step size on a cosine schedule from
// lr to lr_final over the ctr[2] steps of the current train call
__global__ void k_adam(float *w, float *m, float *v, const float *g, size_t n, const int *ctr,
                       float lr0, float lr1, float b1, float b2, float eps)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = ctr[0] + 1, tl = ctr[1], tot = ctr[2];
    const float ph = tot > 1 ? (float)tl / (float)(tot - 1) : 1.0f;
    const float lr = lr1 + 0.5f * (lr0 - lr1) * (1.0f + cosf(3.14159265f * ph));
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)t));
    const float vh = vi / (1.0f - powf(b2, (float)t));
    w[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_step(int *ctr)
{
    ctr[0] += 1;
    ctr[1] += 1;
}

// BF16 copies of one layer: Wb [M][ldw] (zero-padded columns) and WT [ldt]
// rows [K][ldt] (zero-padded columns); master W is [M][K]
__global__ void k_cast(const float *W, int M, int K, bf16 *Wb, int ldw, bf16 *WT, int ldt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int rows = max(M, ldt), cols = max(K, ldw);
    if (i >= rows * cols) return;
    const int a = i / cols, b = i % cols;
    // Wb[a][b], a < M, b < ldw
    if (a < M && b < ldw) Wb[(size_t)a * ldw + b] = (bf16)(b < K ? W[(size_t)a * K + b] : 0.0f);
    // WT[b][a], b < K, a < ldt
    if (b < K && a < ldt) WT[(size_t)b * ldt + a] = (bf16)(a < M ? W[(size_t)a * K + b] : 0.0f);
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// layer l: fan-in K, outputs M, padded K of the BF16 operand, offsets of W, b
__host__ __device__ constexpr int lk(int l) { return l == 0 ? NN_FEAT : NN_W; }
__host__ __device__ constexpr int lm(int l) { return l == 4 ? NN_OUT : NN_W; }
__host__ __device__ constexpr int lkp(int l) { return l == 0 ? NN_KIN : NN_W; }

__host__ __device__ constexpr size_t woff(int l)
{
    size_t o = 0;
    for (int k = 0; k < l; k++) o += (size_t)lm(k) * lk(k) + lm(k);
    return o;
}
__host__ __device__ constexpr size_t boff(int l) { return woff(l) + (size_t)lm(l) * lk(l); }
static_assert(boff(4) + NN_OUT == NN_PARAMS, "parameter layout");

// ----------------------------------------------------- network plumbing --


// columns of the transposed BF16 weight (the M of the backward-data product,
// padded to a multiple of 32 for the output layer)
static int ldt(int l) { return l == 4 ? 32 : lm(l); }

void nn_cast_weights(const NnNet &net, hipStream_t st)
{
    for (int l = 0; l < 5; l++) {
        const int rows = std::max(lm(l), ldt(l)), cols = std::max(lk(l), lkp(l));
        hipLaunchKernelGGL(k_cast, dim3(nblk((size_t)rows * cols, 256)), dim3(256), 0, st,
                           net.w + woff(l), lm(l), lk(l), net.Wb[l], lkp(l), net.WT[l], ldt(l));
    }
}

// forward pass of N points whose features are in X; hidden outputs go to H[l]
// (and Z[l], Ht[l] when given); the output layer uses `out`
static void forward(const NnNet &net, hipStream_t st, const bf16 *X, long N, bf16 *const *H,
                    float *const *Z, bf16 *const *Ht, const EpiOut &out)
{
    const bf16 *in = X;
    for (int l = 0; l < 4; l++) {
        EpiFwd e{net.w + boff(l), Z ? Z[l] : nullptr, H[l], Ht ? Ht[l] : nullptr, NN_W, N};
        gemm(st, net.Wb[l], lkp(l), in, lkp(l), NN_W, (int)N, lkp(l), 1, e);
        in = H[l];
    }
    EpiOut o = out;
    o.bias = net.w + boff(4);
    gemm(st, net.Wb[4], NN_W, in, NN_W, NN_OUT, (int)N, NN_W, 1, o);
}

// ------------------------------------------------------ fused inference --
//
// k_nn_infer evaluates the whole network for 64 points per wave with the
// activations kept in registers.  The accumulator of a 16 x 16 MFMA tile
// gives lane (r = l % 16, q = l / 16) the outputs m = 16 a + 4 q + e
// (e = 0..3) of point n = 16 b + r.  The next layer's B operand for K step
// ks needs 8 values of point n in that lane: it takes e = 0..3 of tiles
// a = 2 ks and 2 ks + 1, i.e. K positions
//     pi(ks, q, j) = 32 ks + (j < 4 ? 4 q + j : 16 + 4 q + j - 4),
// and the weights are stored with their columns in that order, so the sum
// over K is unchanged.  The first layer's B operand is the Fourier features
// in natural order (lane group q, step ks: features 32 ks + 8 q + j).
//
// Weights live in LDS as BF16 (118 784 bytes, one copy per block; blocks are
// persistent): per layer, row m holds its 16-byte K chunks c = 4 ks + q at
// chunk position c ^ (m % 16) (c ^ (m % 8) for the 64-column first layer),
// so the 16 rows a fragment load touches fall in distinct banks.

constexpr int NN_BLOB_W1 = NN_W * NN_KIN;               // 8192
constexpr int NN_BLOB_WH = NN_W * NN_W;                 // 16384
constexpr int NN_BLOB_W5 = 16 * NN_W;                   // 2048 (rows 4..15 zero)
constexpr int NN_BLOB_BF16 = NN_BLOB_W1 + 3 * NN_BLOB_WH + NN_BLOB_W5;
constexpr int NN_BLOB_BIAS = 4 * NN_W + NN_OUT;         // FP32 biases

__device__ __forceinline__ int perm_k(int ks, int q, int j)
{
    return 32 * ks + (j < 4 ? 4 * q + j : 16 + 4 * q + j - 4);
}

// master FP32 parameters -> the swizzled BF16 blob and the bias block
__global__ void k_nn_blob(const float *w, bf16 *blob, float *bias)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < NN_BLOB_BF16) {
        int l, rem;
        if (t < NN_BLOB_W1) {
            l = 0;
            rem = t;
        } else if (t < NN_BLOB_W1 + 3 * NN_BLOB_WH) {
            l = 1 + (t - NN_BLOB_W1) / NN_BLOB_WH;
            rem = (t - NN_BLOB_W1) % NN_BLOB_WH;
        } else {
            l = 4;
            rem = t - NN_BLOB_W1 - 3 * NN_BLOB_WH;
        }
        const int cols = l == 0 ? NN_KIN : NN_W;
        const int m = rem / cols, pos = rem % cols;
        const int cpos = pos / 8, j = pos % 8;
        const int c = l == 0 ? (cpos ^ (m & 7)) : (cpos ^ (m & 15));  // logical chunk
        const int ks = c / 4, q = c % 4;
        float v = 0.0f;
        if (l == 0) {
            const int k = 32 * ks + 8 * q + j;
            if (k < NN_FEAT) v = w[woff(0) + (size_t)m * NN_FEAT + k];
        } else {
            const int k = perm_k(ks, q, j);
            if (m < lm(l)) v = w[woff(l) + (size_t)m * NN_W + k];
        }
        blob[t] = (bf16)v;
    }
    if (t < NN_BLOB_BIAS) {
        const int l = t / NN_W, m = t % NN_W;
        bias[t] = l < 4 ? w[boff(l) + m] : w[boff(4) + (t - 4 * NN_W)];
    }
}

__device__ __forceinline__ float silu_f(float z)
{
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// 8 features of lane group q at K step ks for coordinates u (s, z, x)
__device__ __forceinline__ bf16x8 feat_frag(const float *u, int ks, int q)
{
    bf16x8 f = {};
    if (ks == 1 && q >= 2) return f;
    const float uc = ks == 1 ? u[2] : ((q >> 1) == 0 ? u[0] : u[1]);
    float v[8];
    fourier4(uc, 4 * (q & 1), v);
#pragma unroll
    for (int j = 0; j < 8; j++) f[j] = (bf16)v[j];
    return f;
}

struct NnInferArgs {
    const bf16 *blob;
    const float *bias;
    NnScale sc;
    long npts;
    // points mode: pts [n][3] (s, z, x) -> out [n][4]
    const float *pts;
    float *out;
    // tuple mode: point p = slot * nzx + i at (surf_x of slot, grid point i);
    // T and A of the table path are replaced where |T_nn - T| <= max_dt
    const SlotSpec *spec;
    int nz;
    long nzx;
    float z0, dz, x0, dx, dside, max_dt;
    float *T, *A;
    int *used;          // [slot] points taken from the network
    unsigned *maxres;   // [slot] max |T_nn - T| (float bits)
    // per-slot fallback: the network's T and A go to Tn/An for every point
    // (T and A stay the tables'); k_nn_select then takes them for the slots
    // whose largest residual is within max_dt
    float *Tn, *An;
    // Point p of the launch is grid point p * pstride (1: every point).
    // sample_only: the residuals only (maxres), no T / A / Tn / An / used.
    // skip (per-slot fallback): a tile whose slots all have skip[slot] above
    // skip_lim (a residual already past the limit) is not evaluated.
    long pstride;
    int sample_only;
    const unsigned *skip;
    unsigned skip_lim;
    // per-slot fallback with the tables read at the slot's node ([npos][nzx])
    const float *tab;
};

template <bool Tuples>
__global__ __launch_bounds__(256) void k_nn_infer(const NnInferArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    bf16 *W = (bf16 *)lds_raw;
    float *B = (float *)(lds_raw + NN_BLOB_BF16 * sizeof(bf16));
    for (int k = threadIdx.x; k < NN_BLOB_BF16 / 8; k += blockDim.x)
        ((uint4 *)W)[k] = ((const uint4 *)a.blob)[k];
    for (int k = threadIdx.x; k < NN_BLOB_BIAS; k += blockDim.x) B[k] = a.bias[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const long ntiles = (a.npts + 63) / 64;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = (long)gridDim.x * 4;
    const long stride = Tuples ? a.pstride : 1, nphys = Tuples ? a.npts * stride : a.npts;
    auto run_tile = [&](const long tile) {
        // the lane's four points n = 16 b + r
        float u[4][3], dist[4];
        long pidx[4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            long p = tile * 64 + 16 * b + r;
            if (p >= a.npts) p = a.npts - 1;
            p *= stride;  // the grid point
            pidx[b] = (tile * 64 + 16 * b + r) < a.npts ? p : nphys;
            float s, z, x;
            if (Tuples) {
                const int slot = (int)(p / a.nzx);
                const long i = p % a.nzx;
                s = a.spec[slot].surf_x;
                z = a.z0 + (float)(i % a.nz) * a.dz;
                x = a.x0 + (float)(i / a.nz) * a.dx;
            } else {
                s = a.pts[p * 3];
                z = a.pts[p * 3 + 1];
                x = a.pts[p * 3 + 2];
            }
            u[b][0] = (s - a.sc.s0) * a.sc.s_inv;
            u[b][1] = (z - a.sc.z0) * a.sc.z_inv;
            u[b][2] = (x - a.sc.x0) * a.sc.x_inv;
            const float ddx = x - s, ddz = z - a.sc.zsrc;
            dist[b] = sqrtf(ddx * ddx + ddz * ddz);
        }
        bf16x8 bf[4][4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            bf[b][0] = feat_frag(u[b], 0, q);
            bf[b][1] = feat_frag(u[b], 1, q);
        }
        f32x4 acc[8][4];
        const bf16 *Wl = W;
        for (int l = 0; l < 4; l++) {
            const int nks = l == 0 ? 2 : 4;
            const int cols = l == 0 ? NN_KIN : NN_W;
            const int cmask = l == 0 ? 7 : 15;
#pragma unroll
            for (int am = 0; am < 8; am++)
#pragma unroll
                for (int b = 0; b < 4; b++) acc[am][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                if (ks < nks) {
#pragma unroll
                    for (int am = 0; am < 8; am++) {
                        const int m = 16 * am + r;
                        const int c = (4 * ks + q) ^ (m & cmask);
                        const bf16x8 wa = *(const bf16x8 *)(Wl + (size_t)m * cols + 8 * c);
#pragma unroll
                        for (int b = 0; b < 4; b++)
                            acc[am][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, bf[b][ks],
                                                                                 acc[am][b], 0, 0, 0);
                    }
                }
            }
            Wl += (size_t)NN_W * cols;
            // SiLU in registers -> next layer's B fragments (K order pi)
            const float *bl = B + l * NN_W;
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int ks = 0; ks < 4; ks++) {
                    bf16x8 h;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        h[e] = (bf16)silu_f(acc[2 * ks][b][e] + bl[32 * ks + 4 * q + e]);
                        h[4 + e] = (bf16)silu_f(acc[2 * ks + 1][b][e] + bl[32 * ks + 16 + 4 * q + e]);
                    }
                    bf[b][ks] = h;
                }
        }
        // output layer: one 16-row tile, rows 0..3 real (lane group q = 0)
        f32x4 y[4];
#pragma unroll
        for (int b = 0; b < 4; b++) y[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int c = (4 * ks + q) ^ (r & 15);
            const bf16x8 wa = *(const bf16x8 *)(Wl + (size_t)r * NN_W + 8 * c);
#pragma unroll
            for (int b = 0; b < 4; b++)
                y[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, bf[b][ks], y[b], 0, 0, 0);
        }
        const float *b5 = B + 4 * NN_W;
        int rep[4] = {0, 0, 0, 0};
        float resb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const long p = pidx[b];
            if (q != 0 || p >= nphys) continue;
            const float S0 = a.sc.S0;
            const float T = dist[b] * S0 * (1.0f + y[b][0] + b5[0]);
            const float Ts = S0 * (y[b][3] + b5[3]);
            if (Tuples) {
                const float Tt = a.tab ? a.tab[(size_t)a.spec[p / a.nzx].node * a.nzx + p % a.nzx]
                                       : a.T[p];
                float res = fabsf(T - Tt);
                if (a.sample_only) {
                } else if (a.Tn) {
                    a.Tn[p] = T;
                    a.An[p] = fabsf(Ts * a.dside);
                } else if (res <= a.max_dt) {
                    a.T[p] = T;
                    a.A[p] = fabsf(Ts * a.dside);
                    rep[b] = 1;
                }
                resb[b] = res == res ? res : INFINITY;
            } else {
                float *o = a.out + p * 4;
                o[0] = T;
                o[1] = S0 * (y[b][1] + b5[1]);
                o[2] = S0 * (y[b][2] + b5[2]);
                o[3] = Ts;
            }
        }
        if (Tuples) {
            // per-slot counters: one atomic per wave when its 64 points share
            // a slot, per point otherwise (tiles across a slot boundary)
            const long pf = tile * 64 * stride, pl = min(tile * 64 + 63, a.npts - 1) * stride;
            const int s0 = (int)(pf / a.nzx), s1 = (int)(pl / a.nzx);
            if (s0 == s1) {
                int cnt = rep[0] + rep[1] + rep[2] + rep[3];
                float mx = fmaxf(fmaxf(resb[0], resb[1]), fmaxf(resb[2], resb[3]));
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    cnt += __shfl_xor(cnt, off, 64);
                    mx = fmaxf(mx, __shfl_xor(mx, off, 64));
                }
                if (lane == 0) {
                    atomicAdd(a.used + s0, cnt);
                    atomicMax(a.maxres + s0, __float_as_uint(mx));
                }
            } else if (q == 0) {
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    if (pidx[b] >= nphys) continue;
                    const int slot = (int)(pidx[b] / a.nzx);
                    if (rep[b]) atomicAdd(a.used + slot, 1);
                    atomicMax(a.maxres + slot, __float_as_uint(resb[b]));
                }
            }
        }
    };
    if (!(Tuples && a.skip)) {
        for (long tile = wave; tile < ntiles; tile += nwaves) run_tile(tile);
        return;
    }
    // Only the slots whose residual is still within the limit: the wave
    // reads 64 slots' verdicts at a time and runs its own tiles (tile index =
    // wave mod nwaves) of each slot taken.  A tile across the boundary of two
    // taken slots runs once, with the first; a slot this wave reads as taken
    // may have failed since, which costs only time (its values are not used).
    const int nslots = (int)((a.npts + a.nzx - 1) / a.nzx);
    int last = -2;  // the last slot taken
    for (int sb = 0; sb < nslots; sb += 64) {
        const int sl = sb + lane;
        unsigned long long m = __ballot(sl < nslots && a.skip[sl] <= a.skip_lim);
        while (m) {
            const int s = sb + __builtin_ctzll(m);
            m &= m - 1;
            const long pf = (long)s * a.nzx, pl = min(pf + a.nzx, a.npts) - 1;
            long t0 = pf / 64;
            const long t1 = pl / 64;
            if (last == s - 1 && pf % 64 != 0) t0++;
            last = s;
            long tile = t0 + ((wave - t0) % nwaves + nwaves) % nwaves;
            for (; tile <= t1; tile += nwaves) run_tile(tile);
        }
    }
}

// Per-slot fallback: a slot whose largest residual (maxres, float bits;
// non-negative floats order as their bits) is within max_dt takes the
// network's T and A at every point, any other keeps its tables at every
// point; used[slot] becomes the slot's points taken from the network.
// Grid: x blocks of 256 threads over a slot's nzx points (a thread takes
// points i, i + 256 x, ...), y = slot.
__global__ __launch_bounds__(256) void k_nn_select(const float *__restrict__ Tn,
                                                   const float *__restrict__ An, float *T, float *A,
                                                   const unsigned *__restrict__ maxres,
                                                   unsigned lim_bits, long nzx)
{
    const int slot = blockIdx.y;
    if (maxres[slot] > lim_bits) return;
    const long base = (long)slot * nzx;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nzx;
         i += (long)gridDim.x * blockDim.x) {
        T[base + i] = Tn[base + i];
        A[base + i] = An[base + i];
    }
}

__global__ __launch_bounds__(256) void k_nn_slot_used(int *used, const unsigned *maxres,
                                                      unsigned lim_bits, int nslots, int nzx)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) used[s] = maxres[s] <= lim_bits ? nzx : 0;
}

// Grid-point stride of the residual pre-check of the per-slot fallback: a
// prime, so that the sample walks through every depth row of the grid.
constexpr long kNnSampleStride = 1021;

static size_t infer_lds() { return NN_BLOB_BF16 * sizeof(bf16) + NN_BLOB_BIAS * sizeof(float); }

static int launch_infer(const NnInferArgs &a, bool tuples, hipStream_t st)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    const long tiles = (a.npts + 63) / 64;
    const unsigned grid = (unsigned)std::max<long>(1, std::min<long>((tiles + 3) / 4, cus));
    const int lds = (int)infer_lds();
    hipError_t e;
    if (tuples) {
        e = hipFuncSetAttribute((const void *)k_nn_infer<true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e == hipSuccess)
            hipLaunchKernelGGL(k_nn_infer<true>, dim3(grid), dim3(256), lds, st, a);
    } else {
        e = hipFuncSetAttribute((const void *)k_nn_infer<false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e == hipSuccess)
            hipLaunchKernelGGL(k_nn_infer<false>, dim3(grid), dim3(256), lds, st, a);
    }
    return e != hipSuccess ? (int)e : (int)hipGetLastError();
}

// ---------------------------------------------------------------- host --

}  // namespace kmig

using namespace kmig;

struct kmig_nn {
    kmig_nn_config cfg{};
    NnNet net{};
    hipStream_t stream = nullptr;
    float *m = nullptr, *v = nullptr, *grad = nullptr;
    float *data = nullptr;
    long nrec = 0;
    long N = 0;  // batch capacity
    bf16 *X = nullptr, *Xt = nullptr;
    float *dist = nullptr, *tgt = nullptr, *wt = nullptr, *Y = nullptr, *lossb = nullptr;
    float *Z[4] = {}, *dev_rec = nullptr;
    bf16 *H[4] = {}, *Ht[4] = {}, *dZ[4] = {}, *dZt[4] = {}, *dY = nullptr, *dYt = nullptr;
    long rec_cap = 0;
    int *ctr = nullptr;  // [0] Adam steps taken, [1] step in this call, [2] steps of this call
    float *hist = nullptr;
    int hist_cap = 0;
    float *gloss = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double train_ms = 0.0, train_flops = 0.0;
    bf16 *blob = nullptr;       // inference weights (k_nn_infer layout)
    float *bias = nullptr;
    bool blob_valid = false;
    std::string err;
};

namespace {

int nfail(kmig_nn *n, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (n) n->err = buf;
    return code;
}

#define NHIP(n, call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return nfail((n), KMIG_E_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int nalloc(kmig_nn *n, T **p, size_t count)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess)
        return nfail(n, KMIG_E_NOMEM, "device allocation of %zu bytes failed", count * sizeof(T));
    return KMIG_OK;
}

void nfree(void *p)
{
    if (p) (void)hipFree(p);
}

uint64_t hmix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

NnScale make_scale(const kmig_nn_config &c)
{
    NnScale s;
    s.s0 = c.s_min;
    s.z0 = c.z_min;
    s.x0 = c.x_min;
    s.s_inv = c.s_max > c.s_min ? 1.0f / (c.s_max - c.s_min) : 1.0f;
    s.z_inv = c.z_max > c.z_min ? 1.0f / (c.z_max - c.z_min) : 1.0f;
    s.x_inv = c.x_max > c.x_min ? 1.0f / (c.x_max - c.x_min) : 1.0f;
    s.zsrc = c.z_src;
    s.S0 = c.slowness;
    return s;
}

// batch buffers for N records
int ensure_batch(kmig_nn *n, long N)
{
    if (n->N >= N) return KMIG_OK;
    int rc = KMIG_OK;
    const size_t W = NN_W;
    rc = rc ? rc : nalloc(n, &n->X, (size_t)N * NN_KIN);
    rc = rc ? rc : nalloc(n, &n->Xt, (size_t)N * NN_KIN);
    rc = rc ? rc : nalloc(n, &n->dist, (size_t)N);
    rc = rc ? rc : nalloc(n, &n->tgt, (size_t)N * 4);
    rc = rc ? rc : nalloc(n, &n->wt, (size_t)N);
    rc = rc ? rc : nalloc(n, &n->Y, (size_t)N * 4);
    rc = rc ? rc : nalloc(n, &n->lossb, (size_t)N * 4);
    rc = rc ? rc : nalloc(n, &n->dY, (size_t)N * 32);
    rc = rc ? rc : nalloc(n, &n->dYt, (size_t)N * 4);
    for (int l = 0; l < 4 && !rc; l++) {
        rc = rc ? rc : nalloc(n, &n->Z[l], (size_t)N * W);
        rc = rc ? rc : nalloc(n, &n->H[l], (size_t)N * W);
        rc = rc ? rc : nalloc(n, &n->Ht[l], (size_t)N * W);
        rc = rc ? rc : nalloc(n, &n->dZ[l], (size_t)N * W);
        rc = rc ? rc : nalloc(n, &n->dZt[l], (size_t)N * W);
    }
    if (rc) return rc;
    // dY columns 4..31 are never written: clear once
    hipLaunchKernelGGL(k_fillb, dim3(nblk((size_t)N * 32, 256)), dim3(256), 0, n->stream, n->dY,
                       (size_t)N * 32);
    NHIP(n, hipGetLastError());
    n->N = N;
    if (n->exec) (void)hipGraphExecDestroy(n->exec);
    if (n->graph) (void)hipGraphDestroy(n->graph);
    n->exec = nullptr;
    n->graph = nullptr;
    return KMIG_OK;
}

// forward + loss + backward of the batch already in X/Xt/dist/tgt/wt; the
// gradient (parameter layout) is accumulated into n->grad, which the caller
// clears; loss terms land in n->lossb
void fwd_bwd(kmig_nn *n, long N)
{
    hipStream_t st = n->stream;
    const NnNet &net = n->net;
    const kmig_nn_config &c = n->cfg;
    EpiOut o{};
    o.Y = n->Y;
    o.tgt = n->tgt;
    o.dist = n->dist;
    o.wt = n->wt;
    o.loss = n->lossb;
    o.dY = n->dY;
    o.dYt = n->dYt;
    o.N = N;
    o.S0 = c.slowness;
    o.inv_st2 = 1.0f / (c.sigma_t * c.sigma_t);
    o.inv_sg2 = 1.0f / (c.sigma_g * c.sigma_g);
    forward(net, st, n->X, N, n->H, n->Z, n->Ht, o);
    // split the point reduction of the weight gradients into ~1024-point chunks
    const int ks = (int)std::max<long>(1, N / 1024);
    float *g = n->grad;
    // output layer
    gemm(st, n->dYt, (int)N, n->Ht[3], (int)N, NN_OUT, NN_W, (int)N, ks,
         EpiWgrad{g + woff(4), NN_W});
    hipLaunchKernelGGL(k_rowsum, dim3(NN_OUT), dim3(256), 0, st, n->dYt, N, g + boff(4));
    gemm(st, net.WT[4], 32, n->dY, 32, NN_W, (int)N, 32, 1, EpiBwd{n->Z[3], n->dZ[3], n->dZt[3],
                                                                   NN_W, N});
    for (int l = 3; l >= 1; l--) {
        gemm(st, n->dZt[l], (int)N, n->Ht[l - 1], (int)N, NN_W, NN_W, (int)N, ks,
             EpiWgrad{g + woff(l), NN_W});
        hipLaunchKernelGGL(k_rowsum, dim3(NN_W), dim3(256), 0, st, n->dZt[l], N, g + boff(l));
        gemm(st, net.WT[l], NN_W, n->dZ[l], NN_W, NN_W, (int)N, NN_W, 1,
             EpiBwd{n->Z[l - 1], n->dZ[l - 1], n->dZt[l - 1], NN_W, N});
    }
    gemm(st, n->dZt[0], (int)N, n->Xt, (int)N, NN_W, NN_FEAT, (int)N, ks,
         EpiWgrad{g + woff(0), NN_FEAT});
    hipLaunchKernelGGL(k_rowsum, dim3(NN_W), dim3(256), 0, st, n->dZt[0], N, g + boff(0));
}

double step_flops(long N)
{
    // forward 2 * MACs, backward data + weight gradients 4 * MACs (the input
    // layer has no backward-data product)
    double mac = 0.0, macx = 0.0;
    for (int l = 0; l < 5; l++) mac += (double)lm(l) * lk(l);
    macx = (double)lm(0) * lk(0);
    return (double)N * (2.0 * mac + 4.0 * mac - 2.0 * macx);
}

int capture_step(kmig_nn *n)
{
    hipStream_t st = n->stream;
    NHIP(n, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_fillf, dim3(nblk(NN_PARAMS, 256)), dim3(256), 0, st, n->grad,
                       (size_t)NN_PARAMS, 0.0f);
    hipLaunchKernelGGL(k_sample, dim3(nblk((size_t)n->N, 256)), dim3(256), 0, st, n->dev_rec,
                       n->nrec, (const int *)n->ctr, (uint64_t)n->cfg.seed, make_scale(n->cfg),
                       n->X, n->Xt, n->dist, n->tgt, n->wt, n->N);
    fwd_bwd(n, n->N);
    hipLaunchKernelGGL(k_losssum, dim3(1), dim3(256), 0, st, n->lossb, n->N * 4, n->hist,
                       (const int *)n->ctr, -1);
    const kmig_nn_config &c = n->cfg;
    hipLaunchKernelGGL(k_adam, dim3(nblk(NN_PARAMS, 256)), dim3(256), 0, st, n->net.w, n->m, n->v,
                       (const float *)n->grad, (size_t)NN_PARAMS, (const int *)n->ctr, c.lr,
                       c.lr_final, c.beta1, c.beta2, c.eps);
    nn_cast_weights(n->net, st);
    hipLaunchKernelGGL(k_step, dim3(1), dim3(1), 0, st, n->ctr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess)
        return nfail(n, KMIG_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    n->graph = g;
    NHIP(n, hipGraphInstantiate(&n->exec, g, nullptr, nullptr, 0));
    return KMIG_OK;
}

}  // namespace

namespace kmig {

int nn_prepare(kmig_nn *n)
{
    if (!n) return KMIG_E_PARAM;
    if (n->blob_valid) return KMIG_OK;
    int rc = KMIG_OK;
    if (!n->blob) rc = nalloc(n, &n->blob, NN_BLOB_BF16);
    if (!rc && !n->bias) rc = nalloc(n, &n->bias, NN_BLOB_BIAS);
    if (rc) return rc;
    hipLaunchKernelGGL(k_nn_blob, dim3(nblk(NN_BLOB_BF16, 256)), dim3(256), 0, n->stream,
                       (const float *)n->net.w, n->blob, n->bias);
    NHIP(n, hipGetLastError());
    NHIP(n, hipStreamSynchronize(n->stream));
    n->blob_valid = true;
    return KMIG_OK;
}

int nn_replace_tuples(const kmig_nn *n, const SlotSpec *spec, int nslots, int nz, size_t nzx,
                      float z0, float dz, float x0, float dx, float dside, float max_dt, float *T,
                      float *A, int *used, unsigned *maxres, float *Tn, float *An,
                      const float *tab, hipStream_t st)
{
    if (!n || !n->blob_valid) return (int)hipErrorNotReady;
    if (nslots <= 0 || nzx == 0) return 0;
    NnInferArgs a{};
    a.blob = n->blob;
    a.bias = n->bias;
    a.sc = make_scale(n->cfg);
    a.npts = (long)nslots * (long)nzx;
    a.spec = spec;
    a.nz = nz;
    a.nzx = (long)nzx;
    a.z0 = z0;
    a.dz = dz;
    a.x0 = x0;
    a.dx = dx;
    a.dside = dside;
    a.max_dt = max_dt;
    a.T = T;
    a.A = A;
    a.used = used;
    a.maxres = maxres;
    a.Tn = Tn;
    a.An = An;
    a.tab = Tn ? tab : nullptr;
    a.pstride = 1;
    unsigned lim;
    memcpy(&lim, &max_dt, sizeof(lim));
    const char *pre = getenv("KMIG_NN_PRECHECK");
    if (Tn && !(pre && pre[0] == '0')) {
        // Per-slot fallback: a slot takes the network only if every one of
        // its points is within max_dt, so a slot that misses it at a sample
        // of its points (every kNnSampleStride-th grid point) misses it
        // overall, and the full evaluation runs only for the other slots.
        // (Their maxres is then exact; a skipped slot's is the sample's.)
        NnInferArgs sa = a;
        sa.pstride = kNnSampleStride;
        sa.npts = (a.npts + kNnSampleStride - 1) / kNnSampleStride;
        sa.sample_only = 1;
        int es = launch_infer(sa, true, st);
        if (es) return es;
        a.skip = maxres;
        a.skip_lim = lim;
    }
    int e = launch_infer(a, true, st);
    if (e || !Tn) return e;
    const unsigned sx = std::max(1u, std::min(nblk((size_t)a.nzx, 256), 64u));
    for (int s0 = 0; s0 < nslots; s0 += 65535) {
        const long off = (long)s0 * a.nzx;
        hipLaunchKernelGGL(k_nn_select, dim3(sx, (unsigned)std::min(nslots - s0, 65535)),
                           dim3(256), 0, st, Tn + off, An + off, T + off, A + off, maxres + s0,
                           lim, a.nzx);
    }
    hipLaunchKernelGGL(k_nn_slot_used, dim3(nblk((size_t)nslots, 256)), dim3(256), 0, st, used,
                       maxres, lim, nslots, (int)nzx);
    return (int)hipGetLastError();
}

}  // namespace kmig

extern "C" {

int kmig_nn_config_default(kmig_nn_config *c)
{
    if (!c) return KMIG_E_PARAM;
    *c = kmig_nn_config{};
    c->s_min = 0.0f;
    c->s_max = 1.0f;
    c->z_min = 0.0f;
    c->z_max = 1.0f;
    c->x_min = 0.0f;
    c->x_max = 1.0f;
    c->z_src = 0.0f;
    c->slowness = 0.5f;
    c->sigma_t = 0.001f;
    c->sigma_g = 0.05f;
    c->lr = 2.0e-3f;
    c->lr_final = 5.0e-5f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->eps = 1.0e-8f;
    c->batch = 8192;
    c->seed = 20260404ull;
    return KMIG_OK;
}

int kmig_nn_create(kmig_nn **out, const kmig_nn_config *cfg)
{
    if (!out || !cfg) return KMIG_E_PARAM;
    *out = nullptr;
    kmig_nn *n = new kmig_nn();
    *out = n;
    n->cfg = *cfg;
    if (cfg->batch < 64 || cfg->batch % 64)
        return nfail(n, KMIG_E_PARAM, "batch must be a positive multiple of 64 (got %d)",
                     cfg->batch);
    if (!(cfg->slowness > 0.0f) || !(cfg->sigma_t > 0.0f) || !(cfg->sigma_g > 0.0f))
        return nfail(n, KMIG_E_PARAM, "slowness, sigma_t and sigma_g must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return nfail(n, KMIG_E_DEVICE, "no HIP device");
    NHIP(n, hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    NHIP(n, hipEventCreate(&n->e0));
    NHIP(n, hipEventCreate(&n->e1));
    int rc = KMIG_OK;
    rc = rc ? rc : nalloc(n, &n->net.w, NN_PARAMS);
    rc = rc ? rc : nalloc(n, &n->m, NN_PARAMS);
    rc = rc ? rc : nalloc(n, &n->v, NN_PARAMS);
    rc = rc ? rc : nalloc(n, &n->grad, NN_PARAMS);
    rc = rc ? rc : nalloc(n, &n->ctr, 4);
    rc = rc ? rc : nalloc(n, &n->gloss, 1);
    for (int l = 0; l < 5 && !rc; l++) {
        rc = nalloc(n, &n->net.Wb[l], (size_t)lm(l) * lkp(l));
        if (!rc) rc = nalloc(n, &n->net.WT[l], (size_t)lk(l) * ldt(l));
    }
    if (rc) return rc;
    // initial weights: uniform with variance 1 / fan-in, the output layer
    // scaled by 0.1 (starts near the straight-ray time d * S0); zero biases
    std::vector<float> w(NN_PARAMS, 0.0f);
    uint64_t st = cfg->seed;
    for (int l = 0; l < 5; l++) {
        const float a = std::sqrt(3.0f / (float)lk(l)) * (l == 4 ? 0.1f : 1.0f);
        for (size_t k = 0; k < (size_t)lm(l) * lk(l); k++) {
            st = hmix(st);
            const float u = (float)((st >> 40) * (1.0 / 16777216.0));
            w[woff(l) + k] = a * (2.0f * u - 1.0f);
        }
    }
    return kmig_nn_set_params(n, w.data());
}

int kmig_nn_set_params(kmig_nn *n, const float *params)
{
    if (!n || !params) return KMIG_E_PARAM;
    if (!n->net.w) return nfail(n, KMIG_E_STATE, "network has no device memory");
    NHIP(n, hipMemcpy(n->net.w, params, NN_PARAMS * sizeof(float), hipMemcpyHostToDevice));
    // a new parameter vector restarts Adam
    hipLaunchKernelGGL(k_fillf, dim3(nblk(NN_PARAMS, 256)), dim3(256), 0, n->stream, n->m,
                       (size_t)NN_PARAMS, 0.0f);
    hipLaunchKernelGGL(k_fillf, dim3(nblk(NN_PARAMS, 256)), dim3(256), 0, n->stream, n->v,
                       (size_t)NN_PARAMS, 0.0f);
    nn_cast_weights(n->net, n->stream);
    n->blob_valid = false;
    const int zero[4] = {0, 0, 0, 0};
    NHIP(n, hipMemcpyAsync(n->ctr, zero, sizeof(zero), hipMemcpyHostToDevice, n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    return KMIG_OK;
}

int kmig_nn_get_params(const kmig_nn *n, float *params)
{
    if (!n || !params || !n->net.w) return KMIG_E_PARAM;
    if (hipStreamSynchronize(n->stream) != hipSuccess) return KMIG_E_DEVICE;
    if (hipMemcpy(params, n->net.w, NN_PARAMS * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return KMIG_E_DEVICE;
    return KMIG_OK;
}

int kmig_nn_set_data(kmig_nn *n, const float *rec, long count)
{
    if (!n || !rec || count <= 0) return n ? nfail(n, KMIG_E_PARAM, "no records") : KMIG_E_PARAM;
    if (count > n->rec_cap) {
        int rc = nalloc(n, &n->dev_rec, (size_t)count * 7);
        if (rc) return rc;
        n->rec_cap = count;
        if (n->exec) (void)hipGraphExecDestroy(n->exec);
        if (n->graph) (void)hipGraphDestroy(n->graph);
        n->exec = nullptr;
        n->graph = nullptr;
    }
    NHIP(n, hipMemcpy(n->dev_rec, rec, (size_t)count * 7 * sizeof(float),
                      hipMemcpyHostToDevice));
    if (n->nrec != count && n->exec) {
        (void)hipGraphExecDestroy(n->exec);
        (void)hipGraphDestroy(n->graph);
        n->exec = nullptr;
        n->graph = nullptr;
    }
    n->nrec = count;
    return KMIG_OK;
}

int kmig_nn_train(kmig_nn *n, int steps, float *loss)
{
    if (!n || steps < 0) return KMIG_E_PARAM;
    if (!n->dev_rec || n->nrec <= 0) return nfail(n, KMIG_E_STATE, "no training records");
    if (steps == 0) return KMIG_OK;
    int rc = ensure_batch(n, n->cfg.batch);
    if (rc) return rc;
    if (steps > n->hist_cap) {
        if ((rc = nalloc(n, &n->hist, (size_t)steps))) return rc;
        n->hist_cap = steps;
        if (n->exec) (void)hipGraphExecDestroy(n->exec);
        if (n->graph) (void)hipGraphDestroy(n->graph);
        n->exec = nullptr;
        n->graph = nullptr;
    }
    if (!n->exec && (rc = capture_step(n))) return rc;
    // ctr[1] = 0 (step in this call), ctr[2] = steps; ctr[0] keeps counting
    int host[4];
    NHIP(n, hipStreamSynchronize(n->stream));
    NHIP(n, hipMemcpy(host, n->ctr, sizeof(host), hipMemcpyDeviceToHost));
    host[1] = 0;
    host[2] = steps;
    NHIP(n, hipMemcpy(n->ctr, host, sizeof(host), hipMemcpyHostToDevice));
    n->blob_valid = false;
    NHIP(n, hipEventRecord(n->e0, n->stream));
    for (int s = 0; s < steps; s++) NHIP(n, hipGraphLaunch(n->exec, n->stream));
    NHIP(n, hipEventRecord(n->e1, n->stream));
    NHIP(n, hipEventSynchronize(n->e1));
    float ms = 0.0f;
    NHIP(n, hipEventElapsedTime(&ms, n->e0, n->e1));
    n->train_ms = ms;
    n->train_flops = step_flops(n->N) * steps;
    if (loss)
        NHIP(n, hipMemcpy(loss, n->hist, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost));
    return KMIG_OK;
}

int kmig_nn_grad(kmig_nn *n, const float *rec, long count, float *grad, float *loss)
{
    if (!n || !rec || count <= 0 || !grad) return KMIG_E_PARAM;
    const long N = (count + 63) / 64 * 64;
    int rc = ensure_batch(n, std::max<long>(N, n->N));
    if (rc) return rc;
    float *drec = nullptr;
    if ((rc = nalloc(n, &drec, (size_t)count * 7))) return rc;
    hipError_t e = hipMemcpy(drec, rec, (size_t)count * 7 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipStream_t st = n->stream;
        hipLaunchKernelGGL(k_fillf, dim3(nblk(NN_PARAMS, 256)), dim3(256), 0, st, n->grad,
                           (size_t)NN_PARAMS, 0.0f);
        hipLaunchKernelGGL(k_records, dim3(nblk((size_t)N, 256)), dim3(256), 0, st, drec, count,
                           make_scale(n->cfg), n->X, n->Xt, n->dist, n->tgt, n->wt, N);
        fwd_bwd(n, N);
        hipLaunchKernelGGL(k_losssum, dim3(1), dim3(256), 0, st, n->lossb, N * 4, n->gloss,
                           (const int *)n->ctr, 0);
        e = hipStreamSynchronize(st);
    }
    if (e == hipSuccess)
        e = hipMemcpy(grad, n->grad, NN_PARAMS * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss) e = hipMemcpy(loss, n->gloss, sizeof(float), hipMemcpyDeviceToHost);
    nfree(drec);
    if (e != hipSuccess)
        return nfail(n, KMIG_E_DEVICE, "gradient evaluation failed: %s", hipGetErrorString(e));
    return KMIG_OK;
}

int kmig_nn_eval(kmig_nn *n, const float *pts, long count, float *out)
{
    if (!n || !pts || !out || count <= 0) return KMIG_E_PARAM;
    int rc = nn_prepare(n);
    if (rc) return rc;
    float *dp = nullptr, *dout = nullptr;
    if ((rc = nalloc(n, &dp, (size_t)count * 3))) return rc;
    if ((rc = nalloc(n, &dout, (size_t)count * 4))) {
        nfree(dp);
        return rc;
    }
    hipError_t e = hipMemcpy(dp, pts, (size_t)count * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        NnInferArgs a{};
        a.blob = n->blob;
        a.bias = n->bias;
        a.sc = make_scale(n->cfg);
        a.npts = count;
        a.pts = dp;
        a.out = dout;
        e = (hipError_t)launch_infer(a, false, n->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    if (e == hipSuccess)
        e = hipMemcpy(out, dout, (size_t)count * 4 * sizeof(float), hipMemcpyDeviceToHost);
    nfree(dp);
    nfree(dout);
    if (e != hipSuccess)
        return nfail(n, KMIG_E_DEVICE, "evaluation failed: %s", hipGetErrorString(e));
    return KMIG_OK;
}

int kmig_nn_stats(const kmig_nn *n, double *train_ms, double *train_flops)
{
    if (!n) return KMIG_E_PARAM;
    if (train_ms) *train_ms = n->train_ms;
    if (train_flops) *train_flops = n->train_flops;
    return KMIG_OK;
}

const char *kmig_nn_last_error(const kmig_nn *n) { return n ? n->err.c_str() : "null network"; }

void kmig_nn_destroy(kmig_nn *n)
{
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    if (n->exec) (void)hipGraphExecDestroy(n->exec);
    if (n->graph) (void)hipGraphDestroy(n->graph);
    nfree(n->net.w);
    for (int l = 0; l < 5; l++) {
        nfree(n->net.Wb[l]);
        nfree(n->net.WT[l]);
    }
    void *bufs[] = {n->m,    n->v,     n->grad, n->dev_rec, n->X,  n->Xt,   n->dist,
                    n->tgt,  n->wt,    n->Y,    n->lossb,   n->dY, n->dYt,  n->ctr,
                    n->hist, n->gloss, n->blob, n->bias};
    for (void *p : bufs) nfree(p);
    for (int l = 0; l < 4; l++) {
        nfree(n->Z[l]);
        nfree(n->H[l]);
        nfree(n->Ht[l]);
        nfree(n->dZ[l]);
        nfree(n->dZt[l]);
    }
    if (n->e0) (void)hipEventDestroy(n->e0);
    if (n->e1) (void)hipEventDestroy(n->e1);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

}  // extern "C"
