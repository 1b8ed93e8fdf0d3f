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

// Fourier features of one point, BF16, into row n of X [n][64] (48 used) and
// column n of Xt [64][n] (nullable)
__device__ void features(const NnScale &sc, float s, float z, float x, bf16 *X, bf16 *Xt,
                         long n, long N)
{
    const float u[3] = {(s - sc.s0) * sc.s_inv, (z - sc.z0) * sc.z_inv, (x - sc.x0) * sc.x_inv};
    bf16 *row = X + (size_t)n * NN_KIN;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float f = 3.14159265358979f * u[c];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            float sv, cv;
            sincosf(f, &sv, &cv);
            const int col = c * 16 + 2 * k;
            row[col] = (bf16)sv;
            row[col + 1] = (bf16)cv;
            if (Xt) {
                Xt[(size_t)col * N + n] = (bf16)sv;
                Xt[(size_t)(col + 1) * N + n] = (bf16)cv;
            }
            f *= 2.0f;
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

__global__ void k_points(const float *pts, long n0, long N, NnScale sc, bf16 *X, float *dist)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float *p = pts + (size_t)(n0 + n) * 3;
    features(sc, p[0], p[1], p[2], X, nullptr, n, N);
    dist[n] = src_dist(sc, p[0], p[1], p[2]);
}

// network outputs -> T and derivatives
__global__ void k_outputs(const float *Y, const float *dist, float S0, float *out, long N)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float *y = Y + (size_t)n * 4;
    float *o = out + (size_t)n * 4;
    o[0] = dist[n] * S0 * (1.0f + y[0]);
    o[1] = S0 * y[1];
    o[2] = S0 * y[2];
    o[3] = S0 * y[3];
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

// ----------------------------------------------------- network plumbing --

// layer l: fan-in K, outputs M, padded K of the BF16 operand, offsets of W, b
static const int LK[5] = {NN_FEAT, NN_W, NN_W, NN_W, NN_W};
static const int LM[5] = {NN_W, NN_W, NN_W, NN_W, NN_OUT};
static const int LKP[5] = {NN_KIN, NN_W, NN_W, NN_W, NN_W};

static size_t woff(int l)
{
    size_t o = 0;
    for (int k = 0; k < l; k++) o += (size_t)LM[k] * LK[k] + LM[k];
    return o;
}
static size_t boff(int l) { return woff(l) + (size_t)LM[l] * LK[l]; }

// columns of the transposed BF16 weight (the M of the backward-data product,
// padded to a multiple of 32 for the output layer)
static int ldt(int l) { return l == 4 ? 32 : LM[l]; }

void nn_cast_weights(const NnNet &net, hipStream_t st)
{
    for (int l = 0; l < 5; l++) {
        const int rows = std::max(LM[l], ldt(l)), cols = std::max(LK[l], LKP[l]);
        hipLaunchKernelGGL(k_cast, dim3(nblk((size_t)rows * cols, 256)), dim3(256), 0, st,
                           net.w + woff(l), LM[l], LK[l], net.Wb[l], LKP[l], net.WT[l], ldt(l));
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
        gemm(st, net.Wb[l], LKP[l], in, LKP[l], NN_W, (int)N, LKP[l], 1, e);
        in = H[l];
    }
    EpiOut o = out;
    o.bias = net.w + boff(4);
    gemm(st, net.Wb[4], NN_W, in, NN_W, NN_OUT, (int)N, NN_W, 1, o);
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
    for (int l = 0; l < 5; l++) mac += (double)LM[l] * LK[l];
    macx = (double)LM[0] * LK[0];
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
        rc = nalloc(n, &n->net.Wb[l], (size_t)LM[l] * LKP[l]);
        if (!rc) rc = nalloc(n, &n->net.WT[l], (size_t)LK[l] * ldt(l));
    }
    if (rc) return rc;
    // initial weights: uniform with variance 1 / fan-in, the output layer
    // scaled by 0.1 (starts near the straight-ray time d * S0); zero biases
    std::vector<float> w(NN_PARAMS, 0.0f);
    uint64_t st = cfg->seed;
    for (int l = 0; l < 5; l++) {
        const float a = std::sqrt(3.0f / (float)LK[l]) * (l == 4 ? 0.1f : 1.0f);
        for (size_t k = 0; k < (size_t)LM[l] * LK[l]; k++) {
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
    const long chunk = 65536;
    const long N = std::min<long>((count + 63) / 64 * 64, chunk);
    int rc = ensure_batch(n, std::max<long>(N, n->N));
    if (rc) return rc;
    float *dp = nullptr, *dout = nullptr;
    if ((rc = nalloc(n, &dp, (size_t)count * 3))) return rc;
    if ((rc = nalloc(n, &dout, (size_t)count * 4))) {
        nfree(dp);
        return rc;
    }
    hipError_t e = hipMemcpy(dp, pts, (size_t)count * 3 * sizeof(float), hipMemcpyHostToDevice);
    hipStream_t st = n->stream;
    for (long n0 = 0; n0 < count && e == hipSuccess; n0 += chunk) {
        const long m = std::min(chunk, count - n0);
        const long Np = (m + 63) / 64 * 64;
        hipLaunchKernelGGL(k_points, dim3(nblk((size_t)m, 256)), dim3(256), 0, st, dp, n0, m,
                           make_scale(n->cfg), n->X, n->dist);
        if (Np > m)  // padding rows of X: zeros
            hipLaunchKernelGGL(k_fillb, dim3(nblk((size_t)(Np - m) * NN_KIN, 256)), dim3(256), 0,
                               st, n->X + (size_t)m * NN_KIN, (size_t)(Np - m) * NN_KIN);
        EpiOut o{};
        o.Y = n->Y;
        forward(n->net, st, n->X, Np, n->H, nullptr, nullptr, o);
        hipLaunchKernelGGL(k_outputs, dim3(nblk((size_t)m, 256)), dim3(256), 0, st, n->Y,
                           n->dist, n->cfg.slowness, dout + (size_t)n0 * 4, m);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
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
                    n->hist, n->gloss};
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
