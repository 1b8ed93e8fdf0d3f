// Migration kernel of the kirchmig2d_cig adj=1 path for gfx950 (MI355X).
//
//   k_migrate  per image point, all shots and receivers of a batch: opening
//              angle, weights, angle bins, kirmig_pick, accumulation
//              (src/kirchmig2d_cig.cpp:1203-1299; SERECKIRCH/src/kirmig.cpp:34-58)
//
// One thread owns one image point and adds its updates in the reference's
// order (shot-major, receiver-minor) into its own LDS column of angle bins,
// so no atomics are needed and each (point, bin) sum keeps the reference's
// order.  The column starts from the CIG already in HBM, so splitting the
// shots over several launches keeps that order too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "kmig_index.h"
#include "kmig_internal.h"

namespace kmig {

#define KPI 3.14159265358979323846

// --------------------------------------------------------------- migrate --

// kirmig_pick(adj=1) in float.  The float checks on the floors reject the
// same windows as the reference's integer checks for deltat >= 0 (then
// itm <= it <= itp) and also keep every tap in [0, nt-1] otherwise.
__device__ __forceinline__ float pick(const float *__restrict__ tr, int nt, float t0, float dt,
                                      float inv_dt, float ti, float deltat)
{
    const float lim = (float)(nt - 1);
    float ft = (ti - t0) * inv_dt;
    const float fl = floorf(ft);
    if (!(fl >= 0.0f && fl < lim)) return 0.0f;
    const float tm = ti - deltat - dt;
    float ftm = (tm - t0) * inv_dt;
    const float flm = floorf(ftm);
    if (!(flm >= 0.0f && flm < lim)) return 0.0f;
    const float tp = ti + deltat + dt;
    float ftp = (tp - t0) * inv_dt;
    const float flp = floorf(ftp);
    if (!(flp >= 0.0f && flp < lim)) return 0.0f;
    ft -= fl;
    ftm -= flm;
    ftp -= flp;
    const int it = (int)fl, itm = (int)flm, itp = (int)flp;
    const float imp0 = dt / (dt + tp - tm);
    const float imp = imp0 * imp0;
    const float s = 2.0f * (1.0f - ft) * tr[it] + 2.0f * ft * tr[it + 1] -
                    (1.0f - ftm) * tr[itm] - ftm * tr[itm + 1] - (1.0f - ftp) * tr[itp] -
                    ftp * tr[itp + 1];
    return imp * s;
}

// kirmig_pick(adj=0): the transpose of pick above, the same window checks,
// the six weighted samples of v added into the trace.  Traces are summed by
// many points at once, so the adds are atomic.
__device__ __forceinline__ void spray(float *tr, int nt, float t0, float dt, float inv_dt,
                                      float ti, float deltat, float v)
{
    const float lim = (float)(nt - 1);
    float ft = (ti - t0) * inv_dt;
    const float fl = floorf(ft);
    if (!(fl >= 0.0f && fl < lim)) return;
    const float tm = ti - deltat - dt;
    float ftm = (tm - t0) * inv_dt;
    const float flm = floorf(ftm);
    if (!(flm >= 0.0f && flm < lim)) return;
    const float tp = ti + deltat + dt;
    float ftp = (tp - t0) * inv_dt;
    const float flp = floorf(ftp);
    if (!(flp >= 0.0f && flp < lim)) return;
    ft -= fl;
    ftm -= flm;
    ftp -= flp;
    const int it = (int)fl, itm = (int)flm, itp = (int)flp;
    const float imp0 = dt / (dt + tp - tm);
    const float s = imp0 * imp0 * v;
    atomicAdd(tr + it, 2.0f * (1.0f - ft) * s);
    atomicAdd(tr + it + 1, 2.0f * ft * s);
    atomicAdd(tr + itm, -(1.0f - ftm) * s);
    atomicAdd(tr + itm + 1, -ftm * s);
    atomicAdd(tr + itp, -(1.0f - ftp) * s);
    atomicAdd(tr + itp + 1, -ftp * s);
}

__device__ __forceinline__ float cos_ramp_up(float u) { return 0.5f * (1.0f + cosf((float)KPI * u)); }

// Model = true is adj=0: the LDS column holds the image (read only), and each
// accepted update sprays its weighted image value into the trace of a.spray
// instead of picking from it (src/kirchmig2d_cig.cpp:1300-1312, :1319-1323,
// :1333-1337; the two bins of a linear update are one spray of their
// weighted sum, the reference's two sprays added).
template <int TZ, int TX, bool Model = false>
__global__ __launch_bounds__(TZ *TX) void k_migrate(const MigArgs a)
{
    constexpr int P = TZ * TX;
    extern __shared__ float acc[];
    const int lid = threadIdx.x;
    const int ntz = (a.nz + TZ - 1) / TZ;
    const int bz = blockIdx.x % ntz, bx = blockIdx.x / ntz;
    const int iz = bz * TZ + lid % TZ, ix = bx * TX + lid / TZ;
    const bool valid = iz < a.nz && ix < a.nx;
    const size_t i = valid ? (size_t)ix * a.nz + iz : 0;
    float *my = acc + lid;
    for (int g = 0; g < a.ng; g++) my[g * P] = valid ? a.cig[(size_t)g * a.nzx + i] : 0.0f;

    const float zf = (a.nz > 1) ? (float)iz / (float)(a.nz - 1) : 0.0f;
    const float cutoff = a.oazmin + zf * (a.oazmax - a.oazmin);
    const bool taper_on = !(a.taper_start < 0.0f || a.taper_end <= a.taper_start);
    const float rad2deg = 180.0f / (float)KPI;

    for (int js = 0; js < a.nshots; js++) {
        const size_t si = (size_t)js * a.nzx + i;
        const float sw = valid ? a.sW[si] : 0.0f;
        if (__ballot(sw > 0.0f) == 0) continue;
        const float sT = a.sT[si], sA = a.sA[si];
        float sgx = 0.0f, sgz = 0.0f;
        if (a.mode != 2) {
            sgx = a.sGX[si];
            sgz = a.sGZ[si];
        }
        const int is = a.shot0 + js;
        const TraceMeta *__restrict__ mrow = a.meta + (size_t)is * a.nh;
        const float *__restrict__ trow = a.traces + (size_t)is * a.nh * a.nt;
        for (int ih = 0; ih < a.nh; ih++) {
            const float wap = mrow[ih].wap;
            if (wap < 0.0f) continue;
            const size_t ri = (size_t)mrow[ih].slot * a.nzx + i;
            const float apw = sw * a.rW[ri];
            if (!(apw > 0.0f)) continue;
            const float ti = sT + a.rT[ri] + a.tau;
            const float rA = a.rA[ri];
            const float tx = sA > rA ? sA : rA;
            const float *__restrict__ tr = trow + (size_t)ih * a.nt;
            float *str = Model ? a.spray + ((size_t)is * a.nh + ih) * a.nt : nullptr;
            if (a.mode == 2) {
                if constexpr (Model) {
                    spray(str, a.nt, a.t0, a.dt, a.inv_dt, ti, tx * a.aal,
                          wap * apw * my[mrow[ih].ig * P]);
                    continue;
                }
                const float picked = pick(tr, a.nt, a.t0, a.dt, a.inv_dt, ti, tx * a.aal);
                my[mrow[ih].ig * P] += wap * apw * picked;
                continue;
            }
            const float rgx = a.rGX[ri], rgz = a.rGZ[ri];
            const float dot = sgx * rgx + sgz * rgz;
            const float cross = sgx * rgz - sgz * rgx;
            float angle = atan2f(fabsf(cross), dot) * rad2deg;
            if (a.half_angle) angle *= 0.5f;
            float wt = 1.0f;
            if (taper_on) {
                if (angle >= a.taper_end) continue;
                if (angle > a.taper_start)
                    wt = cos_ramp_up((angle - a.taper_start) / (a.taper_end - a.taper_start));
            }
            float wd;
            if (a.oaztaper <= 0.0f) {
                wd = (angle <= cutoff) ? 1.0f : 0.0f;
            } else if (angle <= cutoff) {
                wd = 1.0f;
            } else if (angle >= cutoff + a.oaztaper) {
                wd = 0.0f;
            } else {
                wd = cos_ramp_up((angle - cutoff) / a.oaztaper);
            }
            const float aw = wt * wd * wap * apw;
            if (!(aw > 0.0f)) continue;
            const float u = (angle - a.g0) / a.dg;
            if (a.mode == 0) {
                if (u < -1.0e-5f || u > (float)(a.ng - 1) + 1.0e-5f) continue;
                int ig0, ig1;
                float w0, w1;
                if (u <= 0.0f) {
                    ig0 = ig1 = 0;
                    w0 = 1.0f;
                    w1 = 0.0f;
                } else if (u >= (float)(a.ng - 1)) {
                    ig0 = ig1 = a.ng - 1;
                    w0 = 1.0f;
                    w1 = 0.0f;
                } else {
                    const float fu = floorf(u);
                    ig0 = (int)fu;
                    ig1 = ig0 + 1;
                    w1 = u - fu;
                    w0 = 1.0f - w1;
                }
                if constexpr (Model) {
                    float m = w0 * my[ig0 * P];
                    if (ig1 != ig0 && w1 > 0.0f) m += w1 * my[ig1 * P];
                    spray(str, a.nt, a.t0, a.dt, a.inv_dt, ti, tx * a.aal, aw * m);
                    continue;
                }
                const float picked = aw * pick(tr, a.nt, a.t0, a.dt, a.inv_dt, ti, tx * a.aal);
                my[ig0 * P] += w0 * picked;
                if (ig1 != ig0 && w1 > 0.0f) my[ig1 * P] += w1 * picked;
            } else {
                const int ig = (int)floorf(u + 0.5f);
                if (ig < 0 || ig >= a.ng) continue;
                if constexpr (Model) {
                    spray(str, a.nt, a.t0, a.dt, a.inv_dt, ti, tx * a.aal, aw * my[ig * P]);
                    continue;
                }
                const float picked = pick(tr, a.nt, a.t0, a.dt, a.inv_dt, ti, tx * a.aal);
                my[ig * P] += aw * picked;
            }
        }
    }
    if (!Model && valid)
        for (int g = 0; g < a.ng; g++) a.cig[(size_t)g * a.nzx + i] = my[g * P];
}

// ------------------------------------------------------ resident variant --
//
// For receivers that do not move with the shot (cmp=0), each thread keeps R
// receiver tuples of its image point in registers and sweeps every shot under
// them: per shot it loads one source tuple (prefetched one shot ahead) and
// runs R independent updates.
// Per point the accumulation order is (receiver chunk, shot, receiver).
// The opening angle is |th_s - th_r| folded into [0, pi] (the angle between
// the two gradient directions, the quantity atan2(|s x r|, s . r) measures),
// and traveltimes arrive split into integer and fractional samples.

// Raised-cosine ramp 0.5 * (1 + cos(pi u)) for u in [0, 1], written as
// 0.5 - 0.5 * sin(pi v) with v = u - 1/2 and the Taylor series of sin to
// x^11 (truncation below 6e-8 on |x| <= pi/2).
__device__ __forceinline__ float ramp_down(float u)
{
    const float x = 3.14159265358979f * (u - 0.5f);
    const float x2 = x * x;
    float p = -2.5052108385441720e-08f;        // -1/11!
    p = p * x2 + 2.7557319223985893e-06f;      //  1/9!
    p = p * x2 - 1.9841269841269841e-04f;      // -1/7!
    p = p * x2 + 8.3333333333333333e-03f;      //  1/5!
    p = p * x2 - 1.6666666666666667e-01f;      // -1/3!
    p = p * x2 + 1.0f;
    return 0.5f - 0.5f * (x * p);
}

// Optional XCD-aware tile order (KMIG_XCD_MAP=1): blocks are dealt
// round-robin to the 8 XCDs, so this gives each XCD a contiguous run of tiles.
// Off by default: measured slower on the salt workload, because contiguous
// regions carry unequal work and the slowest XCD sets the kernel time, while
// the plain order spreads every region over all XCDs.  Clusters of 2 to 16
// z- or x-adjacent tiles per XCD, dealt round-robin (no imbalance), were
// measured too and gained nothing: the tap windows of concurrent neighbours
// do not meet in the L2.
#ifndef KMIG_XCD_MAP
#define KMIG_XCD_MAP 0
#endif

__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    if (!KMIG_XCD_MAP) return b;
    const int xcd = b & 7, k = b >> 3, q = nb >> 3, rem = nb & 7;
    return xcd < rem ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
}

// Phased form of the resident kernel: for a group of G receivers it first
// evaluates every weight, bin and tap index without branches, then issues the
// 3G tap-pair loads together, then accumulates.  Rejected updates keep a zero
// coefficient and read sample 0, so every wave has up to 3G loads in flight
// instead of one update's worth.  Per point the updates land in (receiver
// chunk, shot, receiver) order.
struct Pend {
    float coef, w1, ft, ftm, ftp;
    int ig0;               // first bin; the second is ig0 + 1 (row ng is a spare)
    unsigned o0, om, op;   // byte offsets of the three taps in the trace row
    bool ok;
};

// Tuple loaders: FP32 (Fmt 0); IEEE half fields in the 16-byte layout
// (Fmt 1: the fractional sample decoded to FP32 when the tuple is used, the
// packed (m, -256 q) pair kept as it is, PackedTuple); or 12-byte tuples whose
// 16-bit fields are decoded to the FP32 layout with the slot's scale (Fmt 2:
// bfloat16, 3: UNORM16).  Interpolation and accumulation are FP32 in every
// mode.
__device__ __forceinline__ float from16(uint32_t h, int fmt)
{
    if (fmt == 2) return __uint_as_float(h << 16);
    if (fmt == 3) return (float)h * (1.0f / 65535.0f);
    return (float)__builtin_bit_cast(_Float16, (unsigned short)h);
}

// Halves of a packed f16 pair as FP32 (exact conversions).
__device__ __forceinline__ float half_lo(uint32_t x)
{
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(x & 0xFFFFu));
}
__device__ __forceinline__ float half_hi(uint32_t x)
{
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(x >> 16));
}

// w * -(the high f16 half of x), in FP32 (one v_fma_mix_f32)
__device__ __forceinline__ float mul_neg_half_hi(float w, uint32_t x)
{
    float d;
    asm("v_fma_mix_f32 %0, %1, -%2, 0 op_sel:[0,1,0] op_sel_hi:[0,1,0]" : "=v"(d) : "v"(w), "v"(x));
    return d;
}

// Both halves' maxima of two packed f16 pairs in one instruction: for the
// (m, -256 q) pairs of the FP16 tuples this is (max m, -256 min q).
__device__ __forceinline__ uint32_t pk_max_f16(uint32_t x, uint32_t y)
{
    uint32_t d;
    asm("v_pk_max_f16 %0, %1, %2" : "=v"(d) : "v"(x), "v"(y));
    return d;
}

// f -/+ the low f16 half of x, in FP32 with one rounding (one v_fma_mix_f32:
// the half is an f16 operand of the FP32 fused multiply-add).
__device__ __forceinline__ float sub_half_lo(float f, uint32_t x)
{
    float d;
    asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(d) : "v"(x), "v"(f));
    return d;
}
__device__ __forceinline__ float add_half_lo(float f, uint32_t x)
{
    float d;
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(d) : "v"(x), "v"(f));
    return d;
}

// A tuple as loaded (Raw) and decoded: the loads of the next shot's tuple
// are issued one shot ahead and decoded only when used, so the decode does
// not wait on the load that is meant to be in flight.
struct Raw16 {
    PackedTuple16 q;
    float scale;
};

template <int Fmt> struct RawOf {
    using type = Raw16;
};
template <> struct RawOf<0> {
    using type = PackedTuple;
};
template <> struct RawOf<1> {
    using type = PackedTuple;
};

template <int Fmt>
__device__ __forceinline__ typename RawOf<Fmt>::type load_raw(const MigResArgs &a, bool src,
                                                              size_t slot, size_t i)
{
    const size_t idx = slot * a.nzx + i;
    if constexpr (Fmt <= 1) {
        return src ? a.stup[idx] : a.rtup[idx];
    } else {
        Raw16 r;
        r.q = src ? a.stup16[idx] : a.rtup16[idx];
        r.scale = src ? a.sscale[slot].a_scale : a.rscale[slot].a_scale;
        return r;
    }
}

// Decoded tuples carry the FP32 layout, whose a field is 1 + antialias * a
// (PackedTuple): the 12-byte 16-bit fields fold it in here, once per tuple.
// FP16 tuples decode only the fractional sample (their a field stays the
// packed (m, -256 q) pair that the update reads as it is).
template <int Fmt>
__device__ __forceinline__ PackedTuple decode(const typename RawOf<Fmt>::type &r, float aal)
{
    if constexpr (Fmt == 0) {
        return r;
    } else if constexpr (Fmt == 1) {
        PackedTuple t = r;
        t.frac = half_lo(__float_as_uint(r.frac));
        return t;
    } else {
        PackedTuple t;
        t.iw = r.q.iw;
        t.frac = Fmt == 3 ? (float)(r.q.fa & 0xFFFFu) * (1.0f / 65536.0f)
                          : from16(r.q.fa & 0xFFFFu, Fmt);
        // capped as the FP32 field (an update this wide is rejected anyway)
        t.a = fminf(fmaf(from16(r.q.fa >> 16, Fmt) * r.scale, aal, 1.0f), 1048576.0f);
        t.th = __uint_as_float(r.q.th);
        return t;
    }
}

template <int Fmt>
__device__ __forceinline__ PackedTuple load_tuple(const MigResArgs &a, bool src, size_t slot,
                                                  size_t i)
{
    return decode<Fmt>(load_raw<Fmt>(a, src, slot, i), a.aal);
}

// Opening angle from the two packed gradient directions (output-angle
// degrees): their difference folded into [0, wrap / 2].
__device__ __forceinline__ float pair_angle(const PackedTuple &s, const PackedTuple &r,
                                           float wrap)
{
    const float d = fabsf(s.th - r.th);
    return fminf(d, wrap - d);
}

// floor(x) as an integer in one instruction (the compiler emits v_floor and
// v_cvt_i32 for (int)floorf(x)); x is finite here
__device__ __forceinline__ int cvt_floor(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Everything of one update after the weights: aw is the product of the
// trace-aperture weight, both tuple weights and the angle weights (depth
// mute, taper, bin range), 0 or negative (skipped trace) where rejected;
// u is the bin coordinate.  s_int8, s_frac: the source tuple's integer
// samples (as a byte offset of 8-byte tap pairs) and fractional samples, with
// the constant (tau - t0)/dt already added (once per shot).
// Bin = 0: two linear angle bins from u; Bin = 1: the trace's own bin ibin
// with all of the weight (offset gathers and the stack).
//
// FP16 tuples (Fmt 1) carry m = antialias * a instead of 1 + m, and -256 q
// with q = (1 / (3 + 2 m))^2 beside it: one packed f16 max gives the update's
// larger m and smaller q (q falls as m grows) with no reciprocal per update.
// Its tap offsets are kept one pair low (b8 - 8): the middle tap is then
// loaded 8 bytes and the outer late tap 16 bytes past its offset (immediate
// load offsets, TapOff below), so the outer taps floor(ff -/+ m) -/+ 1 cost
// nothing over floor(ff -/+ m).  The caller's weights carry the 2^-8 of q.
template <int Fmt> struct TapOff {
    static constexpr int o0 = 0, om = 0, op = 0;
};
template <> struct TapOff<1> {
    static constexpr int o0 = 8, om = 0, op = 16;
};

template <int Bin, int Fmt>
__device__ __forceinline__ Pend res_phase1(const PackedTuple &s, int s_int8, float s_frac,
                                           const PackedTuple &r, float aw, float u, int ibin,
                                           const MigResArgs &a)
{
    if constexpr (Fmt == 1) {
        Pend q;
        int ig0;
        float w1;
        if constexpr (Bin == 0) {
            const float uc = __builtin_amdgcn_fmed3f(u, 0.0f, (float)(a.ng - 1));
            ig0 = cvt_floor(uc);
            w1 = __builtin_amdgcn_fractf(uc);
        } else {
            ig0 = ibin;
            w1 = 0.0f;
        }
        int b8 = s_int8 + (((int)r.iw >> 16) << 3) - 8;
        asm("" : "+v"(b8));
        const float ff = s_frac + r.frac;
        const uint32_t mq = pk_max_f16(__float_as_uint(s.a), __float_as_uint(r.a));
        const float fm = sub_half_lo(ff, mq), fp = add_half_lo(ff, mq);
        const int it8 = b8 + (cvt_floor(ff) << 3);   // middle tap at it8 + 8
        const int itm8 = b8 + (cvt_floor(fm) << 3);  // early tap at itm8
        const int itp8 = b8 + (cvt_floor(fp) << 3);  // late tap at itp8 + 16
        // taps inside pairs [0, nt - 2]: itm8 >= 0 and itp8 + 16 < lim8 (the
        // host requires nt >= 3, so lim8 - 16 >= 0); itm8 <= itp8
        const unsigned lim8 = (unsigned)(a.nt - 1) << 3;
        const bool ok = (aw > 0.0f) & ((unsigned)itm8 < lim8 - 16u) & ((unsigned)itp8 < lim8 - 16u);
        q.ok = ok;
        q.coef = mul_neg_half_hi(aw, mq);
        q.w1 = w1;
        q.ft = __builtin_amdgcn_fractf(ff);
        q.ftm = __builtin_amdgcn_fractf(fm);
        q.ftp = __builtin_amdgcn_fractf(fp);
        q.ig0 = ig0;
        // a rejected update's three loads land on the trace's last pair, (0, 0)
        q.o0 = ok ? (unsigned)it8 : lim8 - 8u;
        q.om = ok ? (unsigned)itm8 : lim8;
        q.op = ok ? (unsigned)itp8 : lim8 - 16u;
        return q;
    }
    Pend q;
    // linear bins: u clamped to [0, ng-1] gives the edge bin with w1 = 0 at
    // either end and (floor(u), u - floor(u)) inside
    int ig0;
    float w1;
    if constexpr (Bin == 0) {
        const float uc = __builtin_amdgcn_fmed3f(u, 0.0f, (float)(a.ng - 1));
        ig0 = cvt_floor(uc);
        w1 = __builtin_amdgcn_fractf(uc);
    } else {
        ig0 = ibin;
        w1 = 0.0f;
    }
    // split-sample time: integer parts add as integers (in bytes of 8-byte
    // tap pairs), fractions in float.  Taps at ff and ff -/+ (delta + 1)
    // samples; both a fields are finite and >= 1 (one plus antialias >= 0
    // times a >= 0; the resident variant requires antialias >= 0), so
    // delta + 1 >= 1 and itm < it < itp: the outer taps bound the middle one.
    int b8 = s_int8 + (((int)r.iw >> 16) << 3);
    // kept as one value: re-associated, the three tap offsets cost two
    // integer instructions each instead of one
    asm("" : "+v"(b8));
    const float ff = s_frac + r.frac;
    // 1 + antialias * max(a_s, a_r): the larger folded field (positive floats
    // order as their bit patterns: one integer max)
    const float dA1 = __int_as_float(max(__float_as_int(s.a), __float_as_int(r.a)));
    const float fm = ff - dA1, fp = ff + dA1;
    const int it8 = b8 + (cvt_floor(ff) << 3);
    const int itm8 = b8 + (cvt_floor(fm) << 3), itp8 = b8 + (cvt_floor(fp) << 3);
    const unsigned lim8 = (unsigned)(a.nt - 1) << 3;
    const bool ok = (aw > 0.0f) & ((unsigned)itm8 < lim8) & ((unsigned)itp8 < lim8);
    const float rr = __builtin_amdgcn_rcpf(fmaf(2.0f, dA1, 1.0f));  // 1 / (3 + 2 delta)
    // A rejected update reads the trace's last pair, (0, 0), at all three taps:
    // with its finite fractions and coefficient its picked value is an exact
    // zero, added to bins inside [0, ng] (uc is clamped; row ng is a spare
    // that only ever receives w1 = 0 shares), so only the tap indices are
    // selected.
    q.ok = ok;
    q.coef = aw * (rr * rr);
    q.w1 = w1;
    q.ft = __builtin_amdgcn_fractf(ff);
    q.ftm = __builtin_amdgcn_fractf(fm);
    q.ftp = __builtin_amdgcn_fractf(fp);
    q.ig0 = ig0;
    q.o0 = ok ? (unsigned)it8 : lim8;
    q.om = ok ? (unsigned)itm8 : lim8;
    q.op = ok ? (unsigned)itp8 : lim8;
    return q;
}

// Wave-wide integer min / max for full waves: four DPP steps inside each row
// of 16 lanes (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror: every lane ends with its row's extremum), then the four rows'
// values read as scalars.  No LDS round trips, unlike a shuffle chain.
__device__ __forceinline__ int wave_min_i32(int v)
{
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, true));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, true));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, true));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, true));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int wave_max_i32(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, true));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// LDS ordering among the lanes of one wave: the wave's LDS operations run in
// order, so only the compiler must keep them so.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Survivor masks of the group culls for one launch (MigResArgs::mask).  For
// bound cell c (blockIdx.x) and launch-local shots 64 w .. 64 w + 63
// (blockIdx.y = w; lane = shot), over the receiver chunks of R receivers: a
// group is culled where the cell's source bounds (sbnd) and the union of the
// chunk's receiver bounds in the cell (rbnd) show that no update can add a
// nonzero value:
//  * no point of the cell has source or receiver weight (empty bounds);
//  * time window: an update's last tap lies at s_int + r_int + c_int +
//    floor(ff + delta + 1) >= s_int + r_int + c_int + 1 (ff, delta >= 0), so
//    where the cell's smallest source plus receiver integer samples reach
//    nt - 2 - c_int every tap pair lies past the trace's last pair;
//  * direction: th_s - th_r lies in [dlo, dhi], and the folded opening angle
//    min(|d|, wrap - |d|) is smallest at an end of the |d| interval; where it
//    exceeds tq, the largest angle that receives weight at any row of the
//    cell (plus 1e-3 degrees, far above the rounding of the kernel's angle),
//    every update has zero weight (no_dir: per-trace bins, no direction
//    test);
//  * receiver columns (Cols): the chunk's columns have no trace in the shot.
// All bounds are integers (directions x BND_Q, floor / ceil), so the tests
// are exact; a wave whose points span several cells evaluates a group where
// any of its cells keeps it.
__device__ __forceinline__ bool group_culled(const ThBound &s, int rlo, int rhi, int rtm,
                                             int t_cull, int tq, int wrapq)
{
    if (s.lo > s.hi || rlo > rhi) return true;
    if (s.tmin + rtm >= t_cull) return true;
    const int dlo = s.lo - rhi, dhi = s.hi - rlo;
    const int amin = (dlo <= 0 && dhi >= 0) ? 0 : min(abs(dlo), abs(dhi));
    const int bmax = max(abs(dlo), abs(dhi));
    return min(amin, wrapq - bmax) > tq;
}

template <bool Cols>
__global__ __launch_bounds__(64) void k_cull_mask(const MigResArgs a, int R, int no_dir)
{
    const int c = blockIdx.x, w = blockIdx.y, lane = threadIdx.x;
    const int nwords = gridDim.y;
    const int nrec = Cols ? a.ncols : a.nh;
    const int nchunk = (nrec + R - 1) / R;
    const int j = w * 64 + lane;
    const bool in = j < a.nshots;
    ThBound sb;
    sb.lo = BND_EMPTY;
    sb.hi = -BND_EMPTY;
    sb.tmin = BND_EMPTY;
    sb.pad = 0;
    if (in) sb = a.sbnd[idx::bnd(c, (size_t)a.sbnd_stride, j)];
    // the largest angle that receives weight at any row of the cell: the depth
    // mute's cutoff with its taper band, capped by the bin range (rows past
    // nz are included, which only widens it)
    int tq = 1 << 30;
    if (!no_dir) {
        const int iz = (c % a.ncz) * BND_CZ + (lane % BND_CZ);
        const float zf = (a.nz > 1) ? (float)iz / (float)(a.nz - 1) : 0.0f;
        const float cutoff = a.oazmin + zf * (a.oazmax - a.oazmin);
        const float plim = fminf(cutoff, a.ang_hi);
        const float bhi = fminf(cutoff + a.oaztaper, a.ang_hi_up);
        float tl = a.oaztaper > 0.0f ? bhi : plim;
        for (int o = 32; o > 0; o >>= 1) tl = fmaxf(tl, __shfl_xor(tl, o, 64));
        tl = fmaxf(fminf(tl, 1000.0f), -1000.0f);
        tq = __builtin_amdgcn_readfirstlane((int)ceilf(tl * BND_Q)) + 64;
    }
    const int wrapq = (int)(a.wrap * BND_Q);
    for (int k = 0; k < nchunk; k++) {
        // union of the chunk's receiver bounds in the cell (lanes 0 .. R - 1)
        const int ih = k * R + lane;
        int rlo = BND_EMPTY, rhi = -BND_EMPTY, rtm = BND_EMPTY, cf = 1 << 30, cl = -1;
        if (lane < R && ih < nrec) {
            const int slot = a.rslot[ih];
            if (slot >= 0) {
                const ThBound rb = a.rbnd[idx::bnd(c, (size_t)a.nrslots, slot)];
                rlo = rb.lo;
                rhi = rb.hi;
                rtm = rb.tmin;
            }
            if (Cols) {
                cf = a.cfirst[ih];
                cl = a.clast[ih];
            }
        }
        rlo = wave_min_i32(rlo);
        rhi = wave_max_i32(rhi);
        rtm = wave_min_i32(rtm);
        bool keep = in && !group_culled(sb, rlo, rhi, rtm, a.t_cull, tq, wrapq);
        if (Cols) {
            cf = wave_min_i32(cf);
            cl = wave_max_i32(cl);
            keep = keep && a.scell_off + j >= cf && a.scell_off + j <= cl;
        }
        const unsigned long long m = __ballot(keep);
        if (lane == 0) a.mask[idx::mask_word(c, nchunk, k, nwords, w)] = m;
    }
}

// The survivor masks of a launch with R receivers per chunk (before the
// launch, on its stream).
template <bool Cols>
static int launch_cull_mask(const MigResArgs &a, int R, int no_dir, hipStream_t st)
{
    if (!a.mask || a.nshots <= 0) return 0;
    hipLaunchKernelGGL(k_cull_mask<Cols>, dim3((unsigned)(a.ncz * a.ncx),
                       (unsigned)((a.nshots + 63) / 64)), dim3(64), 0, st, a, R, no_dir);
    return (int)hipGetLastError();
}

// Probe = true only renames the kernel for the autotuner's timing launches,
// so profiler statistics of the production kernel cover production launches.
// One tile of TZ x TX image points of one problem.
// Model = true is the adj=0 transpose: the LDS column holds the image (read
// only) and every accepted update sprays its six weighted samples into the
// traces of a.spray instead of reading its taps.
// Count = true: the same work, but the wave counts the updates it evaluates
// and accepts (a.counts) and writes no CIG.
// Cols = true: receivers that move with the shot (cmp=1).  The chunks run
// over receiver columns (a.ncols; column c keeps tuple slot a.rslot[c]) and
// each chunk sweeps only the shots with a trace in one of its columns, whose
// receiver index and trace-aperture weight come from a.ctab.
// Bin = 1: offset gathers and the stacked image (cig=0): no directions, no
// angle weights; every update goes with all of its weight to its trace's bin
// a.rbin[receiver index] (src/kirchmig2d_cig.cpp:1100-1102, :1327-1331).
template <int TZ, int TX, int R, int G, int Fmt, bool Model = false, bool Count = false,
          bool Cols = false, int Bin = 0>
__device__ __forceinline__ void migrate_tile(const MigResArgs &a, int tile, float *acc,
                                             int sg = 0, int nsplit = 1)
{
    static_assert(!(Model && Cols), "adj=0 with receiver columns runs on the stream kernel");
    static_assert(!(Model && Bin), "adj=0 offset gathers and stacks run on the stream kernel");
    static_assert(!(Model && Fmt), "adj=0 reads FP32 tuples");
    constexpr int P = TZ * TX;
    const int lid = threadIdx.x;
    // tile order z fastest (measured faster than x fastest on the salt
    // workload, where consecutive tiles run together on one XCD)
    const int ntz = (a.nz + TZ - 1) / TZ;
    const int bz = tile % ntz, bx = tile / ntz;
    const int iz = bz * TZ + lid % TZ, ix = bx * TX + lid / TZ;
    const bool valid = iz < a.nz && ix < a.nx;
    const size_t i = valid ? (size_t)ix * a.nz + iz : 0;
    float *my = acc + lid;
    // ng bins and a spare row ng, which takes the zero second share of
    // updates in the last bin, so both bins of an update are one LDS pair
    // the batch's earlier sums enter through shot group 0 only; in the
    // transpose every group reads the image
    for (int g = 0; g <= a.ng; g++)
        my[g * P] = valid && g < a.ng && (Model || sg == 0) ? a.cig[(size_t)g * a.nzx + i] : 0.0f;
    if constexpr (Model) {
        float *win = acc + (a.ng + 1) * P + (lid >> 6) * MODEL_WIN;
        for (int e = lid & 63; e < MODEL_WIN; e += 64) win[e] = 0.0f;
        wave_lds_sync();
    }
    const int jb = (int)((long)sg * a.nshots / nsplit);
    const int je = (int)((long)(sg + 1) * a.nshots / nsplit);
    const float zf = (a.nz > 1) ? (float)iz / (float)(a.nz - 1) : 0.0f;
    const float cutoff = a.oazmin + zf * (a.oazmax - a.oazmin);
    // depth mute: full weight up to cutoff, the ramp below cutoff_hi, zero
    // above; every update also needs angle <= ang_hi (bin range), so the
    // full-weight test is one compare against plim, and the ramp band is
    // (cutoff, bhi) with bhi also capped by the bin range (the band is empty
    // wherever the cutoff lies above the last bin)
    const float cutoff_hi = cutoff + a.oaztaper;
    const float plim = fminf(cutoff, a.ang_hi);
    const float bhi = fminf(cutoff_hi, a.ang_hi_up);
    const bool taper_on = !(a.taper_start < 0.0f || a.taper_end <= a.taper_start);
    // whether any point of the wave has a nonempty depth-mute ramp band below
    // the last bin (cutoff < bhi); where none has, an update's weight is its
    // base weight up to plim and zero beyond, and the band test is skipped
    const bool wave_band = a.oaztaper > 0.0f && __ballot(valid && cutoff < bhi) != 0;
    const float inv32k = 1.0f / 32768.0f;
    const float2 *__restrict__ pairs = (const float2 *)a.pairs;

    // Group culling (survivor masks, k_cull_mask): a (shot, receiver chunk)
    // group is evaluated only where one of the bound cells holding the
    // wave's points has its bit set, i.e. where the group can add a nonzero
    // update; a skipped group's updates would all add exact zeros, so the
    // sums are unchanged.  A wave covers all TZ rows of WX columns; tiles
    // start on cell boundaries, so NC cells hold its points.
    constexpr int WX = 64 / TZ;
    constexpr int NCZ = (TZ + BND_CZ - 1) / BND_CZ;
    constexpr int NCX = (WX + BND_CX - 1) / BND_CX;
    constexpr int NC = NCZ * NCX;
    const bool cull = a.mask != nullptr;
    int cell[NC];
    {
        const int wv = __builtin_amdgcn_readfirstlane(lid >> 6);
        const int cz0 = (bz * TZ) / BND_CZ, cx0 = (bx * TX + wv * WX) / BND_CX;
#pragma unroll
        for (int jx = 0; jx < NCX; jx++)
#pragma unroll
            for (int jz = 0; jz < NCZ; jz++)
                cell[jx * NCZ + jz] = idx::cell(cx0, jx, cz0, jz, a.ncx, a.ncz);
    }

    unsigned long long n_eval = 0, n_acc = 0;  // Count only (wave-uniform)
    const unsigned long long n_valid = Count ? __popcll(__ballot(valid)) : 0;
    // receiver indices (cmp=0) or receiver columns (Cols) of the chunks
    const int nrec = Cols ? a.ncols : a.nh;
    const int nchunk = (nrec + R - 1) / R, nwords = (a.nshots + 63) >> 6;
    for (int r0 = 0; r0 < nrec; r0 += R) {
        // the chunk's shots: the shot group's, and with columns only those
        // with a trace in one of the chunk's columns (the masks hold those
        // ranges already)
        int cjb = jb, cje = je;
        if (Cols && !cull) {
            int f = 1 << 30, l = -1;
            for (int r = 0; r < R && r0 + r < nrec; r++) {
                f = min(f, a.cfirst[r0 + r]);
                l = max(l, a.clast[r0 + r]);
            }
            cjb = max(jb, f - a.scell_off);
            cje = min(je, l + 1 - a.scell_off);
            if (cjb >= cje) continue;
        }
        // Surviving shots in order, from the chunk's mask words (scalar loads,
        // one per 64 shots and cell; the next shot is a scalar bit scan)
        const int kc = r0 / R;
        int wi = -1;
        unsigned long long mw = 0;
        auto next_shot = [&](int j) {
            if (!cull) return j;
            while (j < cje) {
                const int w = j >> 6;
                if (w != wi) {
                    wi = w;
                    mw = 0;
#pragma unroll
                    for (int c = 0; c < NC; c++)
                        mw |= a.mask[idx::mask_word(cell[c], nchunk, kc, nwords, w)];
                }
                const unsigned long long m = mw & (~0ull << (j & 63));
                if (m) return min(cje, (w << 6) + (int)__builtin_ctzll(m));
                j = (w + 1) << 6;
            }
            return cje;
        };
        int js = next_shot(cjb);
        if (js >= cje) continue;  // no group of the chunk survives
        // the chunk's R receiver tuples: all loads issued first and decoded
        // after (a decode next to each load would wait for it, one memory
        // round trip per receiver); an unused receiver has zero fields
        // The chunk's slots are one scalar load (rslot holds -1 past the last
        // receiver); an unused receiver loads slot 0's tuple and gets zero
        // weight, so its updates are rejected as those of a zero tuple.
        int rsl[R];
#pragma unroll
        for (int r = 0; r < R; r++) rsl[r] = a.rslot[r0 + r];
        typename RawOf<Fmt>::type rraw[R];
#pragma unroll
        for (int r = 0; r < R; r++) rraw[r] = load_raw<Fmt>(a, false, (size_t)max(rsl[r], 0), i);
        PackedTuple rt[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            rt[r] = decode<Fmt>(rraw[r], a.aal);
            if (rsl[r] < 0) rt[r].iw = 0u;
        }
        // The shot's traces of the chunk: trace-aperture weights, and with
        // columns the receiver index per column (the column table's rows are
        // padded; a padding column has zero tuple weight).  Scalar loads; for
        // fixed receivers issued with the shot's tuple one surviving shot
        // ahead (with columns, twice the scalar registers of a prefetched row
        // spill, and the row is read when the shot is processed).
        ColTrace tw[R], twn[R];  // fixed receivers only
        auto load_traces = [&](int j) {
            if constexpr (!Cols) {
                const float *__restrict__ wr = a.wap + idx::wap(a.shot0 + j, a.nh, r0, 0);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    twn[r].ih = 0;
                    twn[r].wap = wr[r];
                }
            }
        };
        typename RawOf<Fmt>::type nxt{};
        if (js < cje) {
            nxt = load_raw<Fmt>(a, true, (size_t)js, i);
            load_traces(js);
        }
        while (js < cje) {
            const PackedTuple st = decode<Fmt>(nxt, a.aal);
            const int jc = js;
            if constexpr (!Cols) {
#pragma unroll
                for (int r = 0; r < R; r++) tw[r] = twn[r];
            }
            // the column table's row of this shot (Cols)
            const ColTrace *__restrict__ crow =
                Cols ? a.ctab + idx::ctab(a.scell_off, jc, a.ncols, r0, 0) : nullptr;
            js = next_shot(js + 1);
            if (js < cje) {
                nxt = load_raw<Fmt>(a, true, (size_t)js, i);
                load_traces(js);
            }
            const bool live = valid && (st.iw & 0xFFFFu) != 0u;
            if (__ballot(live) == 0) continue;
            if constexpr (Count) n_eval += n_valid * (unsigned long long)min(R, nrec - r0);
            const int s_int8 = (((int)st.iw >> 16) + a.c_int) << 3;
            const float s_frac = st.frac + a.c_frac;
            // both tuple weights' 2^-15 (and the 2^-8 of the FP16 tuples' q)
            const float s_w = (float)(st.iw & 0xFFFFu) * (Fmt == 1 ? inv32k * inv32k * (1.0f / 256.0f)
                                                                   : inv32k * inv32k);
            const int is = a.shot0 + jc;
            const float2 *__restrict__ trow = pairs + idx::chunk_row(is, a.nh, r0) * a.nt;
#pragma unroll
            for (int g0 = 0; g0 < R; g0 += G) {
                // stage 1: opening angle, bin coordinate and weight.  aw =
                // trace-aperture weight x both tuple weights where the angle
                // passes the depth mute and the bin range, else 0; skipped
                // traces (wap < 0) leave aw < 0, and the receivers past nh
                // that pad the last chunk have zero tuple weight (their wap
                // entries are the next shot's, or the pad after the last
                // shot: the whole array is zeroed before a run writes its
                // rows, so they are finite).  The depth-mute
                // ramp is evaluated only when some lane of the wave has an
                // angle inside the band; outside it the weight is 1 or 0.  (A
                // wave-level skip of rejected groups measured slower: few
                // groups are empty.)
                float ang[G], u[G], aw[G], base[G];
                int ibin[G];  // per-trace bins (offset gathers, stack)
                bool band = false;
#pragma unroll
                for (int k = 0; k < G; k++) {
                    const float wk = Cols ? crow[g0 + k].wap : tw[g0 + k].wap;
                    base[k] = wk * s_w * (float)(rt[g0 + k].iw & 0xFFFFu);
                    if constexpr (Bin == 0) {
                        ang[k] = pair_angle(st, rt[g0 + k], a.wrap);
                        u[k] = fmaf(ang[k], a.inv_dg, a.g0_inv_dg);  // (angle - g0) / dg
                        aw[k] = ang[k] <= plim ? base[k] : 0.0f;
                    } else {
                        // the trace's own bin (a receiver index past nh pads
                        // the chunk with zero weight: any bin of the axis)
                        const int ih = Cols ? crow[g0 + k].ih : idx::imin(r0 + g0 + k, a.nh - 1);
                        ibin[k] = a.rbin[ih];
                        u[k] = 0.0f;
                        aw[k] = base[k];
                    }
                }
                if (Bin == 0 && wave_band) {
#pragma unroll
                    for (int k = 0; k < G; k++)  // bitwise, not short-circuit: no branches
                        band |= (ang[k] > cutoff) & (ang[k] < bhi);
                }
                if (Bin == 0 && wave_band && __ballot(band) != 0) {
#pragma unroll
                    for (int k = 0; k < G; k++) {
                        const float rd = ramp_down(
                            fminf(fmaxf((ang[k] - cutoff) * a.inv_oaztaper, 0.0f), 1.0f));
                        const float wd =
                            ang[k] <= cutoff ? 1.0f : (ang[k] >= cutoff_hi ? 0.0f : rd);
                        aw[k] = ang[k] <= a.ang_hi ? wd * base[k] : 0.0f;
                    }
                }
                if (Bin == 0 && a.lo_check) {
#pragma unroll
                    for (int k = 0; k < G; k++) aw[k] = ang[k] >= a.ang_lo ? aw[k] : 0.0f;
                }
                // the angle taper (angle_taper_start/end) multiplies in here
                if (Bin == 0 && taper_on) {
#pragma unroll
                    for (int k = 0; k < G; k++) {
                        const float rt_ = ramp_down(fminf(
                            fmaxf((ang[k] - a.taper_start) * a.inv_taper, 0.0f), 1.0f));
                        aw[k] *= ang[k] >= a.taper_end ? 0.0f
                                                       : (ang[k] > a.taper_start ? rt_ : 1.0f);
                    }
                }
                // stage 2: bins and tap indices
                Pend q[G];
                bool any = false;
#pragma unroll
                for (int k = 0; k < G; k++) {
                    q[k] = res_phase1<Bin, Fmt>(st, s_int8, s_frac, rt[g0 + k], aw[k], u[k],
                                                ibin[k], a);
                    any |= q[k].ok;
                }
                if constexpr (Count) {
#pragma unroll
                    for (int k = 0; k < G; k++) n_acc += __popcll(__ballot(q[k].ok & valid));
                    continue;
                }
                if (__ballot(any) == 0) continue;
                if constexpr (Model) {
                    // kirmig_pick(adj=0), SERECKIRCH/src/kirmig.cpp:59-67: the two
                    // bins' image values, weighted, sprayed with the transposed
                    // three-tap weights (offsets are in 8-byte pair units: >> 3
                    // gives samples, and ok keeps every tap in [0, nt - 1]).  The
                    // wave's updates of one trace land in a narrow window of it:
                    // they are summed in the wave's LDS window first and the
                    // window is added to the trace with one atomic per touched
                    // sample (windows wider than MODEL_WIN add directly).
                    float *win = acc + (a.ng + 1) * P + (lid >> 6) * MODEL_WIN;
                    const int ln = lid & 63;
#pragma unroll
                    for (int k = 0; k < G; k++) {
                        const bool sp = q[k].ok & valid;
                        // every lane of the wave runs this (uniform control
                        // flow, full waves of 64)
                        int lo = wave_min_i32(sp ? (int)(q[k].om >> 3) : 0x7fffffff);
                        int hi = wave_max_i32(sp ? (int)(q[k].op >> 3) + 1 : -1);
                        if (hi < lo) continue;  // no lane sprays this trace
                        // taps lie in [0, nt - 1]: the window stays inside the trace
                        lo = max(lo, 0);
                        hi = min(hi, a.nt - 1);
                        const float *cell = my + q[k].ig0 * P;
                        const float v =
                            q[k].coef * ((1.0f - q[k].w1) * cell[0] + q[k].w1 * cell[P]);
                        float *tr = a.spray + ((size_t)is * a.nh + r0 + g0 + k) * a.nt;
                        const int it = (int)(q[k].o0 >> 3), itm = (int)(q[k].om >> 3),
                                  itp = (int)(q[k].op >> 3);
                        const float w0 = 2.0f * (1.0f - q[k].ft) * v, w1 = 2.0f * q[k].ft * v;
                        const float wm0 = -(1.0f - q[k].ftm) * v, wm1 = -q[k].ftm * v;
                        const float wp0 = -(1.0f - q[k].ftp) * v, wp1 = -q[k].ftp * v;
                        const int span = hi - lo + 1;
                        if (span <= MODEL_WIN) {
                            // the window is all zero here: cleared at the start
                            // and by every flush
                            if (sp) {
                                atomicAdd(win + (it - lo), w0);
                                atomicAdd(win + (it + 1 - lo), w1);
                                atomicAdd(win + (itm - lo), wm0);
                                atomicAdd(win + (itm + 1 - lo), wm1);
                                atomicAdd(win + (itp - lo), wp0);
                                atomicAdd(win + (itp + 1 - lo), wp1);
                            }
                            wave_lds_sync();
                            for (int e = ln; e < span; e += 64) {
                                const float sv = win[e];
                                win[e] = 0.0f;
                                if (sv != 0.0f) atomicAdd(tr + lo + e, sv);
                            }
                            wave_lds_sync();
                        } else if (sp) {
                            atomicAdd(tr + it, w0);
                            atomicAdd(tr + it + 1, w1);
                            atomicAdd(tr + itm, wm0);
                            atomicAdd(tr + itm + 1, wm1);
                            atomicAdd(tr + itp, wp0);
                            atomicAdd(tr + itp + 1, wp1);
                        }
                    }
                    continue;
                }
                // tap offsets are 32-bit byte offsets from the wave-uniform
                // trace row (indices are in [0, nt - 1])
                float2 v0[G], vm[G], vp[G];
#pragma unroll
                for (int k = 0; k < G; k++) {
                    // receivers past nh (the last chunk's padding) are
                    // rejected and read the sentinel pair of the shot's last
                    // real trace: their own row would lie past this shot
                    // (past the array for the last shot)
                    const char *tr;
                    if constexpr (Cols) {
                        // a column without a trace in this shot (weight < 0)
                        // reads receiver 0's sentinel pair
                        tr = (const char *)(pairs + idx::col_row(is, a.nh, crow[g0 + k].ih) * a.nt);
                    } else {
                        tr = (const char *)(trow + (size_t)idx::tap_row_off(a.nh, r0, g0 + k) * a.nt);
                    }
                    v0[k] = *(const float2 *)(tr + (size_t)q[k].o0 + TapOff<Fmt>::o0);
                    vm[k] = *(const float2 *)(tr + (size_t)q[k].om + TapOff<Fmt>::om);
                    vp[k] = *(const float2 *)(tr + (size_t)q[k].op + TapOff<Fmt>::op);
                }
                // Branch-free accumulation: a rejected update adds an exact
                // zero, and a one-bin update adds its zero second share to the
                // same bin, so the 3G loads above are not sunk into
                // per-update branches and stay in flight together.
#pragma unroll
                for (int k = 0; k < G; k++) {
                    // pairs hold (tr[k], tr[k+1] - tr[k])
                    const float sm = 2.0f * (v0[k].x + q[k].ft * v0[k].y) -
                                     (vm[k].x + q[k].ftm * vm[k].y) -
                                     (vp[k].x + q[k].ftp * vp[k].y);
                    const float picked = q[k].coef * sm;
                    float *cell = my + q[k].ig0 * P;
                    cell[0] += (1.0f - q[k].w1) * picked;
                    cell[P] += q[k].w1 * picked;
                }
            }
        }
    }
    if constexpr (Count) {
        if ((lid & 63) == 0) {
            atomicAdd(a.counts, n_eval);
            atomicAdd(a.counts + 1, n_acc);
        }
        return;
    }
    if (!Model && valid) {
        float *dst = nsplit > 1 ? a.part + (size_t)sg * a.ng * a.nzx : a.cig;
        for (int g = 0; g < a.ng; g++) dst[(size_t)g * a.nzx + i] = my[g * P];
    }
}

// Blocks a resident launch aims at when it splits the shots into groups:
// about 38 rounds of five 16 x 8 tiles per CU, 16 groups on the salt workload
// (measured there: 4 groups 128 ms, 8 121 ms, 16 115 ms, 24 115 ms, 32 116 ms;
// with the time-window and shot-block culling 8 94.8 ms, 11 93.4, 14 91.4,
// 16 90.1, 20 90.4, 24 90.2, 32 92.9); groups keep at least 8 shots, as more
// groups of fewer shots lose on the small workload.
constexpr int kSplitBlocks = 49152;

// cig = part[0] + part[1] + ... (group order), the shot groups' partial CIGs
__global__ __launch_bounds__(256) void k_sum_parts(const float *__restrict__ part,
                                                   float *__restrict__ cig, size_t n, int nsplit)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        float v = part[e];
        for (int s = 1; s < nsplit; s++) v += part[(size_t)s * n + e];
        cig[e] = v;
    }
}

template <int TZ, int TX, int R, int G, bool Probe, int Fmt, bool Cols = false, int Bin = 0>
__global__ __launch_bounds__(TZ *TX, 3) void k_migrate_ph(const MigResArgs a)
{
    extern __shared__ float acc[];
    migrate_tile<TZ, TX, R, G, Fmt, false, false, Cols, Bin>(
        a, xcd_tile(blockIdx.x, gridDim.x), acc, (int)blockIdx.y, (int)gridDim.y);
}

template <int TZ, int TX, int R, int G, bool Cols = false, int Bin = 0>
__global__ __launch_bounds__(TZ *TX, 3) void k_count_ph(const MigResArgs a)
{
    extern __shared__ float acc[];
    migrate_tile<TZ, TX, R, G, 0, false, true, Cols, Bin>(a, xcd_tile(blockIdx.x, gridDim.x),
                                                          acc, (int)blockIdx.y, (int)gridDim.y);
}

template <int TZ, int TX, int R, int G>
__global__ __launch_bounds__(TZ *TX, 3) void k_model_ph(const MigResArgs a)
{
    extern __shared__ float acc[];
    migrate_tile<TZ, TX, R, G, 0, true>(a, xcd_tile(blockIdx.x, gridDim.x), acc, (int)blockIdx.y,
                                        (int)gridDim.y);
}

// Several independent problems (e.g. the depth layers of a recursive
// workflow, each with its own grid, shots, receivers, mute and tuples) in one
// launch: blocks [tile0[p], tile0[p+1]) belong to problem p.  Problems are
// listed heaviest first, so the longest blocks start first.
template <int TZ, int TX, int R, int G, int Fmt>
__global__ __launch_bounds__(TZ *TX, 3) void k_migrate_multi(const MigMultiArgs m)
{
    extern __shared__ float acc[];
    int p = 0;
    while (p + 1 < m.nprob && (int)blockIdx.x >= m.tile0[p + 1]) p++;
    migrate_tile<TZ, TX, R, G, Fmt>(m.prob[p], (int)blockIdx.x - m.tile0[p], acc,
                                    (int)blockIdx.y, (int)gridDim.y);
}

static void tile_dims(int tile, int *tz, int *tx)
{
    switch (tile) {
    case 1: case 13: case 15: case 21: *tz = 8; *tx = 8; break;
    case 3: case 19: *tz = 16; *tx = 16; break;
    case 4: case 16: *tz = 32; *tx = 4; break;
    case 8: case 17: *tz = 64; *tx = 2; break;
    default: *tz = 16; *tx = 8; break;
    }
}

// LDS of a migrate block: ng bin rows plus the resident kernels' spare row
// (one row more than the stream kernel needs; the bound stays conservative)
int migrate_lds_bytes(int ng, int tile)
{
    int tz, tx;
    tile_dims(tile, &tz, &tx);
    return (ng + 1) * tz * tx * (int)sizeof(float);
}

// Shot groups of a resident launch over ntiles tiles: enough blocks for about
// kSplitBlocks in all, at least 8 shots per group, at most the groups the
// partial buffer holds; KMIG_SPLIT forces a count.
static int res_splits(const MigResArgs &a, int ntiles)
{
    int nsplit = 1;
    if (a.part_splits > 1) {
        nsplit = std::min(a.part_splits, std::max(1, (kSplitBlocks + ntiles - 1) / ntiles));
        nsplit = std::min(nsplit, std::max(1, a.nshots / 8));
        if (const char *v = std::getenv("KMIG_SPLIT"))
            nsplit = std::min(std::max(1, std::atoi(v)), std::min(a.part_splits, a.nshots));
    }
    return nsplit;
}

template <int TZ, int TX, bool Model = false>
static int launch_mig_t(const MigArgs &a, hipStream_t st)
{
    const int ntz = (a.nz + TZ - 1) / TZ, ntx = (a.nx + TX - 1) / TX;
    const size_t lds = (size_t)a.ng * TZ * TX * sizeof(float);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_migrate<TZ, TX, Model>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL((k_migrate<TZ, TX, Model>), dim3((unsigned)(ntz * ntx)), dim3(TZ * TX),
                       lds, st, a);
    return (int)hipGetLastError();
}

template <int TZ, int TX, int R, int G, bool Probe = false, int Fmt = 0, bool Cols = false,
          int Bin = 0>
static int launch_ph_t(const MigResArgs &a, hipStream_t st)
{
    const int ntz = (a.nz + TZ - 1) / TZ, ntx = (a.nx + TX - 1) / TX;
    const size_t lds = (size_t)(a.ng + 1) * TZ * TX * sizeof(float);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_migrate_ph<TZ, TX, R, G, Probe, Fmt, Cols, Bin>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    // Shot groups: all points x all shots in one block per tile leaves about
    // 2.4 rounds of long blocks on the salt workload (LDS allows five 16 x 8
    // tiles per CU), and the last, partial round idles most of the GPU; a
    // few thousand blocks per round keep the tail short.  KMIG_SPLIT forces a
    // count (tests).
    const int ntiles = ntz * ntx;
    const int nsplit = res_splits(a, ntiles);
    if (const int e = launch_cull_mask<Cols>(a, R, Bin == 1, st)) return e;
    hipLaunchKernelGGL((k_migrate_ph<TZ, TX, R, G, Probe, Fmt, Cols, Bin>), dim3((unsigned)ntiles,
                       (unsigned)nsplit), dim3(TZ * TX), lds, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nsplit == 1) return (int)e;
    const size_t n = (size_t)a.ng * a.nzx;
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)),
                       dim3(256), 0, st, a.part, a.cig, n, nsplit);
    return (int)hipGetLastError();
}

// Presets: tile shape (TZ x TX points), receivers per chunk R, group G.  The
// 16-bit tuple formats are built for the five presets the autotuner keeps
// for them (16x8 and 8x8 points, R = 8 or 16, G = 4, and 16x8, R = 8, G = 8).
template <bool Probe, int Fmt>
static int launch_res_preset(const MigResArgs &a, int tile, hipStream_t st)
{
    if (a.rbin) {
        // offset gathers and stacks: FP32 tuples, presets 9, 11 and 15
        if constexpr (Fmt != 0) {
            return (int)hipErrorInvalidValue;
        } else if (a.ctab) {
            switch (tile) {
            case 11: return launch_ph_t<16, 8, 8, 8, Probe, 0, true, 1>(a, st);
            case 15: return launch_ph_t<8, 8, 8, 4, Probe, 0, true, 1>(a, st);
            default: return launch_ph_t<16, 8, 8, 4, Probe, 0, true, 1>(a, st);
            }
        } else {
            switch (tile) {
            case 11: return launch_ph_t<16, 8, 8, 8, Probe, 0, false, 1>(a, st);
            case 15: return launch_ph_t<8, 8, 8, 4, Probe, 0, false, 1>(a, st);
            default: return launch_ph_t<16, 8, 8, 4, Probe, 0, false, 1>(a, st);
            }
        }
    }
    if (a.ctab) {
        // receiver columns (cmp=1): the five presets of the 16-bit formats
        switch (tile) {
        case 11: return launch_ph_t<16, 8, 8, 8, Probe, Fmt, true>(a, st);
        case 10: return launch_ph_t<16, 8, 16, 4, Probe, Fmt, true>(a, st);
        case 13: return launch_ph_t<8, 8, 16, 4, Probe, Fmt, true>(a, st);
        case 15: return launch_ph_t<8, 8, 8, 4, Probe, Fmt, true>(a, st);
        default: return launch_ph_t<16, 8, 8, 4, Probe, Fmt, true>(a, st);
        }
    }
    if constexpr (Fmt != 0) {
        switch (tile) {
        case 11: return launch_ph_t<16, 8, 8, 8, Probe, Fmt>(a, st);
        case 10: return launch_ph_t<16, 8, 16, 4, Probe, Fmt>(a, st);
        case 13: return launch_ph_t<8, 8, 16, 4, Probe, Fmt>(a, st);
        case 15: return launch_ph_t<8, 8, 8, 4, Probe, Fmt>(a, st);
        default: return launch_ph_t<16, 8, 8, 4, Probe, Fmt>(a, st);
        }
    }
    switch (tile) {
    case 10: return launch_ph_t<16, 8, 16, 4, Probe>(a, st);
    case 11: return launch_ph_t<16, 8, 8, 8, Probe>(a, st);
    case 13: return launch_ph_t<8, 8, 16, 4, Probe>(a, st);
    case 15: return launch_ph_t<8, 8, 8, 4, Probe>(a, st);
    case 16: return launch_ph_t<32, 4, 8, 4, Probe>(a, st);
    case 17: return launch_ph_t<64, 2, 8, 4, Probe>(a, st);
    case 18: return launch_ph_t<16, 8, 4, 4, Probe>(a, st);
    case 19: return launch_ph_t<16, 16, 8, 4, Probe>(a, st);
    case 20: return launch_ph_t<16, 8, 8, 2, Probe>(a, st);
    case 21: return launch_ph_t<8, 8, 4, 4, Probe>(a, st);
    default: return launch_ph_t<16, 8, 8, 4, Probe>(a, st);  // preset 9
    }
}

template <bool Probe>
static int launch_res_prec(const MigResArgs &a, int tile, int precision, hipStream_t st)
{
    switch (precision) {
    case KMIG_PREC_FP16: return launch_res_preset<Probe, 1>(a, tile, st);
    case KMIG_PREC_BF16: return launch_res_preset<Probe, 2>(a, tile, st);
    case KMIG_PREC_U16: return launch_res_preset<Probe, 3>(a, tile, st);
    default: return launch_res_preset<Probe, 0>(a, tile, st);
    }
}

int launch_migrate_resident(const MigResArgs &a, int tile, int precision, void *stream)
{
    return launch_res_prec<false>(a, tile, precision, (hipStream_t)stream);
}

template <int TZ, int TX, int R, int G>
static int launch_model_t(const MigResArgs &a, hipStream_t st)
{
    const int ntz = (a.nz + TZ - 1) / TZ, ntx = (a.nx + TX - 1) / TX;
    // the image column plus one trace window per wave
    const size_t lds = ((size_t)(a.ng + 1) * TZ * TX + (TZ * TX / 64) * MODEL_WIN) * sizeof(float);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_model_ph<TZ, TX, R, G>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    // shot groups as in the migration (no partial buffers: each group sprays
    // into the traces of its own shots)
    const int ntiles = ntz * ntx;
    int nsplit = std::min(KMIG_MAX_SPLIT, std::max(1, (kSplitBlocks + ntiles - 1) / ntiles));
    nsplit = std::min(nsplit, std::max(1, a.nshots / 8));
    if (const char *v = std::getenv("KMIG_SPLIT"))
        nsplit = std::min(std::max(1, std::atoi(v)), std::min(KMIG_MAX_SPLIT, a.nshots));
    if (const int e = launch_cull_mask<false>(a, R, 0, st)) return e;
    hipLaunchKernelGGL((k_model_ph<TZ, TX, R, G>), dim3((unsigned)ntiles, (unsigned)nsplit),
                       dim3(TZ * TX), lds, st, a);
    return (int)hipGetLastError();
}

int launch_model_resident(const MigResArgs &a, int tile, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    switch (tile) {
    case 11: return launch_model_t<16, 8, 8, 8>(a, st);
    case 15: return launch_model_t<8, 8, 8, 4>(a, st);
    default: return launch_model_t<16, 8, 8, 4>(a, st);  // preset 9
    }
}

int launch_migrate_resident_probe(const MigResArgs &a, int tile, int precision, void *stream)
{
    return launch_res_prec<true>(a, tile, precision, (hipStream_t)stream);
}

template <int TZ, int TX, int R, int G, bool Cols = false, int Bin = 0>
static int launch_count_t(const MigResArgs &a, hipStream_t st)
{
    const int ntz = (a.nz + TZ - 1) / TZ, ntx = (a.nx + TX - 1) / TX;
    const size_t lds = (size_t)(a.ng + 1) * TZ * TX * sizeof(float);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_count_ph<TZ, TX, R, G, Cols, Bin>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    const int ntiles = ntz * ntx;
    if (const int e = launch_cull_mask<Cols>(a, R, Bin == 1, st)) return e;
    hipLaunchKernelGGL((k_count_ph<TZ, TX, R, G, Cols, Bin>), dim3((unsigned)ntiles,
                       (unsigned)res_splits(a, ntiles)), dim3(TZ * TX), lds, st, a);
    return (int)hipGetLastError();
}

int launch_count_resident(const MigResArgs &a, int tile, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (a.rbin && a.ctab) {
        switch (tile) {
        case 11: return launch_count_t<16, 8, 8, 8, true, 1>(a, st);
        case 15: return launch_count_t<8, 8, 8, 4, true, 1>(a, st);
        default: return launch_count_t<16, 8, 8, 4, true, 1>(a, st);
        }
    }
    if (a.rbin) {
        switch (tile) {
        case 11: return launch_count_t<16, 8, 8, 8, false, 1>(a, st);
        case 15: return launch_count_t<8, 8, 8, 4, false, 1>(a, st);
        default: return launch_count_t<16, 8, 8, 4, false, 1>(a, st);
        }
    }
    if (a.ctab) {
        switch (tile) {
        case 11: return launch_count_t<16, 8, 8, 8, true>(a, st);
        case 10: return launch_count_t<16, 8, 16, 4, true>(a, st);
        case 13: return launch_count_t<8, 8, 16, 4, true>(a, st);
        case 15: return launch_count_t<8, 8, 8, 4, true>(a, st);
        default: return launch_count_t<16, 8, 8, 4, true>(a, st);
        }
    }
    switch (tile) {
    case 10: return launch_count_t<16, 8, 16, 4>(a, st);
    case 11: return launch_count_t<16, 8, 8, 8>(a, st);
    case 13: return launch_count_t<8, 8, 16, 4>(a, st);
    case 15: return launch_count_t<8, 8, 8, 4>(a, st);
    case 16: return launch_count_t<32, 4, 8, 4>(a, st);
    case 17: return launch_count_t<64, 2, 8, 4>(a, st);
    case 18: return launch_count_t<16, 8, 4, 4>(a, st);
    case 19: return launch_count_t<16, 16, 8, 4>(a, st);
    case 20: return launch_count_t<16, 8, 8, 2>(a, st);
    case 21: return launch_count_t<8, 8, 4, 4>(a, st);
    default: return launch_count_t<16, 8, 8, 4>(a, st);  // preset 9
    }
}

int migrate_tiles(int nz, int nx, int tile)
{
    int tz, tx;
    tile_dims(tile, &tz, &tx);
    return ((nz + tz - 1) / tz) * ((nx + tx - 1) / tx);
}

template <int TZ, int TX, int R, int G, int Fmt>
static int launch_multi_t(const MigMultiArgs &m, hipStream_t st)
{
    const size_t lds = (size_t)(m.prob[0].ng + 1) * TZ * TX * sizeof(float);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_migrate_multi<TZ, TX, R, G, Fmt>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    // shot groups as in launch_ph_t, one count for every problem (bounded by
    // the smallest problem's shots and partial buffers)
    const int ntiles = m.tile0[m.nprob];
    int splits = KMIG_MAX_SPLIT, min_shots = 1 << 30;
    for (int q = 0; q < m.nprob; q++) {
        splits = std::min(splits, m.prob[q].part_splits);
        min_shots = std::min(min_shots, m.prob[q].nshots);
    }
    int nsplit = 1;
    if (splits > 1) {
        nsplit = std::min(splits, std::max(1, (kSplitBlocks + ntiles - 1) / ntiles));
        nsplit = std::min(nsplit, std::max(1, min_shots / 8));
        if (const char *v = std::getenv("KMIG_SPLIT"))
            nsplit = std::min(std::max(1, std::atoi(v)), std::min(splits, min_shots));
    }
    for (int q = 0; q < m.nprob; q++)
        if (const int e = launch_cull_mask<false>(m.prob[q], R, 0, st)) return e;
    hipLaunchKernelGGL((k_migrate_multi<TZ, TX, R, G, Fmt>), dim3((unsigned)ntiles,
                       (unsigned)nsplit), dim3(TZ * TX), lds, st, m);
    hipError_t e = hipGetLastError();
    for (int q = 0; q < m.nprob && e == hipSuccess && nsplit > 1; q++) {
        const size_t n = (size_t)m.prob[q].ng * m.prob[q].nzx;
        hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)),
                           dim3(256), 0, st, m.prob[q].part, m.prob[q].cig, n, nsplit);
        e = hipGetLastError();
    }
    return (int)e;
}

// presets 9 (16x8 points, R = 8, G = 4), 10 (16x8, R = 16, G = 4), 11 (16x8,
// R = 8, G = 8), 15 (8x8, R = 8, G = 4) and 13 (8x8, R = 16, G = 4); the grid
// in m.tile0 must use the same tile shape
template <int Fmt>
static int launch_multi_fmt(const MigMultiArgs &m, int tile, hipStream_t st)
{
    switch (tile) {
    case 10: return launch_multi_t<16, 8, 16, 4, Fmt>(m, st);
    case 11: return launch_multi_t<16, 8, 8, 8, Fmt>(m, st);
    case 13: return launch_multi_t<8, 8, 16, 4, Fmt>(m, st);
    case 15: return launch_multi_t<8, 8, 8, 4, Fmt>(m, st);
    default: return launch_multi_t<16, 8, 8, 4, Fmt>(m, st);
    }
}

int launch_migrate_multi(const MigMultiArgs &m, int tile, int precision, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    switch (precision) {
    case KMIG_PREC_FP16: return launch_multi_fmt<1>(m, tile, st);
    case KMIG_PREC_BF16: return launch_multi_fmt<2>(m, tile, st);
    case KMIG_PREC_U16: return launch_multi_fmt<3>(m, tile, st);
    default: return launch_multi_fmt<0>(m, tile, st);
    }
}

int launch_migrate(const MigArgs &a, int tile, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    switch (tile) {
    case 1: return launch_mig_t<8, 8>(a, st);
    case 3: return launch_mig_t<16, 16>(a, st);
    case 4: return launch_mig_t<32, 4>(a, st);
    default: return launch_mig_t<16, 8>(a, st);
    }
}

int launch_model(const MigArgs &a, int tile, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    switch (tile) {
    case 1: return launch_mig_t<8, 8, true>(a, st);
    default: return launch_mig_t<16, 8, true>(a, st);
    }
}

}  // namespace kmig
