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

#include <cmath>

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

__device__ __forceinline__ float cos_ramp_up(float u) { return 0.5f * (1.0f + cosf((float)KPI * u)); }

template <int TZ, int TX>
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
            if (a.mode == 2) {
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
                const float picked = aw * pick(tr, a.nt, a.t0, a.dt, a.inv_dt, ti, tx * a.aal);
                my[ig0 * P] += w0 * picked;
                if (ig1 != ig0 && w1 > 0.0f) my[ig1 * P] += w1 * picked;
            } else {
                const int ig = (int)floorf(u + 0.5f);
                if (ig < 0 || ig >= a.ng) continue;
                const float picked = pick(tr, a.nt, a.t0, a.dt, a.inv_dt, ti, tx * a.aal);
                my[ig * P] += aw * picked;
            }
        }
    }
    if (valid)
        for (int g = 0; g < a.ng; g++) a.cig[(size_t)g * a.nzx + i] = my[g * P];
}

static void tile_dims(int tile, int *tz, int *tx)
{
    switch (tile) {
    case 1: *tz = 8; *tx = 8; break;
    case 2: *tz = 16; *tx = 8; break;
    case 3: *tz = 16; *tx = 16; break;
    case 4: *tz = 32; *tx = 4; break;
    default: *tz = 16; *tx = 8; break;
    }
}

int migrate_lds_bytes(int ng, int tile)
{
    int tz, tx;
    tile_dims(tile, &tz, &tx);
    return ng * tz * tx * (int)sizeof(float);
}

template <int TZ, int TX>
static int launch_mig_t(const MigArgs &a, hipStream_t st)
{
    const int ntz = (a.nz + TZ - 1) / TZ, ntx = (a.nx + TX - 1) / TX;
    const size_t lds = (size_t)a.ng * TZ * TX * sizeof(float);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_migrate<TZ, TX>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL((k_migrate<TZ, TX>), dim3((unsigned)(ntz * ntx)), dim3(TZ * TX), lds, st,
                       a);
    return (int)hipGetLastError();
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

}  // namespace kmig
