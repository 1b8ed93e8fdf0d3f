/*
 * CPU oracle for kirchmig2d_cig adj=1 (angle / offset CIG and stacked image).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
 * baseline of bench.py, never by the product library.
 *
 * Restates (file:line in the reference):
 *   src/kirchmig2d_cig.cpp:35-269   bin, taper, mute, FD, smoothing, angle,
 *                                   aperture helpers
 *   src/kirchmig2d_cig.cpp:294-322  gradient-volume linear interpolation
 *   src/kirchmig2d_cig.cpp:381-394  diff2
 *   src/kirchmig2d_cig.cpp:979-1339 shot / receiver / image-point loops (adj=1)
 *   src/kirchmig2d_cig.cpp:979-1351 the same loops with adj=0 (modelling)
 *   SERECKIRCH/src/kirmig.cpp:8-67  doubint, kirmig_pick (adj=1 and adj=0)
 *   SERECKIRCH/src/tinterpsr.cpp:37-188  linear / partial / Hermite
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   whole path (_f32)  bitwise equal to CIGs and QC files written by the
 *                      reference binary (built from its own sources into
 *                      oracle/_ref) on the committed golden cases: every
 *                      option the path has (cmp, mute/QC, trace aperture,
 *                      interp type l/p/h, nearest/linear bins, half angle,
 *                      gradient files, diff, tapers, rmin floor, offset CIG,
 *                      stacked image);
 *   doubint            impulse and constant traces in closed form;
 *   pick               zero on linear traces, impulse weights in closed form;
 *   Hermite            exact on cubic position dependence, identity at nodes;
 *   FD gradient        exact on linear fields, one-sided shrink at edges;
 *   smoothing          brute-force triangular window with edge renormalising;
 *   angle              closed-form opening angle of two known vectors;
 *   bins/tapers/mute   closed-form scalar functions;
 *   _f64 path          agrees with _f32 to float rounding on the golden cases
 *                      and is linear in the data;
 *   model (adj=0)      the reference's adj=0 run crashes, so no golden data:
 *                      pinned as the exact transpose of the pinned migration,
 *                      <L m, d> = <m, L^T d> to 1e-10 in FP64 on ten golden
 *                      geometries (angle, offset, stack, cmp 0/1, nearest,
 *                      half angle), and pick0 against pick sample by sample.
 */
#include "kmig_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define MIN(a, b) ((a) < (b) ? (a) : (b))

/* ---- float build: the reference's arithmetic ---- */
#define R float
#define SFX _f32
#define LIT(x) x##f
#define R_FLOOR floorf
#define R_COS cosf
#define R_ATAN atanf
#define R_ATAN2 atan2f
#define R_SQRT sqrtf
#define R_FABS fabsf
#include "kmig_oracle_impl.h"
#undef R
#undef SFX
#undef LIT
#undef R_FLOOR
#undef R_COS
#undef R_ATAN
#undef R_ATAN2
#undef R_SQRT
#undef R_FABS

/* ---- double build: FP64 oracle ---- */
#define R double
#define SFX _f64
#define LIT(x) x
#define R_FLOOR floor
#define R_COS cos
#define R_ATAN atan
#define R_ATAN2 atan2
#define R_SQRT sqrt
#define R_FABS fabs
#include "kmig_oracle_impl.h"

int kmo_migrate_f32(const kmo_params *p, const kmo_geom *g, const float *data,
                    const float *stab, const float *sder, const float *rtab, const float *rder,
                    const float *sgx, const float *sgz, const float *rgx, const float *rgz,
                    int shot_begin, int shot_end, float *out, float *qc_before, float *qc_after)
{
    return migrate_impl_f32(p, g, data, stab, sder, rtab, rder, sgx, sgz, rgx, rgz,
                            shot_begin, shot_end, out, qc_before, qc_after);
}

int kmo_migrate_f64(const kmo_params *p, const kmo_geom *g, const float *data,
                    const float *stab, const float *sder, const float *rtab, const float *rder,
                    const float *sgx, const float *sgz, const float *rgx, const float *rgz,
                    int shot_begin, int shot_end, double *out, float *qc_before, float *qc_after)
{
    return migrate_impl_f64(p, g, data, stab, sder, rtab, rder, sgx, sgz, rgx, rgz,
                            shot_begin, shot_end, out, qc_before, qc_after);
}

int kmo_model_f32(const kmo_params *p, const kmo_geom *g, const float *image,
                  const float *stab, const float *sder, const float *rtab, const float *rder,
                  const float *sgx, const float *sgz, const float *rgx, const float *rgz,
                  int shot_begin, int shot_end, float *data)
{
    return model_impl_f32(p, g, image, stab, sder, rtab, rder, sgx, sgz, rgx, rgz, shot_begin,
                          shot_end, data);
}

int kmo_model_f64(const kmo_params *p, const kmo_geom *g, const float *image,
                  const float *stab, const float *sder, const float *rtab, const float *rder,
                  const float *sgx, const float *sgz, const float *rgx, const float *rgz,
                  int shot_begin, int shot_end, double *data)
{
    return model_impl_f64(p, g, image, stab, sder, rtab, rder, sgx, sgz, rgx, rgz, shot_begin,
                          shot_end, data);
}

/* the transposed pick: sample sprayed into trace (added) */
void kmo_pick0_f64(int nt, double t0, double dt, double ti, double deltat, double sample,
                   double *trace)
{
    pick0_f64(nt, t0, dt, ti, deltat, sample, trace);
}

void kmo_doubint_f32(int nt, float *trace) { doubint__f32(nt, trace); }
void kmo_doubint_f64(int nt, double *trace) { doubint__f64(nt, trace); }

float kmo_pick_f32(int nt, float t0, float dt, float ti, float deltat, const float *trace)
{
    float s = 0.0f;
    pick_f32(nt, t0, dt, ti, deltat, &s, trace);
    return s;
}

double kmo_pick_f64(int nt, double t0, double dt, double ti, double deltat, const double *trace)
{
    double s = 0.0;
    pick_f64(nt, t0, dt, ti, deltat, &s, trace);
    return s;
}

void kmo_interp_side_f32(int type, int nz, int nx, float dx_grid, int npos, float o, float d,
                         float pos, const float *tab, const float *der, float *T, float *Tx,
                         int *inode)
{
    int k = interp_side_f32(type, nz, nx, dx_grid, npos, o, d, pos, tab, der, T, Tx);
    if (inode) *inode = k;
}

void kmo_grad_plane_f32(const float *tab, int nz, int nx, float dz, float dx, int radius,
                        float *gx, float *gz)
{
    grad_plane_f32(tab, nz, nx, dz, dx, radius, gx, gz);
}

void kmo_smooth_f32(float *f, int nz, int nx, int r)
{
    size_t n = (size_t)nz * nx;
    float *t1 = (float *)malloc(sizeof(float) * n), *t2 = (float *)malloc(sizeof(float) * n);
    smooth_f32(f, nz, nx, r, t1, t2);
    free(t1);
    free(t2);
}

float kmo_angle_f32(float sx, float sz, float rx, float rz, int half_angle)
{
    return angle_f32(sx, sz, rx, rz, half_angle);
}

int kmo_linear_bins_f32(float value, float g0, float dg, int ng, int *ig0, int *ig1, float *w0,
                        float *w1)
{
    return linear_bins_f32(value, g0, dg, ng, ig0, ig1, w0, w1);
}

void kmo_mute_f32(float *trace, int nt, float t0, float dt, float sx, float rx, float slope,
                  float intercept, float taper, int inverted)
{
    mute_f32(trace, nt, t0, dt, sx, rx, slope, intercept, taper, inverted);
}

float kmo_leg_weight_f32(float leg_deg, float aper, float band)
{
    return leg_weight_f32(leg_deg, aper, band);
}

float kmo_floor_weight_f32(float d, float rmin, float rt) { return floor_weight_f32(d, rmin, rt); }

float kmo_depth_mute_f32(float angle, int iz, int nz, float amin, float amax, float band)
{
    return depth_mute_f32(angle, iz, nz, amin, amax, band);
}

float kmo_angle_taper_f32(float angle, float start, float end)
{
    return angle_taper_f32(angle, start, end);
}
