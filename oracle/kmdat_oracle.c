/*
 * CPU oracle for kirchdat2d_freq: frequency-domain Kirchhoff datuming of a
 * (t, h, s) data cube through a receiver-side and a source-side operator,
 * the direct (use_bf=0) path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ and the CPU baseline of bench.py,
 * never by the product library.
 *
 * Restates (file:line in the reference):
 *   src/kirchdat2d_frequency_cs.cpp:79-98    next_pow2, choose_nfft
 *   src/kirchdat2d_frequency_cs.cpp:100-184  build_filter_F, build_H_iw
 *   src/kirchdat2d_frequency_cs.cpp:186-279  pair accumulation, table slope,
 *                                            antialias, aperture taper and
 *                                            data-edge taper weights
 *   src/kirchdat2d_frequency_cs.cpp:281-313  read_traces_to_freq
 *   src/kirchdat2d_frequency_cs.cpp:470-491  write_traces_from_freq
 *   src/kirchdat2d_frequency_cs.cpp:493-672  receiver / source operators
 *                                            (cmp 0 and 1)
 *   src/kirchdat2d_frequency_cs.cpp:1177-1413 main's order of the stages
 *   src/greensol_auto.cpp:23-101             the datum traveltime table
 *   src/kirchdat2d_frequency_cs.cpp:674-1121 the butterfly operators
 *                                            (use_bf=1, float build only):
 *                                            per frequency the phase /
 *                                            amplitude matrix of each side
 *                                            (once for cmp=0, per shot or
 *                                            per receiver gather for cmp=1)
 *                                            through kbf_oracle.c's
 *                                            segmented butterfly
 * The FFTs (FFTW in the reference, which is not built here) are a textbook
 * radix-2 Cooley-Tukey transform in double, with FFTW's sign and scaling
 * conventions (forward e^{-i}, unnormalised backward).
 *
 * How each part is pinned (tests/test_datum_*.py):
 *   whole path         against the reference's time-domain datuming program
 *                      kirchdat2d_auto_cs (built from its own sources into
 *                      oracle/_ref), which applies the same filter, delays,
 *                      weights, apertures and cmp=1 regrouping sample by
 *                      sample in time; the two agree when the options only
 *                      the frequency program has are off (antialias=0,
 *                      data_taper=0, taper=0), to the FFT's rounding and the
 *                      one boundary term the time program drops;
 *   greensol           bitwise against greensol_auto's output;
 *   FFT                against numpy's FFT, and round trips;
 *   H                  against the DFT of the explicit time-domain filter;
 *   taper, data-edge taper, antialias weights
 *                      closed-form scalar functions (raised cosines: the
 *                      end values, midpoints and symmetry);
 *   f64 path           agrees with the f32 path to float rounding, linear in
 *                      the data;
 *   use_bf=1           the butterfly is pinned in kbf_oracle.c; the matrices
 *                      carry the direct path's weights, so with no legal
 *                      butterfly panel (bf_panel_levels=0) the two paths
 *                      agree to float rounding.
 */
#include "kmdat_oracle.h"
#include "kbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* In-place radix-2 complex FFT of length n (a power of two): forward
   X[k] = sum_j x[j] e^{-2 pi i jk/n}, inverse (inverse != 0) with e^{+},
   both unnormalised. */
void kmdo_fft_f64(int n, double *re, double *im, int inverse)
{
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    const double sgn = inverse ? 1.0 : -1.0;
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int k = 0; k < half; k++) {
            const double ang = sgn * 2.0 * M_PI * (double)k / (double)len;
            const double wr = cos(ang), wi = sin(ang);
            for (int i = k; i < n; i += len) {
                const int j = i + half;
                const double xr = re[j] * wr - im[j] * wi;
                const double xi = re[j] * wi + im[j] * wr;
                re[j] = re[i] - xr;
                im[j] = im[i] - xi;
                re[i] += xr;
                im[i] += xi;
            }
        }
    }
}

/* next_pow2 + choose_nfft, :79-98 */
int kmdo_nfft(int nt, int nsam, float dt, int nrg, int nsg, const float *rtable,
              const float *stable)
{
    float max_tau = 0.0f;
    for (size_t i = 0; i < (size_t)nrg * nrg; i++)
        if (rtable[i] > max_tau) max_tau = rtable[i];
    for (size_t i = 0; i < (size_t)nsg * nsg; i++)
        if (stable[i] > max_tau) max_tau = stable[i];
    const int pmax = (int)ceilf(max_tau / dt) + (nsam - 2);
    int p = 1;
    while (p < nt + pmax + 2) p <<= 1;
    return p;
}

int kmdo_greensol(const float *ttable, int nz, int nx, int ngx, float oz, float dz, float ox,
                  float dx, float ogx, float dgx, float datum, float *tgreen)
{
    const int idatum = (int)((datum - oz) / dz + 0.5);
    if (idatum < 0 || idatum >= nz) return -1;
    for (int i1 = 0; i1 < ngx; i1++) {
        for (int i2 = 0; i2 < ngx; i2++) {
            const float gx1 = ogx + i1 * dgx, gx2 = ogx + i2 * dgx;
            const int ix1 = (int)((gx1 - ox) / dx + 0.5);
            const int ix2 = (int)((gx2 - ox) / dx + 0.5);
            if (ix1 < 0 || ix1 >= nx || ix2 < 0 || ix2 >= nx) return -2;
            tgreen[(size_t)i1 * ngx + i2] = ttable[((size_t)i1 * nx + ix2) * nz + idatum];
        }
    }
    return 0;
}

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SFX)

#define R float
#define SFX _f32
#define R_SQRT sqrtf
#define R_CEIL ceilf
#define R_SIN sinf
#define R_COS cosf
#define R_FABS fabsf
#define R_FMIN fminf
#define R_FMAX fmaxf
#define KMDO_WITH_BF 1
#include "kmdat_oracle_impl.h"
#undef KMDO_WITH_BF
#undef R
#undef SFX
#undef R_SQRT
#undef R_CEIL
#undef R_SIN
#undef R_COS
#undef R_FABS
#undef R_FMIN
#undef R_FMAX

#define R double
#define SFX _f64
#define R_SQRT sqrt
#define R_CEIL ceil
#define R_SIN sin
#define R_COS cos
#define R_FABS fabs
#define R_FMIN fmin
#define R_FMAX fmax
#include "kmdat_oracle_impl.h"

void kmdo_filter_f32(float tau, float dt, int nsam, float *Fk) { filter_f32(tau, dt, nsam, Fk); }

void kmdo_H_f32(float tau, float dt, int nsam, int nfft, int iw, float *hr, float *hi)
{
    float *Fk = (float *)calloc(nsam > 1 ? nsam : 2, sizeof(float));
    Hw_f32(tau, dt, nsam, nfft, iw, Fk, hr, hi);
    free(Fk);
}

void kmdo_H_f64(double tau, double dt, int nsam, int nfft, int iw, double *hr, double *hi)
{
    double *Fk = (double *)calloc(nsam > 1 ? nsam : 2, sizeof(double));
    Hw_f64(tau, dt, nsam, nfft, iw, Fk, hr, hi);
    free(Fk);
}

float kmdo_taper_weight(int left, int center, int right, int tap, int stride)
{
    return taper_f32(left, center, right, tap, stride);
}

float kmdo_edge_taper(int i, int n, int tap) { return edge_f32(i, n, tap); }

float kmdo_antialias_weight(float antialias, int nfft, float dt, int iw, float dtau)
{
    return aa_f32(antialias, nfft, dt, iw, dtau);
}
