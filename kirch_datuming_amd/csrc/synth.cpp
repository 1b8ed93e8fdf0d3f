// libkmsynth: synthetic inputs for the adj=1 imaging path (velocity models
// and seismic data), all seeded and deterministic.  The traveltime tables
// come from the fast-marching table source (eikonal.cpp, libkmeik).
//
// Surrogate training models: Gaussian-random-field velocities (white noise
// smoothed by a separable Gaussian filter, rescaled, added to a random linear
// v(z) trend, clipped).
//
// Data: 15-20 Hz Ricker wavelets at analytic two-way times (flat reflectors
// in v(z) = v0 + g z, or point scatterers with table times), plus seeded
// Gaussian noise.  The RNG is splitmix64 (Steele, Lea, Flood 2014) with a
// Box-Muller transform, so every machine builds the same inputs.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

double ricker(double t, double f)
{
    const double a = (M_PI * f * t) * (M_PI * f * t);
    return (1.0 - 2.0 * a) * std::exp(-a);
}

void smooth5(std::vector<double> &v, int nz, int nx)
{
    // separable 5-point Gaussian [1 4 6 4 1]/16 with edge clamping, z then x
    static const double w[5] = {1.0 / 16, 4.0 / 16, 6.0 / 16, 4.0 / 16, 1.0 / 16};
    std::vector<double> t(v.size());
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            double s = 0;
            for (int k = -2; k <= 2; k++) {
                const int jz = std::min(std::max(iz + k, 0), nz - 1);
                s += w[k + 2] * v[(size_t)ix * nz + jz];
            }
            t[(size_t)ix * nz + iz] = s;
        }
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            double s = 0;
            for (int k = -2; k <= 2; k++) {
                const int jx = std::min(std::max(ix + k, 0), nx - 1);
                s += w[k + 2] * t[(size_t)jx * nz + iz];
            }
            v[(size_t)ix * nz + iz] = s;
        }
}

// separable Gaussian filter of standard deviation sig (cells), radius 3 sig,
// edge-normalised
void gauss2(std::vector<double> &v, int nz, int nx, double sig_z, double sig_x)
{
    auto kern = [](double sig) {
        const int r = std::max(1, (int)std::ceil(3.0 * sig));
        std::vector<double> k(2 * r + 1);
        for (int i = -r; i <= r; i++) k[i + r] = std::exp(-0.5 * i * i / (sig * sig));
        return k;
    };
    const std::vector<double> kz = kern(sig_z), kx = kern(sig_x);
    const int rz = (int)kz.size() / 2, rx = (int)kx.size() / 2;
    std::vector<double> t(v.size());
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            double s = 0.0, w = 0.0;
            for (int k = -rz; k <= rz; k++) {
                const int jz = iz + k;
                if (jz < 0 || jz >= nz) continue;
                s += kz[k + rz] * v[(size_t)ix * nz + jz];
                w += kz[k + rz];
            }
            t[(size_t)ix * nz + iz] = s / w;
        }
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            double s = 0.0, w = 0.0;
            for (int k = -rx; k <= rx; k++) {
                const int jx = ix + k;
                if (jx < 0 || jx >= nx) continue;
                s += kx[k + rx] * t[(size_t)jx * nz + iz];
                w += kx[k + rx];
            }
            v[(size_t)ix * nz + iz] = s / w;
        }
}

}  // namespace

extern "C" {

// Gaussian-random-field model: v = v0 + g z + amp * f(z, x), f a unit-variance
// field with Gaussian correlation exp(-r^2 / corr^2) (white noise through a
// Gaussian filter of standard deviation corr / 2), clipped to [vmin, vmax].
// v0 in [1.5, 2.0], g in [0.2, 0.6] 1/s and amp in [0.15, 0.45] km/s are
// drawn from the seed.
void kms_velocity_grf(float *v, int nz, int nx, float dz, float dx, float corr, float vmin,
                      float vmax, uint64_t seed)
{
    Rng rng(seed);
    const double v0 = 1.5 + 0.5 * rng.uniform(), g = 0.2 + 0.4 * rng.uniform();
    const double amp = 0.15 + 0.3 * rng.uniform();
    std::vector<double> f((size_t)nz * nx);
    for (double &q : f) q = rng.normal();
    gauss2(f, nz, nx, 0.5 * corr / dz, 0.5 * corr / dx);
    double m = 0.0, e = 0.0;
    for (double q : f) m += q;
    m /= (double)f.size();
    for (double q : f) e += (q - m) * (q - m);
    const double sd = std::sqrt(e / (double)f.size());
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            const size_t k = (size_t)ix * nz + iz;
            double vel = v0 + g * iz * (double)dz + amp * (f[k] - m) / (sd > 0.0 ? sd : 1.0);
            vel = std::min(std::max(vel, (double)vmin), (double)vmax);
            v[k] = (float)vel;
        }
}

// v(z) = v0 + g z on (nx, nz), z fastest.
void kms_velocity_vz(float *v, int nz, int nx, float dz, float v0, float g)
{
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) v[(size_t)ix * nz + iz] = v0 + g * (float)(iz * (double)dz);
}

// Five-layer salt-like model: 1.5 / 2.0 / 2.6 / 4.5 (salt body with a
// Gaussian-bump top) / 3.2 km/s, smoothed with a 5-point Gaussian.
void kms_velocity_salt(float *v, int nz, int nx, float dz, float dx)
{
    std::vector<double> m((size_t)nz * nx);
    const double zmax = (nz - 1) * (double)dz, xmax = (nx - 1) * (double)dx;
    for (int ix = 0; ix < nx; ix++) {
        const double x = ix * (double)dx, xr = x / std::max(xmax, 1e-9);
        const double z1 = 0.2 * zmax, z2 = 0.38 * zmax, zbase = 0.8 * zmax;
        const double ztop = 0.62 * zmax - 0.22 * zmax * std::exp(-std::pow((xr - 0.5) / 0.14, 2));
        const bool salt_x = xr > 0.18 && xr < 0.82;
        for (int iz = 0; iz < nz; iz++) {
            const double z = iz * (double)dz;
            double vel;
            if (z < z1) vel = 1.5;
            else if (z < z2) vel = 2.0;
            else if (salt_x && z >= ztop && z < zbase) vel = 4.5;
            else if (z < zbase) vel = 2.6;
            else vel = 3.2;
            m[(size_t)ix * nz + iz] = vel;
        }
    }
    smooth5(m, nz, nx);
    for (size_t k = 0; k < m.size(); k++) v[k] = (float)m[k];
}

// Analytic first-arrival time in v(z) = v0 + g z between (0, xs) and (z, x).
double kms_time_vz(double v0, double g, double xs, double x, double z)
{
    const double r2 = (x - xs) * (x - xs) + z * z;
    if (g == 0.0) return std::sqrt(r2) / v0;
    return std::acosh(1.0 + g * g * r2 / (2.0 * v0 * (v0 + g * z))) / g;
}

// Flat-reflector data in v(z): for each (shot, receiver) a Ricker of peak
// frequency f at the specular two-way time of each reflector (the midpoint
// is the reflection point in a laterally invariant medium), then Gaussian
// noise with sigma = rms / 10^(snr_db/20).  cmp selects offset (1) or
// absolute receiver (0) coordinates for axis 2.  data is (ns, nh, nt).
void kms_data_reflectors(float *data, int nt, float t0, float dt, int nh, float h0, float dh,
                         int ns, float s0, float ds, int cmp, float v0, float g, int nref,
                         const float *zref, const float *amp, float freq, float snr_db,
                         uint64_t seed)
{
    const size_t ntr = (size_t)ns * nh;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long tr = 0; tr < (long)ntr; tr++) {
        const int is = (int)(tr / nh), ih = (int)(tr % nh);
        const double s = s0 + is * (double)ds, h = h0 + ih * (double)dh;
        const double r = cmp ? s + h : h;
        float *out = data + (size_t)tr * nt;
        for (int it = 0; it < nt; it++) out[it] = 0.0f;
        for (int k = 0; k < nref; k++) {
            const double half = 0.5 * std::fabs(r - s);
            const double tt = 2.0 * kms_time_vz(v0, g, 0.0, half, zref[k]);
            const int c = (int)std::floor((tt - t0) / dt);
            const int w = (int)std::ceil(2.5 / (freq * dt));
            for (int it = std::max(0, c - w); it <= std::min(nt - 1, c + w); it++)
                out[it] += (float)(amp[k] * ricker(t0 + it * (double)dt - tt, freq));
        }
    }
    if (snr_db > 0.0f) {
        double e = 0.0;
        for (size_t k = 0; k < ntr * nt; k++) e += (double)data[k] * data[k];
        const double sigma = std::sqrt(e / (double)(ntr * nt)) / std::pow(10.0, snr_db / 20.0);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (long tr = 0; tr < (long)ntr; tr++) {
            Rng rng(seed ^ (0x5851F42D4C957F2Dull * (uint64_t)(tr + 1)));
            float *out = data + (size_t)tr * nt;
            for (int it = 0; it < nt; it++) out[it] += (float)(sigma * rng.normal());
        }
    }
}

// Point-scatterer (Born-style) data with table times plus a direct arrival:
// nsc scatterers at seeded random grid nodes with amplitudes in [-1, 1];
// source and receiver times are read from the position tables at the nearest
// position node.  data is (ns, nh, nt).
void kms_data_scatterers(float *data, int nt, float t0, float dt, int nh, float h0, float dh,
                         int ns, float s0, float ds, int cmp, const float *ST, int snpos, float so,
                         float sd, const float *RT, int rnpos, float ro, float rd, int nz, int nx,
                         int nsc, float freq, float vdirect, uint64_t seed)
{
    const size_t n = (size_t)nz * nx;
    Rng rng(seed);
    std::vector<size_t> pt(nsc);
    std::vector<double> a(nsc);
    for (int k = 0; k < nsc; k++) {
        const int iz = (int)(nz * (0.15 + 0.8 * rng.uniform()));
        const int ix = (int)(nx * (0.05 + 0.9 * rng.uniform()));
        pt[k] = (size_t)std::min(ix, nx - 1) * nz + std::min(iz, nz - 1);
        a[k] = 2.0 * rng.uniform() - 1.0;
    }
    const int w = (int)std::ceil(2.5 / (freq * dt));
    const long ntr = (long)ns * nh;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long tr = 0; tr < ntr; tr++) {
        const int is = (int)(tr / nh), ih = (int)(tr % nh);
        const double s = s0 + is * (double)ds, h = h0 + ih * (double)dh;
        const double r = cmp ? s + h : h;
        const int ks = std::min(std::max((int)std::lround((s - so) / sd), 0), snpos - 1);
        const int kr = std::min(std::max((int)std::lround((r - ro) / rd), 0), rnpos - 1);
        float *out = data + (size_t)tr * nt;
        for (int it = 0; it < nt; it++) out[it] = 0.0f;
        auto add = [&](double tt, double amp) {
            const int c = (int)std::floor((tt - t0) / dt);
            for (int it = std::max(0, c - w); it <= std::min(nt - 1, c + w); it++)
                out[it] += (float)(amp * ricker(t0 + it * (double)dt - tt, freq));
        };
        if (vdirect > 0.0f) add(std::fabs(r - s) / vdirect, 2.0);
        for (int k = 0; k < nsc; k++)
            add((double)ST[(size_t)ks * n + pt[k]] + (double)RT[(size_t)kr * n + pt[k]], a[k]);
    }
}

}  // extern "C"
