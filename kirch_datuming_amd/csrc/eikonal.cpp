// Fast-marching traveltime tables with source derivatives (the eikods2d
// table source) on the host, one position per OpenMP thread.
//
//   march      first-arrival traveltime by fast marching: analytic times in a
//              box around the source, then the first/second-order upwind
//              quadratic update of each neighbour of the point taken off the
//              front (SERECKIRCH/src/eikods.cpp:65-112,
//              SERECKIRCH/src/neighbors.cpp:34-53, :74-88, :106-188, :271-302,
//              :418-484; SERECKIRCH/src/pqueue.cpp)
//   derivs     dT/dl and dT/ds by the upwind recursion in traveltime order
//              (SERECKIRCH/src/eikods.cpp:120-218)
//   upwind     branch-aware spatial gradients (src/eikods2d.cpp:30-66)
//
// The tables must be the reference's bit for bit, so every float/double
// operation follows the reference's types and order, and the front is a
// binary heap with the same sift rules: a point whose time drops while it
// waits on the front keeps its heap position (its key is read afresh at each
// comparison), and marching stops once no point is left unreached, leaving
// the last front points at their tentative times.  The derivative pass
// visits points in ascending time; each point reads only neighbours of
// strictly smaller time, so any such order gives the same result, and a
// radix sort on the time bits replaces the comparison sort.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "kmig_eik.h"

namespace {

enum : uint8_t { kIn = 0, kFront = 1, kOut = 2 };

// the upwind neighbour value of one axis in the quadratic update
struct Leg {
    double stencil, value, delta;
};

// binary min-heap of grid indices keyed by t[index], 1-based
struct Front {
    std::vector<int> h;
    int n = 0;
    const float *t = nullptr;

    void push(int i)
    {
        if ((size_t)n + 1 >= h.size()) h.resize(2 * h.size() + 2);  // never shrinks
        int pos = ++n;
        const float v = t[i];
        for (int q = pos >> 1; q > 0; q >>= 1) {
            if (v > t[h[q]]) break;
            h[pos] = h[q];
            pos = q;
        }
        h[pos] = i;
    }
    int pop()
    {
        if (n == 0) return -1;
        const int top = h[1];
        const int last = h[n--];
        const float v = t[last];
        int pos = 1;
        for (int c = 2; c <= n; c <<= 1) {
            if (c < n && t[h[c]] > t[h[c + 1]]) c++;
            if (v <= t[h[c]]) break;
            h[pos] = h[c];
            pos = c;
        }
        h[pos] = last;
        return top;
    }
};

struct Marcher {
    int n[3], s[3], order;
    float rdx[3];
    double v1 = 0.0;
    const float *vv = nullptr;   // slowness squared
    float *t = nullptr;
    std::vector<uint8_t> flag;
    Front front;

    Marcher(int n1, int n2, int n3, float d1, float d2, float d3, int ord)
    {
        n[0] = n1; n[1] = n2; n[2] = n3;
        s[0] = 1; s[1] = n1; s[2] = n1 * n2;
        order = ord;
        const float d[3] = {d1, d2, d3};
        for (int j = 0; j < 3; j++) rdx[j] = (float)(1. / (d[j] * d[j]));
        int band = 0;
        if (n1 > 1) band += 2 * n2 * n3;
        if (n2 > 1) band += 2 * n1 * n3;
        if (n3 > 1) band += 2 * n1 * n2;
        front.h.assign((size_t)10 * band + 1, 0);
        flag.assign((size_t)n1 * n2 * n3, kOut);
    }

    static void clamp(int *i, const int *nn)
    {
        for (int j = 0; j < 3; j++) i[j] = i[j] < 0 ? 0 : (i[j] >= nn[j] ? nn[j] - 1 : i[j]);
    }

    // analytic times in the box of b samples around the source, whose
    // boundary starts the front; returns the points left unreached
    int near_source(const float *xs, const int *b, const float *d, const bool *plane)
    {
        const int n123 = n[0] * n[1] * n[2];
        for (int i = 0; i < n123; i++) {
            flag[i] = kOut;
            t[i] = FLT_MAX;
        }
        int lo[3], hi[3];
        for (int j = 0; j < 3; j++) {
            const int c = (int)((double)(xs[j] / d[j]) + 0.5);
            lo[j] = c - b[j];
            hi[j] = c + b[j];
        }
        clamp(lo, n);
        clamp(hi, n);
        const int ic = (lo[0] + hi[0]) / 2 + n[0] * ((lo[1] + hi[1]) / 2 + n[1] * (lo[2] + hi[2]) / 2);
        v1 = vv[ic];
        int left = n123;
        for (int i3 = lo[2]; i3 <= hi[2]; i3++)
            for (int i2 = lo[1]; i2 <= hi[1]; i2++)
                for (int i1 = lo[0]; i1 <= hi[0]; i1++) {
                    left--;
                    const int i = i1 + n[0] * (i2 + n[1] * i3);
                    double del[3];
                    del[0] = xs[0] - i1 * d[0];
                    del[1] = xs[1] - i2 * d[1];
                    del[2] = xs[2] - i3 * d[2];
                    double r2 = 0.0;
                    for (int j = 0; j < 3; j++)
                        if (!plane[2 - j]) r2 += del[j] * del[j];
                    t[i] = sqrtf((float)(v1 * r2));
                    flag[i] = kIn;
                    if ((n[0] > 1 && (i1 == lo[0] || i1 == hi[0])) ||
                        (n[1] > 1 && (i2 == lo[1] || i2 == hi[1])) ||
                        (n[2] > 1 && (i3 == lo[2] || i3 == hi[2])))
                        front.push(i);
                }
        return left;
    }

    // one axis of the quadratic update: the smaller neighbour along the axis
    // (coordinate c of n, stride st), second-order stencil where the point
    // two steps out is known and not later
    void leg(Leg &l, int i, int c, int nc, int st, float rd) const
    {
        const float a = c > 0 ? t[i - st] : FLT_MAX;
        const float b = c < nc - 1 ? t[i + st] : FLT_MAX;
        l.delta = rd;
        l.stencil = l.value = a < b ? a : b;
        if (order > 1) {
            float tt;
            if (a < b && c - 2 >= 0) {
                const int k = i - 2 * st;
                if (flag[k] != kOut && a >= (tt = t[k])) {
                    l.delta *= 2.25;
                    l.stencil = (4.0 * l.value - tt) / 3.0;
                }
            }
            if (a > b && c + 2 <= nc - 1) {
                const int k = i + 2 * st;
                if (flag[k] != kOut && b >= (tt = t[k])) {
                    l.delta *= 2.25;
                    l.stencil = (4.0 * l.value - tt) / 3.0;
                }
            }
        }
    }

    bool solve(int m, float *res, Leg *const *v) const
    {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int j = 0; j < m; j++) {
            a += v[j]->delta;
            b += v[j]->stencil * v[j]->delta;
            c += v[j]->stencil * v[j]->stencil * v[j]->delta;
        }
        b /= a;
        const double disc = b * b + (v1 - c) / a;
        if (disc < 0.0) return false;
        const double r = b + sqrt(disc);
        if (r <= v[m - 1]->value) return false;
        *res = (float)r;
        return true;
    }

    // new time at grid point i = (iz, ix) from its upwind neighbours; the
    // models are 2-D (n[2] = 1), so the third leg is always empty
    float quad(int i, int iz, int ix)
    {
        Leg x[3];
        leg(x[0], i, iz, n[0], 1, rdx[0]);
        leg(x[1], i, ix, n[1], s[1], rdx[1]);
        x[2].stencil = x[2].value = FLT_MAX;
        x[2].delta = rdx[2];
        // the legs in ascending value, ties as the reference's network
        Leg *v[3];
        if (x[0].value <= x[1].value) {
            if (x[1].value <= x[2].value) {
                v[0] = x; v[1] = x + 1; v[2] = x + 2;
            } else if (x[2].value <= x[0].value) {
                v[0] = x + 2; v[1] = x; v[2] = x + 1;
            } else {
                v[0] = x; v[1] = x + 2; v[2] = x + 1;
            }
        } else {
            if (x[0].value <= x[2].value) {
                v[0] = x + 1; v[1] = x; v[2] = x + 2;
            } else if (x[2].value <= x[1].value) {
                v[0] = x + 2; v[1] = x + 1; v[2] = x;
            } else {
                v[0] = x + 1; v[1] = x + 2; v[2] = x;
            }
        }
        v1 = vv[i];
        float res;
        if (v[2]->value < FLT_MAX) {
            if (solve(3, &res, v) || solve(2, &res, v) || solve(1, &res, v)) return res;
        } else if (v[1]->value < FLT_MAX) {
            if (solve(2, &res, v) || solve(1, &res, v)) return res;
        } else if (v[0]->value < FLT_MAX) {
            if (solve(1, &res, v)) return res;
        }
        return FLT_MAX;
    }

    int relax(float value, int i)
    {
        if (value < t[i]) {
            t[i] = value;
            if (flag[i] == kOut) {
                flag[i] = kFront;
                front.push(i);
                return 1;
            }
        }
        return 0;
    }

    // the four neighbours in the reference's order (+z, -z, +x, -x)
    int update_neighbours(int i)
    {
        const int ix = i / n[0], iz = i - ix * n[0];
        int added = 0;
        if (iz + 1 <= n[0] - 1 && flag[i + 1] != kIn) added += relax(quad(i + 1, iz + 1, ix), i + 1);
        if (iz - 1 >= 0 && flag[i - 1] != kIn) added += relax(quad(i - 1, iz - 1, ix), i - 1);
        const int sx = s[1];
        if (ix + 1 <= n[1] - 1 && flag[i + sx] != kIn)
            added += relax(quad(i + sx, iz, ix + 1), i + sx);
        if (ix - 1 >= 0 && flag[i - sx] != kIn) added += relax(quad(i - sx, iz, ix - 1), i - sx);
        return added;
    }

    // returns false when the front ran empty before every point was reached
    // (the reference warns and goes on to the derivatives, as this does)
    bool march(float *time, const float *slow2, const float *xs, const int *b, const float *d,
               const bool *plane)
    {
        t = time;
        vv = slow2;
        front.t = time;
        front.n = 0;
        int left = near_source(xs, b, d, plane);
        while (left > 0) {
            const int i = front.pop();
            if (i < 0) return false;
            flag[i] = kIn;
            left -= update_neighbours(i);
        }
        return true;
    }
};

// ascending order of the times (radix sort on the float bits, stable, so ties
// stay in index order): ord[k] is the k-th point.  buf is reused.
void time_order(const float *t, int n, std::vector<uint64_t> &buf, std::vector<int> &ord)
{
    buf.resize(2 * (size_t)n);
    uint64_t *a = buf.data(), *b = buf.data() + n;
    for (int i = 0; i < n; i++) {
        uint32_t u;
        std::memcpy(&u, &t[i], 4);
        // order-preserving map of the float bits (negative times never occur)
        u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
        a[i] = ((uint64_t)u << 32) | (uint32_t)i;
    }
    for (int shift = 32; shift < 64; shift += 11) {
        int cnt[2049] = {0};
        for (int i = 0; i < n; i++) cnt[((a[i] >> shift) & 2047u) + 1]++;
        for (int k = 0; k < 2048; k++) cnt[k + 1] += cnt[k];
        for (int i = 0; i < n; i++) b[cnt[(a[i] >> shift) & 2047u]++] = a[i];
        std::swap(a, b);
    }
    ord.resize(n);
    for (int i = 0; i < n; i++) ord[i] = (int)(a[i] & 0xFFFFFFFFu);
}

struct Derivs {
    int nn[3], ss[3], ndim;
    const float *t0, *v0;
    float d[3];

    void first(int l, float *dl1, float *ds1, int i0) const
    {
        int ii[3];
        ii[1] = i0 / nn[0];
        ii[0] = i0 - ii[1] * nn[0];
        ii[2] = 0;  // 2-D
        double rhs = 0.0, den = 0.0, dw = 0.0, dt = 0.0;
        for (int j = 0; j < ndim; j++) {
            const int a = i0 - ss[j], b = i0 + ss[j];
            const int i1 = (ii[j] == 0 || (ii[j] != nn[j] - 1 && t0[a] > t0[b])) ? b : a;
            if (t0[i1] < t0[i0]) {
                rhs += (t0[i0] - t0[i1]) * dl1[i1] / (d[j] * d[j]);
                den += (t0[i0] - t0[i1]) / (d[j] * d[j]);
                if (l == j) {
                    if (i1 == b) {
                        const int i2 = i1 + ss[j];
                        if (ii[j] < nn[j] - 2 && t0[i1] > t0[i2])
                            dt = (-t0[i2] + 4. * t0[i1] - 3. * t0[i0]) / (2. * d[j]);
                        else
                            dt = (t0[i1] - t0[i0]) / d[j];
                    } else {
                        const int i2 = i1 - ss[j];
                        if (ii[j] > 1 && t0[i1] > t0[i2])
                            dt = (3. * t0[i0] - 4. * t0[i1] + t0[i2]) / (2. * d[j]);
                        else
                            dt = (t0[i0] - t0[i1]) / d[j];
                    }
                }
            }
        }
        if (den == 0.0) {
            dl1[i0] = 0.0f;
            ds1[i0] = 0.0f;
            return;
        }
        const int sl = ss[l];
        if (0 < ii[l] && ii[l] < nn[l] - 1)
            dw = (v0[i0 + sl] - v0[i0 - sl]) / (2. * d[l]);
        else if (ii[l] == 0)
            dw = (-v0[i0 + 2 * sl] + 4. * v0[i0 + sl] - 3. * v0[i0]) / (2. * d[l]);
        else
            dw = (3. * v0[i0] - 4. * v0[i0 - sl] + v0[i0 - 2 * sl]) / (2. * d[l]);
        dl1[i0] = (float)((0.5 * dw + rhs) / den);
        ds1[i0] = (float)(dl1[i0] - dt);
    }
};

// branch-aware one-sided gradient of t along one axis (z: 0, x: 1)
float upwind(const float *t, int i0, int axis, const int *nn, const int *ss, const float *dd)
{
    int ii[3];
    ii[0] = i0 % nn[0];
    ii[1] = (i0 / ss[1]) % nn[1];
    ii[2] = i0 / ss[2];
    const int a = i0 - ss[axis], b = i0 + ss[axis];
    int i1;
    if (ii[axis] == 0)
        i1 = b;
    else if (ii[axis] != nn[axis] - 1 && t[b] < t[a])
        i1 = b;
    else
        i1 = a;
    if (!(t[i1] < t[i0])) return 0.0f;
    if (i1 == b) {
        const int i2 = i1 + ss[axis];
        if (ii[axis] < nn[axis] - 2 && t[i2] < t[i1])
            return (-t[i2] + 4.0f * t[i1] - 3.0f * t[i0]) / (2.0f * dd[axis]);
        return (t[i1] - t[i0]) / dd[axis];
    }
    const int i2 = i1 - ss[axis];
    if (ii[axis] > 1 && t[i2] < t[i1])
        return (3.0f * t[i0] - 4.0f * t[i1] + t[i2]) / (2.0f * dd[axis]);
    return (t[i0] - t[i1]) / dd[axis];
}

}  // namespace

extern "C" {

void kmig_eik_config_default(kmig_eik_config *c)
{
    std::memset(c, 0, sizeof(*c));
    c->order = 2;
    c->b1 = 1;
    c->b2 = 1;
    c->l = 1;
    c->nthreads = 0;
}

int kmig_eikods2d(const kmig_eik_config *cfg, const float *vel, int n1, int n2, float o1,
                  float o2, float d1, float d2, const float *shots, int nshot, float *time,
                  float *tdl1, float *tds1, float *gradz, float *gradx)
{
    if (!cfg || !vel || !shots || !time || !tds1) return KMIG_EIK_E_PARAM;
    if (n1 <= 0 || n2 <= 0 || nshot <= 0) return KMIG_EIK_E_PARAM;
    if (!(d1 > 0.0f) || !(d2 > 0.0f)) return KMIG_EIK_E_PARAM;
    if (cfg->order < 1 || cfg->order > 2 || cfg->l < 0 || cfg->l > 1) return KMIG_EIK_E_PARAM;
    const long long n123ll = (long long)n1 * n2;
    if (n123ll > 0x7fffffffLL) return KMIG_EIK_E_PARAM;
    const int n123 = (int)n123ll;
    const int n3 = 1;
    const float d3 = d2, o3 = 0.0f;  // a 2-D model has one sample along axis 3
    const int b1 = cfg->b1 < 1 ? 1 : cfg->b1, b2 = cfg->b2 < 1 ? 1 : cfg->b2, b3 = 1;
    // slowness squared, as the reference forms it from velocity
    std::vector<float> slow2(n123);
    for (int i = 0; i < n123; i++) {
        const float v = vel[i];
        slow2[i] = (float)(1. / (v * v));
    }
    int status = KMIG_EIK_OK;
#ifdef _OPENMP
    const int nth = cfg->nthreads > 0 ? cfg->nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nth)
#endif
    {
        Marcher m(n1, n2, n3, d1, d2, d3, cfg->order);
        std::vector<int> ord;
        std::vector<uint64_t> sortbuf;
        std::vector<float> dl1_scratch(tdl1 ? 0 : n123);  // the recursion needs dT/dl
        Derivs dv;
        dv.nn[0] = n1; dv.nn[1] = n2; dv.nn[2] = n3;
        dv.ss[0] = 1; dv.ss[1] = n1; dv.ss[2] = n1 * n2;
        dv.ndim = n3 == 1 ? 2 : 3;
        dv.v0 = slow2.data();
        dv.d[0] = d1; dv.d[1] = d2; dv.d[2] = d3;
        const bool plane[3] = {false, false, false};
        const int b[3] = {b1, b2, b3};
        const float d[3] = {d1, d2, d3};
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
        for (int is = 0; is < nshot; is++) {
            const size_t off = (size_t)is * n123;
            float *t = time + off;
            // shots: (z, lateral, cross-line) per position, as a shot file
            const float xs[3] = {shots[3 * is] - o1, shots[3 * is + 1] - o2,
                                 shots[3 * is + 2] - o3};
            if (!m.march(t, slow2.data(), xs, b, d, plane)) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
                status = KMIG_EIK_W_FRONT;
            }
            dv.t0 = t;
            time_order(t, n123, sortbuf, ord);
            float *dl1 = tdl1 ? tdl1 + off : dl1_scratch.data(), *ds1 = tds1 + off;
            for (int k = 0; k < n123; k++) dv.first(cfg->l, dl1, ds1, ord[k]);
            if (gradz || gradx) {
                const int nn[3] = {n1, n2, n3};
                const int ss[3] = {1, n1, n1 * n2};
                const float dd[3] = {d1, d2, d3};
                for (int q = 0; q < n123; q++) {
                    if (gradz) gradz[off + q] = upwind(t, q, 0, nn, ss, dd);
                    if (gradx) gradx[off + q] = upwind(t, q, 1, nn, ss, dd);
                }  // (a small share of the time: one pass of two stencils)
            }
        }
    }
    return status;
}

}  // extern "C"
