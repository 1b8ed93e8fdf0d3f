/*
 * Body of the CPU oracle, included twice by kmig_oracle.c:
 *   R = float  (SFX = _f32): the reference's own mixed float/double arithmetic;
 *   R = double (SFX = _f64): everything promoted to double.
 * Every function names the reference lines it restates.  TEST INFRASTRUCTURE.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SFX)

/* value_to_nearest_bin, src/kirchmig2d_cig.cpp:35-41 */
static int F(nearest_bin)(R value, R g0, R dg, int ng)
{
    if (ng <= 0 || dg <= 0 || !isfinite(value)) return -1;
    int ig = (int)R_FLOOR((value - g0) / dg + LIT(0.5));
    if (ig < 0 || ig >= ng) return -1;
    return ig;
}

/* value_to_linear_bins, src/kirchmig2d_cig.cpp:54-79 */
static int F(linear_bins)(R value, R g0, R dg, int ng, int *ig0, int *ig1, R *w0, R *w1)
{
    if (ng <= 0 || dg <= 0 || !isfinite(value)) return 0;
    const R u = (value - g0) / dg;
    if (u < -LIT(1.0e-5) || u > (R)(ng - 1) + LIT(1.0e-5)) return 0;
    if (u <= 0) {
        *ig0 = 0; *ig1 = 0; *w0 = 1; *w1 = 0;
        return 1;
    }
    if (u >= (R)(ng - 1)) {
        *ig0 = ng - 1; *ig1 = ng - 1; *w0 = 1; *w1 = 0;
        return 1;
    }
    *ig0 = (int)R_FLOOR(u);
    *ig1 = *ig0 + 1;
    *w1 = u - (R)(*ig0);
    *w0 = LIT(1.0) - *w1;
    return 1;
}

/* angle_taper_weight, src/kirchmig2d_cig.cpp:88-96 */
static R F(angle_taper)(R angle, R start, R end)
{
    if (!isfinite(angle)) return 0;
    if (start < 0 || end <= start) return 1;
    if (angle <= start) return 1;
    if (angle >= end) return 0;
    const R x = (angle - start) / (end - start);
    return LIT(0.5) * (LIT(1.0) + R_COS((R)M_PI * x));
}

/* apply_direct_wave_mute, src/kirchmig2d_cig.cpp:110-129 */
static void F(mute)(R *trace, int nt, R t0, R dt, R source_x, R receiver_x,
                    R slope, R intercept, R taper_length, int inverted)
{
    const R offset_time = slope * R_FABS(receiver_x - source_x);
    const R boundary = inverted ? intercept - offset_time : intercept + offset_time;
    for (int it = 0; it < nt; it++) {
        const R time = t0 + (R)it * dt;
        R weight = 0;
        if (time >= boundary + taper_length) {
            weight = 1;
        } else if (time > boundary) {
            const R u = (time - boundary) / taper_length;
            weight = LIT(0.5) * (LIT(1.0) - R_COS((R)M_PI * u));
        }
        trace[it] *= weight;
    }
}

/* finite_diff_x_radius / finite_diff_z_radius, src/kirchmig2d_cig.cpp:130-152,
   compute_grad_plane :424-435 (table index i = ix*nz + iz, :27-30) */
static void F(grad_plane)(const R *tab, int nz, int nx, R dz, R dx, int radius, R *gx, R *gz)
{
    if (radius < 1) radius = 1;
    for (int ix = 0; ix < nx; ix++) {
        for (int iz = 0; iz < nz; iz++) {
            const size_t i = (size_t)ix * nz + iz;
            R vx = 0, vz = 0;
            if (!(nx <= 1 || dx == 0)) {
                const int il = MAX(0, ix - radius);
                const int ir = MIN(nx - 1, ix + radius);
                if (ir != il)
                    vx = (tab[(size_t)ir * nz + iz] - tab[(size_t)il * nz + iz]) /
                         ((R)(ir - il) * dx);
            }
            if (!(nz <= 1 || dz == 0)) {
                const int it = MAX(0, iz - radius);
                const int ib = MIN(nz - 1, iz + radius);
                if (ib != it)
                    vz = (tab[(size_t)ix * nz + ib] - tab[(size_t)ix * nz + it]) /
                         ((R)(ib - it) * dz);
            }
            gx[i] = vx;
            gz[i] = vz;
        }
    }
}

/* boxsmooth_axis_z, src/kirchmig2d_cig.cpp:438-451 */
static void F(box_z)(const R *in, R *out, int nz, int nx, int r)
{
    for (int ix = 0; ix < nx; ix++) {
        const R *c = in + (size_t)ix * nz;
        R *o = out + (size_t)ix * nz;
        for (int iz = 0; iz < nz; iz++) {
            const int a = (iz - r < 0) ? 0 : iz - r;
            const int b = (iz + r > nz - 1) ? nz - 1 : iz + r;
            R s = 0;
            for (int k = a; k <= b; k++) s += c[k];
            o[iz] = s / (R)(b - a + 1);
        }
    }
}

/* boxsmooth_axis_x, src/kirchmig2d_cig.cpp:454-465 */
static void F(box_x)(const R *in, R *out, int nz, int nx, int r)
{
    for (int iz = 0; iz < nz; iz++) {
        for (int ix = 0; ix < nx; ix++) {
            const int a = (ix - r < 0) ? 0 : ix - r;
            const int b = (ix + r > nx - 1) ? nx - 1 : ix + r;
            R s = 0;
            for (int k = a; k <= b; k++) s += in[(size_t)k * nz + iz];
            out[(size_t)ix * nz + iz] = s / (R)(b - a + 1);
        }
    }
}

/* smooth_one_pass :468-482 and smooth_grad_component :486-492 (rz = rx = r) */
static void F(smooth)(R *f, int nz, int nx, int r, R *tmp1, R *tmp2)
{
    const size_t n = (size_t)nz * nx;
    if (r <= 0) return;
    /* first box: f -> tmp1 (through tmp2) */
    F(box_z)(f, tmp2, nz, nx, r);
    F(box_x)(tmp2, tmp1, nz, nx, r);
    /* second box: tmp1 -> f (through tmp2) */
    F(box_z)(tmp1, tmp2, nz, nx, r);
    F(box_x)(tmp2, f, nz, nx, r);
    (void)n;
}

/* scattering_angle_deg_from_grad, src/kirchmig2d_cig.cpp:171-187 */
static R F(angle)(R sx, R sz, R rx, R rz, int half_angle)
{
    R ns = R_SQRT(sx * sx + sz * sz);
    R nr = R_SQRT(rx * rx + rz * rz);
    if (ns <= LIT(1.0e-12) || nr <= LIT(1.0e-12)) return (R)NAN;
    const R dot = sx * rx + sz * rz;
    const R cross = sx * rz - sz * rx;
    const R opening = R_ATAN2(R_FABS(cross), dot);
    R deg = opening * LIT(180.0) / (R)M_PI;
    return half_angle ? LIT(0.5) * deg : deg;
}

/* angle_depth_mute_weight, src/kirchmig2d_cig.cpp:200-211 */
static R F(depth_mute)(R angle, int iz, int nz, R amin, R amax, R band)
{
    if (!isfinite(angle)) return 0;
    const R zf = (nz > 1) ? (R)iz / (R)(nz - 1) : 0;
    const R cutoff = amin + zf * (amax - amin);
    if (band <= 0) return (angle <= cutoff) ? 1 : 0;
    if (angle <= cutoff) return 1;
    if (angle >= cutoff + band) return 0;
    const R u = (angle - cutoff) / band;
    return LIT(0.5) * (LIT(1.0) + R_COS((R)M_PI * u));
}

/* aperture_leg_weight, src/kirchmig2d_cig.cpp:226-235 */
static R F(leg_weight)(R leg_angle, R aper, R band)
{
    if (!isfinite(leg_angle)) return 0;
    if (band <= 0) return (leg_angle <= aper) ? 1 : 0;
    const R inner = aper - band;
    if (leg_angle <= inner) return 1;
    if (leg_angle >= aper) return 0;
    const R u = (leg_angle - inner) / band;
    return LIT(0.5) * (LIT(1.0) + R_COS((R)M_PI * u));
}

/* aperture_floor_weight, src/kirchmig2d_cig.cpp:255-269 */
static R F(floor_weight)(R d, R rmin, R rt)
{
    if (rmin <= 0) return 0;
    if (rt <= 0) return (d <= rmin) ? 1 : 0;
    const R inner = rmin - rt;
    if (inner <= 0) {
        if (d >= rmin) return 0;
        const R u = d / rmin;
        return LIT(0.5) * (LIT(1.0) + R_COS((R)M_PI * u));
    }
    if (d <= inner) return 1;
    if (d >= rmin) return 0;
    const R u = (d - inner) / rt;
    return LIT(0.5) * (LIT(1.0) + R_COS((R)M_PI * u));
}

/* diff2, src/kirchmig2d_cig.cpp:381-394 */
static void F(diff2)(R *trace, int n, R d, R *tmp)
{
    for (int i = 0; i < n; i++) tmp[i] = trace[i];
    for (int i = 1; i < n - 1; i++) trace[i] = (tmp[i + 1] - tmp[i - 1]) / (LIT(2.0) * d);
    trace[0] = (tmp[1] - tmp[0]) / d;
    trace[n - 1] = (tmp[n - 1] - tmp[n - 2]) / d;
}

/* doubint, SERECKIRCH/src/kirmig.cpp:8-24 */
static void F(doubint_)(int nt, R *trace)
{
    R tt = trace[0];
    for (int it = 1; it < nt; it++) {
        tt += trace[it];
        trace[it] = tt;
    }
    tt = trace[nt - 1];
    for (int it = nt - 2; it >= 0; it--) {
        tt += trace[it];
        trace[it] = tt;
    }
}

/* kirmig_pick(adj=1), SERECKIRCH/src/kirmig.cpp:34-58.  The reference writes
   its sum with double literals (2., 1.), so the sum is formed in double and
   rounded once when added to the float sample. */
static void F(pick)(int nt, R t0, R dt, R ti, R deltat, R *sample, const R *trace)
{
    int it, itm, itp;
    R ft, tm, tp, ftm, ftp, imp;
    ft = (ti - t0) / dt; it = (int)R_FLOOR(ft); ft -= (R)it;
    if (it < 0 || it >= nt - 1) return;
    tm = ti - deltat - dt;
    ftm = (tm - t0) / dt; itm = (int)R_FLOOR(ftm); ftm -= (R)itm;
    if (itm < 0) return;
    tp = ti + deltat + dt;
    ftp = (tp - t0) / dt; itp = (int)R_FLOOR(ftp); ftp -= (R)itp;
    if (itp >= nt - 1) return;
    imp = dt / (dt + tp - tm);
    imp *= imp;
    sample[0] += imp * (2. * (1. - ft) * trace[it] + 2. * ft * trace[it + 1] -
                        (1. - ftm) * trace[itm] - ftm * trace[itm + 1] -
                        (1. - ftp) * trace[itp] - ftp * trace[itp + 1]);
}

/* kirmig_pick(adj=0), SERECKIRCH/src/kirmig.cpp:34-53, :59-67: the transpose
   of the pick above.  Each trace update is formed in double (the reference's
   2. and 1. literals) and rounded once into the float trace sample. */
static void F(pick0)(int nt, R t0, R dt, R ti, R deltat, R sample, R *trace)
{
    int it, itm, itp;
    R ft, tm, tp, ftm, ftp, imp, s;
    ft = (ti - t0) / dt; it = (int)R_FLOOR(ft); ft -= (R)it;
    if (it < 0 || it >= nt - 1) return;
    tm = ti - deltat - dt;
    ftm = (tm - t0) / dt; itm = (int)R_FLOOR(ftm); ftm -= (R)itm;
    if (itm < 0) return;
    tp = ti + deltat + dt;
    ftp = (tp - t0) / dt; itp = (int)R_FLOOR(ftp); ftp -= (R)itp;
    if (itp >= nt - 1) return;
    imp = dt / (dt + tp - tm);
    imp *= imp;
    s = imp * sample;
    trace[it] += 2. * (1. - ft) * s;
    trace[it + 1] += 2. * ft * s;
    trace[itm] -= (1. - ftm) * s;
    trace[itm + 1] -= ftm * s;
    trace[itp] -= (1. - ftp) * s;
    trace[itp + 1] -= ftp * s;
}

/* tinterp_linear / dinterp_linear, SERECKIRCH/src/tinterpsr.cpp:37-72 */
static void F(lin)(size_t nn, float dx_, float x, const float *p0, const float *p1, R *T, R *Tx)
{
    const R dx = (R)dx_;
    const R t = (R)x / dx;
    for (size_t i = 0; i < nn; i++) {
        const R a = (R)p0[i], b = (R)p1[i];
        T[i] = (R)((1. - t) * a + (R)(t * b));
        Tx[i] = (-a + b) / dx;
    }
}

/* tinterp_partial / dinterp_partial, SERECKIRCH/src/tinterpsr.cpp:74-140
   (nn = nz*nx, n1 = nz; k1/k2 truncate toward zero as the C conversion does) */
static void F(partial)(size_t nn_, int n1, float dd_, float dx_, float x, const float *p0,
                       const float *p1, R *T, R *Tx)
{
    const long nn = (long)nn_;
    const R dd = (R)dd_, dx = (R)dx_;
    const R t = (R)x / dx;
    /* shifts are geometry: float as in the reference */
    const int k1 = (int)((float)(x / dd_) + 0.5);
    const int k2 = (int)((float)((dx_ - x) / dd_) + 0.5);
    for (long i = 0; i < nn; i++) {
        if (i - (long)k1 * n1 >= 0 && i + (long)k2 * n1 < nn) {
            T[i] = (R)((1. - t) * (R)p0[i - (long)k1 * n1] + (R)(t * (R)p1[i + (long)k2 * n1]));
        } else {
            T[i] = (R)((1. - t) * (R)p0[i] + (R)(t * (R)p1[i]));
        }
    }
    for (long i = 0; i < nn; i++) {
        R v;
        if (i - (long)k1 * n1 >= 0 && i + (long)k2 * n1 < nn) {
            v = (-(R)p0[i - (long)k1 * n1] + (R)p1[i + (long)k2 * n1]) / dx;
            if (i - (long)(k1 + 1) * n1 >= 0 && i - (long)(k1 - 1) * n1 < nn)
                v = (R)(v - (1. - t) * (R)(-(R)p0[i - (long)(k1 + 1) * n1] +
                                          (R)p0[i - (long)(k1 - 1) * n1]) / (2. * dd));
            else if (i - (long)(k1 + 1) * n1 < 0)
                v = (R)(v - (1. - t) * (R)(-(R)p0[i - (long)k1 * n1] +
                                          (R)p0[i - (long)(k1 - 1) * n1]) / dd);
            else
                v = (R)(v - (1. - t) * (R)(-(R)p0[i - (long)(k1 + 1) * n1] +
                                          (R)p0[i - (long)k1 * n1]) / dd);
            if (i - (long)(k2 + 1) * n1 >= 0 && i - (long)(k2 - 1) * n1 < nn)
                v = (R)(v - t * (R)(-(R)p1[i - (long)(k2 + 1) * n1] +
                                   (R)p1[i - (long)(k2 - 1) * n1]) / (2. * dd));
            else if (i - (long)(k2 + 1) * n1 < 0)
                v = (R)(v - t * (R)(-(R)p1[i - (long)k2 * n1] +
                                   (R)p1[i - (long)(k2 - 1) * n1]) / dd);
            else
                v = (R)(v - t * (R)(-(R)p1[i - (long)(k2 + 1) * n1] +
                                   (R)p1[i - (long)k2 * n1]) / dd);
        } else {
            v = (-(R)p0[i] + (R)p1[i]) / dx;
        }
        Tx[i] = v;
    }
}

/* tinterp_hermite / dinterp_hermite, SERECKIRCH/src/tinterpsr.cpp:142-188:
   float t, double basis coefficients, double sum rounded once to the store. */
static void F(hermite)(size_t nn, float dx_, float x, const float *p0, const float *p1,
                       const float *m0, const float *m1, R *T, R *Tx)
{
    const R dx = (R)dx_;
    const R t = (R)x / dx;
    const double h00 = 2. * t * t * t - 3. * t * t + 1.;
    const double h10 = t * t * t - 2. * t * t + t;
    const double h01 = -2. * t * t * t + 3. * t * t;
    const double h11 = t * t * t - t * t;
    const double d00 = -6. * t * (1. - t);
    const double d10 = (1. - t) * (1. - 3. * t);
    const double d01 = 6. * t * (1. - t);
    const double d11 = 3. * t * t - 2. * t;
    for (size_t i = 0; i < nn; i++) {
        T[i] = (R)(h00 * p0[i] + h10 * dx * m0[i] + h01 * p1[i] + h11 * dx * m1[i]);
        Tx[i] = (R)((d00 * p0[i] + d10 * dx * m0[i] + d01 * p1[i] + d11 * dx * m1[i]) / dx);
    }
}

/* One side's table at coordinate pos: the slice choice and clamping of
   src/kirchmig2d_cig.cpp:987-1028 (source) and :1148-1180 (receiver).
   Positions, node choice and the in-interval coordinate are geometry and are
   evaluated in float in both builds, as the reference does; only the
   interpolation arithmetic follows R.  Returns the clamped node index.
   type 'p' at image columns ix < k2 reads p1 before its slice (the
   reference's dinterp_partial indexes p1[i - k2*n1] there,
   SERECKIRCH/src/tinterpsr.cpp:128-133); here the tables are one contiguous
   volume, so that read lands in the previous slice, a defined value. */
static int F(interp_side)(int type, int nz, int nx, float dx_grid, int npos, float o, float d,
                          float pos, const float *tab, const float *der, R *T, R *Tx)
{
    const size_t nzx = (size_t)nz * nx;
    int inode;
    if (npos <= 1 || d == 0.0f) {
        inode = 0;
    } else {
        const float f = (pos - o) / d;
        inode = (int)floorf(f);
        if (f <= 0.0f) {
            inode = 0;
        } else if (f >= (float)(npos - 1)) {
            inode = npos - 1;
        } else {
            const float coord = pos - (o + (float)inode * d);
            const float *p0 = tab + (size_t)inode * nzx, *p1 = tab + (size_t)(inode + 1) * nzx;
            const float *m0 = der + (size_t)inode * nzx, *m1 = der + (size_t)(inode + 1) * nzx;
            switch (type) {
            case 'l': F(lin)(nzx, d, coord, p0, p1, T, Tx); break;
            case 'p': F(partial)(nzx, nz, dx_grid, d, coord, p0, p1, T, Tx); break;
            default: F(hermite)(nzx, d, coord, p0, p1, m0, m1, T, Tx); break;
            }
            return inode;
        }
    }
    for (size_t i = 0; i < nzx; i++) {
        T[i] = (R)tab[(size_t)inode * nzx + i];
        Tx[i] = (R)der[(size_t)inode * nzx + i];
    }
    return inode;
}

/* interp_volume_linear_stream, src/kirchmig2d_cig.cpp:294-322 */
static void F(interp_volume_linear)(R *out, const float *vol, int npos, size_t nxy,
                                    float coord, float o, float d)
{
    if (npos <= 1 || d == 0.0f) {
        for (size_t i = 0; i < nxy; i++) out[i] = (R)vol[i];
        return;
    }
    const float f = (coord - o) / d;
    if (f <= 0.0f) {
        for (size_t i = 0; i < nxy; i++) out[i] = (R)vol[i];
        return;
    }
    if (f >= (float)(npos - 1)) {
        for (size_t i = 0; i < nxy; i++) out[i] = (R)vol[(size_t)(npos - 1) * nxy + i];
        return;
    }
    const int i0 = (int)floorf(f);
    const R w = (R)(f - (float)i0); /* interpolation weight: geometry, float */
    const float *s0 = vol + (size_t)i0 * nxy, *s1 = vol + (size_t)(i0 + 1) * nxy;
    for (size_t i = 0; i < nxy; i++) out[i] = (LIT(1.0) - w) * (R)s0[i] + w * (R)s1[i];
}

/* main() of src/kirchmig2d_cig.cpp, adj=1 branch: shot loop :979-1053,
   receiver loop :1054-1201, image-point loop :1203-1339. */
static int F(migrate_impl)(const kmo_params *p, const kmo_geom *g_in, const float *data,
                           const float *stab, const float *sder, const float *rtab,
                           const float *rder, const float *sgxv, const float *sgzv,
                           const float *rgxv, const float *rgzv, int shot_begin, int shot_end,
                           R *out, float *qc_before, float *qc_after)
{
    kmo_geom g = *g_in;
    if (p->adj != 1) return -1;
    if (g.nh == 1) g.dh = 0; /* :700 */
    if (g.ns == 1) g.ds = 0; /* :701 */
    const int nz = g.nz, nx = g.nx, nt = g.nt, nh = g.nh, ns = g.ns;
    const size_t nzx = (size_t)nz * nx;
    const int use_grad = (sgxv && sgzv && rgxv && rgzv);
    const int angle_mode = p->cig && p->cig_is_angle;
    const int ng = p->cig ? p->ng : 1;
    const R cig0 = p->cig ? (R)p->g0 : 0, dcig = p->cig ? (R)p->dg : 1;
    if ((sgxv || sgzv || rgxv || rgzv) && !use_grad) return -2;
    if (nz <= 0 || nx <= 0 || nt < 2 || nh <= 0 || ns <= 0 || ng <= 0) return -3;
    if (shot_begin < 0) shot_begin = 0;
    if (shot_end > ns) shot_end = ns;

    memset(out, 0, sizeof(R) * nzx * (size_t)ng);
    R *stable = (R *)malloc(sizeof(R) * nzx), *stablex = (R *)malloc(sizeof(R) * nzx);
    R *rtable = (R *)malloc(sizeof(R) * nzx), *rtablex = (R *)malloc(sizeof(R) * nzx);
    R *gsx = (R *)malloc(sizeof(R) * nzx), *gsz = (R *)malloc(sizeof(R) * nzx);
    R *grx = (R *)malloc(sizeof(R) * nzx), *grz = (R *)malloc(sizeof(R) * nzx);
    R *tmp1 = (R *)malloc(sizeof(R) * nzx), *tmp2 = (R *)malloc(sizeof(R) * nzx);
    R *trace = (R *)malloc(sizeof(R) * (size_t)nt), *ttmp = (R *)malloc(sizeof(R) * (size_t)nt);

    /* acquisition geometry (positions, node and trace-aperture indices) is
       evaluated in float in both builds, as the reference does; the values
       that enter the summation follow R */
    const float h0 = g.h0, s0 = g.s0;
    const float sy0 = g.sy0, sdy = g.sdy, ry0 = g.ry0, rdy = g.rdy;
    const R t0 = g.t0, dt = g.dt, dh = g.dh, ds = g.ds;
    const R z0 = g.z0, dz = g.dz, x0 = g.x0, dx = g.dx;
    const int sny = g.sny, rny = g.rny;
    const R tau = p->tau, aal = p->antialias, aper = p->aperture;
    const R ap_taper = p->aperture_taper, ap_rmin = p->aperture_rmin, ap_rt = p->aperture_rtaper;
    (void)z0;

    for (int is = shot_begin; is < shot_end; is++) {
        const float s = s0 + (float)is * g.ds; /* :980 */
        int ist = F(interp_side)(p->interp_type, nz, nx, g.dx, sny, sy0, sdy, s, stab, sder,
                                 stable, stablex);
        if (angle_mode) {
            if (use_grad) {
                F(interp_volume_linear)(gsx, sgxv, sny, nzx, s, sy0, sdy);
                F(interp_volume_linear)(gsz, sgzv, sny, nzx, s, sy0, sdy);
            } else {
                F(grad_plane)(stable, nz, nx, dz, dx, p->angle_grad_radius, gsx, gsz);
            }
            if (p->angle_smooth > 0) {
                F(smooth)(gsx, nz, nx, p->angle_smooth, tmp1, tmp2);
                F(smooth)(gsz, nz, nx, p->angle_smooth, tmp1, tmp2);
            }
        }
        for (int ih = 0; ih < nh; ih++) {
            const float h = h0 + (float)ih * g.dh;
            const float receiver_pos = p->cmp ? (s + h) : h;
            float rf = 0.0f;
            if (rny > 1 && rdy != 0.0f) rf = (receiver_pos - ry0) / rdy;
            const int iht = (rny > 1 && rdy != 0.0f) ? (int)floorf(rf) : 0;
            R trace_ap_weight = 1;
            const float *src = data + ((size_t)is * nh + ih) * (size_t)nt;
            for (int it = 0; it < nt; it++) trace[it] = (R)src[it];
            if (p->direct_mute && is == p->direct_mute_qc_shot && qc_before)
                for (int it = 0; it < nt; it++) qc_before[(size_t)ih * nt + it] = (float)trace[it];
            if (p->direct_mute)
                F(mute)(trace, nt, t0, dt, s, p->cmp ? (s + h) : h, p->direct_mute_slope,
                        p->direct_mute_intercept, p->direct_mute_taper, p->direct_mute_invert);
            if (p->direct_mute && is == p->direct_mute_qc_shot && qc_after)
                for (int it = 0; it < nt; it++) qc_after[(size_t)ih * nt + it] = (float)trace[it];
            if (p->aperture_trace != -1) { /* :1082-1099 */
                const R ad = (R)fabsf((float)(iht - ist));
                if (ad >= (R)p->aperture_trace) continue;
                if (p->aperture_trace_taper > 0) {
                    const R taper_begin =
                        MAX((R)0, (R)(p->aperture_trace - p->aperture_trace_taper));
                    if (ad > taper_begin) {
                        const R denom = (R)p->aperture_trace - taper_begin;
                        if (denom > 0) {
                            const R x = (ad - taper_begin) / denom;
                            trace_ap_weight = LIT(0.5) * (LIT(1.0) + R_COS((R)M_PI * x));
                        }
                    }
                }
            }
            int ig_offset = 0;
            if (p->cig && !p->cig_is_angle) {
                ig_offset = F(nearest_bin)((R)h, cig0, dcig, ng);
                if (ig_offset < 0) continue;
            }
            F(doubint_)(nt, trace);
            if (p->isdiff == 1) F(diff2)(trace, nt, dt, ttmp);

            F(interp_side)(p->interp_type, nz, nx, g.dx, rny, ry0, rdy, receiver_pos, rtab, rder,
                           rtable, rtablex);
            if (angle_mode) {
                if (use_grad) {
                    F(interp_volume_linear)(grx, rgxv, rny, nzx, receiver_pos, ry0, rdy);
                    F(interp_volume_linear)(grz, rgzv, rny, nzx, receiver_pos, ry0, rdy);
                } else {
                    F(grad_plane)(rtable, nz, nx, dz, dx, p->angle_grad_radius, grx, grz);
                }
                if (p->angle_smooth > 0) {
                    F(smooth)(grx, nz, nx, p->angle_smooth, tmp1, tmp2);
                    F(smooth)(grz, nz, nx, p->angle_smooth, tmp1, tmp2);
                }
            }
            const R rec_surface_x = (R)(p->cmp ? (s + h) : h);
            long i;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
            for (i = 0; i < (long)nzx; i++) {
                const int iz = (int)(i % nz);
                const int ix = (int)((i - iz) / nz);
                const R zdist = MAX(R_FABS((R)iz * dz), MAX(LIT(1.0e-12), LIT(1.0e-6) * R_FABS(dz)));
                const R ximg = x0 + (R)ix * dx;
                const R dsx = R_FABS(s - ximg);
                const R drx = R_FABS(rec_surface_x - ximg);
                const R as_deg = R_ATAN(dsx / zdist) * LIT(180.0) / (R)M_PI;
                const R ar_deg = R_ATAN(drx / zdist) * LIT(180.0) / (R)M_PI;
                const R ws_leg = MAX(F(leg_weight)(as_deg, aper, ap_taper),
                                     F(floor_weight)(dsx, ap_rmin, ap_rt));
                const R wr_leg = MAX(F(leg_weight)(ar_deg, aper, ap_taper),
                                     F(floor_weight)(drx, ap_rmin, ap_rt));
                const R apw = ws_leg * wr_leg;
                if (apw <= 0) continue;
                const R ti = stable[i] + rtable[i] + tau;
                const R tx = MAX(R_FABS(stablex[i] * ds), R_FABS(rtablex[i] * dh));
                if (angle_mode) {
                    const R angle = F(angle)(gsx[i], gsz[i], grx[i], grz[i], p->half_angle);
                    const R aw = F(angle_taper)(angle, p->angle_taper_start, p->angle_taper_end) *
                                 F(depth_mute)(angle, iz, nz, p->oazmin, p->oazmax, p->oaztaper) *
                                 trace_ap_weight * apw;
                    if (aw <= 0) continue;
                    if (p->angle_interp) {
                        int ig0 = -1, ig1 = -1;
                        R w0 = 0, w1 = 0;
                        if (!F(linear_bins)(angle, cig0, dcig, ng, &ig0, &ig1, &w0, &w1)) continue;
                        R picked = 0;
                        F(pick)(nt, t0, dt, ti, tx * aal, &picked, trace);
                        picked *= aw;
                        out[(size_t)ig0 * nzx + i] += w0 * picked;
                        if (ig1 != ig0 && w1 > 0) out[(size_t)ig1 * nzx + i] += w1 * picked;
                    } else {
                        const int ig = F(nearest_bin)(angle, cig0, dcig, ng);
                        if (ig < 0) continue;
                        R picked = 0;
                        F(pick)(nt, t0, dt, ti, tx * aal, &picked, trace);
                        out[(size_t)ig * nzx + i] += aw * picked;
                    }
                } else {
                    const int ig = p->cig ? ig_offset : 0;
                    R picked = 0;
                    F(pick)(nt, t0, dt, ti, tx * aal, &picked, trace);
                    out[(size_t)ig * nzx + i] += trace_ap_weight * apw * picked;
                }
            }
        }
    }
    free(stable); free(stablex); free(rtable); free(rtablex);
    free(gsx); free(gsz); free(grx); free(grz); free(tmp1); free(tmp2);
    free(trace); free(ttmp);
    return 0;
}

/* adj=0 modelling, src/kirchmig2d_cig.cpp:979-1351 with adj = 0: for every
   trace the image values are sprayed with the transposed pick and the
   weights, bins and apertures of the migration, then the trace is doubly
   integrated (:1341-1350) and written; traces outside the trace aperture or
   the offset axis are written as zeros (:1106-1139).  No mute, no diff (the
   reference applies both to input traces only).  The point loop is
   sequential, the order of the reference's own run on one thread (it sums
   per-thread trace buffers in thread order, :1342-1347). */
static int F(model_impl)(const kmo_params *p, const kmo_geom *g_in, const float *image,
                         const float *stab, const float *sder, const float *rtab,
                         const float *rder, const float *sgxv, const float *sgzv,
                         const float *rgxv, const float *rgzv, int shot_begin, int shot_end,
                         R *data)
{
    kmo_geom g = *g_in;
    if (g.nh == 1) g.dh = 0;
    if (g.ns == 1) g.ds = 0;
    const int nz = g.nz, nx = g.nx, nt = g.nt, nh = g.nh, ns = g.ns;
    const size_t nzx = (size_t)nz * nx;
    const int use_grad = (sgxv && sgzv && rgxv && rgzv);
    const int angle_mode = p->cig && p->cig_is_angle;
    const int ng = p->cig ? p->ng : 1;
    const R cig0 = p->cig ? (R)p->g0 : 0, dcig = p->cig ? (R)p->dg : 1;
    if ((sgxv || sgzv || rgxv || rgzv) && !use_grad) return -2;
    if (nz <= 0 || nx <= 0 || nt < 2 || nh <= 0 || ns <= 0 || ng <= 0) return -3;
    if (shot_begin < 0) shot_begin = 0;
    if (shot_end > ns) shot_end = ns;

    R *stable = (R *)malloc(sizeof(R) * nzx), *stablex = (R *)malloc(sizeof(R) * nzx);
    R *rtable = (R *)malloc(sizeof(R) * nzx), *rtablex = (R *)malloc(sizeof(R) * nzx);
    R *gsx = (R *)malloc(sizeof(R) * nzx), *gsz = (R *)malloc(sizeof(R) * nzx);
    R *grx = (R *)malloc(sizeof(R) * nzx), *grz = (R *)malloc(sizeof(R) * nzx);
    R *tmp1 = (R *)malloc(sizeof(R) * nzx), *tmp2 = (R *)malloc(sizeof(R) * nzx);
    R *img = (R *)malloc(sizeof(R) * nzx * (size_t)ng);
    for (size_t k = 0; k < nzx * (size_t)ng; k++) img[k] = (R)image[k];

    const float h0 = g.h0, s0 = g.s0;
    const float sy0 = g.sy0, sdy = g.sdy, ry0 = g.ry0, rdy = g.rdy;
    const R t0 = g.t0, dt = g.dt, dh = g.dh, ds = g.ds;
    const R dz = g.dz, x0 = g.x0, dx = g.dx;
    const int sny = g.sny, rny = g.rny;
    const R tau = p->tau, aal = p->antialias, aper = p->aperture;
    const R ap_taper = p->aperture_taper, ap_rmin = p->aperture_rmin, ap_rt = p->aperture_rtaper;

    for (int is = shot_begin; is < shot_end; is++) {
        const float s = s0 + (float)is * g.ds;
        int ist = F(interp_side)(p->interp_type, nz, nx, g.dx, sny, sy0, sdy, s, stab, sder,
                                 stable, stablex);
        if (angle_mode) {
            if (use_grad) {
                F(interp_volume_linear)(gsx, sgxv, sny, nzx, s, sy0, sdy);
                F(interp_volume_linear)(gsz, sgzv, sny, nzx, s, sy0, sdy);
            } else {
                F(grad_plane)(stable, nz, nx, dz, dx, p->angle_grad_radius, gsx, gsz);
            }
            if (p->angle_smooth > 0) {
                F(smooth)(gsx, nz, nx, p->angle_smooth, tmp1, tmp2);
                F(smooth)(gsz, nz, nx, p->angle_smooth, tmp1, tmp2);
            }
        }
        for (int ih = 0; ih < nh; ih++) {
            const float h = h0 + (float)ih * g.dh;
            const float receiver_pos = p->cmp ? (s + h) : h;
            float rf = 0.0f;
            if (rny > 1 && rdy != 0.0f) rf = (receiver_pos - ry0) / rdy;
            const int iht = (rny > 1 && rdy != 0.0f) ? (int)floorf(rf) : 0;
            R *trace = data + ((size_t)is * nh + ih) * (size_t)nt;
            for (int it = 0; it < nt; it++) trace[it] = 0;
            R trace_ap_weight = 1;
            if (p->aperture_trace != -1) {
                const R ad = (R)fabsf((float)(iht - ist));
                if (ad >= (R)p->aperture_trace) continue;
                if (p->aperture_trace_taper > 0) {
                    const R taper_begin =
                        MAX((R)0, (R)(p->aperture_trace - p->aperture_trace_taper));
                    if (ad > taper_begin) {
                        const R denom = (R)p->aperture_trace - taper_begin;
                        if (denom > 0) {
                            const R x = (ad - taper_begin) / denom;
                            trace_ap_weight = LIT(0.5) * (LIT(1.0) + R_COS((R)M_PI * x));
                        }
                    }
                }
            }
            int ig_offset = 0;
            if (p->cig && !p->cig_is_angle) {
                ig_offset = F(nearest_bin)((R)h, cig0, dcig, ng);
                if (ig_offset < 0) continue;
            }
            F(interp_side)(p->interp_type, nz, nx, g.dx, rny, ry0, rdy, receiver_pos, rtab, rder,
                           rtable, rtablex);
            if (angle_mode) {
                if (use_grad) {
                    F(interp_volume_linear)(grx, rgxv, rny, nzx, receiver_pos, ry0, rdy);
                    F(interp_volume_linear)(grz, rgzv, rny, nzx, receiver_pos, ry0, rdy);
                } else {
                    F(grad_plane)(rtable, nz, nx, dz, dx, p->angle_grad_radius, grx, grz);
                }
                if (p->angle_smooth > 0) {
                    F(smooth)(grx, nz, nx, p->angle_smooth, tmp1, tmp2);
                    F(smooth)(grz, nz, nx, p->angle_smooth, tmp1, tmp2);
                }
            }
            const R rec_surface_x = (R)(p->cmp ? (s + h) : h);
            for (size_t i = 0; i < nzx; i++) {
                const int iz = (int)(i % nz);
                const int ix = (int)((i - iz) / nz);
                const R zdist = MAX(R_FABS((R)iz * dz), MAX(LIT(1.0e-12), LIT(1.0e-6) * R_FABS(dz)));
                const R ximg = x0 + (R)ix * dx;
                const R dsx = R_FABS(s - ximg);
                const R drx = R_FABS(rec_surface_x - ximg);
                const R as_deg = R_ATAN(dsx / zdist) * LIT(180.0) / (R)M_PI;
                const R ar_deg = R_ATAN(drx / zdist) * LIT(180.0) / (R)M_PI;
                const R ws_leg = MAX(F(leg_weight)(as_deg, aper, ap_taper),
                                     F(floor_weight)(dsx, ap_rmin, ap_rt));
                const R wr_leg = MAX(F(leg_weight)(ar_deg, aper, ap_taper),
                                     F(floor_weight)(drx, ap_rmin, ap_rt));
                const R apw = ws_leg * wr_leg;
                if (apw <= 0) continue;
                const R ti = stable[i] + rtable[i] + tau;
                const R tx = MAX(R_FABS(stablex[i] * ds), R_FABS(rtablex[i] * dh));
                if (angle_mode) {
                    const R angle = F(angle)(gsx[i], gsz[i], grx[i], grz[i], p->half_angle);
                    const R aw = F(angle_taper)(angle, p->angle_taper_start, p->angle_taper_end) *
                                 F(depth_mute)(angle, iz, nz, p->oazmin, p->oazmax, p->oaztaper) *
                                 trace_ap_weight * apw;
                    if (aw <= 0) continue;
                    if (p->angle_interp) {
                        int ig0 = -1, ig1 = -1;
                        R w0 = 0, w1 = 0;
                        if (!F(linear_bins)(angle, cig0, dcig, ng, &ig0, &ig1, &w0, &w1)) continue;
                        F(pick0)(nt, t0, dt, ti, tx * aal, aw * w0 * img[(size_t)ig0 * nzx + i],
                                 trace);
                        if (ig1 != ig0 && w1 > 0)
                            F(pick0)(nt, t0, dt, ti, tx * aal,
                                     aw * w1 * img[(size_t)ig1 * nzx + i], trace);
                    } else {
                        const int ig = F(nearest_bin)(angle, cig0, dcig, ng);
                        if (ig < 0) continue;
                        F(pick0)(nt, t0, dt, ti, tx * aal, aw * img[(size_t)ig * nzx + i], trace);
                    }
                } else {
                    const int ig = p->cig ? ig_offset : 0;
                    F(pick0)(nt, t0, dt, ti, tx * aal,
                             trace_ap_weight * apw * img[(size_t)ig * nzx + i], trace);
                }
            }
            F(doubint_)(nt, trace);
        }
    }
    free(stable); free(stablex); free(rtable); free(rtablex);
    free(gsx); free(gsz); free(grx); free(grz); free(tmp1); free(tmp2); free(img);
    return 0;
}

#undef F
