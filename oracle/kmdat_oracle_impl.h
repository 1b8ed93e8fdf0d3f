/* Body of the kirchdat2d_freq oracle, included twice by kmdat_oracle.c with
   R = float (the reference's own arithmetic: float everywhere except the
   operator weight, which the reference forms in double through M_PI) and
   R = double (everything in double).  F(name) appends the precision suffix. */

/* build_filter_F, src/kirchdat2d_frequency_cs.cpp:100-117: the second
   difference of sqrt(((tau + k dt) / tau)^2 - 1), k = 0 .. nsam-1 (the k = 0
   value is 0), nsam-1 taps. */
static void F(filter)(R tau, R dt, int nsam, R *Fk)
{
    if (nsam <= 1) return;
    R prev2 = 0, prev1 = R_SQRT(((tau + dt) / tau) * ((tau + dt) / tau) - (R)1);
    Fk[0] = prev1 - prev2;
    for (int k = 1; k < nsam - 1; k++) {
        const R t = (tau + (R)(k + 1) * dt) / tau;
        const R curr = R_SQRT(t * t - (R)1);
        Fk[k] = curr - (R)2 * prev1 + prev2;
        prev2 = prev1;
        prev1 = curr;
    }
}

/* build_H_iw, src/kirchdat2d_frequency_cs.cpp:119-184: the spectrum at bin iw
   of sum_k F[k]/dt [(1-delta) x(n-m-k) + delta x(n-m-k+1)], m = ceil(tau/dt),
   delta = m - tau/dt, in the reference's factored form. */
static void F(Hw)(R tau, R dt, int nsam, int nfft, int iw, R *Fk, R *hr, R *hi)
{
    if (tau <= (R)0 || !isfinite(tau)) {
        *hr = 0;
        *hi = 0;
        return;
    }
    F(filter)(tau, dt, nsam, Fk);
    const R q = tau / dt;
    const int m = (int)R_CEIL(q);
    const R delta = (R)m - q;
    const R omega = (R)2 * (R)M_PI * (R)iw / (R)nfft;
    const R theta = -omega;
    const R step_i = R_SIN(theta), step_r = R_COS(theta);
    R z_r = 1, z_i = 0, sum_r = 0, sum_i = 0;
    for (int k = 0; k <= nsam - 2; k++) {
        const R fk = Fk[k] / dt;
        sum_r += fk * z_r;
        sum_i += fk * z_i;
        const R next_r = z_r * step_r - z_i * step_i;
        const R next_i = z_r * step_i + z_i * step_r;
        z_r = next_r;
        z_i = next_i;
    }
    const R phase_m = -omega * (R)m, phase_m1 = -omega * (R)(m - 1);
    const R sin_m = R_SIN(phase_m), cos_m = R_COS(phase_m);
    const R sin_m1 = R_SIN(phase_m1), cos_m1 = R_COS(phase_m1);
    const R lin_r = ((R)1 - delta) * cos_m + delta * cos_m1;
    const R lin_i = ((R)1 - delta) * sin_m + delta * sin_m1;
    *hr = lin_r * sum_r - lin_i * sum_i;
    *hi = lin_r * sum_i + lin_i * sum_r;
}

/* table_slope_first_index, :192-201 */
static R F(slope)(const float *table, int n, int row, int col, R dcoord)
{
    if (n <= 1 || dcoord == (R)0) return 0;
    const float *t = table;
#define TB(r_, c_) ((R)t[(size_t)(r_) * n + (c_)])
    if (row <= 0) return (TB(1, col) - TB(0, col)) / dcoord;
    if (row >= n - 1) return (TB(n - 1, col) - TB(n - 2, col)) / dcoord;
    return (R)0.5 * (TB(row + 1, col) - TB(row - 1, col)) / dcoord;
#undef TB
}

/* antialias_weight_iw, :203-233 */
static R F(aa)(R antialias, int nfft, R dt, int iw, R dtau)
{
    if (antialias <= (R)0) return 1;
    if (iw <= 0) return 1;
    dtau = R_FABS(dtau) * antialias;
    if (dtau <= (R)1e-12 || !isfinite(dtau)) return 1;
    const R freq = (R)iw / ((R)nfft * dt);
    const R fnyq = (R)0.5 / dt;
    const R fmax = (R)0.5 / dtau;
    if (fmax >= fnyq) return 1;
    if (fmax <= (R)0) return 0;
    const R fpass = (R)0.8 * fmax;
    if (freq <= fpass) return 1;
    if (freq >= fmax) return 0;
    const R x = (freq - fpass) / (fmax - fpass);
    return (R)0.5 * ((R)1 + R_COS((R)M_PI * x));
}

/* taper_weight, :236-248 */
static R F(taper)(int left, int center, int right, int tap, int stride)
{
    if (center < left || center > right) return 0;
    if (tap <= 0) return 1;
    if (stride < 1) stride = 1;
    const R width = (R)tap * (R)stride;
    const R ul = R_FMIN((R)1, R_FMAX((R)0, (R)(center - left) / width));
    const R ur = R_FMIN((R)1, R_FMAX((R)0, (R)(right - center) / width));
    return (R)0.5 * ((R)1 - R_COS((R)M_PI * ul)) * (R)0.5 * ((R)1 - R_COS((R)M_PI * ur));
}

/* data_edge_taper_weight, :254-279 */
static R F(edge)(int i, int n, int tap)
{
    if (tap <= 0 || n <= 1) return 1;
    int ntap = tap;
    const int max_tap = (n - 1) / 2;
    if (ntap > max_tap) ntap = max_tap;
    if (ntap <= 0) return 1;
    if (i < ntap) return (R)0.5 * ((R)1 - R_COS((R)M_PI * ((R)i / (R)ntap)));
    if (i > n - 1 - ntap) return (R)0.5 * ((R)1 - R_COS((R)M_PI * ((R)(n - 1 - i) / (R)ntap)));
    return 1;
}

/* The operator weight of one summation pair (:512-519, :585-591, :638-641):
   the reference writes coef / M_PI * ... with a double M_PI, so the float
   build forms it in double and rounds once. */
static R F(weight)(R coef, R dstep, R datum, R tau, R dist)
{
    return (R)((double)coef / M_PI * (double)dstep * (double)datum * (double)tau / (double)dist);
}

typedef struct {
    R re, im;
} F(cpx);

/* One operator pair: out += w * aa * H(tau) * in, H cached per table key
   (the reference's hr_cache / hi_cache: a pure function of (key, iw)). */
static void F(pair)(const kmdo_cfg *cfg, int nfft, R dt, int nsam, int iw, R *Fk, R *hcache,
                    unsigned char *cached, size_t key, R tau, R w, F(cpx) a, F(cpx) *out)
{
    if (R_FABS(w) < (R)1e-20) return;
    if (!cached[key]) {
        F(Hw)(tau, dt, nsam, nfft, iw, Fk, &hcache[2 * key], &hcache[2 * key + 1]);
        cached[key] = 1;
    }
    const R br = hcache[2 * key], bi = hcache[2 * key + 1];
    out->re += w * (a.re * br - a.im * bi);
    out->im += w * (a.re * bi + a.im * br);
    (void)cfg;
}

#define IDX3(is_, ih_, iw_) ((((size_t)(is_) * nh + (ih_)) * nw) + (iw_))

/* apply_receiver_operator_freq_iw, :493-546 */
static int F(recv_op)(const kmdo_cfg *cfg, int nh, int ns, int nw, int nfft, int nsam, R dt, R h0,
                      R dh, R s0, R ds, const float *rtable, int nrg, R rg0, R drg,
                      const F(cpx) *Uin, F(cpx) *Utmp, int iw, R *Fk, R *hc, unsigned char *cached)
{
    memset(cached, 0, (size_t)nrg * nrg);
    for (int is = 0; is < ns; is++) {
        for (int ih = 0; ih < nh; ih++) {
            const R base = cfg->cmp ? (s0 + (R)is * ds) : (R)0;
            const int c = (int)(((base + h0 + (R)ih * dh - rg0) / drg) + (R)0.5);
            const int left = (ih - cfg->aper < 0) ? 0 : ih - cfg->aper;
            const int right = (ih + cfg->aper > nh - 1) ? nh - 1 : ih + cfg->aper;
            for (int ic = left; ic <= right; ic++) {
                const int cc = (int)(((base + h0 + (R)ic * dh - rg0) / drg) + (R)0.5);
                if (c < 0 || c >= nrg || cc < 0 || cc >= nrg) return -2;
                const R coef = F(taper)(left, ic, right, cfg->tap, 1);
                const R tau = (R)rtable[(size_t)cc * nrg + c];
                const R dist = (R)cfg->rdatum * (R)cfg->rdatum +
                               (R)((ic - ih) * (ic - ih)) * dh * dh;
                R w = F(weight)(coef, dh, (R)cfg->rdatum, tau, dist);
                const R sl = F(slope)(rtable, nrg, cc, c, drg);
                w *= F(aa)((R)cfg->antialias, nfft, dt, iw, sl * R_FABS(dh));
                F(pair)(cfg, nfft, dt, nsam, iw, Fk, hc, cached, (size_t)cc * nrg + c, tau, w,
                        Uin[IDX3(is, ic, iw)], &Utmp[IDX3(is, ih, iw)]);
            }
        }
    }
    return 0;
}

/* apply_source_operator_freq_iw, :548-672 (both cmp branches) */
static int F(src_op)(const kmdo_cfg *cfg, int nh, int ns, int nw, int nfft, int nsam, R dt, R h0,
                     R dh, R s0, R ds, const float *stable, int nsg, R sg0, R dsg,
                     const F(cpx) *Utmp, F(cpx) *Uout, int iw, R *Fk, R *hc,
                     unsigned char *cached)
{
    memset(cached, 0, (size_t)nsg * nsg);
    (void)h0;
    if (cfg->cmp == 1) {
        const R s = R_FABS((R)(ns - 1) * ds), h = R_FABS((R)(nh - 1) * dh);
        const int jump = (R_FABS(ds) >= R_FABS(dh)) ? 1 : (int)(dh / ds + (R)0.5);
        const R dr = (R_FABS(ds) >= R_FABS(dh)) ? R_FABS(dh) : R_FABS(ds);
        const int nr = (int)((s + h) / dr + (R)1.5);
        const R sgn_s = (ds <= (R)0) ? (R)-1 : (R)0, sgn_h = (dh <= (R)0) ? (R)-1 : (R)0;
        for (int ir = 0; ir < nr; ir++) {
            const R r = (R)ir * dr + sgn_s * s + sgn_h * h;
            int sleft = (int)(((R)ir * dr + sgn_s * s + ((ds <= (R)0) ? (R)0 : (R)-1) * h) / ds +
                              (R)0.5);
            int sright = (int)(((R)ir * dr + sgn_s * s + ((ds <= (R)0) ? (R)-1 : (R)0) * h) / ds +
                               (R)0.5);
            if (sleft < 0) sleft = 0;
            if (sright > ns - 1) sright = ns - 1;
            int left = (int)((r - (R)sleft * ds) / dh + (R)0.5);
            if (left < 0 || left > nh - 1) sleft++;
            int right = (int)((r - (R)sright * ds) / dh + (R)0.5);
            if (right < 0 || right > nh - 1) sright--;
            for (int is = sleft; is <= sright; is += jump) {
                const int c = (int)((s0 + (R)is * ds - sg0) / dsg + (R)0.5);
                if (c < 0 || c >= nsg) return -3;
                const int ih = (int)((r - (R)is * ds) / dh + (R)0.5);
                left = (is - jump * cfg->aper < sleft) ? sleft : is - jump * cfg->aper;
                right = (is + jump * cfg->aper > sright) ? sright : is + jump * cfg->aper;
                for (int ic = left; ic <= right; ic += jump) {
                    const int cc = (int)((s0 + (R)ic * ds - sg0) / dsg + (R)0.5);
                    if (cc < 0 || cc >= nsg) return -3;
                    const int hh = (int)((r - (R)ic * ds) / dh + (R)0.5);
                    if (ih < 0 || ih >= nh || hh < 0 || hh >= nh) return -4;
                    const R coef = F(taper)(left, ic, right, cfg->tap, jump);
                    const R tau = (R)stable[(size_t)cc * nsg + c];
                    const R dist = (R)cfg->sdatum * (R)cfg->sdatum +
                                   (R)((ic - is) * (ic - is)) * ds * ds;
                    R w = F(weight)(coef, ds, (R)cfg->sdatum, tau, dist);
                    const R sl = F(slope)(stable, nsg, cc, c, dsg);
                    w *= F(aa)((R)cfg->antialias, nfft, dt, iw, sl * R_FABS((R)jump * ds));
                    F(pair)(cfg, nfft, dt, nsam, iw, Fk, hc, cached, (size_t)cc * nsg + c, tau, w,
                            Utmp[IDX3(ic, hh, iw)], &Uout[IDX3(is, ih, iw)]);
                }
            }
        }
    } else {
        for (int ih = 0; ih < nh; ih++) {
            for (int is = 0; is < ns; is++) {
                const int c = (int)((s0 + (R)is * ds - sg0) / dsg + (R)0.5);
                if (c < 0 || c >= nsg) return -3;
                const int left = (is - cfg->aper < 0) ? 0 : is - cfg->aper;
                const int right = (is + cfg->aper > ns - 1) ? ns - 1 : is + cfg->aper;
                for (int ic = left; ic <= right; ic++) {
                    const int cc = (int)((s0 + (R)ic * ds - sg0) / dsg + (R)0.5);
                    if (cc < 0 || cc >= nsg) return -3;
                    const R coef = F(taper)(left, ic, right, cfg->tap, 1);
                    const R tau = (R)stable[(size_t)cc * nsg + c];
                    const R dist = (R)cfg->sdatum * (R)cfg->sdatum +
                                   (R)((ic - is) * (ic - is)) * ds * ds;
                    R w = F(weight)(coef, ds, (R)cfg->sdatum, tau, dist);
                    const R sl = F(slope)(stable, nsg, cc, c, dsg);
                    w *= F(aa)((R)cfg->antialias, nfft, dt, iw, sl * R_FABS(ds));
                    F(pair)(cfg, nfft, dt, nsam, iw, Fk, hc, cached, (size_t)cc * nsg + c, tau, w,
                            Utmp[IDX3(ic, ih, iw)], &Uout[IDX3(is, ih, iw)]);
                }
            }
        }
    }
    return 0;
}

#ifdef KMDO_WITH_BF
/* One entry of a butterfly operator's matrix (:680-745, :747-812,
   :1034-1060): the direct path's weight w (with the antialias factor), tau,
   and amp = w H(tau) exp(i omega tau) with omega the float angular frequency
   2 pi iw / (nfft dt); entries whose weight is below 1e-20 stay zero. */
static void F(bf_entry)(int nfft, R dt, int nsam, int iw, float omega, R *Fk, R *hc,
                        unsigned char *cached, size_t key, R tau, R w, float *tau_m, float *amp_m,
                        size_t e)
{
    if (R_FABS(w) < (R)1e-20) return;
    if (!cached[key]) {
        F(Hw)(tau, dt, nsam, nfft, iw, Fk, &hc[2 * key], &hc[2 * key + 1]);
        cached[key] = 1;
    }
    const R hr = hc[2 * key], hi = hc[2 * key + 1];
    const float ph = omega * (float)tau;
    const float sp = sinf(ph), cp = cosf(ph);
    tau_m[e] = (float)tau;
    amp_m[2 * e] = (float)(w * (hr * cp - hi * sp));
    amp_m[2 * e + 1] = (float)(w * (hr * sp + hi * cp));
}

typedef struct {
    float *tau, *amp; /* n x n, n = max(nh, ns) */
    float *vin, *vout; /* n x max(nh, ns) vectors */
} F(bfwork);

static float F(bf_omega)(int nfft, R dt, int iw)
{
    return 2.0f * (float)M_PI * (float)iw / ((float)nfft * (float)dt);
}

static kbo_cfg F(bf_cfg)(const kmdo_cfg *cfg)
{
    kbo_cfg b;
    b.p = cfg->bf_p;
    b.leaf = cfg->bf_leaf;
    b.levels = cfg->bf_levels;
    b.amp_eps = cfg->bf_amp_eps;
    b.phase_tol = cfg->bf_phase_tol;
    return b;
}

/* apply_receiver_operator_freq_iw_bf, :814-945: cmp=0 builds the nh x nh
   matrix once and applies it to every shot; cmp=1 builds it per shot. */
static int F(recv_op_bf)(const kmdo_cfg *cfg, int nh, int ns, int nw, int nfft, int nsam, R dt,
                         R h0, R dh, R s0, R ds, const float *rtable, int nrg, R rg0, R drg,
                         const F(cpx) *Uin, F(cpx) *Utmp, int iw, R *Fk, R *hc,
                         unsigned char *cached, F(bfwork) *W)
{
    memset(cached, 0, (size_t)nrg * nrg);
    const float omega = F(bf_omega)(nfft, dt, iw);
    const kbo_cfg bc = F(bf_cfg)(cfg);
    const int nmat = cfg->cmp ? ns : 1;
    for (int m = 0; m < nmat; m++) {
        const R base = cfg->cmp ? (s0 + (R)m * ds) : (R)0;
        memset(W->tau, 0, sizeof(float) * (size_t)nh * nh);
        memset(W->amp, 0, sizeof(float) * 2 * (size_t)nh * nh);
        for (int ih = 0; ih < nh; ih++) {
            const int c = (int)(((base + h0 + (R)ih * dh - rg0) / drg) + (R)0.5);
            const int left = (ih - cfg->aper < 0) ? 0 : ih - cfg->aper;
            const int right = (ih + cfg->aper > nh - 1) ? nh - 1 : ih + cfg->aper;
            for (int ic = left; ic <= right; ic++) {
                const int cc = (int)(((base + h0 + (R)ic * dh - rg0) / drg) + (R)0.5);
                if (c < 0 || c >= nrg || cc < 0 || cc >= nrg) return -2;
                const R coef = F(taper)(left, ic, right, cfg->tap, 1);
                const R tau = (R)rtable[(size_t)cc * nrg + c];
                const R dist = (R)cfg->rdatum * (R)cfg->rdatum +
                               (R)((ic - ih) * (ic - ih)) * dh * dh;
                R w = F(weight)(coef, dh, (R)cfg->rdatum, tau, dist);
                const R sl = F(slope)(rtable, nrg, cc, c, drg);
                w *= F(aa)((R)cfg->antialias, nfft, dt, iw, sl * R_FABS(dh));
                F(bf_entry)(nfft, dt, nsam, iw, omega, Fk, hc, cached, (size_t)cc * nrg + c, tau,
                            w, W->tau, W->amp, (size_t)ih * nh + ic);
            }
        }
        const int is0 = cfg->cmp ? m : 0, nvec = cfg->cmp ? 1 : ns;
        for (int v = 0; v < nvec; v++)
            for (int ic = 0; ic < nh; ic++) {
                const F(cpx) x = Uin[IDX3(is0 + v, ic, iw)];
                W->vin[2 * ((size_t)v * nh + ic)] = (float)x.re;
                W->vin[2 * ((size_t)v * nh + ic) + 1] = (float)x.im;
            }
        if (kbo_seg_apply(nh, W->tau, W->amp, omega, &bc, nvec, W->vin, W->vout, NULL, 0) < 0)
            return -1;
        for (int v = 0; v < nvec; v++)
            for (int ih = 0; ih < nh; ih++) {
                F(cpx) *y = &Utmp[IDX3(is0 + v, ih, iw)];
                y->re += W->vout[2 * ((size_t)v * nh + ih)];
                y->im += W->vout[2 * ((size_t)v * nh + ih) + 1];
            }
    }
    return 0;
}

/* apply_source_operator_freq_iw_bf, :947-1121: cmp=0 builds the ns x ns
   matrix once and applies it to every receiver; cmp=1 builds one per
   common-receiver gather ir (the shots sleft .. sright in steps of jump). */
static int F(src_op_bf)(const kmdo_cfg *cfg, int nh, int ns, int nw, int nfft, int nsam, R dt,
                        R h0, R dh, R s0, R ds, const float *stable, int nsg, R sg0, R dsg,
                        const F(cpx) *Utmp, F(cpx) *Uout, int iw, R *Fk, R *hc,
                        unsigned char *cached, F(bfwork) *W)
{
    memset(cached, 0, (size_t)nsg * nsg);
    (void)h0;
    const float omega = F(bf_omega)(nfft, dt, iw);
    const kbo_cfg bc = F(bf_cfg)(cfg);
    const size_t nn = (size_t)ns * ns;
    if (cfg->cmp == 1) {
        const R s = R_FABS((R)(ns - 1) * ds), h = R_FABS((R)(nh - 1) * dh);
        const int jump = (R_FABS(ds) >= R_FABS(dh)) ? 1 : (int)(dh / ds + (R)0.5);
        const R dr = (R_FABS(ds) >= R_FABS(dh)) ? R_FABS(dh) : R_FABS(ds);
        const int nr = (int)((s + h) / dr + (R)1.5);
        const R sgn_s = (ds <= (R)0) ? (R)-1 : (R)0, sgn_h = (dh <= (R)0) ? (R)-1 : (R)0;
        for (int ir = 0; ir < nr; ir++) {
            const R r = (R)ir * dr + sgn_s * s + sgn_h * h;
            int sleft = (int)(((R)ir * dr + sgn_s * s + ((ds <= (R)0) ? (R)0 : (R)-1) * h) / ds +
                              (R)0.5);
            int sright = (int)(((R)ir * dr + sgn_s * s + ((ds <= (R)0) ? (R)-1 : (R)0) * h) / ds +
                               (R)0.5);
            if (sleft < 0) sleft = 0;
            if (sright > ns - 1) sright = ns - 1;
            const int left0 = (int)((r - (R)sleft * ds) / dh + (R)0.5);
            if (left0 < 0 || left0 > nh - 1) sleft++;
            const int right0 = (int)((r - (R)sright * ds) / dh + (R)0.5);
            if (right0 < 0 || right0 > nh - 1) sright--;
            memset(W->vin, 0, sizeof(float) * 2 * (size_t)ns);
            memset(W->tau, 0, sizeof(float) * nn);
            memset(W->amp, 0, sizeof(float) * 2 * nn);
            for (int ic = sleft; ic <= sright; ic += jump) {
                const int hh = (int)((r - (R)ic * ds) / dh + (R)0.5);
                if (hh < 0 || hh >= nh) continue;
                const F(cpx) x = Utmp[IDX3(ic, hh, iw)];
                W->vin[2 * (size_t)ic] = (float)x.re;
                W->vin[2 * (size_t)ic + 1] = (float)x.im;
            }
            for (int is = sleft; is <= sright; is += jump) {
                const int c = (int)((s0 + (R)is * ds - sg0) / dsg + (R)0.5);
                if (c < 0 || c >= nsg) return -3;
                const int ih = (int)((r - (R)is * ds) / dh + (R)0.5);
                if (ih < 0 || ih >= nh) continue;
                const int ileft = (is - jump * cfg->aper < sleft) ? sleft : is - jump * cfg->aper;
                const int iright = (is + jump * cfg->aper > sright) ? sright : is + jump * cfg->aper;
                for (int ic = ileft; ic <= iright; ic += jump) {
                    const int cc = (int)((s0 + (R)ic * ds - sg0) / dsg + (R)0.5);
                    if (cc < 0 || cc >= nsg) return -3;
                    const int hh = (int)((r - (R)ic * ds) / dh + (R)0.5);
                    if (hh < 0 || hh >= nh) continue;
                    const R coef = F(taper)(ileft, ic, iright, cfg->tap, jump);
                    const R tau = (R)stable[(size_t)cc * nsg + c];
                    const R dist = (R)cfg->sdatum * (R)cfg->sdatum +
                                   (R)((ic - is) * (ic - is)) * ds * ds;
                    R w = F(weight)(coef, ds, (R)cfg->sdatum, tau, dist);
                    const R sl = F(slope)(stable, nsg, cc, c, dsg);
                    w *= F(aa)((R)cfg->antialias, nfft, dt, iw, sl * R_FABS((R)jump * ds));
                    F(bf_entry)(nfft, dt, nsam, iw, omega, Fk, hc, cached, (size_t)cc * nsg + c,
                                tau, w, W->tau, W->amp, (size_t)is * ns + ic);
                }
            }
            if (kbo_seg_apply(ns, W->tau, W->amp, omega, &bc, 1, W->vin, W->vout, NULL, 0) < 0)
                return -1;
            for (int is = sleft; is <= sright; is += jump) {
                const int ih = (int)((r - (R)is * ds) / dh + (R)0.5);
                if (ih < 0 || ih >= nh) continue;
                F(cpx) *y = &Uout[IDX3(is, ih, iw)];
                y->re += W->vout[2 * (size_t)is];
                y->im += W->vout[2 * (size_t)is + 1];
            }
        }
        return 0;
    }
    memset(W->tau, 0, sizeof(float) * nn);
    memset(W->amp, 0, sizeof(float) * 2 * nn);
    for (int is = 0; is < ns; is++) {
        const int c = (int)((s0 + (R)is * ds - sg0) / dsg + (R)0.5);
        if (c < 0 || c >= nsg) return -3;
        const int left = (is - cfg->aper < 0) ? 0 : is - cfg->aper;
        const int right = (is + cfg->aper > ns - 1) ? ns - 1 : is + cfg->aper;
        for (int ic = left; ic <= right; ic++) {
            const int cc = (int)((s0 + (R)ic * ds - sg0) / dsg + (R)0.5);
            if (cc < 0 || cc >= nsg) return -3;
            const R coef = F(taper)(left, ic, right, cfg->tap, 1);
            const R tau = (R)stable[(size_t)cc * nsg + c];
            const R dist = (R)cfg->sdatum * (R)cfg->sdatum + (R)((ic - is) * (ic - is)) * ds * ds;
            R w = F(weight)(coef, ds, (R)cfg->sdatum, tau, dist);
            const R sl = F(slope)(stable, nsg, cc, c, dsg);
            w *= F(aa)((R)cfg->antialias, nfft, dt, iw, sl * R_FABS(ds));
            F(bf_entry)(nfft, dt, nsam, iw, omega, Fk, hc, cached, (size_t)cc * nsg + c, tau, w,
                        W->tau, W->amp, (size_t)is * ns + ic);
        }
    }
    for (int ih = 0; ih < nh; ih++)
        for (int ic = 0; ic < ns; ic++) {
            const F(cpx) x = Utmp[IDX3(ic, ih, iw)];
            W->vin[2 * ((size_t)ih * ns + ic)] = (float)x.re;
            W->vin[2 * ((size_t)ih * ns + ic) + 1] = (float)x.im;
        }
    if (kbo_seg_apply(ns, W->tau, W->amp, omega, &bc, nh, W->vin, W->vout, NULL, 0) < 0) return -1;
    for (int ih = 0; ih < nh; ih++)
        for (int is = 0; is < ns; is++) {
            F(cpx) *y = &Uout[IDX3(is, ih, iw)];
            y->re += W->vout[2 * ((size_t)ih * ns + is)];
            y->im += W->vout[2 * ((size_t)ih * ns + is) + 1];
        }
    return 0;
}
#endif

/* Spectra [nw] of every trace -> nt output samples: the c2r transform
   (unnormalised, Hermitian), divided by nfft, optionally reversed
   (write_traces_from_freq, :470-491). */
static void F(to_time)(const F(cpx) *U, int ntr, int nt, int nw, int nfft, int reverse,
                       float *out)
{
    double *re = (double *)malloc(sizeof(double) * nfft);
    double *im = (double *)malloc(sizeof(double) * nfft);
    for (int t = 0; t < ntr; t++) {
        const F(cpx) *u = U + (size_t)t * nw;
        re[0] = (double)u[0].re;
        im[0] = 0.0;
        for (int k = 1; k < nfft / 2; k++) {
            re[k] = (double)u[k].re;
            im[k] = (double)u[k].im;
            re[nfft - k] = (double)u[k].re;
            im[nfft - k] = -(double)u[k].im;
        }
        re[nfft / 2] = (double)u[nfft / 2].re;
        im[nfft / 2] = 0.0;
        kmdo_fft_f64(nfft, re, im, 1);
        float *o = out + (size_t)t * nt;
        for (int it = 0; it < nt; it++) {
            const R v = (R)re[it] / (R)nfft;
            o[reverse ? nt - 1 - it : it] = (float)v;
        }
    }
    free(re);
    free(im);
}

int F(kmdo_datum)(const kmdo_cfg *cfg, const float *data, int nt, int nh, int ns, float dt_,
                  float h0_, float dh_, float s0_, float ds_, const float *stable, int nsg,
                  float sg0_, float dsg_, const float *rtable, int nrg, float rg0_, float drg_,
                  float *out, float *interm, int iw_begin, int iw_end)
{
    if (!cfg || !data || !out || nt < 1 || nh < 1 || ns < 1 || nsg < 1 || nrg < 1) return -1;
#ifndef KMDO_WITH_BF
    if (cfg->use_bf) return -1;
#endif
    const R dt = dt_, h0 = h0_, dh = dh_, s0 = s0_, ds = ds_;
    const R sg0 = sg0_, dsg = dsg_, rg0 = rg0_, drg = drg_;
    const int nsam = (int)(cfg->length / dt_) + 2;
    const int nfft = kmdo_nfft(nt, nsam, dt_, nrg, nsg, rtable, stable);
    const int nw = nfft / 2 + 1;
    if (iw_end <= 0 || iw_end > nw) iw_end = nw;
    if (iw_begin < 0) iw_begin = 0;
    const size_t ntr = (size_t)ns * nh, nU = ntr * nw;
    F(cpx) *Uin = (F(cpx) *)calloc(nU, sizeof(F(cpx)));
    F(cpx) *Utmp = (F(cpx) *)calloc(nU, sizeof(F(cpx)));
    F(cpx) *Uout = (F(cpx) *)calloc(nU, sizeof(F(cpx)));
    if (!Uin || !Utmp || !Uout) {
        free(Uin); free(Utmp); free(Uout);
        return -1;
    }
    /* read_traces_to_freq, :281-313: reverse, edge taper, zero pad, r2c */
    {
        double *re = (double *)malloc(sizeof(double) * nfft);
        double *im = (double *)malloc(sizeof(double) * nfft);
        for (int is = 0; is < ns; is++) {
            const R ws = F(edge)(is, ns, cfg->data_tap);
            for (int ih = 0; ih < nh; ih++) {
                const R wd = ws * F(edge)(ih, nh, cfg->data_tap);
                const float *tr = data + ((size_t)is * nh + ih) * nt;
                for (int k = 0; k < nfft; k++) re[k] = im[k] = 0.0;
                for (int it = 0; it < nt; it++) {
                    const float v = tr[nt - 1 - it];
                    re[it] = wd == (R)1 ? (double)v : (double)(R)(wd * (R)v);
                }
                kmdo_fft_f64(nfft, re, im, 0);
                F(cpx) *u = Uin + ((size_t)is * nh + ih) * nw;
                for (int k = 0; k < nw; k++) {
                    u[k].re = (R)re[k];
                    u[k].im = (R)im[k];
                }
            }
        }
        free(re);
        free(im);
    }
    int err = 0;
    const int nkey = nrg > nsg ? nrg : nsg;
    const int nmax = nh > ns ? nh : ns;
#pragma omp parallel reduction(min : err)
    {
        R *Fk = (R *)calloc(nsam > 1 ? nsam : 2, sizeof(R));
        R *hc = (R *)malloc(sizeof(R) * 2 * (size_t)nkey * nkey);
        unsigned char *cached = (unsigned char *)malloc((size_t)nkey * nkey);
#ifdef KMDO_WITH_BF
        F(bfwork) W = {NULL, NULL, NULL, NULL};
        if (cfg->use_bf) {
            W.tau = (float *)malloc(sizeof(float) * (size_t)nmax * nmax);
            W.amp = (float *)malloc(sizeof(float) * 2 * (size_t)nmax * nmax);
            W.vin = (float *)malloc(sizeof(float) * 2 * (size_t)nmax * nmax);
            W.vout = (float *)malloc(sizeof(float) * 2 * (size_t)nmax * nmax);
        }
#endif
        (void)nmax;
#pragma omp for schedule(static, 8)
        for (int iw = iw_begin; iw < iw_end; iw++) {
            if (err < 0) continue;
#ifdef KMDO_WITH_BF
            const int e = cfg->use_bf
                              ? F(recv_op_bf)(cfg, nh, ns, nw, nfft, nsam, dt, h0, dh, s0, ds, rtable,
                                              nrg, rg0, drg, Uin, Utmp, iw, Fk, hc, cached, &W)
                              : F(recv_op)(cfg, nh, ns, nw, nfft, nsam, dt, h0, dh, s0, ds, rtable,
                                           nrg, rg0, drg, Uin, Utmp, iw, Fk, hc, cached);
#else
            const int e = F(recv_op)(cfg, nh, ns, nw, nfft, nsam, dt, h0, dh, s0, ds, rtable, nrg,
                                     rg0, drg, Uin, Utmp, iw, Fk, hc, cached);
#endif
            if (e < err) err = e;
        }
#pragma omp barrier
#pragma omp for schedule(static, 8)
        for (int iw = iw_begin; iw < iw_end; iw++) {
            if (err < 0) continue;
#ifdef KMDO_WITH_BF
            const int e = cfg->use_bf
                              ? F(src_op_bf)(cfg, nh, ns, nw, nfft, nsam, dt, h0, dh, s0, ds, stable,
                                             nsg, sg0, dsg, Utmp, Uout, iw, Fk, hc, cached, &W)
                              : F(src_op)(cfg, nh, ns, nw, nfft, nsam, dt, h0, dh, s0, ds, stable,
                                          nsg, sg0, dsg, Utmp, Uout, iw, Fk, hc, cached);
#else
            const int e = F(src_op)(cfg, nh, ns, nw, nfft, nsam, dt, h0, dh, s0, ds, stable, nsg,
                                    sg0, dsg, Utmp, Uout, iw, Fk, hc, cached);
#endif
            if (e < err) err = e;
        }
#ifdef KMDO_WITH_BF
        free(W.tau);
        free(W.amp);
        free(W.vin);
        free(W.vout);
#endif
        free(Fk);
        free(hc);
        free(cached);
    }
    if (!err) {
        if (interm) F(to_time)(Utmp, (int)ntr, nt, nw, nfft, 0, interm);
        F(to_time)(Uout, (int)ntr, nt, nw, nfft, 1, out);
    }
    free(Uin);
    free(Utmp);
    free(Uout);
    return err;
}
#undef IDX3
