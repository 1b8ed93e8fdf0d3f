// libkmdat host side: keys, operator plans, rocFFT transforms and the
// datuming pipeline of kirchdat2d_freq (src/kirchdat2d_frequency_cs.cpp).
//
// Index and weight arithmetic follows the reference's expressions in float
// (this file is built without contraction): the trace and table indices
// round with (int)(x + 0.5f) exactly as the reference does, which matters
// where positions fall on .5 ties.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kmdat_internal.h"
#include "kmig_dat.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

using namespace kmdat;

struct Side {
    std::vector<float> tab;  // [n][n]
    int n = 0;
    float o = 0.0f, d = 1.0f;
    bool set = false;
};

// One operator (receiver or source side) planned on the host.
struct OpPlan {
    bool banded = false;
    // table keys: (summation position cc, output position c)
    std::vector<float> tau, aarg;
    // banded form: [n_o][nb]
    std::vector<int> pkey;
    std::vector<float> pw;
    int n_o = 0, n_i = 0, n_b = 0, nb = 0, aper = 0;
    size_t in_b = 0, in_i = 0, out_b = 0, out_o = 0;
    // pair lists per output trace
    std::vector<int> rowptr;
    std::vector<PairRec> pairs;
    long long npairs = 0;
    // butterfly panels (use_bf=1)
    bool bf = false;
    BfPlan bfp;
    // cmp=1 source side: the rows of every common-receiver gather ir, row is:
    // output trace and pair range ([ir * ns + is], -1 where the gather has no row)
    std::vector<int> g_o, g_q0, g_q1;
    int ngather = 0;
};

// Device copies of one operator's butterfly plan.
struct BfDev {
    void *nodes = nullptr, *items = nullptr, *tau = nullptr, *lag = nullptr, *sws = nullptr,
         *omega = nullptr, *grp_ptr = nullptr, *grp_row0 = nullptr, *grp_rows = nullptr,
         *grp_chunk = nullptr, *grp = nullptr, *quads = nullptr, *node_pair0 = nullptr,
         *node_out = nullptr;
    void **all() { return &nodes; }
    static constexpr int N = 14;
};

struct kmdat_ctx {
    kmdat_config cfg{};
    std::string err;
    hipStream_t stream = nullptr;
    Side side[2];
    // data geometry
    int nt = 0, nh = 0, ns = 0;
    float dt = 0, h0 = 0, dh = 0, s0 = 0, ds = 0;
    bool data_set = false, planned = false, ran = false;
    int nfft = 0, nw = 0, nsam = 0;
    OpPlan op[2];  // 0 receiver side, 1 source side
    // device buffers
    float *d_data = nullptr, *d_out = nullptr, *d_interm = nullptr, *d_wtr = nullptr;
    float *d_pad = nullptr;
    float2 *d_u1 = nullptr, *d_u2 = nullptr, *d_ktab = nullptr;
    size_t cap_data = 0, cap_out = 0, cap_interm = 0, cap_wtr = 0, cap_pad = 0, cap_u1 = 0,
           cap_u2 = 0, cap_ktab = 0;
    // per-side device plan arrays
    float *d_tau[2] = {nullptr, nullptr}, *d_aarg[2] = {nullptr, nullptr};
    int *d_pkey[2] = {nullptr, nullptr}, *d_rowptr[2] = {nullptr, nullptr};
    float *d_pw[2] = {nullptr, nullptr};
    PairRec *d_pairs[2] = {nullptr, nullptr};
    // rocFFT plans for chunks of `chunk` traces and for the tail
    int chunk = 0, plan_nfft = 0, plan_chunk = 0, plan_tail = 0;
    rocfft_plan pf = nullptr, pb = nullptr, pf_tail = nullptr, pb_tail = nullptr;
    rocfft_execution_info info = nullptr;
    void *work = nullptr;
    size_t work_cap = 0;
    hipEvent_t ev[12] = {};
    // butterfly plans on the device and the work buffers both sides share
    BfDev bfd[2];
    bool bf_ran[2] = {false, false};  // the butterfly kernels of a side ran (events recorded)
    unsigned long long *d_bf_mask = nullptr;
    float2 *d_bf_sw = nullptr, *d_bf_d = nullptr;
    size_t cap_bf_mask = 0, cap_bf_sw = 0, cap_bf_d = 0;
    kmdat_stats st{};
    bool mfma = false;  // KMDAT_MFMA=1: the matrix-core banded kernel
    int super_o = 16, super_b = 8;  // KMDAT_SUPER=o,b: tile grouping of the VALU kernel
};

static int fail(kmdat_ctx *c, int code, const char *fmt, ...)
{
    char buf[768];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, x)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            return fail(c, KMDAT_E_DEVICE, "%s: %s (%s:%d)", #x, hipGetErrorString(e_),     \
                        __FILE__, __LINE__);                                                 \
    } while (0)

#define FFTCHK(c, x)                                                                         \
    do {                                                                                     \
        rocfft_status s_ = (x);                                                              \
        if (s_ != rocfft_status_success)                                                     \
            return fail(c, KMDAT_E_DEVICE, "%s: rocfft status %d", #x, (int)s_);             \
    } while (0)

template <typename T>
static int ensure(kmdat_ctx *c, T **p, size_t *cap, size_t n)
{
    if (n <= *cap && *p) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (n == 0) return 0;
    HIPCHK(c, hipMalloc((void **)p, n * sizeof(T)));
    *cap = n;
    return 0;
}

// ------------------------------------------------------ reference helpers --

// taper_weight (:236-248)
static float taper_weight(int left, int center, int right, int tap, int stride)
{
    if (center < left || center > right) return 0.0f;
    if (tap <= 0) return 1.0f;
    if (stride < 1) stride = 1;
    const float width = (float)tap * (float)stride;
    const float ul = fminf(1.0f, fmaxf(0.0f, (float)(center - left) / width));
    const float ur = fminf(1.0f, fmaxf(0.0f, (float)(right - center) / width));
    return 0.5f * (1.0f - cosf((float)M_PI * ul)) * 0.5f * (1.0f - cosf((float)M_PI * ur));
}

// data_edge_taper_weight (:254-279)
static float edge_weight(int i, int n, int tap)
{
    if (tap <= 0 || n <= 1) return 1.0f;
    int ntap = tap;
    const int max_tap = (n - 1) / 2;
    if (ntap > max_tap) ntap = max_tap;
    if (ntap <= 0) return 1.0f;
    if (i < ntap) return 0.5f * (1.0f - cosf((float)M_PI * ((float)i / (float)ntap)));
    if (i > n - 1 - ntap)
        return 0.5f * (1.0f - cosf((float)M_PI * ((float)(n - 1 - i) / (float)ntap)));
    return 1.0f;
}

// table_slope_first_index (:192-201)
static float table_slope(const std::vector<float> &t, int n, int row, int col, float dcoord)
{
    if (n <= 1 || dcoord == 0.0f) return 0.0f;
    auto T = [&](int r, int cc) { return t[(size_t)r * n + cc]; };
    if (row <= 0) return (T(1, col) - T(0, col)) / dcoord;
    if (row >= n - 1) return (T(n - 1, col) - T(n - 2, col)) / dcoord;
    return 0.5f * (T(row + 1, col) - T(row - 1, col)) / dcoord;
}

// the pair weight coef / M_PI * step * datum * tau / dist, formed in double
// (M_PI is a double) and rounded once, as the reference does
static float pair_weight(float coef, float step, float datum, float tau, float dist)
{
    return (float)((double)coef / M_PI * (double)step * (double)datum * (double)tau /
                   (double)dist);
}

// Dense key index over (summation position cc, output position c) of one
// table: key attributes are the traveltime and the antialias argument.
struct KeyMap {
    std::vector<int> idx;
    int n = 0;
    void reset(int n_)
    {
        n = n_;
        idx.assign((size_t)n * n, -1);
    }
    int get(OpPlan &op, const Side &sd, int cc, int c, float astep)
    {
        int &k = idx[(size_t)cc * n + c];
        if (k < 0) {
            k = (int)op.tau.size();
            op.tau.push_back(sd.tab[(size_t)cc * n + c]);
            op.aarg.push_back(table_slope(sd.tab, n, cc, c, sd.d) * astep);
        }
        return k;
    }
};

// ------------------------------------------------------------- planning --

static int plan_receiver(kmdat_ctx *c)
{
    const kmdat_config &p = c->cfg;
    const Side &R = c->side[1];
    OpPlan &op = c->op[0];
    op = OpPlan{};
    KeyMap km;
    km.reset(R.n);
    const int nh = c->nh, ns = c->ns, aper = p.aperture;
    const float h0 = c->h0, dh = c->dh, s0 = c->s0, ds = c->ds, rg0 = R.o, drg = R.d;
    const float astep = fabsf(dh);
    auto pos = [&](int is, int ih) {
        const float base = p.cmp ? (s0 + is * ds) : 0.0f;
        return (int)(((base + h0 + ih * dh - rg0) / drg) + 0.5f);
    };
    if (!p.cmp) {
        // apply_receiver_operator_freq_iw (:493-546) with cmp=0: the pairs do
        // not depend on the shot
        op.banded = true;
        op.n_o = op.n_i = nh;
        op.n_b = ns;
        op.aper = std::min(aper, nh - 1);  // |ic - ih| never exceeds either
        op.nb = 2 * op.aper + 1;
        op.pkey.assign((size_t)op.n_o * op.nb, -1);
        op.pw.assign((size_t)op.n_o * op.nb, 0.0f);
        for (int ih = 0; ih < nh; ih++) {
            const int cc0 = pos(0, ih);
            const int left = (ih - aper < 0) ? 0 : ih - aper;
            const int right = (ih + aper > nh - 1) ? nh - 1 : ih + aper;
            for (int ic = left; ic <= right; ic++) {
                const int cc = pos(0, ic);
                if (cc0 < 0 || cc0 >= R.n || cc < 0 || cc >= R.n)
                    return fail(c, KMDAT_E_DIM, "Receiver table too small.");
                const float coef = taper_weight(left, ic, right, p.taper, 1);
                const float tau = R.tab[(size_t)cc * R.n + cc0];
                const float dist = p.rdatum * p.rdatum + (ic - ih) * (ic - ih) * dh * dh;
                const float w = pair_weight(coef, dh, p.rdatum, tau, dist);
                if (w == 0.0f) continue;
                const int j = ic - ih + op.aper;
                op.pkey[(size_t)ih * op.nb + j] = km.get(op, R, cc, cc0, astep);
                op.pw[(size_t)ih * op.nb + j] = w;
                op.npairs += ns;
            }
        }
        op.in_b = (size_t)nh;   // in/out [is][ih], strides in spectra (x nw below)
        op.in_i = 1;
        op.out_b = (size_t)nh;
        op.out_o = 1;
        return KMDAT_OK;
    }
    // cmp=1: the receivers move with the shot, so do the keys
    op.rowptr.assign((size_t)ns * nh + 1, 0);
    for (int is = 0; is < ns; is++) {
        for (int ih = 0; ih < nh; ih++) {
            const int o = is * nh + ih;
            const int cpos = pos(is, ih);
            const int left = (ih - aper < 0) ? 0 : ih - aper;
            const int right = (ih + aper > nh - 1) ? nh - 1 : ih + aper;
            for (int ic = left; ic <= right; ic++) {
                const int cc = pos(is, ic);
                if (cpos < 0 || cpos >= R.n || cc < 0 || cc >= R.n)
                    return fail(c, KMDAT_E_DIM, "Receiver table too small.");
                const float coef = taper_weight(left, ic, right, p.taper, 1);
                const float tau = R.tab[(size_t)cc * R.n + cpos];
                const float dist = p.rdatum * p.rdatum + (ic - ih) * (ic - ih) * dh * dh;
                const float w = pair_weight(coef, dh, p.rdatum, tau, dist);
                if (w == 0.0f) continue;
                PairRec q{};
                q.tr = is * nh + ic;
                q.key = km.get(op, R, cc, cpos, astep);
                q.w = w;
                op.pairs.push_back(q);
            }
            if (op.pairs.size() > (size_t)INT32_MAX)
                return fail(c, KMDAT_E_DIM, "more than 2^31 receiver-side summation pairs");
            op.rowptr[(size_t)o + 1] = (int)op.pairs.size();
        }
    }
    op.npairs = (long long)op.pairs.size();
    return KMDAT_OK;
}

static int plan_source(kmdat_ctx *c)
{
    const kmdat_config &p = c->cfg;
    const Side &S = c->side[0];
    OpPlan &op = c->op[1];
    op = OpPlan{};
    KeyMap km;
    km.reset(S.n);
    const int nh = c->nh, ns = c->ns, aper = p.aperture;
    const float dh = c->dh, s0 = c->s0, ds = c->ds, sg0 = S.o, dsg = S.d;
    auto spos = [&](int is) { return (int)((s0 + is * ds - sg0) / dsg + 0.5f); };
    if (!p.cmp) {
        // apply_source_operator_freq_iw (:639-667): for each receiver the
        // shots are summed with pairs that do not depend on the receiver
        op.banded = true;
        op.n_o = op.n_i = ns;
        op.n_b = nh;
        op.aper = std::min(aper, std::max(ns - 1, 0));
        op.nb = 2 * op.aper + 1;
        op.pkey.assign((size_t)op.n_o * op.nb, -1);
        op.pw.assign((size_t)op.n_o * op.nb, 0.0f);
        const float astep = fabsf(ds);
        for (int is = 0; is < ns; is++) {
            const int cs = spos(is);
            if (cs < 0 || cs >= S.n) return fail(c, KMDAT_E_DIM, "Source table too small.");
            const int left = (is - aper < 0) ? 0 : is - aper;
            const int right = (is + aper > ns - 1) ? ns - 1 : is + aper;
            for (int ic = left; ic <= right; ic++) {
                const int cc = spos(ic);
                if (cc < 0 || cc >= S.n) return fail(c, KMDAT_E_DIM, "Source table too small.");
                const float coef = taper_weight(left, ic, right, p.taper, 1);
                const float tau = S.tab[(size_t)cc * S.n + cs];
                const float dist = p.sdatum * p.sdatum + (ic - is) * (ic - is) * ds * ds;
                const float w = pair_weight(coef, ds, p.sdatum, tau, dist);
                if (w == 0.0f) continue;
                const int j = ic - is + op.aper;
                op.pkey[(size_t)is * op.nb + j] = km.get(op, S, cc, cs, astep);
                op.pw[(size_t)is * op.nb + j] = w;
                op.npairs += nh;
            }
        }
        op.in_b = 1;  // in/out [is][ih]: batch = receiver, output / summation = shot
        op.in_i = (size_t)nh;
        op.out_b = 1;
        op.out_o = (size_t)nh;
        return KMDAT_OK;
    }
    // cmp=1 (:566-626): common-receiver regrouping r = source + offset, with
    // a shot step `jump` when the shots are finer than the offsets
    const float s = fabsf((ns - 1) * ds), h = fabsf((nh - 1) * dh);
    const int jump = (fabsf(ds) >= fabsf(dh)) ? 1 : (int)(dh / ds + 0.5f);
    const float dr = (fabsf(ds) >= fabsf(dh)) ? fabsf(dh) : fabsf(ds);
    const int nr = (int)((s + h) / dr + 1.5f);
    const float astep = fabsf((float)jump * ds);
    std::vector<std::vector<PairRec>> rows((size_t)ns * nh);
    op.ngather = nr;
    op.g_o.assign((size_t)nr * ns, -1);
    op.g_q0.assign((size_t)nr * ns, 0);
    op.g_q1.assign((size_t)nr * ns, 0);
    for (int ir = 0; ir < nr; ir++) {
        const float r = ir * dr + ((ds <= 0.f) ? -1.f : 0.f) * s + ((dh <= 0.f) ? -1.f : 0.f) * h;
        int sleft = (int)((ir * dr + ((ds <= 0.f) ? -1.f : 0.f) * s +
                           ((ds <= 0.f) ? 0.f : -1.f) * h) / ds + 0.5f);
        int sright = (int)((ir * dr + ((ds <= 0.f) ? -1.f : 0.f) * s +
                            ((ds <= 0.f) ? -1.f : 0.f) * h) / ds + 0.5f);
        if (sleft < 0) sleft = 0;
        if (sright > ns - 1) sright = ns - 1;
        int left = (int)((r - sleft * ds) / dh + 0.5f);
        if (left < 0 || left > nh - 1) sleft++;
        int right = (int)((r - sright * ds) / dh + 0.5f);
        if (right < 0 || right > nh - 1) sright--;
        for (int is = sleft; is <= sright; is += jump) {
            const int cs = spos(is);
            if (cs < 0 || cs >= S.n) return fail(c, KMDAT_E_DIM, "Source table too small.");
            const int ih = (int)((r - is * ds) / dh + 0.5f);
            left = (is - jump * aper < sleft) ? sleft : is - jump * aper;
            right = (is + jump * aper > sright) ? sright : is + jump * aper;
            const size_t gi = (size_t)ir * ns + is;
            const bool row_ok = ih >= 0 && ih < nh;
            if (row_ok) {
                op.g_o[gi] = is * nh + ih;
                op.g_q0[gi] = (int)rows[(size_t)is * nh + ih].size();  // local, made absolute below
            }
            for (int ic = left; ic <= right; ic += jump) {
                const int cc = spos(ic);
                if (cc < 0 || cc >= S.n) return fail(c, KMDAT_E_DIM, "Source table too small.");
                const int hh = (int)((r - ic * ds) / dh + 0.5f);
                if (ih < 0 || ih >= nh || hh < 0 || hh >= nh)
                    return fail(c, KMDAT_E_DIM,
                                "cmp=1 regrouping of receiver gather %d reaches offset %d outside "
                                "the %d traces of the data (the reference reads outside its "
                                "arrays there)",
                                ir, ih < 0 || ih >= nh ? ih : hh, nh);
                const float coef = taper_weight(left, ic, right, p.taper, jump);
                const float tau = S.tab[(size_t)cc * S.n + cs];
                const float dist = p.sdatum * p.sdatum + (ic - is) * (ic - is) * ds * ds;
                const float w = pair_weight(coef, ds, p.sdatum, tau, dist);
                if (w == 0.0f) continue;
                PairRec q{};
                q.tr = ic * nh + hh;
                q.key = km.get(op, S, cc, cs, astep);
                q.w = w;
                rows[(size_t)is * nh + ih].push_back(q);
            }
            if (row_ok) op.g_q1[gi] = (int)rows[(size_t)is * nh + ih].size();
        }
    }
    op.rowptr.assign((size_t)ns * nh + 1, 0);
    size_t tot = 0;
    for (size_t o = 0; o < rows.size(); o++) {
        tot += rows[o].size();
        if (tot > (size_t)INT32_MAX)
            return fail(c, KMDAT_E_DIM, "more than 2^31 source-side summation pairs");
        op.rowptr[o + 1] = (int)tot;
    }
    op.pairs.reserve(tot);
    for (auto &rw : rows) op.pairs.insert(op.pairs.end(), rw.begin(), rw.end());
    op.npairs = (long long)tot;
    for (size_t gi = 0; gi < op.g_o.size(); gi++)
        if (op.g_o[gi] >= 0) {
            op.g_q0[gi] += op.rowptr[(size_t)op.g_o[gi]];
            op.g_q1[gi] += op.rowptr[(size_t)op.g_o[gi]];
        }
    return KMDAT_OK;
}

static void destroy_plans(kmdat_ctx *c)
{
    if (c->pf) rocfft_plan_destroy(c->pf);
    if (c->pb) rocfft_plan_destroy(c->pb);
    if (c->pf_tail) rocfft_plan_destroy(c->pf_tail);
    if (c->pb_tail) rocfft_plan_destroy(c->pb_tail);
    c->pf = c->pb = c->pf_tail = c->pb_tail = nullptr;
    c->plan_nfft = c->plan_chunk = c->plan_tail = 0;
}

static int make_fft_plans(kmdat_ctx *c, int ntr)
{
    const int chunk = std::min(ntr, c->chunk);
    const int tail = ntr % chunk;
    if (c->pf && c->plan_nfft == c->nfft && c->plan_chunk == chunk && c->plan_tail == tail)
        return KMDAT_OK;
    destroy_plans(c);
    const size_t len = (size_t)c->nfft;
    FFTCHK(c, rocfft_plan_create(&c->pf, rocfft_placement_notinplace,
                                 rocfft_transform_type_real_forward, rocfft_precision_single, 1,
                                 &len, (size_t)chunk, nullptr));
    FFTCHK(c, rocfft_plan_create(&c->pb, rocfft_placement_notinplace,
                                 rocfft_transform_type_real_inverse, rocfft_precision_single, 1,
                                 &len, (size_t)chunk, nullptr));
    if (tail) {
        FFTCHK(c, rocfft_plan_create(&c->pf_tail, rocfft_placement_notinplace,
                                     rocfft_transform_type_real_forward, rocfft_precision_single,
                                     1, &len, (size_t)tail, nullptr));
        FFTCHK(c, rocfft_plan_create(&c->pb_tail, rocfft_placement_notinplace,
                                     rocfft_transform_type_real_inverse, rocfft_precision_single,
                                     1, &len, (size_t)tail, nullptr));
    }
    size_t wmax = 0, w = 0;
    for (rocfft_plan pl : {c->pf, c->pb, c->pf_tail, c->pb_tail}) {
        if (!pl) continue;
        FFTCHK(c, rocfft_plan_get_work_buffer_size(pl, &w));
        wmax = std::max(wmax, w);
    }
    if (wmax > c->work_cap) {
        if (c->work) (void)hipFree(c->work);
        c->work = nullptr;
        c->work_cap = 0;
        HIPCHK(c, hipMalloc(&c->work, wmax));
        c->work_cap = wmax;
    }
    c->plan_nfft = c->nfft;
    c->plan_chunk = chunk;
    c->plan_tail = tail;
    return KMDAT_OK;
}

static int upload_op(kmdat_ctx *c, int k)
{
    OpPlan &op = c->op[k];
    const size_t nk = op.tau.size();
    size_t cap = 0;
    if (c->d_tau[k]) (void)hipFree(c->d_tau[k]);
    if (c->d_aarg[k]) (void)hipFree(c->d_aarg[k]);
    if (c->d_pkey[k]) (void)hipFree(c->d_pkey[k]);
    if (c->d_pw[k]) (void)hipFree(c->d_pw[k]);
    if (c->d_rowptr[k]) (void)hipFree(c->d_rowptr[k]);
    if (c->d_pairs[k]) (void)hipFree(c->d_pairs[k]);
    c->d_tau[k] = c->d_aarg[k] = c->d_pw[k] = nullptr;
    c->d_pkey[k] = c->d_rowptr[k] = nullptr;
    c->d_pairs[k] = nullptr;
    int rc;
    if ((rc = ensure(c, &c->d_tau[k], &cap, std::max<size_t>(nk, 1)))) return rc;
    cap = 0;
    if ((rc = ensure(c, &c->d_aarg[k], &cap, std::max<size_t>(nk, 1)))) return rc;
    if (nk) {
        HIPCHK(c, hipMemcpy(c->d_tau[k], op.tau.data(), nk * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_aarg[k], op.aarg.data(), nk * 4, hipMemcpyHostToDevice));
    }
    if (op.banded) {
        const size_t n = op.pkey.size();
        cap = 0;
        if ((rc = ensure(c, &c->d_pkey[k], &cap, n))) return rc;
        cap = 0;
        if ((rc = ensure(c, &c->d_pw[k], &cap, n))) return rc;
        HIPCHK(c, hipMemcpy(c->d_pkey[k], op.pkey.data(), n * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_pw[k], op.pw.data(), n * 4, hipMemcpyHostToDevice));
    } else {
        cap = 0;
        if ((rc = ensure(c, &c->d_rowptr[k], &cap, op.rowptr.size()))) return rc;
        HIPCHK(c, hipMemcpy(c->d_rowptr[k], op.rowptr.data(), op.rowptr.size() * 4,
                            hipMemcpyHostToDevice));
        cap = 0;
        if ((rc = ensure(c, &c->d_pairs[k], &cap, std::max<size_t>(op.pairs.size(), 1))))
            return rc;
        if (!op.pairs.empty())
            HIPCHK(c, hipMemcpy(c->d_pairs[k], op.pairs.data(), op.pairs.size() * sizeof(PairRec),
                                hipMemcpyHostToDevice));
    }
    return KMDAT_OK;
}

template <typename T>
static int upload_vec(kmdat_ctx *c, void **dst, const std::vector<T> &v)
{
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    const size_t n = std::max<size_t>(v.size(), 1);
    HIPCHK(c, hipMalloc(dst, n * sizeof(T)));
    if (!v.empty())
        HIPCHK(c, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return KMDAT_OK;
}

static void free_bf(BfDev &d)
{
    void **p = d.all();
    for (int i = 0; i < BfDev::N; i++) {
        if (p[i]) (void)hipFree(p[i]);
        p[i] = nullptr;
    }
}

static int upload_bf(kmdat_ctx *c, int k)
{
    const BfPlan &b = c->op[k].bfp;
    BfDev &d = c->bfd[k];
    int rc;
    if ((rc = upload_vec(c, &d.nodes, b.nodes)) || (rc = upload_vec(c, &d.items, b.items)) ||
        (rc = upload_vec(c, &d.tau, b.tau)) || (rc = upload_vec(c, &d.lag, b.lag)) ||
        (rc = upload_vec(c, &d.sws, b.sws)) || (rc = upload_vec(c, &d.omega, b.omega)) ||
        (rc = upload_vec(c, &d.grp_ptr, b.grp_ptr)) ||
        (rc = upload_vec(c, &d.grp_row0, b.grp_row0)) ||
        (rc = upload_vec(c, &d.grp_rows, b.grp_rows)) ||
        (rc = upload_vec(c, &d.grp_chunk, b.grp_chunk)) || (rc = upload_vec(c, &d.grp, b.grp)) ||
        (rc = upload_vec(c, &d.quads, b.quads)) ||
        (rc = upload_vec(c, &d.node_pair0, b.node_pair0)) ||
        (rc = upload_vec(c, &d.node_out, b.node_out)))
        return rc;
    if ((rc = ensure(c, &c->d_bf_mask, &c->cap_bf_mask, std::max<size_t>(b.items.size(), 1))))
        return rc;
    if ((rc = ensure(c, &c->d_bf_sw, &c->cap_bf_sw, std::max<size_t>((size_t)b.sw_elems, 1))))
        return rc;
    if ((rc = ensure(c, &c->d_bf_d, &c->cap_bf_d, std::max<size_t>((size_t)b.d_elems, 1))))
        return rc;
    return KMDAT_OK;
}

// choose_nfft (:87-98) over both tables
static int choose_nfft(const kmdat_ctx *c)
{
    float max_tau = 0.0f;
    for (const Side &s : c->side)
        for (float v : s.tab)
            if (v > max_tau) max_tau = v;
    const int pmax = (int)ceilf(max_tau / c->dt) + (c->nsam - 2);
    int p = 1;
    while (p < c->nt + pmax + 2) p <<= 1;
    return p;
}

static int plan(kmdat_ctx *c)
{
    if (c->planned) return KMDAT_OK;
    // buffers of an earlier run may be replaced below
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const kmdat_config &p = c->cfg;
    c->nsam = (int)(p.length / c->dt) + 2;
    if (c->nsam - 1 > KMDAT_MAX_NSAM)
        return fail(c, KMDAT_E_PARAM, "length=%g gives a filter of %d samples (at most %d)",
                    (double)p.length, c->nsam, KMDAT_MAX_NSAM + 1);
    c->nfft = choose_nfft(c);
    c->nw = c->nfft / 2 + 1;
    int rc;
    if ((rc = plan_receiver(c))) return rc;
    if ((rc = plan_source(c))) return rc;
    for (int k = 0; k < 2; k++)
        if ((rc = upload_op(c, k))) return rc;
    // use_bf=1: the butterfly panels of both operators: one matrix per side
    // (cmp=0, banded), one per shot (cmp=1 receiver side) or per receiver
    // gather (cmp=1 source side)
    std::vector<int> ro, rq0, rq1;
    for (int k = 0; k < 2; k++) {
        OpPlan &op = c->op[k];
        op.bf = p.use_bf != 0;
        op.bfp = BfPlan{};
        if (!op.bf) continue;
        BfPlanIn in{};
        in.tau_key = op.tau.data();
        in.aarg = op.aarg.data();
        in.nkey = (int)op.tau.size();
        in.nw = c->nw;
        in.nfft = c->nfft;
        in.dt = c->dt;
        in.antialias = p.antialias;
        in.p = p.bf_p;
        in.leaf = p.bf_n_leaf;
        in.levels = p.bf_panel_levels;
        // bf1d.cpp:949: a non-positive or non-finite tolerance reads as 2
        in.phase_tol = (p.bf_phase_tol <= 0.0f || !std::isfinite(p.bf_phase_tol))
                           ? 2.0f
                           : p.bf_phase_tol;
        if (op.banded) {
            in.n = op.n_o;
            in.nb = op.nb;
            in.aper = op.aper;
            in.pkey = op.pkey.data();
            in.pw = op.pw.data();
        } else if (k == 0) {
            // shot is: row ih = output trace is nh + ih, column ic = tr % nh
            const int nh = c->nh, ns = c->ns;
            ro.assign((size_t)ns * nh, 0);
            rq0.assign(ro.size(), 0);
            rq1.assign(ro.size(), 0);
            for (size_t o = 0; o < ro.size(); o++) {
                ro[o] = (int)o;
                rq0[o] = op.rowptr[o];
                rq1[o] = op.rowptr[o + 1];
            }
            in.n = nh;
            in.ninst = ns;
            in.pairs = op.pairs.data();
            in.row_o = ro.data();
            in.row_q0 = rq0.data();
            in.row_q1 = rq1.data();
            in.col_base = nh;
            in.col_by_div = 0;
        } else {
            // receiver gather ir: row is, column ic = tr / nh
            in.n = c->ns;
            in.ninst = op.ngather;
            in.pairs = op.pairs.data();
            in.row_o = op.g_o.data();
            in.row_q0 = op.g_q0.data();
            in.row_q1 = op.g_q1.data();
            in.col_base = c->nh;
            in.col_by_div = 1;
        }
        std::string e;
        if ((rc = bf_plan(in, op.bfp, e))) return fail(c, rc, "%s", e.c_str());
        if ((rc = upload_bf(c, k))) return rc;
    }
    // per-trace data-edge taper weight ws * wh (:290-295)
    const size_t ntr = (size_t)c->ns * c->nh;
    std::vector<float> wtr(ntr);
    for (int is = 0; is < c->ns; is++) {
        const float ws = edge_weight(is, c->ns, p.data_taper);
        for (int ih = 0; ih < c->nh; ih++)
            wtr[(size_t)is * c->nh + ih] = ws * edge_weight(ih, c->nh, p.data_taper);
    }
    if ((rc = ensure(c, &c->d_wtr, &c->cap_wtr, ntr))) return rc;
    HIPCHK(c, hipMemcpy(c->d_wtr, wtr.data(), ntr * 4, hipMemcpyHostToDevice));
    // spectra, tables, transform scratch: the chunk keeps the padded real
    // traces near 1 GiB
    const size_t nu = ntr * (size_t)c->nw;
    if ((rc = ensure(c, &c->d_u1, &c->cap_u1, nu))) return rc;
    if ((rc = ensure(c, &c->d_u2, &c->cap_u2, nu))) return rc;
    const size_t nk = std::max(c->op[0].tau.size(), c->op[1].tau.size());
    if ((rc = ensure(c, &c->d_ktab, &c->cap_ktab, std::max<size_t>(nk, 1) * c->nw))) return rc;
    c->chunk = (int)std::max<size_t>(1, std::min<size_t>(ntr, (size_t)(1u << 28) / c->nfft));
    if ((rc = ensure(c, &c->d_pad, &c->cap_pad, (size_t)c->chunk * c->nfft))) return rc;
    if ((rc = ensure(c, &c->d_out, &c->cap_out, ntr * (size_t)c->nt))) return rc;
    if (p.interm && (rc = ensure(c, &c->d_interm, &c->cap_interm, ntr * (size_t)c->nt)))
        return rc;
    if ((rc = make_fft_plans(c, (int)ntr))) return rc;
    c->st = kmdat_stats{};
    c->st.nfft = c->nfft;
    c->st.nw = c->nw;
    c->st.nsam = c->nsam;
    c->st.keys_receiver = (long long)c->op[0].tau.size();
    c->st.keys_source = (long long)c->op[1].tau.size();
    c->st.pairs_receiver = c->op[0].npairs;
    c->st.pairs_source = c->op[1].npairs;
    c->st.cmacs = (double)(c->op[0].npairs + c->op[1].npairs) * (double)c->nw;
    c->st.banded = c->op[0].banded ? 1 : 0;
    c->st.butterfly = (c->op[0].bf || c->op[1].bf) ? 1 : 0;
    c->st.bf_panels_receiver = (long long)c->op[0].bfp.nodes.size();
    c->st.bf_panels_source = (long long)c->op[1].bfp.nodes.size();
    c->st.bf_items_receiver = (long long)c->op[0].bfp.items.size();
    c->st.bf_items_source = (long long)c->op[1].bfp.items.size();
    c->planned = true;
    return KMDAT_OK;
}

// ------------------------------------------------------------------ API --

extern "C" {

void kmdat_config_default(kmdat_config *cfg)
{
    if (!cfg) return;
    memset(cfg, 0, sizeof *cfg);
    cfg->cmp = 1;
    cfg->aperture = 50;
    cfg->taper = 10;
    cfg->data_taper = 0;
    cfg->antialias = 1.0f;
    cfg->length = 0.025f;
    cfg->use_bf = 1;
    cfg->bf_p = 12;
    cfg->bf_n_leaf = 16;
    cfg->bf_panel_levels = 1;
    cfg->bf_amp_eps = 1e-20f;
    cfg->bf_phase_tol = 2.0f;
}

int kmdat_config_check(const kmdat_config *p, char *err, size_t n)
{
    auto bad = [&](const char *m) {
        if (err && n) snprintf(err, n, "%s", m);
        return KMDAT_E_PARAM;
    };
    if (!p) return bad("no configuration");
    if (p->cmp != 0 && p->cmp != 1) return bad("cmp must be 0 or 1.");
    if (p->bf_p <= 0 || p->bf_n_leaf <= 1) return bad("Need bf_p > 0 and bf_n_leaf > 1.");
    if (p->bf_p >= p->bf_n_leaf)
        return bad("Need bf_p < bf_n_leaf for meaningful butterfly interpolation / exact "
                   "leaf-stage settings.");
    if (p->data_taper < 0) return bad("data_taper must be >= 0.");
    if (p->antialias < 0.0f) return bad("antialias must be >= 0.");
    if (!(p->length > 0.0f)) return bad("length must be > 0.");
    if (p->aperture < 0) return bad("aperture must be >= 0.");
    return KMDAT_OK;
}

int kmdat_create(kmdat_ctx **out, const kmdat_config *cfg)
{
    if (!out || !cfg) return KMDAT_E_PARAM;
    *out = nullptr;
    char err[256];
    int rc = kmdat_config_check(cfg, err, sizeof err);
    kmdat_ctx *c = new kmdat_ctx();
    c->cfg = *cfg;
    *out = c;
    if (rc) return fail(c, rc, "%s", err);
    static std::once_flag once;
    std::call_once(once, [] { rocfft_setup(); });
    if (const char *v = getenv("KMDAT_MFMA")) c->mfma = atoi(v) != 0;
    if (const char *v = getenv("KMDAT_SUPER")) {
        int so = 1, sb = 1;
        if (sscanf(v, "%d,%d", &so, &sb) == 2 && so >= 1 && sb >= 1) {
            c->super_o = so;
            c->super_b = sb;
        }
    }
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) HIPCHK(c, hipEventCreate(&e));
    return KMDAT_OK;
}

void kmdat_destroy(kmdat_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    destroy_plans(c);
    if (c->info) rocfft_execution_info_destroy(c->info);
    for (void *p : {(void *)c->d_data, (void *)c->d_out, (void *)c->d_interm, (void *)c->d_wtr,
                    (void *)c->d_pad, (void *)c->d_u1, (void *)c->d_u2, (void *)c->d_ktab,
                    c->work})
        if (p) (void)hipFree(p);
    for (int k = 0; k < 2; k++) {
        for (void *p : {(void *)c->d_tau[k], (void *)c->d_aarg[k], (void *)c->d_pkey[k],
                        (void *)c->d_pw[k], (void *)c->d_rowptr[k], (void *)c->d_pairs[k]})
            if (p) (void)hipFree(p);
        free_bf(c->bfd[k]);
    }
    for (void *p : {(void *)c->d_bf_mask, (void *)c->d_bf_sw, (void *)c->d_bf_d})
        if (p) (void)hipFree(p);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *kmdat_last_error(const kmdat_ctx *c) { return c ? c->err.c_str() : "no context"; }

int kmdat_set_green(kmdat_ctx *c, int side, const float *table, int n, float o, float d)
{
    if (!c) return KMDAT_E_PARAM;
    if (side != 0 && side != 1) return fail(c, KMDAT_E_PARAM, "side must be 0 (source) or 1");
    if (!table || n <= 0) return fail(c, KMDAT_E_DIM, "empty datum table");
    if (!(d != 0.0f) || !std::isfinite(d) || !std::isfinite(o))
        return fail(c, KMDAT_E_PARAM, "datum table spacing must be finite and nonzero");
    Side &s = c->side[side];
    s.tab.assign(table, table + (size_t)n * n);
    s.n = n;
    s.o = o;
    s.d = d;
    s.set = true;
    c->planned = false;
    return KMDAT_OK;
}

int kmdat_set_data(kmdat_ctx *c, const float *data, int nt, float dt, int nh, float h0, float dh,
                   int ns, float s0, float ds)
{
    if (!c || !data) return KMDAT_E_PARAM;
    if (nt < 1 || nh < 1 || ns < 1) return fail(c, KMDAT_E_DIM, "Input must be 3D.");
    if ((long long)ns * nh > INT32_MAX)
        return fail(c, KMDAT_E_DIM, "more than 2^31 traces in the cube");
    if (!(dt > 0.0f)) return fail(c, KMDAT_E_PARAM, "time sampling must be > 0");
    if (c->cfg.cmp == 1 && (!(dh != 0.0f) || !(ds != 0.0f)))
        return fail(c, KMDAT_E_PARAM, "cmp=1 needs nonzero shot and offset spacings");
    if (c->nt != nt || c->nh != nh || c->ns != ns || c->dt != dt || c->h0 != h0 ||
        c->dh != dh || c->s0 != s0 || c->ds != ds)
        c->planned = false;
    c->nt = nt;
    c->dt = dt;
    c->nh = nh;
    c->h0 = h0;
    c->dh = dh;
    c->ns = ns;
    c->s0 = s0;
    c->ds = ds;
    const size_t n = (size_t)nt * nh * ns;
    int rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));  // an earlier run may still read d_data
    if ((rc = ensure(c, &c->d_data, &c->cap_data, n))) return rc;
    HIPCHK(c, hipMemcpy(c->d_data, data, n * sizeof(float), hipMemcpyHostToDevice));
    c->data_set = true;
    c->ran = false;
    return KMDAT_OK;
}

static int transforms(kmdat_ctx *c, bool forward, float2 *U, float *dst, int reverse)
{
    const int ntr = c->ns * c->nh;
    if (!c->info) FFTCHK(c, rocfft_execution_info_create(&c->info));
    FFTCHK(c, rocfft_execution_info_set_stream(c->info, c->stream));
    if (c->work_cap)
        FFTCHK(c, rocfft_execution_info_set_work_buffer(c->info, c->work, c->work_cap));
    for (int t0 = 0; t0 < ntr; t0 += c->plan_chunk) {
        const int nb = std::min(c->plan_chunk, ntr - t0);
        rocfft_plan pl = forward ? (nb == c->plan_chunk ? c->pf : c->pf_tail)
                                 : (nb == c->plan_chunk ? c->pb : c->pb_tail);
        float2 *u = U + (size_t)t0 * c->nw;
        if (forward) {
            HIPCHK(c, (hipError_t)launch_pad(c->d_data, c->d_wtr, c->d_pad, t0, nb, c->nt,
                                              c->nfft, c->stream));
            void *in[1] = {c->d_pad}, *out[1] = {u};
            FFTCHK(c, rocfft_execute(pl, in, out, c->info));
        } else {
            void *in[1] = {u}, *out[1] = {c->d_pad};
            FFTCHK(c, rocfft_execute(pl, in, out, c->info));
            HIPCHK(c, (hipError_t)launch_post(c->d_pad, dst, t0, nb, c->nt, c->nfft, reverse,
                                               c->stream));
        }
    }
    return KMDAT_OK;
}

// The butterfly kernels' arguments for operator k (the batch strides of the
// banded form are the caller's).
static bf::Args bf_args(kmdat_ctx *c, int k, const float2 *in, float2 *out)
{
    const OpPlan &op = c->op[k];
    const BfPlan &bp = op.bfp;
    const BfDev &d = c->bfd[k];
    bf::Args b{};
    b.pkey = c->d_pkey[k];
    b.pw = c->d_pw[k];
    b.tau_key = c->d_tau[k];
    b.aarg = c->d_aarg[k];
    b.ktab = c->d_ktab;
    b.omega = (const float *)d.omega;
    b.nb = op.nb;
    b.aper = op.aper;
    b.nw = c->nw;
    b.nfft = c->nfft;
    b.dt = c->dt;
    b.antialias = c->cfg.antialias;
    const float eps = c->cfg.bf_amp_eps < 0.0f ? 0.0f : c->cfg.bf_amp_eps;
    b.eps2 = eps * eps;
    b.nodes = (const bf::Node *)d.nodes;
    b.items = (const bf::Item *)d.items;
    b.nitems = (int)bp.items.size();
    b.tau = (const double *)d.tau;
    b.lag = (const float *)d.lag;
    b.sws = (const bf::SwSample *)d.sws;
    for (int L = 0; L <= bf::MAX_L; L++) {
        b.lag_off[L] = bp.lag_off[L];
        b.sws_off[L] = bp.sws_off[L];
    }
    b.p = c->cfg.bf_p;
    b.leaf = c->cfg.bf_n_leaf;
    b.maxL = bp.maxL;
    b.mask = c->d_bf_mask;
    b.sw = c->d_bf_sw;
    b.d = c->d_bf_d;
    b.grp_ptr = (const int *)d.grp_ptr;
    b.grp_row0 = (const int *)d.grp_row0;
    b.grp_rows = (const int *)d.grp_rows;
    b.grp_chunk = (const int *)d.grp_chunk;
    b.grp = (const int *)d.grp;
    b.quads = (const bf::Quad *)d.quads;
    b.nquads = (int)bp.quads.size();
    b.ngroup = (int)bp.grp_row0.size();
    b.node_pair0 = (const int *)d.node_pair0;
    b.node_out = (const int *)d.node_out;
    b.in = in;
    b.out = out;
    return b;
}

static int apply_op(kmdat_ctx *c, int k, const float2 *in, float2 *out)
{
    OpPlan &op = c->op[k];
    const int nw = c->nw;
    if (op.banded) {
        BandArgs a{};
        a.in = in;
        a.out = out;
        a.ktab = c->d_ktab;
        a.pkey = c->d_pkey[k];
        a.pw = c->d_pw[k];
        a.n_o = op.n_o;
        a.n_i = op.n_i;
        a.n_b = op.n_b;
        a.nb = op.nb;
        a.aper = op.aper;
        a.nw = nw;
        a.in_b = op.in_b * nw;
        a.in_i = op.in_i * nw;
        a.out_b = op.out_b * nw;
        a.out_o = op.out_o * nw;
        a.super_o = c->super_o;
        a.super_b = c->super_b;
        HIPCHK(c, (hipError_t)(c->mfma ? launch_band_mfma(a, c->stream)
                                        : launch_band(a, c->stream)));
        if (op.bf && !op.bfp.items.empty()) {
            bf::Args b = bf_args(c, k, in, out);
            b.n_b = op.n_b;
            b.in_b = a.in_b;
            b.in_i = a.in_i;
            b.out_b = a.out_b;
            b.out_o = a.out_o;
            HIPCHK(c, hipEventRecord(c->ev[7 + 2 * k], c->stream));
            HIPCHK(c, (hipError_t)launch_bf(b, c->stream));
            HIPCHK(c, hipEventRecord(c->ev[8 + 2 * k], c->stream));
            c->bf_ran[k] = true;
        }
    } else {
        PairArgs a{};
        a.in = in;
        a.out = out;
        a.ktab = c->d_ktab;
        a.rowptr = c->d_rowptr[k];
        a.pairs = c->d_pairs[k];
        a.nw = nw;
        HIPCHK(c, (hipError_t)launch_pairs(a, c->ns * c->nh, c->stream));
        if (op.bf && !op.bfp.items.empty()) {
            bf::Args b = bf_args(c, k, in, out);
            b.pairs = c->d_pairs[k];
            HIPCHK(c, hipEventRecord(c->ev[7 + 2 * k], c->stream));
            HIPCHK(c, (hipError_t)launch_bf_pairs(b, c->stream));
            HIPCHK(c, hipEventRecord(c->ev[8 + 2 * k], c->stream));
            c->bf_ran[k] = true;
        }
    }
    return KMDAT_OK;
}

int kmdat_run(kmdat_ctx *c)
{
    if (!c) return KMDAT_E_PARAM;
    if (!c->data_set) return fail(c, KMDAT_E_STATE, "kmdat_set_data must come first");
    if (!c->side[0].set || !c->side[1].set)
        return fail(c, KMDAT_E_STATE, "both datum tables must be set before kmdat_run");
    int rc;
    if ((rc = plan(c))) return rc;
    const kmdat_config &p = c->cfg;
    hipEvent_t *e = c->ev;
    c->bf_ran[0] = c->bf_ran[1] = false;
    HIPCHK(c, hipEventRecord(e[0], c->stream));
    // forward transforms into u1
    if ((rc = transforms(c, true, c->d_u1, nullptr, 0))) return rc;
    HIPCHK(c, hipEventRecord(e[1], c->stream));
    // receiver side: u1 -> u2
    HIPCHK(c, (hipError_t)launch_ktab(c->d_ktab, c->d_tau[0], c->d_aarg[0],
                                       (int)c->op[0].tau.size(), c->nw, c->nfft, c->dt, c->nsam,
                                       p.antialias, c->stream));
    HIPCHK(c, hipEventRecord(e[2], c->stream));
    if ((rc = apply_op(c, 0, c->d_u1, c->d_u2))) return rc;
    HIPCHK(c, hipEventRecord(e[3], c->stream));
    // source side: u2 -> u1 (the input spectra are no longer needed)
    HIPCHK(c, (hipError_t)launch_ktab(c->d_ktab, c->d_tau[1], c->d_aarg[1],
                                       (int)c->op[1].tau.size(), c->nw, c->nfft, c->dt, c->nsam,
                                       p.antialias, c->stream));
    HIPCHK(c, hipEventRecord(e[4], c->stream));
    if ((rc = apply_op(c, 1, c->d_u2, c->d_u1))) return rc;
    HIPCHK(c, hipEventRecord(e[5], c->stream));
    // inverse transforms (c2r may overwrite its input: u2 is only read
    // after the source side)
    if ((rc = transforms(c, false, c->d_u1, c->d_out, 1))) return rc;
    if (p.interm && (rc = transforms(c, false, c->d_u2, c->d_interm, 0))) return rc;
    HIPCHK(c, hipEventRecord(e[6], c->stream));
    c->ran = true;
    return KMDAT_OK;
}

int kmdat_sync(kmdat_ctx *c)
{
    if (!c) return KMDAT_E_PARAM;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->ran) {
        float t[6];
        for (int k = 0; k < 6; k++) HIPCHK(c, hipEventElapsedTime(&t[k], c->ev[k], c->ev[k + 1]));
        c->st.ms_forward = t[0];
        c->st.ms_tables = t[1] + t[3];
        c->st.ms_receiver = t[2];
        c->st.ms_source = t[4];
        c->st.ms_inverse = t[5];
        HIPCHK(c, hipEventElapsedTime(&c->st.ms_total, c->ev[0], c->ev[6]));
        c->st.ms_bf_receiver = c->st.ms_bf_source = 0.0f;
        if (c->bf_ran[0])
            HIPCHK(c, hipEventElapsedTime(&c->st.ms_bf_receiver, c->ev[7], c->ev[8]));
        if (c->bf_ran[1])
            HIPCHK(c, hipEventElapsedTime(&c->st.ms_bf_source, c->ev[9], c->ev[10]));
    }
    return KMDAT_OK;
}

int kmdat_get_output(kmdat_ctx *c, float *out)
{
    if (!c || !out) return KMDAT_E_PARAM;
    if (!c->ran) return fail(c, KMDAT_E_STATE, "no datuming run yet");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->d_out, (size_t)c->ns * c->nh * c->nt * sizeof(float),
                        hipMemcpyDeviceToHost));
    return KMDAT_OK;
}

int kmdat_get_interm(kmdat_ctx *c, float *out)
{
    if (!c || !out) return KMDAT_E_PARAM;
    if (!c->cfg.interm) return fail(c, KMDAT_E_STATE, "the context was created without interm");
    if (!c->ran) return fail(c, KMDAT_E_STATE, "no datuming run yet");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->d_interm, (size_t)c->ns * c->nh * c->nt * sizeof(float),
                        hipMemcpyDeviceToHost));
    return KMDAT_OK;
}

int kmdat_get_stats(const kmdat_ctx *c, kmdat_stats *st)
{
    if (!c || !st) return KMDAT_E_PARAM;
    *st = c->st;
    return KMDAT_OK;
}

int kmdat_datum(kmdat_ctx *c, const float *data, int nt, float dt, int nh, float h0, float dh,
                int ns, float s0, float ds, float *out, float *interm)
{
    int rc = kmdat_set_data(c, data, nt, dt, nh, h0, dh, ns, s0, ds);
    if (rc) return rc;
    if ((rc = kmdat_run(c))) return rc;
    if ((rc = kmdat_sync(c))) return rc;
    if ((rc = kmdat_get_output(c, out))) return rc;
    if (interm && (rc = kmdat_get_interm(c, interm))) return rc;
    return KMDAT_OK;
}

int kmdat_greensol(const float *ttable, int nz, int nx, int ngx, float oz, float dz, float ox,
                   float dx, float ogx, float dgx, float datum, float *tgreen)
{
    if (!ttable || !tgreen || nz < 1 || nx < 1 || ngx < 1) return KMDAT_E_PARAM;
    // src/greensol_auto.cpp:70-95: the datum row of every table position at
    // the x sample nearest each position
    const int idatum = (int)((datum - oz) / dz + 0.5);
    if (idatum < 0 || idatum >= nz) return KMDAT_E_PARAM;
    for (int i1 = 0; i1 < ngx; i1++) {
        for (int i2 = 0; i2 < ngx; i2++) {
            const float gx1 = ogx + i1 * dgx, gx2 = ogx + i2 * dgx;
            const int ix1 = (int)((gx1 - ox) / dx + 0.5);
            const int ix2 = (int)((gx2 - ox) / dx + 0.5);
            if (ix1 < 0 || ix1 >= nx || ix2 < 0 || ix2 >= nx) return KMDAT_E_DIM;
            tgreen[(size_t)i1 * ngx + i2] = ttable[((size_t)i1 * nx + ix2) * nz + idatum];
        }
    }
    return KMDAT_OK;
}

}  // extern "C"
