// C ABI of libkmig: context, HBM residency, per-run geometry, launches.
//
// The per-shot / per-trace geometry of the reference (source and receiver
// positions, table node choice, Hermite coordinate, trace aperture, offset
// bin) is evaluated here on the host with the reference's float arithmetic
// (src/kirchmig2d_cig.cpp:979-1105, 1148-1165) and handed to the kernels as
// SlotSpec / TraceMeta records.  Everything per image point runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "kmig.h"
#include "kmig_internal.h"
#include "kmig_nn_internal.h"

using namespace kmig;

namespace {

struct Side {
    float *T = nullptr, *D = nullptr, *GX = nullptr, *GZ = nullptr;
    int npos = 0;
    float o = 0.0f, d = 0.0f;
    bool set = false;
    size_t cap = 0;  // elements allocated per volume
    // T is finite and never -0.0, D finite (k_table_check): Hermite
    // interpolation at coordinate 0 then gives the node's values exactly
    bool plain = false;
    // learned surrogate of the table interpolation (kmig_set_surrogate)
    kmig_nn *nn = nullptr;
    float nn_max_dt = 0.0f;
    int nn_fallback = KMIG_NN_FALLBACK_SLOT;
    int *nn_used = nullptr;          // [slot] points taken from the network
    unsigned *nn_maxres = nullptr;   // [slot] max |T_nn - T_table| (float bits)
    size_t nn_cap = 0, nn_cap_m = 0; // slots allocated
    size_t nn_slots = 0;             // slots of the last plan
};

struct Batch {
    int s0, s1;       // shot range
    int nrcv;         // receiver slots
    size_t spec_r0;   // offset of this batch's receiver specs
};

}  // namespace

struct kmig_ctx {
    kmig_params p;
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // trace preparation, overlapped with the tuple preparation
    int nz = 0, nx = 0;
    float z0 = 0, dz = 0, x0 = 0, dx = 0;
    size_t nzx = 0;
    Side side[2];
    int nt = 0, nh = 0, ns = 0;
    float t0 = 0, dt = 0, h0 = 0, dh = 0, s0 = 0, ds = 0;
    bool data_set = false;
    float *raw = nullptr, *prep = nullptr;
    size_t raw_cap = 0;
    TraceMeta *meta = nullptr;
    size_t meta_cap = 0;
    SlotSpec *spec_s = nullptr, *spec_r = nullptr;
    size_t spec_s_cap = 0, spec_r_cap = 0;
    Tuples ts, tr;
    size_t ts_cap = 0, tr_cap = 0;  // slots
    float *tmp1 = nullptr, *tmp2 = nullptr;
    size_t tmp1_cap = 0, tmp2_cap = 0;  // elements
    float *cig = nullptr;
    size_t cig_cap = 0;
    float *qcb = nullptr, *qca = nullptr;
    size_t qc_cap = 0;
    bool qc_valid = false;
    std::vector<Batch> batches;
    // resident-variant state
    PackedTuple *pk_s = nullptr, *pk_r = nullptr;
    size_t pk_s_cap = 0, pk_r_cap = 0;
    PackedTuple16 *pk16_s = nullptr, *pk16_r = nullptr;
    size_t pk16_s_cap = 0, pk16_r_cap = 0;
    SlotScale *sc_s = nullptr, *sc_r = nullptr;
    size_t sc_s_cap = 0, sc_r_cap = 0;
    ThBound *bnd_s = nullptr, *bnd_r = nullptr;  // direction bounds per (slot, cell)
    unsigned long long *mask = nullptr;  // survivor masks of the group culls (MigResArgs)
    size_t mask_cap = 0;
    size_t bnd_s_cap = 0, bnd_r_cap = 0;
    float *pairs = nullptr;
    size_t pairs_cap = 0;
    float *tmpd = nullptr;          // diff2 scratch (diff=1 only)
    size_t tmpd_cap = 0;
    float *wap = nullptr;
    size_t wap_cap = 0;
    float *part = nullptr;  // shot groups' partial CIGs [KMIG_MAX_SPLIT][ng][nzx]
    size_t part_cap = 0;
    int *rslot = nullptr;
    size_t rslot_cap = 0;
    unsigned long long *cnt = nullptr;  // kmig_count_updates: evaluated, accepted
    size_t cnt_cap = 0;
    bool rcv_invariant = false;     // every shot uses receiver slot rslot[ih]
    // receivers that move with the shot (cmp=1), one batch: receiver columns
    // (rslot then holds each column's slot; see MigResArgs)
    bool rcv_columns = false;
    // every source (receiver) slot of the plan copies a table node: the fused
    // tuple pass then reads the tables directly and k_interp does not run
    bool direct_s = false, direct_r = false;
    unsigned *chk = nullptr;  // k_table_check's flag
    size_t chk_cap = 0;
    int n_direct = 0;         // fused tuple passes of the run that read nodes directly
    int ncols = 0;
    ColTrace *ctab = nullptr;
    size_t ctab_cap = 0;
    int *cfirst = nullptr, *clast = nullptr;
    size_t cfirst_cap = 0, clast_cap = 0;
    // offset gathers and the stack on the resident kernel: the output bin of
    // each receiver index (MigResArgs::rbin)
    int *rbin = nullptr;
    size_t rbin_cap = 0;
    int variant_used = 0;
    int tuned_tile = 0;             // autotuned resident preset (0: not tuned yet)
    long long tuned_version = -1;
    std::vector<std::pair<int, float>> tune_log;
    std::vector<hipEvent_t> events;
    std::vector<int> ev_kind;       // per event pair: 0 trace, 1 tuple, 2 migrate
    int ev_used = 0;
    int last_sb = -1, last_se = -1;
    long long geo_version = 0, plan_version = -1;
    kmig_stats st{};
    bool ran = false, pending = false;
    // captured step (kmig_set_graph): the graph of the last eager resident
    // run and the plan it was captured for
    bool graph_on = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::string gkey;
    hipEvent_t gfork = nullptr, gjoin = nullptr;
};

namespace {

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

// Tuple layouts of the resident kernel: FP32 and FP16 fields in 16 bytes
// (PackedTuple: pk_s, pk_r), BF16 and UNORM16 fields in 12 bytes with
// per-slot scales (PackedTuple16: pk16_s, pk16_r, sc_s, sc_r).
bool full_layout(int precision)
{
    return precision == KMIG_PREC_FP32 || precision == KMIG_PREC_FP16;
}

int fail(kmig_ctx *c, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
    return code;
}

#define HIPCHK(c, call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((c), KMIG_E_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int ensure(kmig_ctx *c, T **ptr, size_t *cap, size_t n)
{
    if (*cap >= n && *ptr) return KMIG_OK;
    if (*ptr) dfree(*ptr);
    *ptr = nullptr;
    *cap = 0;
    if (n == 0) return KMIG_OK;
    hipError_t e = hipMalloc((void **)ptr, n * sizeof(T));
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(c, KMIG_E_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                    hipGetErrorString(e));
    }
    *cap = n;
    // KMIG_POISON=1 fills every new float buffer with NaN bytes, so a read
    // of memory that no kernel or copy has written shows up as NaN in the
    // results instead of depending on what the allocator hands back.  Only
    // float buffers: their contents are values, never indices.
    if constexpr (std::is_same<T, float>::value) {
        static const bool poison = [] {
            const char *v = getenv("KMIG_POISON");
            return v && atoi(v) != 0;
        }();
        if (poison) HIPCHK(c, hipMemset(*ptr, 0xFF, n * sizeof(T)));
    }
    return KMIG_OK;
}

int ensure_tuples(kmig_ctx *c, Tuples &t, size_t *cap, size_t slots)
{
    if (*cap >= slots && t.T) return KMIG_OK;
    float **f[5] = {&t.T, &t.A, &t.GX, &t.GZ, &t.W};
    for (float **q : f) {
        if (*q) dfree(*q);
        *q = nullptr;
    }
    *cap = 0;
    for (float **q : f) {
        size_t dummy = 0;
        int rc = ensure(c, q, &dummy, slots * c->nzx);
        if (rc) return rc;
    }
    *cap = slots;
    t.cap = (int)slots;
    return KMIG_OK;
}

// Table node choice and interpolation coefficients for one position, as
// src/kirchmig2d_cig.cpp:987-1028 (source) / :1148-1180 (receiver) and
// SERECKIRCH/src/tinterpsr.cpp:37-188 compute them (float geometry, double
// Hermite basis from the float coordinate).  Returns the clamped node.
int make_spec(const kmig_ctx *c, const Side &sd, float pos, SlotSpec *sp)
{
    memset(sp, 0, sizeof(*sp));
    sp->surf_x = pos;
    int node = 0, interp = 0;
    float coord = 0.0f;
    if (!(sd.npos <= 1 || sd.d == 0.0f)) {
        const float f = (pos - sd.o) / sd.d;
        node = (int)floorf(f);
        if (f <= 0.0f) {
            node = 0;
        } else if (f >= (float)(sd.npos - 1)) {
            node = sd.npos - 1;
        } else {
            interp = 1;
            coord = pos - (sd.o + (float)node * sd.d);
        }
    }
    sp->node = node;
    sp->interp = interp;
    sp->coord = coord;
    if (interp) {
        const float dx = sd.d;
        const float t = coord / dx;
        sp->t = t;
        // identical expressions (and float/double promotions) to tinterpsr.cpp:150-184
        const double h00 = 2. * t * t * t - 3. * t * t + 1.;
        const double h10 = t * t * t - 2. * t * t + t;
        const double h01 = -2. * t * t * t + 3. * t * t;
        const double h11 = t * t * t - t * t;
        const double d00 = -6. * t * (1. - t);
        const double d10 = (1. - t) * (1. - 3. * t);
        const double d01 = 6. * t * (1. - t);
        const double d11 = 3. * t * t - 2. * t;
        sp->h[0] = h00;
        sp->h[1] = h10 * dx;
        sp->h[2] = h01;
        sp->h[3] = h11 * dx;
        sp->h[4] = d00;
        sp->h[5] = d10 * dx;
        sp->h[6] = d01;
        sp->h[7] = d11 * dx;
        sp->k1 = (int)(coord / c->dx + 0.5);
        sp->k2 = (int)((dx - coord) / c->dx + 0.5);
    }
    // gradient volumes: interp_volume_linear_stream, :294-322
    sp->gnode = 0;
    sp->ginterp = 0;
    if (!(sd.npos <= 1 || sd.d == 0.0f)) {
        const float f = (pos - sd.o) / sd.d;
        if (f <= 0.0f) {
            sp->gnode = 0;
        } else if (f >= (float)(sd.npos - 1)) {
            sp->gnode = sd.npos - 1;
        } else {
            const int i0 = (int)floorf(f);
            sp->gnode = i0;
            sp->ginterp = 1;
            sp->gw = f - (float)i0;
        }
    }
    return node;
}

float nearest_bin_h(float value, float g0, float dg, int ng)
{
    if (ng <= 0 || dg <= 0.0f || !std::isfinite(value)) return -1;
    int ig = (int)floorf((value - g0) / dg + 0.5f);
    if (ig < 0 || ig >= ng) return -1;
    return (float)ig;
}

size_t free_device_bytes()
{
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)8 << 30;
    return fr;
}

// Receiver columns of a run over shots [sb, se) whose receivers move with
// the shot (one batch): every migrating trace (is, ih) goes to a column that
// keeps its receiver slot for all shots.  A slot gets one column per trace of
// one shot at that position (normally one; more where a shot has coincident
// receivers).  Columns are ordered by receiver position, so the columns of a
// chunk lie side by side and share their shots.
struct Columns {
    std::vector<int> colslot, cfirst, clast;
    std::vector<ColTrace> ctab;  // [(is - sb) * ncols + c], then RES_WAP_PAD zeros
};

void build_columns(const kmig_ctx *c, int sb, int se, const std::vector<TraceMeta> &meta,
                   const std::vector<SlotSpec> &rspec, Columns &out)
{
    const int nh = c->nh, nslot = (int)rspec.size(), nsb = se - sb;
    std::vector<std::vector<int>> cols_of(nslot);
    std::vector<int> colslot, dup, used(nslot, 0), touched;
    std::vector<int> col_of((size_t)nsb * nh, -1);
    for (int js = 0; js < nsb; js++) {
        for (int ih = 0; ih < nh; ih++) {
            const int slot = meta[(size_t)js * nh + ih].slot;
            if (slot < 0) continue;
            const int k = used[slot]++;
            if (k == 0) touched.push_back(slot);
            std::vector<int> &v = cols_of[slot];
            if (k == (int)v.size()) {
                v.push_back((int)colslot.size());
                colslot.push_back(slot);
                dup.push_back(k);
            }
            col_of[(size_t)js * nh + ih] = v[k];
        }
        for (int sl : touched) used[sl] = 0;
        touched.clear();
    }
    const int ncols = (int)colslot.size();
    std::vector<int> order(ncols), rank(ncols);
    for (int k = 0; k < ncols; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const float px = rspec[colslot[x]].surf_x, py = rspec[colslot[y]].surf_x;
        if (px != py) return px < py;
        if (colslot[x] != colslot[y]) return colslot[x] < colslot[y];
        return dup[x] < dup[y];
    });
    out.colslot.resize(ncols);
    for (int r = 0; r < ncols; r++) {
        rank[order[r]] = r;
        out.colslot[r] = colslot[order[r]];
    }
    out.ctab.assign((size_t)nsb * ncols, ColTrace{0, -1.0f});
    out.ctab.resize((size_t)nsb * ncols + RES_WAP_PAD, ColTrace{0, 0.0f});
    out.cfirst.assign(ncols, 1 << 30);
    out.clast.assign(ncols, -1);
    for (int js = 0; js < nsb; js++)
        for (int ih = 0; ih < nh; ih++) {
            const int col = col_of[(size_t)js * nh + ih];
            if (col < 0) continue;
            const int r = rank[col];
            out.ctab[(size_t)js * ncols + r] = ColTrace{ih, meta[(size_t)js * nh + ih].wap};
            out.cfirst[r] = std::min(out.cfirst[r], js);
            out.clast[r] = std::max(out.clast[r], js);
        }
}

// Build TraceMeta for shots [sb, se), the batches and their slot specs.
int plan_run(kmig_ctx *c, int sb, int se, std::vector<TraceMeta> &meta,
             std::vector<SlotSpec> &sspec, std::vector<SlotSpec> &rspec,
             std::vector<float> &wapv, std::vector<int> &rslot, Columns &cols)
{
    const kmig_params &p = c->p;
    const Side &S = c->side[0], &Rt = c->side[1];
    const int nh = c->nh;
    const float dh = (nh == 1) ? 0.0f : c->dh;        // :700
    const float ds = (c->ns == 1) ? 0.0f : c->ds;     // :701
    const bool angle = p.cig && p.cig_is_angle;
    const size_t fields = angle ? 5 : 3;
    const size_t per_slot = c->nzx * sizeof(float) * (fields + 2);
    size_t budget = (size_t)(0.6 * (double)free_device_bytes());
    // already-allocated tuple/tmp memory can be reused
    budget += (c->ts_cap + c->tr_cap) * c->nzx * sizeof(float) * 5 +
              (c->tmp1_cap + c->tmp2_cap) * sizeof(float);
    size_t max_slots = std::max<size_t>(2, budget / per_slot);
    // KMIG_MAX_SLOTS caps the tuple slots per batch (forces shot batching)
    if (const char *env = getenv("KMIG_MAX_SLOTS")) {
        const long v = atol(env);
        if (v > 0) max_slots = std::max<size_t>(2, (size_t)v);
    }

    meta.assign((size_t)(se - sb) * nh, TraceMeta{});
    wapv.assign((size_t)(se - sb) * nh, -1.0f);
    rslot.assign((size_t)nh, -1);
    bool invariant = true;
    sspec.clear();
    rspec.clear();
    c->batches.clear();
    Batch cur{sb, sb, 0, 0};
    std::unordered_map<uint32_t, int> slot_of;
    std::vector<SlotSpec> cur_r;
    // Receivers that move with the shot, planned for the resident kernel:
    // s + h in float gives one physical receiver position several values a
    // few ulps apart (0.0125 k + 0.0125 j rounds differently for different
    // k, j), and each value would be its own tuple slot and receiver column,
    // splitting the shots of one position over several columns.  Positions
    // within 8 ulps of the largest |s| + |h| of the run (the size of the
    // sum's rounding error) of the smallest of their group therefore share
    // that smallest value's slot: a position change of at most 8 ulps (7.6e-6
    // km for |s| + |h| up to 16 km), whose traveltime change (about 5e-9 s)
    // lies far below the float rounding of the traveltimes themselves.  The
    // trace records (aperture index, mute) keep each trace's own position.
    std::unordered_map<uint32_t, float> pos_rep;
    const bool group_pos = p.adj == 1 && p.cmp &&
                           (angle ? p.angle_interp == 1 : p.precision == KMIG_PREC_FP32) &&
                           p.antialias >= 0.0f && p.variant != KMIG_VARIANT_STREAM;
    if (group_pos) {
        std::vector<float> all;
        all.reserve((size_t)(se - sb) * nh);
        float mag = 0.0f;
        for (int is = sb; is < se; is++) {
            const float s = c->s0 + (float)is * ds;
            for (int ih = 0; ih < nh; ih++) {
                const float h = c->h0 + (float)ih * dh;
                all.push_back(s + h);
                mag = std::max(mag, std::fabs(s) + std::fabs(h));
            }
        }
        const float tol =
            8.0f * (std::nextafter(mag, std::numeric_limits<float>::infinity()) - mag);
        std::sort(all.begin(), all.end());
        all.erase(std::unique(all.begin(), all.end()), all.end());
        float first = 0.0f;
        for (size_t k = 0; k < all.size(); k++) {
            const float x = all[k];
            if (k == 0 || !(x - first <= tol)) first = x;
            uint32_t key;
            memcpy(&key, &x, 4);
            pos_rep[key] = first;
        }
    }
    auto slot_pos = [&](float pos) {
        if (!group_pos) return pos;
        uint32_t key;
        memcpy(&key, &pos, 4);
        auto it = pos_rep.find(key);
        return it == pos_rep.end() ? pos : it->second;
    };
    auto close = [&](int s1) {
        cur.s1 = s1;
        cur.nrcv = (int)cur_r.size();
        cur.spec_r0 = rspec.size();
        rspec.insert(rspec.end(), cur_r.begin(), cur_r.end());
        c->batches.push_back(cur);
        cur = Batch{s1, s1, 0, 0};
        cur_r.clear();
        slot_of.clear();
    };
    for (int is = sb; is < se; is++) {
        const float s = c->s0 + (float)is * ds;  // :980
        SlotSpec ss;
        const int ist = make_spec(c, S, s, &ss);
        // trace records and the receivers this shot needs
        std::vector<std::pair<int, float>> need;  // (ih, receiver_pos) of migrating traces
        for (int ih = 0; ih < nh; ih++) {
            TraceMeta &m = meta[(size_t)(is - sb) * nh + ih];
            const float h = c->h0 + (float)ih * dh;
            const float receiver_pos = p.cmp ? (s + h) : h;
            float rf = 0.0f;
            if (Rt.npos > 1 && Rt.d != 0.0f) rf = (receiver_pos - Rt.o) / Rt.d;
            const int iht = (Rt.npos > 1 && Rt.d != 0.0f) ? (int)floorf(rf) : 0;
            m.src_x = s;
            m.rcv_x = p.cmp ? (s + h) : h;
            m.slot = -1;
            m.ig = 0;
            float wap = 1.0f;
            bool skip = false;
            if (p.aperture_trace != -1) {  // :1082-1099
                const float ad = fabsf((float)(iht - ist));
                if (ad >= (float)p.aperture_trace) {
                    skip = true;
                } else if (p.aperture_trace_taper > 0) {
                    const float taper_begin =
                        std::max(0.0f, (float)(p.aperture_trace - p.aperture_trace_taper));
                    if (ad > taper_begin) {
                        const float denom = (float)p.aperture_trace - taper_begin;
                        if (denom > 0.0f) {
                            const float x = (ad - taper_begin) / denom;
                            wap = 0.5f * (1.0f + cosf((float)M_PI * x));
                        }
                    }
                }
            }
            if (!skip && p.cig && !p.cig_is_angle) {  // :1100-1102
                const float ig = nearest_bin_h(h, p.g0, p.dg, p.ng);
                if (ig < 0) skip = true;
                else m.ig = (int)ig;
            }
            m.wap = skip ? -1.0f : wap;
            wapv[(size_t)(is - sb) * nh + ih] = m.wap;
            if (!skip) need.push_back({ih, slot_pos(receiver_pos)});
        }
        // new receiver positions of this shot
        int fresh = 0;
        for (auto &q : need) {
            uint32_t key;
            memcpy(&key, &q.second, 4);
            if (!slot_of.count(key)) fresh++;
        }
        const size_t nsrc_now = (size_t)(is - cur.s0);
        if (is > cur.s0 && (nsrc_now + 1 + cur_r.size() + fresh) > max_slots) close(is);
        for (auto &q : need) {
            uint32_t key;
            memcpy(&key, &q.second, 4);
            auto it = slot_of.find(key);
            int slot;
            if (it == slot_of.end()) {
                slot = (int)cur_r.size();
                SlotSpec rs;
                make_spec(c, Rt, q.second, &rs);
                cur_r.push_back(rs);
                slot_of[key] = slot;
            } else {
                slot = it->second;
            }
            meta[(size_t)(is - sb) * nh + q.first].slot = slot;
            int &rs = rslot[(size_t)q.first];
            if (rs < 0) rs = slot;
            else if (rs != slot) invariant = false;
        }
        sspec.push_back(ss);
    }
    close(se);
    c->rcv_invariant = invariant && c->batches.size() == 1;
    // receivers that move with the shot, all in one batch, for the resident
    // kernel's bins (angle gathers, linear bins): receiver columns
    c->rcv_columns = false;
    cols = Columns{};
    // (migration only: the adj=0 transpose keeps receiver indices)
    if (p.adj == 1 && !invariant && c->batches.size() == 1 && (!angle || p.angle_interp == 1)) {
        build_columns(c, sb, se, meta, rspec, cols);
        c->rcv_columns = !cols.colslot.empty();
    }
    return KMIG_OK;
}

int check_tables_type_p(kmig_ctx *c, const std::vector<SlotSpec> &v)
{
    if (c->p.interp_type != 'p') return KMIG_OK;
    if (c->nx < 2)
        return fail(c, KMIG_E_PARAM, "type=p needs at least two image columns (n2 >= 2).");
    for (const SlotSpec &s : v)
        if (s.interp && (s.k1 > c->nx || s.k2 > c->nx || s.k1 < 0 || s.k2 < 0))
            return fail(c, KMIG_E_PARAM,
                        "type=p: the table position spacing spans more than the image width.");
    return KMIG_OK;
}

hipEvent_t next_event(kmig_ctx *c)
{
    if (c->ev_used >= (int)c->events.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->events.push_back(e);
    }
    return c->events[c->ev_used++];
}


// Angle thresholds of the linear-bin range test of the resident kernel,
// which evaluates u = fmaf(angle, inv_dg, g0_inv_dg) (correctly rounded on
// host and device alike).  u is non-decreasing in the angle and non-negative
// floats order as their bit patterns, so a bisection over the bits of
// [0, amax] (amax = wrap / 2, the largest folded angle) finds the exact
// boundaries:
//   u <= ng - 1 + 1e-5  <=>  angle <= ang_hi  (-1: no angle passes)
//   u >= -1e-5          <=>  angle >= ang_lo  (lo_check = 0: all of [0, amax])
static void bin_thresholds(float amax, float inv_dg, float g0_inv_dg, int ng, float *ang_hi,
                           float *ang_lo, int *lo_check)
{
    auto val = [](uint32_t bits) {
        float f;
        std::memcpy(&f, &bits, 4);
        return f;
    };
    auto u_of = [&](uint32_t bits) { return std::fmaf(val(bits), inv_dg, g0_inv_dg); };
    uint32_t hb;
    std::memcpy(&hb, &amax, 4);
    const float u_hi = (float)(ng - 1) + 1.0e-5f, u_lo = -1.0e-5f;
    // largest angle with u <= u_hi
    if (u_of(hb) <= u_hi) {
        *ang_hi = amax;
    } else if (!(u_of(0) <= u_hi)) {
        *ang_hi = -1.0f;
    } else {
        uint32_t lo = 0, hi = hb;  // u(lo) <= u_hi < u(hi)
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            (u_of(mid) <= u_hi ? lo : hi) = mid;
        }
        *ang_hi = val(lo);
    }
    // smallest angle with u >= u_lo
    if (u_of(0) >= u_lo) {
        *ang_lo = 0.0f;
        *lo_check = 0;
    } else if (!(u_of(hb) >= u_lo)) {
        *ang_lo = std::numeric_limits<float>::infinity();
        *lo_check = 1;
    } else {
        uint32_t lo = 0, hi = hb;  // u(lo) < u_lo <= u(hi)
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            (u_of(mid) >= u_lo ? hi : lo) = mid;
        }
        *ang_lo = val(hi);
        *lo_check = 1;
    }
}

MigResArgs res_args(kmig_ctx *c, int shot0, int nshots, int ng)
{
    const kmig_params &p = c->p;
    MigResArgs R{};
    R.nz = c->nz;
    R.nx = c->nx;
    R.nzx = c->nzx;
    R.nt = c->nt;
    R.nh = c->nh;
    R.shot0 = shot0;
    R.nshots = nshots;
    R.stup = c->pk_s;
    R.rtup = c->pk_r;
    R.stup16 = c->pk16_s;
    R.rtup16 = c->pk16_r;
    R.sscale = c->sc_s;
    R.rscale = c->sc_r;
    R.rslot = c->rslot;
    R.wap = c->wap;
    R.pairs = c->pairs;
    // KMIG_CULL=0 switches the group culling off (same results, for timing)
    const char *cull_env = std::getenv("KMIG_CULL");
    const bool cull = !(cull_env && std::atoi(cull_env) == 0);
    R.sbnd = cull ? c->bnd_s : nullptr;
    R.mask = cull ? c->mask : nullptr;
    R.rbnd = cull ? c->bnd_r : nullptr;
    R.scell_off = 0;
    R.ncz = (c->nz + BND_CZ - 1) / BND_CZ;
    R.ncx = (c->nx + BND_CX - 1) / BND_CX;
    R.sbnd_stride = nshots;
    R.nrslots = c->batches.empty() ? 0 : c->batches[0].nrcv;
    if (c->rcv_columns) {
        R.ctab = c->ctab;
        R.cfirst = c->cfirst;
        R.clast = c->clast;
        R.ncols = c->ncols;
    }
    if (!(p.cig && p.cig_is_angle)) R.rbin = c->rbin;
    R.cig = c->cig;
    const double cst = ((double)p.tau - (double)c->t0) / (double)c->dt;
    const double ci = std::floor(cst);
    R.c_int = (int)ci;
    R.c_frac = (float)(cst - ci);
    R.t_cull = R.nt - 2 - R.c_int;
    R.aal = p.antialias;
    R.wrap = p.half_angle ? 180.0f : 360.0f;
    R.inv_dg = 1.0f / p.dg;
    R.g0_inv_dg = -p.g0 / p.dg;
    bin_thresholds(0.5f * R.wrap, R.inv_dg, R.g0_inv_dg, ng, &R.ang_hi, &R.ang_lo, &R.lo_check);
    R.ang_hi_up = std::nextafter(R.ang_hi, std::numeric_limits<float>::infinity());
    R.ng = ng;
    R.oazmin = p.oazmin;
    R.oazmax = p.oazmax;
    R.oaztaper = p.oaztaper;
    R.inv_oaztaper = p.oaztaper > 0.0f ? 1.0f / p.oaztaper : 0.0f;
    R.taper_start = p.angle_taper_start;
    R.taper_end = p.angle_taper_end;
    R.inv_taper = p.angle_taper_end > p.angle_taper_start
                      ? 1.0f / (p.angle_taper_end - p.angle_taper_start)
                      : 0.0f;
    return R;
}

// Project autotuner for the resident kernel: times every tile / receiver-chunk
// preset on a slice of the real shots (about 5e9 updates, twice each, HIP
// events on the context stream) and keeps the fastest.  Runs once per context
// and input set; it writes into the CIG, which the run clears afterwards.
int autotune(kmig_ctx *c, const MigResArgs &full)
{
    if (c->tuned_tile > 0 && c->tuned_version == c->geo_version) return KMIG_OK;
    static const int presets32[] = {9, 15, 16, 20, 11, 10, 18};
    static const int presets16[] = {9, 15, 10, 13, 11};
    // the 16-bit formats and receiver columns are built for five presets,
    // offset gathers and stacks for three
    static const int presets_bin[] = {9, 15, 11};
    const bool p16 = c->p.precision != KMIG_PREC_FP32 || c->rcv_columns;
    const bool pbin = full.rbin != nullptr;
    const int *presets = pbin ? presets_bin : (p16 ? presets16 : presets32);
    const int npresets = pbin ? 3 : (p16 ? 5 : 7);
    const double per_shot = (double)c->nh * (double)c->nzx;
    int S = (int)std::max(1.0, std::min((double)full.nshots, 5e9 / std::max(per_shot, 1.0)));
    // two half slices, centred at 1/4 and 3/4 of the batch (source tuples are
    // batch-local): how much of the work the group culling removes varies
    // along the shot line, so one middle slice can misjudge the presets
    const int S1 = std::max(1, S / 2), S2 = std::max(0, S - S1);
    MigResArgs Rs[2];
    for (int h = 0; h < 2; h++) {
        const int n = h ? S2 : S1;
        const int off = std::min(std::max(0, (full.nshots * (1 + 2 * h)) / 4 - n / 2),
                                 std::max(0, full.nshots - n));
        MigResArgs &R = Rs[h];
        R = full;
        R.nshots = n;
        R.shot0 = full.shot0 + off;
        if (full.stup) R.stup = full.stup + (size_t)off * c->nzx;
        if (full.stup16) R.stup16 = full.stup16 + (size_t)off * c->nzx;
        if (full.sscale) R.sscale = full.sscale + off;
        if (full.sbnd) R.sbnd = full.sbnd + off;  // rows keep their full stride
        R.scell_off = full.scell_off + off;
    }
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    float best = 1e30f;
    int best_tile = 2;
    c->tune_log.clear();
    for (int ip = 0; ip < npresets; ip++) {
        const int t = presets[ip];
        if (migrate_lds_bytes(full.ng, t) > 64 * 1024) continue;
        float tmin = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            HIPCHK(c, hipEventRecord(e0, c->stream));
            for (int h = 0; h < 2; h++)
                if (Rs[h].nshots > 0)
                    HIPCHK(c, (hipError_t)launch_migrate_resident_probe(Rs[h], t, c->p.precision,
                                                                        c->stream));
            HIPCHK(c, hipEventRecord(e1, c->stream));
            HIPCHK(c, hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
            tmin = std::min(tmin, ms);
        }
        c->tune_log.push_back({t, tmin});
        if (tmin < best) {
            best = tmin;
            best_tile = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (getenv("KMIG_TUNE_LOG")) {
        fprintf(stderr, "kmig autotune (%d shots x %d receivers):", S, c->nh);
        for (auto &q : c->tune_log) fprintf(stderr, " tile%d=%.2fms", q.first, q.second);
        fprintf(stderr, " -> tile%d\n", best_tile);
    }
    c->tuned_tile = best_tile;
    c->tuned_version = c->geo_version;
    // the tuning launches accumulated into the CIG: clear it again
    const int ng = full.ng;
    HIPCHK(c, (hipError_t)launch_fill(c->cig, (size_t)ng * c->nzx, 0.0f, c->stream));
    return KMIG_OK;
}

// Host geometry, allocations and uploads for shots [sb, se); cached while the
// inputs and the shot range are unchanged.
int plan_if_needed(kmig_ctx *c, int sb, int se)
{
    const kmig_params &p = c->p;
    const bool angle = p.cig && p.cig_is_angle;
    int rc;
    for (Side &sd : c->side)
        if (sd.nn && nn_prepare(sd.nn) != KMIG_OK)
            return fail(c, KMIG_E_DEVICE, "surrogate: %s", kmig_nn_last_error(sd.nn));
    if (c->plan_version != c->geo_version || sb != c->last_sb || se != c->last_se) {
        std::vector<TraceMeta> meta;
        std::vector<SlotSpec> sspec, rspec;
        std::vector<float> wapv;
        std::vector<int> rslot;
        Columns cols;
        if ((rc = plan_run(c, sb, se, meta, sspec, rspec, wapv, rslot, cols))) return rc;
        if ((rc = check_tables_type_p(c, sspec))) return rc;
        if ((rc = check_tables_type_p(c, rspec))) return rc;
        // meta is indexed by global trace number
        if ((rc = ensure(c, &c->meta, &c->meta_cap, (size_t)c->ns * c->nh))) return rc;
        if (!meta.empty())
            HIPCHK(c, hipMemcpy(c->meta + (size_t)sb * c->nh, meta.data(),
                                meta.size() * sizeof(TraceMeta), hipMemcpyHostToDevice));
        if ((rc = ensure(c, &c->spec_s, &c->spec_s_cap, std::max<size_t>(1, sspec.size()))))
            return rc;
        if ((rc = ensure(c, &c->spec_r, &c->spec_r_cap, std::max<size_t>(1, rspec.size()))))
            return rc;
        // slots that copy a node: no interpolation, or Hermite interpolation
        // at coordinate 0 of plain tables (k_interp gives the node's T and
        // dT/ds bit for bit there)
        auto all_direct = [&](const std::vector<SlotSpec> &v, const Side &sd) {
            for (const SlotSpec &sp : v)
                if (sp.interp && !(p.interp_type == 'h' && sp.coord == 0.0f && sd.plain))
                    return false;
            return true;
        };
        c->direct_s = all_direct(sspec, c->side[0]);
        c->direct_r = all_direct(rspec, c->side[1]);
        if (!sspec.empty())
            HIPCHK(c, hipMemcpy(c->spec_s, sspec.data(), sspec.size() * sizeof(SlotSpec),
                                hipMemcpyHostToDevice));
        if (!rspec.empty())
            HIPCHK(c, hipMemcpy(c->spec_r, rspec.data(), rspec.size() * sizeof(SlotSpec),
                                hipMemcpyHostToDevice));
        for (int k = 0; k < 2; k++) {
            Side &sd = c->side[k];
            sd.nn_slots = k == 0 ? sspec.size() : rspec.size();
            if (!sd.nn) continue;
            if ((rc = ensure(c, &sd.nn_used, &sd.nn_cap, std::max<size_t>(1, sd.nn_slots))))
                return rc;
            if ((rc = ensure(c, &sd.nn_maxres, &sd.nn_cap_m, std::max<size_t>(1, sd.nn_slots))))
                return rc;
        }
        size_t maxs = 0, maxr = 0;
        for (const Batch &b : c->batches) {
            maxs = std::max(maxs, (size_t)(b.s1 - b.s0));
            maxr = std::max(maxr, (size_t)b.nrcv);
        }
        if ((rc = ensure_tuples(c, c->ts, &c->ts_cap, std::max<size_t>(1, maxs)))) return rc;
        if ((rc = ensure_tuples(c, c->tr, &c->tr_cap, std::max<size_t>(1, maxr)))) return rc;
        const size_t tslots = std::max<size_t>(1, std::max(maxs, maxr));
        // (the per-slot surrogate fallback keeps the network's values there)
        if ((angle && p.angle_smooth > 0) || c->side[0].nn || c->side[1].nn) {
            if ((rc = ensure(c, &c->tmp1, &c->tmp1_cap, tslots * c->nzx))) return rc;
            if ((rc = ensure(c, &c->tmp2, &c->tmp2_cap, tslots * c->nzx))) return rc;
        }
        // resident variant: angle CIG, two-bin interpolation, receivers that do
        // not move with the shot, one batch, times within the int16 sample range
        const double cst = ((double)p.tau - (double)c->t0) / (double)c->dt;
        // bins the resident kernel forms: two linear angle bins, or the
        // trace's own bin (offset gathers, the stack) for migration with FP32
        // tuples and an axis whose LDS columns fit 8 x 8 tiles in 64 KiB
        const int ng_out = p.cig ? p.ng : 1;
        const bool bins_ok = angle ? p.angle_interp == 1
                                   : (p.adj == 1 && p.precision == KMIG_PREC_FP32 &&
                                      migrate_lds_bytes(ng_out, 15) <= 64 * 1024);
        // (FP16 tuples read their late tap 16 bytes past a pair offset: nt >= 3)
        const bool res_ok = bins_ok && (c->rcv_invariant || c->rcv_columns) &&
                            p.antialias >= 0.0f && std::fabs(cst) + (double)c->nt < 30000.0 &&
                            (p.precision != KMIG_PREC_FP16 || c->nt >= 3);
        int variant = KMIG_VARIANT_STREAM;
        if (p.variant == KMIG_VARIANT_RESIDENT) {
            if (!res_ok)
                return fail(c, KMIG_E_PARAM,
                            "variant=resident needs linear angle bins (mode=angle, angle_interp=1) "
                            "or migration with FP32 tables into offset gathers or a stack, "
                            "antialias >= 0 and tuples that fit one batch");
            variant = KMIG_VARIANT_RESIDENT;
        } else if (p.variant == KMIG_VARIANT_AUTO && res_ok) {
            variant = KMIG_VARIANT_RESIDENT;
        }
        if (p.precision != KMIG_PREC_FP32 && variant != KMIG_VARIANT_RESIDENT)
            return fail(c, KMIG_E_PARAM,
                        "precision=fp16/bf16 applies to the resident kernel: it needs "
                        "mode=angle, angle_interp=1, antialias >= 0, tuples that fit one "
                        "batch and (fp16) nt >= 3");
        c->variant_used = variant;
        if (variant == KMIG_VARIANT_RESIDENT && !full_layout(p.precision)) {
            if ((rc = ensure(c, &c->pk16_s, &c->pk16_s_cap, std::max<size_t>(1, maxs) * c->nzx)))
                return rc;
            if ((rc = ensure(c, &c->pk16_r, &c->pk16_r_cap, std::max<size_t>(1, maxr) * c->nzx)))
                return rc;
            if ((rc = ensure(c, &c->sc_s, &c->sc_s_cap, std::max<size_t>(1, maxs)))) return rc;
            if ((rc = ensure(c, &c->sc_r, &c->sc_r_cap, std::max<size_t>(1, maxr)))) return rc;
        }
        if (variant == KMIG_VARIANT_RESIDENT && full_layout(p.precision)) {
            if ((rc = ensure(c, &c->pk_s, &c->pk_s_cap, std::max<size_t>(1, maxs) * c->nzx)))
                return rc;
            if ((rc = ensure(c, &c->pk_r, &c->pk_r_cap, std::max<size_t>(1, maxr) * c->nzx)))
                return rc;
        }
        if (variant == KMIG_VARIANT_RESIDENT) {
            const size_t ncell = (size_t)((c->nz + BND_CZ - 1) / BND_CZ) *
                                 ((c->nx + BND_CX - 1) / BND_CX);
            if ((rc = ensure(c, &c->bnd_s, &c->bnd_s_cap, std::max<size_t>(1, maxs) * ncell)))
                return rc;
            if ((rc = ensure(c, &c->bnd_r, &c->bnd_r_cap, std::max<size_t>(1, maxr) * ncell)))
                return rc;
            // survivor masks for the smallest receiver chunk of any preset (4)
            const size_t nrec = c->rcv_columns ? cols.colslot.size() : (size_t)c->nh;
            if ((rc = ensure(c, &c->mask, &c->mask_cap,
                             ncell * ((nrec + 3) / 4) * ((std::max<size_t>(1, maxs) + 63) / 64))))
                return rc;
            if ((rc = ensure(c, &c->pairs, &c->pairs_cap,
                             trace_alloc((size_t)c->ns * c->nh, c->nt) * 2)))
                return rc;
            // The resident kernels read a whole receiver chunk of wap
            // entries: past nh in the last chunk of a shot, i.e. the first
            // entries of the next shot's row (of the pad after the last
            // shot).  Those receivers have zero tuple weight, so any finite
            // value there adds an exact zero; a run over a shot range writes
            // only its own rows, so the whole array (pad included) is zeroed
            // first and never holds uninitialised memory.
            const size_t nwap = (size_t)c->ns * c->nh;
            if ((rc = ensure(c, &c->wap, &c->wap_cap, nwap + RES_WAP_PAD))) return rc;
            HIPCHK(c, hipMemset(c->wap, 0, (nwap + RES_WAP_PAD) * sizeof(float)));
            if (!wapv.empty())
                HIPCHK(c, hipMemcpy(c->wap + (size_t)sb * c->nh, wapv.data(),
                                    wapv.size() * sizeof(float), hipMemcpyHostToDevice));
            // receiver slot per receiver index (cmp=0) or per receiver column
            // (-1 past the last, so a chunk's R slots load without bounds tests)
            std::vector<int> rs = c->rcv_columns ? cols.colslot : rslot;
            rs.resize(rs.size() + RSLOT_PAD, -1);
            if ((rc = ensure(c, &c->rslot, &c->rslot_cap, rs.size()))) return rc;
            HIPCHK(c, hipMemcpy(c->rslot, rs.data(), rs.size() * sizeof(int),
                                hipMemcpyHostToDevice));
            c->ncols = c->rcv_columns ? (int)cols.colslot.size() : 0;
            if (!angle) {
                // per receiver index: its offset bin (offset gathers; a
                // receiver outside the axis has no migrating trace), or 0
                std::vector<int> rb((size_t)c->nh, 0);
                const float dh = (c->nh == 1) ? 0.0f : c->dh;
                for (int ih = 0; ih < c->nh && p.cig; ih++) {
                    const float ig = nearest_bin_h(c->h0 + (float)ih * dh, p.g0, p.dg, p.ng);
                    rb[(size_t)ih] = ig < 0 ? 0 : (int)ig;
                }
                if ((rc = ensure(c, &c->rbin, &c->rbin_cap, rb.size()))) return rc;
                HIPCHK(c, hipMemcpy(c->rbin, rb.data(), rb.size() * sizeof(int),
                                    hipMemcpyHostToDevice));
            }
            if (c->rcv_columns) {
                if ((rc = ensure(c, &c->ctab, &c->ctab_cap, cols.ctab.size()))) return rc;
                if ((rc = ensure(c, &c->cfirst, &c->cfirst_cap, cols.cfirst.size()))) return rc;
                if ((rc = ensure(c, &c->clast, &c->clast_cap, cols.clast.size()))) return rc;
                HIPCHK(c, hipMemcpy(c->ctab, cols.ctab.data(), cols.ctab.size() * sizeof(ColTrace),
                                    hipMemcpyHostToDevice));
                HIPCHK(c, hipMemcpy(c->cfirst, cols.cfirst.data(),
                                    cols.cfirst.size() * sizeof(int), hipMemcpyHostToDevice));
                HIPCHK(c, hipMemcpy(c->clast, cols.clast.data(), cols.clast.size() * sizeof(int),
                                    hipMemcpyHostToDevice));
            }
        }
        c->plan_version = c->geo_version;
        c->last_sb = sb;
        c->last_se = se;
    }
    return KMIG_OK;
}

// CIG and QC buffers of a run over shots [sb, se).
int ensure_outputs(kmig_ctx *c, int sb, int se)
{
    const kmig_params &p = c->p;
    const int ng = p.cig ? p.ng : 1;
    int rc;
    if ((rc = ensure(c, &c->cig, &c->cig_cap, (size_t)ng * c->nzx))) return rc;
    const bool want_qc = p.direct_mute && p.direct_mute_qc_shot >= sb && p.direct_mute_qc_shot < se;
    if (want_qc) {
        size_t cap2 = c->qc_cap;
        if ((rc = ensure(c, &c->qcb, &c->qc_cap, (size_t)c->nt * c->nh))) return rc;
        if ((rc = ensure(c, &c->qca, &cap2, (size_t)c->nt * c->nh))) return rc;
    }
    c->qc_valid = want_qc;
    return KMIG_OK;
}

// CIG clear and trace preparation of shots [sb, se) on stream st.
int enqueue_trace_prep(kmig_ctx *c, hipStream_t st, int sb, int se)
{
    const kmig_params &p = c->p;
    const int ng = p.cig ? p.ng : 1;
    const bool resident = c->variant_used == KMIG_VARIANT_RESIDENT;
    int rc;

    HIPCHK(c, (hipError_t)launch_fill(c->cig, (size_t)ng * c->nzx, 0.0f, st));

    TracePrepArgs ta{};
    ta.raw = c->raw;
    ta.prep = c->prep;
    ta.meta = c->meta;
    ta.qcb = c->qcb;
    ta.qca = c->qca;
    ta.nt = c->nt;
    ta.nh = c->nh;
    ta.shot_begin = sb;
    ta.shot_end = se;
    ta.t0 = c->t0;
    ta.dt = c->dt;
    ta.mute = p.direct_mute;
    ta.invert = p.direct_mute_invert;
    ta.qc_shot = c->qc_valid ? p.direct_mute_qc_shot : -1;
    ta.diff = p.diff;
    ta.slope = p.direct_mute_slope;
    ta.intercept = p.direct_mute_intercept;
    ta.taper = p.direct_mute_taper;
    if (p.diff == 1) {
        if ((rc = ensure(c, &c->tmpd, &c->tmpd_cap, (size_t)c->ns * c->nh * c->nt))) return rc;
    }
    HIPCHK(c, (hipError_t)launch_trace_prep2(ta, resident ? c->pairs : nullptr, c->tmpd,
                                             st));
    return KMIG_OK;
}

// Table interpolation, then (for a side with a surrogate) the network step
// on the slots [slot0, slot0 + a.nslots) of that side's counters, then
// gradients, smoothing and weights.  With an output given (resident path:
// `packed` for FP32, `packed16` + `scale` for the 16-bit formats), the last
// part and the packing run as one fused pass where its radii allow
// (KMIG_TUPLE_FUSED=0 turns the fused pass off); returns in *fused whether
// it did.
struct PackOut {
    PackedTuple *packed = nullptr;
    PackedTuple16 *packed16 = nullptr;
    SlotScale *scale = nullptr;
    ThBound *bnd = nullptr;  // the fused pass also writes the slots' bounds here
    int precision = KMIG_PREC_FP32;
    float th_scale = 0.0f;
};

int side_tuples(kmig_ctx *c, int side, const TuplePrepArgs &a, Tuples &o, size_t slot0,
                hipStream_t st, const PackOut &po, bool *fused)
{
    Side &sd = c->side[side];
    const char *env = getenv("KMIG_TUPLE_FUSED");
    const bool use_fused = (po.packed || po.packed16) && tuple_fused_ok(a) &&
                           !(env && env[0] == '0');
    if (fused) *fused = use_fused;
    // the fused pass reads copied table nodes itself (not with gradient
    // files, interpolated with the tables; with a surrogate only under the
    // per-slot fallback, whose taken slots' values go to the planes);
    // KMIG_TUPLE_DIRECT=0 keeps k_interp
    const char *denv = getenv("KMIG_TUPLE_DIRECT");
    const bool per_slot = sd.nn_fallback == KMIG_NN_FALLBACK_SLOT;
    TuplePrepArgs af = a;
    af.direct = use_fused && (!sd.nn || per_slot) && !a.use_grad_files &&
                (side == 0 ? c->direct_s : c->direct_r) && !(denv && denv[0] == '0');
    if (!af.direct) HIPCHK(c, (hipError_t)launch_tuple_interp(a, o, st));
    c->n_direct += af.direct ? 1 : 0;
    if (sd.nn && a.nslots > 0 && af.direct) {
        af.nn_maxres = sd.nn_maxres + slot0;
        memcpy(&af.nn_lim, &sd.nn_max_dt, sizeof(af.nn_lim));
    }
    if (sd.nn && a.nslots > 0) {
        HIPCHK(c, (hipError_t)launch_fill((float *)(sd.nn_used + slot0), (size_t)a.nslots, 0.0f,
                                          st));
        HIPCHK(c, (hipError_t)launch_fill((float *)(sd.nn_maxres + slot0), (size_t)a.nslots,
                                          0.0f, st));
        HIPCHK(c, (hipError_t)nn_replace_tuples(sd.nn, a.spec, a.nslots, c->nz, c->nzx, c->z0,
                                                c->dz, c->x0, c->dx, a.dside, sd.nn_max_dt, o.T,
                                                o.A, sd.nn_used + slot0, sd.nn_maxres + slot0,
                                                per_slot ? c->tmp1 : nullptr,
                                                per_slot ? c->tmp2 : nullptr,
                                                af.direct ? a.tab : nullptr, st));
    }
    if (use_fused && po.packed) {
        HIPCHK(c, (hipError_t)launch_tuple_fused(af, o, po.packed, c->dt, po.th_scale,
                                                 po.precision, po.bnd, st));
        return KMIG_OK;
    }
    if (use_fused) {
        HIPCHK(c, (hipError_t)launch_tuple_fused16(af, o, po.packed16, po.scale, c->dt,
                                                   po.th_scale, po.precision, po.bnd, st));
        return KMIG_OK;
    }
    HIPCHK(c, (hipError_t)launch_tuple_finish(a, o, c->tmp1, c->tmp2, st));
    return KMIG_OK;
}

// Tuple preparation (and packing) of one shot batch on stream st.
int enqueue_tuple_prep(kmig_ctx *c, hipStream_t st, const Batch &b, size_t spec_s_off)
{
    const kmig_params &p = c->p;
    const bool angle = p.cig && p.cig_is_angle;
    const bool gfiles = c->side[0].GX != nullptr;
    const bool resident = c->variant_used == KMIG_VARIANT_RESIDENT;
    const float dh = (c->nh == 1) ? 0.0f : c->dh;
    const float ds = (c->ns == 1) ? 0.0f : c->ds;
    const int nsb = b.s1 - b.s0;
    TuplePrepArgs A{};
    A.nz = c->nz;
    A.nx = c->nx;
    A.nzx = c->nzx;
    A.x0 = c->x0;
    A.dx = c->dx;
    A.dz = c->dz;
    A.dd = c->dx;
    A.type = p.interp_type;
    A.use_grad_files = gfiles ? 1 : 0;
    A.radius = p.angle_grad_radius;
    A.smooth = p.angle_smooth;
    A.angle_mode = angle ? 1 : 0;
    A.aper = p.aperture;
    A.aper_taper = p.aperture_taper;
    A.rmin = p.aperture_rmin;
    A.rtaper = p.aperture_rtaper;
    A.aal = p.antialias;
    // packed directions in output-angle degrees (the opening angle is then
    // a difference of two directions folded at 360, or 180 for half angles)
    const float th_scale = angle ? (float)((p.half_angle ? 90.0 : 180.0) / M_PI) : 0.0f;
    // FP32 and FP16 fields use the 16-byte layout, BF16 and UNORM16 the 12-byte one
    const bool pack32 = resident && full_layout(p.precision);
    bool fused_s = false, fused_r = false;
    // source side
    A.tab = c->side[0].T;
    A.der = c->side[0].D;
    A.gxv = c->side[0].GX;
    A.gzv = c->side[0].GZ;
    A.spec = c->spec_s + spec_s_off;
    A.nslots = nsb;
    A.dpos = c->side[0].d;
    A.dside = ds;
    PackOut po;
    po.th_scale = th_scale;
    po.precision = p.precision;
    if (pack32) po.packed = c->pk_s;
    if (resident && !pack32) {
        po.packed16 = c->pk16_s;
        po.scale = c->sc_s;
    }
    if (resident) po.bnd = c->bnd_s;
    int rc = side_tuples(c, 0, A, c->ts, spec_s_off, st, po, &fused_s);
    if (rc) return rc;
    // receiver side
    A.tab = c->side[1].T;
    A.der = c->side[1].D;
    A.gxv = c->side[1].GX;
    A.gzv = c->side[1].GZ;
    A.spec = c->spec_r + b.spec_r0;
    A.nslots = b.nrcv;
    A.dpos = c->side[1].d;
    A.dside = dh;
    if (pack32) po.packed = c->pk_r;
    if (resident && !pack32) {
        po.packed16 = c->pk16_r;
        po.scale = c->sc_r;
    }
    if (resident) po.bnd = c->bnd_r;
    if ((rc = side_tuples(c, 1, A, c->tr, b.spec_r0, st, po, &fused_r))) return rc;
    if (pack32) {
        if (!fused_s)
            HIPCHK(c, (hipError_t)launch_pack(c->ts, c->pk_s, nsb, c->nzx, c->dt, th_scale,
                                              p.antialias, p.precision, st));
        if (!fused_r)
            HIPCHK(c, (hipError_t)launch_pack(c->tr, c->pk_r, b.nrcv, c->nzx, c->dt, th_scale,
                                              p.antialias, p.precision, st));
    } else if (resident) {
        if (!fused_s)
            HIPCHK(c, (hipError_t)launch_pack16(c->ts, c->pk16_s, c->sc_s, nsb, c->nzx, c->dt,
                                                th_scale, p.precision, st));
        if (!fused_r)
            HIPCHK(c, (hipError_t)launch_pack16(c->tr, c->pk16_r, c->sc_r, b.nrcv, c->nzx,
                                                c->dt, th_scale, p.precision, st));
    }
    if (resident) {
        // the fused pass writes the bounds of its tuples itself
        if (!fused_s)
            HIPCHK(c, (hipError_t)launch_th_bounds(pack32 ? c->pk_s : nullptr,
                                                   pack32 ? nullptr : c->pk16_s, nsb, c->nz,
                                                   c->nx, c->bnd_s, st));
        if (!fused_r)
            HIPCHK(c, (hipError_t)launch_th_bounds(pack32 ? c->pk_r : nullptr,
                                                   pack32 ? nullptr : c->pk16_r, b.nrcv, c->nz,
                                                   c->nx, c->bnd_r, st));
    }
    return KMIG_OK;
}

}  // namespace

extern "C" {

const char *kmig_version(void) { return "kmig 0.1 (gfx950)"; }

int kmig_create(kmig_ctx **out, const kmig_params *p)
{
    if (!out || !p) return KMIG_E_PARAM;
    *out = nullptr;
    char err[512] = {0};
    int rc = kmig_params_check(p, -1, err, sizeof(err));
    kmig_ctx *c = new kmig_ctx();
    c->p = *p;
    if (rc) {
        c->err = err;
        *out = c;
        return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        *out = c;
        return fail(c, KMIG_E_DEVICE, "no HIP device is available");
    }
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
        *out = c;
        return fail(c, KMIG_E_DEVICE, "hipStreamCreate failed");
    }
    *out = c;
    return KMIG_OK;
}

void kmig_destroy(kmig_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (Side &s : c->side) {
        dfree(s.T);
        dfree(s.D);
        dfree(s.GX);
        dfree(s.GZ);
        dfree(s.nn_used);
        dfree(s.nn_maxres);
    }
    dfree(c->raw);
    dfree(c->prep);
    dfree(c->meta);
    dfree(c->spec_s);
    dfree(c->spec_r);
    for (Tuples *t : {&c->ts, &c->tr}) {
        dfree(t->T);
        dfree(t->A);
        dfree(t->GX);
        dfree(t->GZ);
        dfree(t->W);
    }
    dfree(c->tmp1);
    dfree(c->tmp2);
    dfree(c->cig);
    dfree(c->qcb);
    dfree(c->qca);
    dfree(c->pk_s);
    dfree(c->pk_r);
    dfree(c->pk16_s);
    dfree(c->bnd_s);
    dfree(c->mask);
    dfree(c->chk);
    dfree(c->bnd_r);
    dfree(c->pk16_r);
    dfree(c->sc_s);
    dfree(c->sc_r);
    dfree(c->pairs);
    dfree(c->tmpd);
    dfree(c->wap);
    dfree(c->part);
    dfree(c->rslot);
    dfree(c->cnt);
    dfree(c->ctab);
    dfree(c->cfirst);
    dfree(c->clast);
    dfree(c->rbin);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->gfork) (void)hipEventDestroy(c->gfork);
    if (c->gjoin) (void)hipEventDestroy(c->gjoin);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *kmig_last_error(const kmig_ctx *c) { return c ? c->err.c_str() : "null context"; }

void *kmig_stream(kmig_ctx *c) { return c ? (void *)c->stream : nullptr; }

int kmig_set_surrogate(kmig_ctx *c, int side, kmig_nn *nn, float max_dt)
{
    if (!c) return KMIG_E_PARAM;
    if (side != 0 && side != 1) return fail(c, KMIG_E_PARAM, "side must be 0 or 1");
    if (nn && !(max_dt >= 0.0f))
        return fail(c, KMIG_E_PARAM, "surrogate residual limit must be >= 0");
    c->side[side].nn = nn;
    c->side[side].nn_max_dt = max_dt;
    c->geo_version++;  // the next run plans its counters
    return KMIG_OK;
}

int kmig_set_surrogate_fallback(kmig_ctx *c, int side, int mode)
{
    if (!c) return KMIG_E_PARAM;
    if (side != 0 && side != 1) return fail(c, KMIG_E_PARAM, "side must be 0 or 1");
    if (mode != KMIG_NN_FALLBACK_SLOT && mode != KMIG_NN_FALLBACK_POINT)
        return fail(c, KMIG_E_PARAM, "fallback mode must be KMIG_NN_FALLBACK_SLOT or _POINT");
    c->side[side].nn_fallback = mode;
    c->geo_version++;
    return KMIG_OK;
}

int kmig_get_surrogate_stats(kmig_ctx *c, int side, double *points, double *used,
                             float *max_residual)
{
    if (!c) return KMIG_E_PARAM;
    if (side != 0 && side != 1) return fail(c, KMIG_E_PARAM, "side must be 0 or 1");
    Side &sd = c->side[side];
    if (!sd.nn || !sd.nn_used || !c->ran)
        return fail(c, KMIG_E_STATE, "no surrogate run on this side");
    int rc = kmig_sync(c);
    if (rc) return rc;
    std::vector<int> u(sd.nn_slots);
    std::vector<unsigned> m(sd.nn_slots);
    if (!u.empty()) {
        HIPCHK(c, hipMemcpy(u.data(), sd.nn_used, u.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(m.data(), sd.nn_maxres, m.size() * sizeof(unsigned),
                            hipMemcpyDeviceToHost));
    }
    double su = 0.0;
    float mx = 0.0f;
    for (size_t k = 0; k < u.size(); k++) {
        su += u[k];
        float f;
        memcpy(&f, &m[k], sizeof(f));
        mx = std::max(mx, f);
    }
    if (points) *points = (double)sd.nn_slots * (double)c->nzx;
    if (used) *used = su;
    if (max_residual) *max_residual = mx;
    return KMIG_OK;
}

int kmig_set_tables(kmig_ctx *c, int side, const float *T, const float *dTds, const float *gx,
                    const float *gz, int nz, int nx, int npos, float z0, float dz, float x0,
                    float dx, float o_pos, float d_pos)
{
    if (!c) return KMIG_E_PARAM;
    if (!c->stream) return fail(c, KMIG_E_STATE, "context has no device");
    if (side != 0 && side != 1) return fail(c, KMIG_E_PARAM, "side must be 0 or 1");
    if (!T || !dTds) return fail(c, KMIG_E_PARAM, "T and dTds are required");
    if ((gx == nullptr) != (gz == nullptr))
        return fail(c, KMIG_E_PARAM,
                    "For mode=angle, please provide all of sgradx=, sgradz=, rgradx=, and "
                    "rgradz=, or provide none of them.");
    if (nz <= 0 || nx <= 0 || npos <= 0)
        return fail(c, KMIG_E_DIM, "Need 3D data volume for %s= (z,x,%s)",
                    side ? "rtable" : "stable", side ? "r" : "s");
    const Side &other = c->side[1 - side];
    if (other.set && (nz != c->nz || nx != c->nx)) {
        if (side == 1)
            return fail(c, KMIG_E_DIM, "Dimension mismatch in rtable=. Expected n1=%d n2=%d",
                        c->nz, c->nx);
        return fail(c, KMIG_E_DIM, "Dimension mismatch in stable=. Expected n1=%d n2=%d", c->nz,
                    c->nx);
    }
    if (side == 0 || !other.set) {
        c->nz = nz;
        c->nx = nx;
        c->z0 = z0;
        c->dz = dz;
        c->x0 = x0;
        c->dx = dx;
        c->nzx = (size_t)nz * nx;
    }
    Side &s = c->side[side];
    const size_t n = (size_t)npos * c->nzx;
    int rc;
    if ((rc = ensure(c, &s.T, &s.cap, n))) return rc;
    size_t cap2 = s.cap;
    if ((rc = ensure(c, &s.D, &cap2, n))) return rc;
    HIPCHK(c, hipMemcpy(s.T, T, n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(s.D, dTds, n * sizeof(float), hipMemcpyHostToDevice));
    {
        if ((rc = ensure(c, &c->chk, &c->chk_cap, 1))) return rc;
        HIPCHK(c, (hipError_t)launch_table_check(s.T, s.D, n, c->chk, c->stream));
        unsigned bad = 1;
        HIPCHK(c, hipMemcpyAsync(&bad, c->chk, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        s.plain = bad == 0;
    }
    if (gx) {
        size_t cg = 0;
        dfree(s.GX);
        dfree(s.GZ);
        s.GX = s.GZ = nullptr;
        if ((rc = ensure(c, &s.GX, &cg, n))) return rc;
        cg = 0;
        if ((rc = ensure(c, &s.GZ, &cg, n))) return rc;
        HIPCHK(c, hipMemcpy(s.GX, gx, n * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(s.GZ, gz, n * sizeof(float), hipMemcpyHostToDevice));
    } else {
        dfree(s.GX);
        dfree(s.GZ);
        s.GX = s.GZ = nullptr;
    }
    s.npos = npos;
    s.o = o_pos;
    s.d = d_pos;
    s.set = true;
    c->geo_version++;
    return KMIG_OK;
}

int kmig_set_data(kmig_ctx *c, const float *data, int nt, float t0, float dt, int nh, float h0,
                  float dh, int ns, float s0, float ds)
{
    if (!c) return KMIG_E_PARAM;
    if (!c->stream) return fail(c, KMIG_E_STATE, "context has no device");
    if (!data || nt < 2 || nh <= 0 || ns <= 0)
        return fail(c, KMIG_E_DIM, "Need 3D data volume for seismic_data= (t,h/r,s)");
    char err[512] = {0};
    int rc = kmig_params_check(&c->p, ns, err, sizeof(err));
    if (rc) return fail(c, rc, "%s", err);
    const size_t n = (size_t)nt * nh * ns, npad = trace_alloc((size_t)nh * ns, nt);
    size_t cap = c->raw_cap;
    // a run still in flight may read the traces being replaced
    HIPCHK(c, hipStreamSynchronize(c->stream2));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = ensure(c, &c->raw, &c->raw_cap, npad))) return rc;
    if ((rc = ensure(c, &c->prep, &cap, npad))) return rc;
    HIPCHK(c, hipMemcpy(c->raw, data, n * sizeof(float), hipMemcpyHostToDevice));
    c->nt = nt;
    c->t0 = t0;
    c->dt = dt;
    c->nh = nh;
    c->h0 = h0;
    c->dh = dh;
    c->ns = ns;
    c->s0 = s0;
    c->ds = ds;
    c->data_set = true;
    c->geo_version++;
    return KMIG_OK;
}

// Arguments of the stream migrate kernel for one shot batch.
static MigArgs stream_args(const kmig_ctx *c, const Batch &b)
{
    const kmig_params &p = c->p;
    const bool angle = p.cig && p.cig_is_angle;
    MigArgs M{};
    M.nz = c->nz;
    M.nx = c->nx;
    M.nzx = c->nzx;
    M.nt = c->nt;
    M.t0 = c->t0;
    M.dt = c->dt;
    M.inv_dt = 1.0f / c->dt;
    M.nh = c->nh;
    M.shot0 = b.s0;
    M.nshots = b.s1 - b.s0;
    M.sT = c->ts.T;
    M.sA = c->ts.A;
    M.sGX = c->ts.GX;
    M.sGZ = c->ts.GZ;
    M.sW = c->ts.W;
    M.rT = c->tr.T;
    M.rA = c->tr.A;
    M.rGX = c->tr.GX;
    M.rGZ = c->tr.GZ;
    M.rW = c->tr.W;
    M.meta = c->meta;
    M.traces = c->prep;
    M.cig = c->cig;
    M.tau = p.tau;
    M.aal = p.antialias;
    M.mode = angle ? (p.angle_interp ? 0 : 1) : 2;
    M.half_angle = p.half_angle;
    M.g0 = p.cig ? p.g0 : 0.0f;
    M.dg = p.cig ? p.dg : 1.0f;
    M.oazmin = p.oazmin;
    M.oazmax = p.oazmax;
    M.oaztaper = p.oaztaper;
    M.taper_start = p.angle_taper_start;
    M.taper_end = p.angle_taper_end;
    M.ng = p.cig ? p.ng : 1;
    return M;
}

// The resident migrate launch (and its partial sums) of batch b on stream
// st; tune = autotune first where the preset is automatic (never under graph
// capture: the autotuner synchronises).
static int enqueue_resident(kmig_ctx *c, const Batch &b, int ng, bool tune, hipStream_t st)
{
    const kmig_params &p = c->p;
    const int nsb = b.s1 - b.s0;
    int rc;
    MigResArgs R = res_args(c, b.s0, nsb, ng);
    // KMIG_SPLIT forces a shot-group count (tests): then short shot ranges
    // split too, so their partial sums are in the checked path
    const char *fs = getenv("KMIG_SPLIT");
    if (nsb >= 16 || (fs && atoi(fs) > 1 && nsb >= 2)) {
        // shot groups for the launch's tail (the launcher picks the count)
        if ((rc = ensure(c, &c->part, &c->part_cap, (size_t)KMIG_MAX_SPLIT * ng * c->nzx)))
            return rc;
        R.part = c->part;
        R.part_splits = KMIG_MAX_SPLIT;
    }
    int tile = p.tile;
    if (tile <= 0) {
        if (tune && (rc = autotune(c, R))) return rc;
        tile = c->tuned_tile;
    }
    HIPCHK(c, (hipError_t)launch_migrate_resident(R, tile, p.precision, st));
    return KMIG_OK;
}

// What a captured step depends on besides the buffers of its plan: the shot
// range, the preset, the QC shot and the switches read while enqueueing.
static std::string graph_key(const kmig_ctx *c, int sb, int se)
{
    char buf[256];
    const char *e1 = getenv("KMIG_CULL"), *e2 = getenv("KMIG_SPLIT"),
               *e3 = getenv("KMIG_TUPLE_FUSED");
    snprintf(buf, sizeof(buf), "%lld %d %d %d %d %d|%s|%s|%s", c->plan_version, sb, se,
             c->p.tile > 0 ? c->p.tile : c->tuned_tile, c->qc_valid ? 1 : 0,
             c->variant_used, e1 ? e1 : "", e2 ? e2 : "", e3 ? e3 : "");
    return buf;
}

// Capture one resident step over shots [sb, se) as a graph: the trace
// preparation on stream2 as one branch (forked from and joined back into the
// context stream by events), the tuple preparation on the context stream as
// the other, then the migrate launch and its partial sums.  The capture
// executes nothing; the graph is instantiated for replays.
static int capture_step(kmig_ctx *c, int sb, int se, int ng)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->gkey.clear();
    if (!c->gfork) HIPCHK(c, hipEventCreateWithFlags(&c->gfork, hipEventDisableTiming));
    if (!c->gjoin) HIPCHK(c, hipEventCreateWithFlags(&c->gjoin, hipEventDisableTiming));
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = KMIG_OK;
    hipError_t e = hipEventRecord(c->gfork, c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->stream2, c->gfork, 0);
    if (e == hipSuccess) rc = enqueue_trace_prep(c, c->stream2, sb, se);
    if (e == hipSuccess && rc == KMIG_OK) e = hipEventRecord(c->gjoin, c->stream2);
    if (e == hipSuccess && rc == KMIG_OK) rc = enqueue_tuple_prep(c, c->stream, c->batches[0], 0);
    if (e == hipSuccess && rc == KMIG_OK) e = hipStreamWaitEvent(c->stream, c->gjoin, 0);
    if (e == hipSuccess && rc == KMIG_OK && c->batches[0].s1 > c->batches[0].s0)
        rc = enqueue_resident(c, c->batches[0], ng, false, c->stream);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || rc != KMIG_OK || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != KMIG_OK) return rc;
        return fail(c, KMIG_E_DEVICE, "graph capture of the step failed: %s",
                    hipGetErrorString(e != hipSuccess ? e : ec));
    }
    c->graph = g;
    HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    c->gkey = graph_key(c, sb, se);
    return KMIG_OK;
}

int kmig_run(kmig_ctx *c, int sb, int se)
{
    if (!c) return KMIG_E_PARAM;
    if (!c->stream) return fail(c, KMIG_E_STATE, "context has no device");
    if (!c->side[0].set || !c->side[1].set)
        return fail(c, KMIG_E_STATE, "source and receiver tables must be set before kmig_run");
    if (!c->data_set) return fail(c, KMIG_E_STATE, "data must be set before kmig_run");
    if (c->p.adj != 1)
        return fail(c, KMIG_E_PARAM, "kmig_run migrates (adj=1); an adj=0 context models with "
                                     "kmig_model");
    const kmig_params &p = c->p;
    const bool angle = p.cig && p.cig_is_angle;
    const bool gfiles = c->side[0].GX != nullptr;
    if (angle && gfiles != (c->side[1].GX != nullptr))
        return fail(c, KMIG_E_PARAM,
                    "For mode=angle, please provide all of sgradx=, sgradz=, rgradx=, and "
                    "rgradz=, or provide none of them.");
    if (sb < 0) sb = 0;
    if (se > c->ns || se < 0) se = c->ns;
    if (se < sb) se = sb;
    int rc;
    const int ng = p.cig ? p.ng : 1;
    c->ev_used = 0;
    c->ev_kind.clear();

    if ((rc = plan_if_needed(c, sb, se))) return rc;
    const bool resident = c->variant_used == KMIG_VARIANT_RESIDENT;
    if ((rc = ensure_outputs(c, sb, se))) return rc;
    // Captured step: a resident run with one batch and no surrogate, whose
    // plan was captured after an eager run of it, replays the graph.
    const bool graphable = c->graph_on && resident && c->batches.size() == 1 &&
                           !c->side[0].nn && !c->side[1].nn;
    if (graphable && c->gexec && c->gkey == graph_key(c, sb, se)) {
        hipEvent_t g0 = next_event(c), g1 = next_event(c);
        HIPCHK(c, hipEventRecord(g0, c->stream));
        HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        HIPCHK(c, hipEventRecord(g1, c->stream));
        // the phase times stay those of the eager run the graph was captured
        // from; the total is the replay's
        c->st.graph = 1;
        c->pending = true;
        c->ran = true;
        return KMIG_OK;
    }
    // The trace preparation (data) runs on stream2 while the tuple preparation
    // (tables) runs on the context stream; the first migrate launch waits for
    // both.  stream2 starts after everything earlier on the context stream,
    // so a new run never rewrites traces an earlier launch still reads.
    hipEvent_t e0 = next_event(c);
    HIPCHK(c, hipEventRecord(e0, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->stream2, e0, 0));
    if ((rc = enqueue_trace_prep(c, c->stream2, sb, se))) return rc;
    hipEvent_t e1 = next_event(c);
    HIPCHK(c, hipEventRecord(e1, c->stream2));
    c->ev_kind.push_back(0);
    bool joined = false;

    size_t spec_s_off = 0;
    int launches = 0;
    c->n_direct = 0;
    for (const Batch &b : c->batches) {
        const int nsb = b.s1 - b.s0;
        hipEvent_t a0 = next_event(c);
        HIPCHK(c, hipEventRecord(a0, c->stream));
        if ((rc = enqueue_tuple_prep(c, c->stream, b, spec_s_off))) return rc;
        hipEvent_t a1 = next_event(c);
        HIPCHK(c, hipEventRecord(a1, c->stream));
        c->ev_kind.push_back(1);

        const MigArgs M = stream_args(c, b);
        int tile = p.tile;
        if (tile <= 0) {
            tile = resident ? 9 : 2;  // 16 x 8 points
            if (migrate_lds_bytes(ng, tile) > 64 * 1024) tile = resident ? 15 : 1;
        }
        if (migrate_lds_bytes(ng, tile) > 160 * 1024)
            return fail(c, KMIG_E_PARAM, "output axis of %d samples does not fit in LDS", ng);
        if (!joined) {
            HIPCHK(c, hipStreamWaitEvent(c->stream, e1, 0));
            joined = true;
        }
        if (nsb > 0 && resident) {
            if ((rc = enqueue_resident(c, b, ng, true, c->stream))) return rc;
            launches++;
        } else if (nsb > 0) {
            HIPCHK(c, (hipError_t)launch_migrate(M, tile, c->stream));
            launches++;
        }
        hipEvent_t m1 = next_event(c);
        HIPCHK(c, hipEventRecord(m1, c->stream));
        c->ev_kind.push_back(2);
        spec_s_off += (size_t)nsb;
    }
    if (!joined) HIPCHK(c, hipStreamWaitEvent(c->stream, e1, 0));
    c->st = kmig_stats{};
    c->st.updates = (double)(se - sb) * (double)c->nh * (double)c->nzx;
    c->st.migrate_launches = launches;
    c->st.shot_batches = (int)c->batches.size();
    c->st.variant = c->variant_used;
    c->st.tile = resident ? (p.tile > 0 ? p.tile : c->tuned_tile) : p.tile;
    int maxr = 0;
    for (const Batch &b : c->batches) maxr = std::max(maxr, b.nrcv);
    c->st.receiver_slots = maxr;
    c->st.tuple_direct = c->n_direct;
    c->pending = true;
    c->ran = true;
    // the next run with this plan replays a graph of this step
    if (graphable && (rc = capture_step(c, sb, se, ng))) return rc;
    return KMIG_OK;
}

int kmig_set_graph(kmig_ctx *c, int enable)
{
    if (!c) return KMIG_E_PARAM;
    c->graph_on = enable != 0;
    if (!c->graph_on && c->gexec) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipGraphExecDestroy(c->gexec);
        (void)hipGraphDestroy(c->graph);
        c->gexec = nullptr;
        c->graph = nullptr;
        c->gkey.clear();
    }
    return KMIG_OK;
}

int kmig_sync(kmig_ctx *c)
{
    if (!c) return KMIG_E_PARAM;
    if (!c->stream) return fail(c, KMIG_E_STATE, "context has no device");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->pending && c->st.graph && c->ev_used >= 2) {
        // a replayed graph: events around the launch give the step's time
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->events[0], c->events[1]);
        c->st.ms_total = ms;
        c->pending = false;
    } else if (c->pending && c->ev_used >= 2) {
        // events: e0, e1 (trace prep), then per batch a0, a1 (tuples), m1 (migrate)
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->events[0], c->events[1]);
        c->st.ms_trace_prep = ms;
        // then per batch: a0 .. a1 (tuple prep), a1 .. m1 (migrate)
        int k = 2;
        double tup = 0.0, mig = 0.0;
        for (size_t b = 0; b < c->batches.size() && k + 2 < c->ev_used; b++, k += 3) {
            float a = 0.0f, m = 0.0f;
            (void)hipEventElapsedTime(&a, c->events[k], c->events[k + 1]);
            (void)hipEventElapsedTime(&m, c->events[k + 1], c->events[k + 2]);
            tup += a;
            mig += m;
        }
        c->st.ms_tuple_prep = tup;
        c->st.ms_migrate = mig;
        (void)hipEventElapsedTime(&ms, c->events[0], c->events[c->ev_used - 1]);
        c->st.ms_total = ms;
        c->pending = false;
    }
    return KMIG_OK;
}

int kmig_get_cig(kmig_ctx *c, float *cig)
{
    if (!c || !cig) return KMIG_E_PARAM;
    if (!c->ran) return fail(c, KMIG_E_STATE, "kmig_run has not been called");
    int rc = kmig_sync(c);
    if (rc) return rc;
    const int ng = c->p.cig ? c->p.ng : 1;
    HIPCHK(c, hipMemcpy(cig, c->cig, (size_t)ng * c->nzx * sizeof(float), hipMemcpyDeviceToHost));
    return KMIG_OK;
}

int kmig_get_qc(kmig_ctx *c, float *before, float *after)
{
    if (!c) return KMIG_E_PARAM;
    if (!c->qc_valid) return fail(c, KMIG_E_STATE, "no direct-mute QC gather in the last run");
    int rc = kmig_sync(c);
    if (rc) return rc;
    const size_t n = (size_t)c->nt * c->nh * sizeof(float);
    if (before) HIPCHK(c, hipMemcpy(before, c->qcb, n, hipMemcpyDeviceToHost));
    if (after) HIPCHK(c, hipMemcpy(after, c->qca, n, hipMemcpyDeviceToHost));
    return KMIG_OK;
}

int kmig_get_stats(const kmig_ctx *c, kmig_stats *st)
{
    if (!c || !st) return KMIG_E_PARAM;
    *st = c->st;
    return KMIG_OK;
}

int kmig_count_updates(kmig_ctx *c, double *nominal, double *evaluated, double *accepted)
{
    if (!c) return KMIG_E_PARAM;
    if (!c->ran || c->last_sb < 0) return fail(c, KMIG_E_STATE, "kmig_run has not been called");
    if (c->variant_used != KMIG_VARIANT_RESIDENT || c->batches.size() != 1 ||
        c->p.precision != KMIG_PREC_FP32)
        return fail(c, KMIG_E_PARAM,
                    "update counts are kept for the resident kernel with FP32 tuples");
    int rc = kmig_sync(c);
    if (rc) return rc;
    if ((rc = ensure(c, &c->cnt, &c->cnt_cap, 2))) return rc;
    const Batch &b = c->batches[0];
    const int nsb = b.s1 - b.s0;
    const int ng = c->p.cig ? c->p.ng : 1;
    MigResArgs R = res_args(c, b.s0, nsb, ng);
    const char *fs = getenv("KMIG_SPLIT");
    if (c->part && (nsb >= 16 || (fs && atoi(fs) > 1 && nsb >= 2)))
        R.part_splits = KMIG_MAX_SPLIT;  // the production launch's shot groups
    R.counts = c->cnt;
    HIPCHK(c, (hipError_t)launch_fill((float *)c->cnt, 4, 0.0f, c->stream));
    if (nsb > 0)
        HIPCHK(c, (hipError_t)launch_count_resident(R, c->st.tile, c->stream));
    unsigned long long h[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(h, c->cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (nominal) *nominal = (double)nsb * (double)c->nh * (double)c->nzx;
    if (evaluated) *evaluated = (double)h[0];
    if (accepted) *accepted = (double)h[1];
    return KMIG_OK;
}

int kmig_model(kmig_ctx *c, const float *image, int nt, float t0, float dt, int nh, float h0,
               float dh, int ns, float s0, float ds, float *data)
{
    if (!c || !image || !data) return KMIG_E_PARAM;
    if (!c->stream) return fail(c, KMIG_E_STATE, "context has no device");
    const kmig_params &p = c->p;
    if (p.adj != 0) return fail(c, KMIG_E_PARAM, "kmig_model needs adj=0");
    if (!c->side[0].set || !c->side[1].set)
        return fail(c, KMIG_E_STATE, "source and receiver tables must be set before kmig_model");
    if (nt < 2 || nh <= 0 || ns <= 0)
        return fail(c, KMIG_E_DIM, "adj=0 needs nt>1, nh>0 and ns>0 for the data volume");
    char err[512] = {0};
    int rc = kmig_params_check(&p, ns, err, sizeof(err));
    if (rc) return fail(c, rc, "%s", err);
    const bool angle = p.cig && p.cig_is_angle;
    if (angle && (c->side[0].GX != nullptr) != (c->side[1].GX != nullptr))
        return fail(c, KMIG_E_PARAM,
                    "For mode=angle, please provide all of sgradx=, sgradz=, rgradx=, and "
                    "rgradz=, or provide none of them.");
    // the data axes of the adj=0 keys (src/kirchmig2d_cig.cpp:663-701) and the
    // trace buffers: raw takes the spray, prep the doubly integrated traces
    const size_t n = (size_t)nt * nh * ns, npad = trace_alloc((size_t)nh * ns, nt);
    size_t cap = c->raw_cap;
    if ((rc = ensure(c, &c->raw, &c->raw_cap, npad))) return rc;
    if ((rc = ensure(c, &c->prep, &cap, npad))) return rc;
    c->nt = nt;
    c->t0 = t0;
    c->dt = dt;
    c->nh = nh;
    c->h0 = h0;
    c->dh = dh;
    c->ns = ns;
    c->s0 = s0;
    c->ds = ds;
    c->data_set = false;  // no input traces: kmig_run stays refused
    c->geo_version++;
    if ((rc = plan_if_needed(c, 0, ns))) return rc;
    if (p.precision != KMIG_PREC_FP32)
        return fail(c, KMIG_E_PARAM, "adj=0 on the GPU computes with precision=fp32 tables");
    const bool resident = c->variant_used == KMIG_VARIANT_RESIDENT;
    if ((rc = ensure_outputs(c, 0, ns))) return rc;
    const int ng = p.cig ? p.ng : 1;
    HIPCHK(c, hipMemcpy(c->cig, image, (size_t)ng * c->nzx * sizeof(float),
                        hipMemcpyHostToDevice));
    HIPCHK(c, (hipError_t)launch_fill(c->raw, n, 0.0f, c->stream));
    c->ev_used = 0;
    c->ev_kind.clear();
    hipEvent_t e0 = next_event(c);
    HIPCHK(c, hipEventRecord(e0, c->stream));
    size_t spec_s_off = 0;
    // resident presets 9 (16 x 8), 11, 15 (8 x 8); stream tiles 2 (16 x 8), 1 (8 x 8)
    int tile = resident ? (p.tile == 11 || p.tile == 15 ? p.tile : 9) : (p.tile == 1 ? 1 : 2);
    if (migrate_lds_bytes(ng, tile) > 64 * 1024) tile = resident ? 15 : 1;
    if (migrate_lds_bytes(ng, tile) > 160 * 1024)
        return fail(c, KMIG_E_PARAM, "output axis of %d samples does not fit in LDS", ng);
    for (const Batch &b : c->batches) {
        const int nsb = b.s1 - b.s0;
        if ((rc = enqueue_tuple_prep(c, c->stream, b, spec_s_off))) return rc;
        if (nsb > 0 && resident) {
            MigResArgs R = res_args(c, b.s0, nsb, ng);
            R.spray = c->raw;
            HIPCHK(c, (hipError_t)launch_model_resident(R, tile, c->stream));
        } else if (nsb > 0) {
            MigArgs M = stream_args(c, b);
            M.spray = c->raw;
            HIPCHK(c, (hipError_t)launch_model(M, tile, c->stream));
        }
        spec_s_off += (size_t)nsb;
    }
    // doubint of every trace (:1341-1349): the trace kernels with no mute,
    // no diff and no tap pairs
    TracePrepArgs ta{};
    ta.raw = c->raw;
    ta.prep = c->prep;
    ta.meta = c->meta;
    ta.nt = nt;
    ta.nh = nh;
    ta.shot_begin = 0;
    ta.shot_end = ns;
    ta.t0 = t0;
    ta.dt = dt;
    ta.qc_shot = -1;
    HIPCHK(c, (hipError_t)launch_trace_prep2(ta, nullptr, nullptr, c->stream));
    hipEvent_t e1 = next_event(c);
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipMemcpyAsync(data, c->prep, n * sizeof(float), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    c->st = kmig_stats{};
    c->st.updates = (double)ns * (double)nh * (double)c->nzx;
    c->st.migrate_launches = (int)c->batches.size();
    c->st.shot_batches = (int)c->batches.size();
    c->st.variant = c->variant_used;
    c->st.tile = tile;
    c->st.ms_total = ms;
    c->st.ms_migrate = ms;
    c->ev_used = 0;
    return KMIG_OK;
}

int kmig_migrate(kmig_ctx *c, const float *data, int nt, float t0, float dt, int nh, float h0,
                 float dh, int ns, float s0, float ds, float *cig, float *qc_before,
                 float *qc_after)
{
    int rc = kmig_set_data(c, data, nt, t0, dt, nh, h0, dh, ns, s0, ds);
    if (rc) return rc;
    if ((rc = kmig_run(c, 0, ns))) return rc;
    if ((rc = kmig_sync(c))) return rc;
    if (cig && (rc = kmig_get_cig(c, cig))) return rc;
    if ((qc_before || qc_after) && c->qc_valid) rc = kmig_get_qc(c, qc_before, qc_after);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------------ batches --

struct kmig_batch {
    std::vector<kmig_ctx *> ctx;
    std::vector<long long> versions;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    size_t nodes = 0;
    std::string err;
    double updates = 0.0, ms = 0.0;
    bool pending = false;
    int tile = 0;                              // preset of the captured graph
    std::vector<std::pair<int, float>> tune_log;
};

namespace {

int bfail(kmig_batch *b, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    b->err = buf;
    return code;
}

#define BHIP(b, call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return bfail((b), KMIG_E_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Build (or rebuild) the captured graph of one batch step with migrate
// preset `tile` (9, 10, 15 or 13).
int batch_capture(kmig_batch *b, int tile)
{
    const int n = (int)b->ctx.size();
    MigMultiArgs m{};
    m.nprob = n;
    // heaviest problems first (longest blocks start first)
    std::vector<int> order(n);
    for (int k = 0; k < n; k++) order[k] = k;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        const kmig_ctx *a = b->ctx[x], *c = b->ctx[y];
        return (double)a->ns * a->nh > (double)c->ns * c->nh;
    });
    const kmig_params &p0 = b->ctx[0]->p;
    const int ng = p0.ng;
    m.tile0[0] = 0;
    b->updates = 0.0;
    for (int q = 0; q < n; q++) {
        kmig_ctx *c = b->ctx[order[q]];
        m.prob[q] = res_args(c, 0, c->ns, ng);
        if (c->ns >= 16) {  // shot groups (the launcher picks the count)
            const int rc0 = ensure(c, &c->part, &c->part_cap, (size_t)KMIG_MAX_SPLIT * ng * c->nzx);
            if (rc0) return bfail(b, rc0, "%s", c->err.c_str());
            m.prob[q].part = c->part;
            m.prob[q].part_splits = KMIG_MAX_SPLIT;
        }
        m.tile0[q + 1] = m.tile0[q] + migrate_tiles(c->nz, c->nx, tile);
        b->updates += (double)c->ns * c->nh * (double)c->nzx;
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = KMIG_OK;
    for (int k = 0; k < n && rc == KMIG_OK; k++) {
        kmig_ctx *c = b->ctx[k];
        rc = enqueue_trace_prep(c, b->stream, 0, c->ns);
        if (rc == KMIG_OK) rc = enqueue_tuple_prep(c, b->stream, c->batches[0], 0);
        if (rc != KMIG_OK) b->err = c->err;
    }
    if (rc == KMIG_OK) {
        const int e = launch_migrate_multi(m, tile, p0.precision, b->stream);
        if (e) rc = bfail(b, KMIG_E_DEVICE, "multi-problem migrate launch failed: %s",
                          hipGetErrorString((hipError_t)e));
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (rc != KMIG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess)
        return bfail(b, KMIG_E_DEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
    b->graph = g;
    BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->nodes = 0;
    (void)hipGraphGetNodes(b->graph, nullptr, &b->nodes);
    b->tile = tile;
    return KMIG_OK;
}

// Batch autotuner: captures the step with each multi-problem preset whose
// accumulator fits 64 KiB of LDS, replays each twice (HIP events on the
// batch stream) and keeps the fastest.  Every replay recomputes the CIGs
// from zero, so the timing replays leave no trace in the results.
int batch_tune(kmig_batch *b)
{
    static const int presets[] = {9, 10, 11, 15, 13};
    const int ng = b->ctx[0]->p.ng;
    b->tune_log.clear();
    int best = 0;
    float tbest = 1e30f;
    for (int t : presets) {
        if (migrate_lds_bytes(ng, t) > 64 * 1024) continue;
        int rc = batch_capture(b, t);
        if (rc) return rc;
        float tmin = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            BHIP(b, hipEventRecord(b->e0, b->stream));
            BHIP(b, hipGraphLaunch(b->exec, b->stream));
            BHIP(b, hipEventRecord(b->e1, b->stream));
            BHIP(b, hipEventSynchronize(b->e1));
            float ms = 0.0f;
            BHIP(b, hipEventElapsedTime(&ms, b->e0, b->e1));
            tmin = std::min(tmin, ms);
        }
        b->tune_log.push_back({t, tmin});
        if (tmin < tbest) {
            tbest = tmin;
            best = t;
        }
    }
    if (best == 0) best = 15;  // nothing fits 64 KiB: the smallest tile, more LDS
    if (getenv("KMIG_TUNE_LOG")) {
        fprintf(stderr, "kmig batch autotune:");
        for (auto &q : b->tune_log) fprintf(stderr, " tile%d=%.2fms", q.first, q.second);
        fprintf(stderr, " -> tile%d\n", best);
    }
    return b->tile == best ? KMIG_OK : batch_capture(b, best);
}

}  // namespace

extern "C" {

int kmig_batch_create(kmig_batch **out, kmig_ctx *const *ctx, int n)
{
    if (!out) return KMIG_E_PARAM;
    *out = nullptr;
    kmig_batch *b = new kmig_batch();
    *out = b;
    if (!ctx || n <= 0 || n > KMIG_MAX_PROBLEMS)
        return bfail(b, KMIG_E_PARAM, "a batch holds 1..%d contexts", KMIG_MAX_PROBLEMS);
    for (int k = 0; k < n; k++) {
        if (!ctx[k] || !ctx[k]->stream)
            return bfail(b, KMIG_E_STATE, "context %d has no device", k);
        b->ctx.push_back(ctx[k]);
    }
    BHIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BHIP(b, hipEventCreate(&b->e0));
    BHIP(b, hipEventCreate(&b->e1));
    return KMIG_OK;
}

int kmig_batch_run(kmig_batch *b)
{
    if (!b) return KMIG_E_PARAM;
    if (!b->stream) return bfail(b, KMIG_E_STATE, "batch has no stream");
    const int n = (int)b->ctx.size();
    bool changed = b->exec == nullptr || (int)b->versions.size() != n;
    b->versions.resize(n, -2);
    for (int k = 0; k < n; k++) {
        kmig_ctx *c = b->ctx[k];
        if (!c->side[0].set || !c->side[1].set || !c->data_set)
            return bfail(b, KMIG_E_STATE, "context %d: tables and data must be set", k);
        int rc = plan_if_needed(c, 0, c->ns);
        if (rc == KMIG_OK) rc = ensure_outputs(c, 0, c->ns);
        if (rc) return bfail(b, rc, "context %d: %s", k, c->err.c_str());
        if (c->variant_used != KMIG_VARIANT_RESIDENT || c->batches.size() != 1 ||
            c->rcv_columns)
            return bfail(b, KMIG_E_PARAM,
                         "context %d: batches need the resident kernel (mode=angle, "
                         "angle_interp=1, cmp=0) and tuples that fit one shot batch", k);
        const kmig_params &p = c->p, &q = b->ctx[0]->p;
        if (!p.cig || p.ng != q.ng || p.precision != q.precision)
            return bfail(b, KMIG_E_PARAM,
                         "context %d: every context of a batch needs the same angle axis "
                         "length and precision", k);
        if (b->versions[k] != c->plan_version) changed = true;
        b->versions[k] = c->plan_version;
    }
    if (changed) {
        BHIP(b, hipStreamSynchronize(b->stream));
        const int t = b->ctx[0]->p.tile;
        const int rc = t > 0 ? batch_capture(b, t) : batch_tune(b);
        if (rc) return rc;
    }
    BHIP(b, hipEventRecord(b->e0, b->stream));
    BHIP(b, hipGraphLaunch(b->exec, b->stream));
    BHIP(b, hipEventRecord(b->e1, b->stream));
    for (kmig_ctx *c : b->ctx) {
        c->ran = true;
        c->pending = false;
        c->st = kmig_stats{};
        c->st.updates = (double)c->ns * c->nh * (double)c->nzx;
        c->st.migrate_launches = 1;
        c->st.shot_batches = 1;
        c->st.variant = c->variant_used;
    }
    b->pending = true;
    return KMIG_OK;
}

int kmig_batch_sync(kmig_batch *b)
{
    if (!b) return KMIG_E_PARAM;
    BHIP(b, hipStreamSynchronize(b->stream));
    if (b->pending) {
        float ms = 0.0f;
        BHIP(b, hipEventElapsedTime(&ms, b->e0, b->e1));
        b->ms = ms;
        b->pending = false;
    }
    return KMIG_OK;
}

int kmig_batch_stats(const kmig_batch *b, double *updates, double *ms, int *graph_nodes)
{
    if (!b) return KMIG_E_PARAM;
    if (updates) *updates = b->updates;
    if (ms) *ms = b->ms;
    if (graph_nodes) *graph_nodes = (int)b->nodes;
    return KMIG_OK;
}

const char *kmig_batch_last_error(const kmig_batch *b) { return b ? b->err.c_str() : "null batch"; }

}  // extern "C"

namespace {

// The checks and messages of src/layercom_cig.cpp:245-305 and :424-426, and
// the stitch arguments of the batch's layer CIGs (all but the lateral
// weights).  g_count <= 0 selects every angle sample.
int stitch_setup(kmig_batch *b, const int *start, const int *end, const float *scale,
                 int stitch_blend, int scalex_aper, double scalex_scale, int *g_first,
                 int *g_count, StitchArgs *s)
{
    const int n = (int)b->ctx.size();
    if (stitch_blend < -1) return bfail(b, KMIG_E_PARAM, "stitch_blend must be -1 or nonnegative");
    if (scalex_aper < 0) return bfail(b, KMIG_E_PARAM, "scalex_aper must be nonnegative");
    if (!std::isfinite(scalex_scale)) return bfail(b, KMIG_E_PARAM, "Invalid scalex_scale");
    if (n < 1) return bfail(b, KMIG_E_PARAM, "No layers found");
    if (start[0] != 0) return bfail(b, KMIG_E_PARAM, "First layer must start at depth sample 0");
    for (int i = 0; i < n; i++) {
        if (start[i] < 0 || end[i] < start[i])
            return bfail(b, KMIG_E_PARAM, "Invalid layer range [%d,%d]", start[i], end[i]);
        if (i > 0 && start[i] > end[i - 1])
            return bfail(b, KMIG_E_PARAM, "Layer %d does not overlap previous layer", i + 1);
        if (i > 0 && end[i] <= end[i - 1])
            return bfail(b, KMIG_E_PARAM, "Layer %d must extend deeper than previous layer",
                         i + 1);
    }
    const kmig_ctx *c0 = b->ctx[0];
    const int nx = c0->nx, ng = c0->p.cig ? c0->p.ng : 1;
    if (*g_count <= 0) {
        *g_first = 0;
        *g_count = ng;
    }
    if (*g_first < 0 || *g_first + *g_count > ng)
        return bfail(b, KMIG_E_PARAM, "angle window [%d, %d) outside the %d angle samples",
                     *g_first, *g_first + *g_count, ng);
    *s = StitchArgs{};
    for (int i = 0; i < n; i++) {
        const kmig_ctx *c = b->ctx[i];
        const int ngi = c->p.cig ? c->p.ng : 1;
        if (!c->ran || !c->cig) return bfail(b, KMIG_E_STATE, "layer %d has no CIG yet", i + 1);
        if (c->nz != end[i] - start[i] + 1 || c->nx != nx || ngi != ng)
            return bfail(b, KMIG_E_DIM,
                         "layer %d dimensions are %d,%d,%d; expected %d,%d,%d", i + 1, c->nz,
                         c->nx, ngi, end[i] - start[i] + 1, nx, ng);
        s->layer[i] = c->cig;
        s->start[i] = start[i];
        s->end[i] = end[i];
        s->scale[i] = scale ? scale[i] : 1.0f;
    }
    s->nlayers = n;
    s->nz = end[n - 1] + 1;
    s->nx = nx;
    s->g0 = *g_first;
    s->ng_out = *g_count;
    s->stitch_blend = stitch_blend;
    return KMIG_OK;
}

}  // namespace

extern "C" {

int kmig_batch_stitch(kmig_batch *b, const int *start, const int *end, const float *scale,
                      int stitch_blend, int scalex_aper, double scalex_scale, int g_first,
                      int g_count, float *out)
{
    if (!b || !start || !end || !out) return b ? bfail(b, KMIG_E_PARAM, "null argument")
                                               : KMIG_E_PARAM;
    StitchArgs s{};
    int rc = stitch_setup(b, start, end, scale, stitch_blend, scalex_aper, scalex_scale, &g_first,
                          &g_count, &s);
    if (rc) return rc;
    const int nx = s.nx;
    // lateral edge scale (src/layercom_cig.cpp:445-456)
    std::vector<float> lat(nx, 1.0f);
    for (int i2 = 0; i2 < nx && scalex_aper > 0; i2++) {
        const int edge = std::min(i2, nx - 1 - i2);
        if (edge < scalex_aper) {
            const double taper =
                scalex_aper == 1 ? 1.0 : 1.0 - (double)edge / (double)(scalex_aper - 1);
            lat[i2] = (float)(1.0 + (scalex_scale - 1.0) * taper);
        }
    }
    const size_t nout = (size_t)s.nz * nx * g_count;
    float *dlat = nullptr, *dout = nullptr;
    unsigned *dbad = nullptr;
    BHIP(b, hipStreamSynchronize(b->stream));
    hipError_t e = hipMalloc((void **)&dlat, nx * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&dout, nout * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&dbad, sizeof(unsigned));
    if (e == hipSuccess) e = hipMemcpy(dlat, lat.data(), nx * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dbad, 0, sizeof(unsigned));
    s.lat = dlat;
    if (e == hipSuccess) e = (hipError_t)launch_stitch(s, dout, dbad, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    unsigned bad = 0;
    if (e == hipSuccess) e = hipMemcpy(&bad, dbad, sizeof(unsigned), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out, dout, nout * sizeof(float), hipMemcpyDeviceToHost);
    dfree(dlat);
    dfree(dout);
    dfree(dbad);
    if (e != hipSuccess)
        return bfail(b, KMIG_E_DEVICE, "stitch failed: %s", hipGetErrorString(e));
    if (bad) return bfail(b, KMIG_E_PARAM, "layers contain non-finite samples after scaling");
    return KMIG_OK;
}

int kmig_batch_stitch_noalloc(kmig_batch *b, const int *start, const int *end, const float *scale,
                              int stitch_blend, int scalex_aper, double scalex_scale, int g_first,
                              int g_count, float *out)
{
    if (!b || !start || !end || !out) return b ? bfail(b, KMIG_E_PARAM, "null argument")
                                               : KMIG_E_PARAM;
    StitchArgs s{};
    int rc = stitch_setup(b, start, end, scale, stitch_blend, scalex_aper, scalex_scale, &g_first,
                          &g_count, &s);
    if (rc) return rc;
    if (s.nx > KMIG_STITCH_NOALLOC_MAX_NX)
        return bfail(b, KMIG_E_PARAM,
                     "the allocation-free stitch takes at most %d lateral samples (n2 = %d); "
                     "use kmig_batch_stitch", KMIG_STITCH_NOALLOC_MAX_NX, s.nx);
    // the output must be device memory the kernel can write
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, out) != hipSuccess ||
        !(at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged)) {
        (void)hipGetLastError();  // clear the lookup's error state
        return bfail(b, KMIG_E_PARAM,
                     "the allocation-free stitch writes device memory: out is not a device "
                     "pointer");
    }
    if (at.type == hipMemoryTypeDevice && at.device != b->ctx[0]->device)
        return bfail(b, KMIG_E_PARAM, "out lies on device %d, the layers on device %d",
                     at.device, b->ctx[0]->device);
    // and it must hold the whole window: g_count x nx x nz floats
    const size_t nout = (size_t)s.nz * s.nx * g_count;
    void *abase = nullptr;
    size_t asize = 0;
    if (at.type == hipMemoryTypeDevice &&
        hipMemGetAddressRange(&abase, &asize, out) == hipSuccess &&
        (char *)(out + nout) > (char *)abase + asize)
        return bfail(b, KMIG_E_DIM,
                     "out holds %zu bytes past its address, the stitch writes %zu",
                     (size_t)((char *)abase + asize - (char *)out), nout * sizeof(float));
    BHIP(b, hipStreamSynchronize(b->stream));
    unsigned bad = 0;
    const int e = launch_stitch_noalloc(s, scalex_aper, scalex_scale, out, &bad, b->stream);
    if (e) return bfail(b, KMIG_E_DEVICE, "stitch failed: %s", hipGetErrorString((hipError_t)e));
    if (bad) return bfail(b, KMIG_E_PARAM, "layers contain non-finite samples after scaling");
    return KMIG_OK;
}

void kmig_batch_destroy(kmig_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
