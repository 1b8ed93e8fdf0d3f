// Host planning of the segmented butterfly (use_bf=1) for the operators of
// kirchdat2d_freq: the banded (cmp=0) ones, one matrix per side and
// frequency, and the pair-list (cmp=1) ones, one matrix per shot (receiver
// side, :882-935) or per common-receiver gather (source side, :975-1087).
//
// The reference factors each frequency's matrix K = amp exp(-i omega tau)
// (SERECKIRCH/src/bf1d.cpp:935-961): top panels of leaf 2^levels, each split
// adaptively (:879-933) until a panel is fully active, square, leaf 2^L
// (L >= 1) and its non-separable phase is at most phase_tol (:816-877) --
// then the panel is a butterfly -- or it is a leaf, evaluated exactly.  The
// split positions depend on sizes only, so the panels that can ever become
// butterflies are known before any frequency: this file enumerates them
// ("nodes"), with
//   * the frequencies their residual admits: the residual is
//     max |omega (tau(r, c) - tau(r, c1) - tau(r1, c) + tau(r1, c1))| over the
//     reference's 5 x 5 samples, and since rounding is monotone it equals
//     omega times the largest |difference|, so the test passes for every
//     frequency up to a last index;
//   * an upper bound from the antialias weight, which is zero from the first
//     frequency at or above 0.5 / |slope step| on (kirchdat2d_frequency_cs.cpp:
//     203-233), so a panel holding that key cannot be fully active there;
//   * the traveltime samples of every phase of its factor (bilinear, with
//     the exact-sample shortcut, bf1d.cpp:169-206), in double;
//   * its candidate ancestors: a node is the butterfly at a frequency only
//     if no panel containing it already was.
// Whether a node is fully active at a frequency (|w aa| >= 1e-20 and
// |amp|^2 > amp_eps^2 for every entry, :703-727) and so its final mask is
// decided on the GPU from the operator table.
//
// Arithmetic: built without contraction; positions, Lagrange values and
// samples in double, as the reference forms them.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <unordered_map>
#include <vector>

#include "kmdat_internal.h"
#include "kmig_dat.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace kmdat {

namespace {

using bf::Node;

// Boxes of one block of nloc = leaf 2^L points (bf1d.cpp:61-96): level l has
// 2^l boxes, entry (2^l - 1) + b.
struct Tree {
    int L = 0, p = 0;
    std::vector<int> i1;
    std::vector<double> x0, node;  // node: [box][p]
    void build(int nloc, int leaf, int p_)
    {
        p = p_;
        L = 0;
        while ((leaf << L) < nloc) L++;
        const int nbox = (2 << L) - 1;
        i1.assign(nbox, 0);
        x0.assign(nbox, 0.0);
        node.assign((size_t)nbox * p, 0.0);
        std::vector<double> cheb(p);
        for (int k = 0; k < p; k++)
            cheb[k] = std::cos(((2.0 * (double)(k + 1) - 1.0) * M_PI) / (2.0 * (double)p));
        for (int l = 0; l <= L; l++) {
            const int nb = 1 << l, size = nloc / nb;
            for (int b = 0; b < nb; b++) {
                const int e = nb - 1 + b;
                const double lo = (double)(b * size), hi = (double)((b + 1) * size - 1);
                const double c = 0.5 * (lo + hi), width = hi - lo;
                i1[e] = b * size;
                x0[e] = c;
                for (int k = 0; k < p; k++) node[(size_t)e * p + k] = c + 0.5 * width * cheb[k];
            }
        }
    }
    int box(int l, int b) const { return (1 << l) - 1 + b; }
    double nd(int l, int b, int k) const { return node[(size_t)box(l, b) * p + k]; }
    double ctr(int l, int b) const { return x0[box(l, b)]; }
};

// Barycentric Lagrange basis (bf1d.cpp:98-151).
void lagrange(int p, const double *nodes, double x, double *out)
{
    for (int j = 0; j < p; j++) out[j] = 0.0;
    for (int j = 0; j < p; j++)
        if (std::fabs(x - nodes[j]) < 1e-13) {
            out[j] = 1.0;
            return;
        }
    double sum = 0.0;
    for (int j = 0; j < p; j++) {
        double wj = 1.0;
        for (int k = 0; k < p; k++)
            if (k != j) wj /= (nodes[j] - nodes[k]);
        out[j] = wj / (x - nodes[j]);
        sum += out[j];
    }
    for (int j = 0; j < p; j++) out[j] /= sum;
}

// Bilinear sample position (bf1d.cpp:169-206): the exact-sample shortcut,
// clamps, four entries and their weights in the reference's order.
bf::SwSample bilinear_at(int n, double x, double z)
{
    bf::SwSample s{};
    x = std::min(std::max(x, 0.0), (double)(n - 1));
    z = std::min(std::max(z, 0.0), (double)(n - 1));
    const long jr = std::lround(x), ir = std::lround(z);
    if (std::fabs(x - (double)jr) < 1e-12 && std::fabs(z - (double)ir) < 1e-12) {
        for (int q = 0; q < 4; q++) s.e[q] = (int)(ir * n + jr);
        s.w[0] = 1.0;
        return s;
    }
    int j = (int)std::floor(x), i = (int)std::floor(z);
    j = std::max(0, std::min(j, n - 2));
    i = std::max(0, std::min(i, n - 2));
    const double fx = x - (double)j, fz = z - (double)i;
    s.e[0] = i * n + j;
    s.e[1] = i * n + j + 1;
    s.e[2] = (i + 1) * n + j;
    s.e[3] = (i + 1) * n + j + 1;
    s.w[0] = (1.0 - fx) * (1.0 - fz);
    s.w[1] = fx * (1.0 - fz);
    s.w[2] = (1.0 - fx) * fz;
    s.w[3] = fx * fz;
    return s;
}

// The traveltime sample at a position: the exact sample, or the bilinear
// sum evaluated as the reference writes it (bf1d.cpp:202-206).
double sample(const std::vector<float> &t, int n, double x, double z)
{
    const bf::SwSample s = bilinear_at(n, x, z);
    if (s.w[1] == 0.0 && s.w[2] == 0.0 && s.w[3] == 0.0 && s.e[0] == s.e[3])
        return (double)t[(size_t)s.e[0]];
    return s.w[0] * (double)t[(size_t)s.e[0]] + s.w[1] * (double)t[(size_t)s.e[1]] +
           s.w[2] * (double)t[(size_t)s.e[2]] + s.w[3] * (double)t[(size_t)s.e[3]];
}

bool pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

// One matrix as the planner reads it: the table keys of a row's column span,
// or false when a column of the span has no summation pair; pair0 receives
// the pair index of the span's first column (pair lists) or -1.
struct RowSource {
    virtual ~RowSource() = default;
    virtual bool span(int r, int c0, int n, int *keys, int *pair0) const = 0;
    virtual int out_trace(int r) const = 0;
};

struct BandedRows : RowSource {
    const BfPlanIn &in;
    explicit BandedRows(const BfPlanIn &i) : in(i) {}
    bool span(int r, int c0, int n, int *keys, int *pair0) const override
    {
        *pair0 = -1;
        for (int c = 0; c < n; c++) {
            const int j = c0 + c - r + in.aper;
            if (j < 0 || j >= in.nb) return false;
            keys[c] = in.pkey[(size_t)r * in.nb + j];
            if (keys[c] < 0) return false;
        }
        return true;
    }
    int out_trace(int) const override { return -1; }
};

// Rows of one instance of a pair-list operator: row r holds the pairs
// [q0, q1) of output trace o, in increasing column.
struct PairRows : RowSource {
    const BfPlanIn &in;
    int inst;
    PairRows(const BfPlanIn &i, int k) : in(i), inst(k) {}
    int col(int q) const
    {
        const int tr = in.pairs[q].tr;
        return in.col_by_div ? tr / in.col_base : tr % in.col_base;
    }
    bool span(int r, int c0, int n, int *keys, int *pair0) const override
    {
        const size_t e = (size_t)inst * in.n + r;
        if (in.row_o[e] < 0) return false;
        int lo = in.row_q0[e], hi = in.row_q1[e];
        while (lo < hi) {  // first pair of column >= c0
            const int m = (lo + hi) / 2;
            if (col(m) < c0)
                lo = m + 1;
            else
                hi = m;
        }
        if (lo + n > in.row_q1[e]) return false;
        for (int c = 0; c < n; c++) {
            if (col(lo + c) != c0 + c) return false;
            keys[c] = in.pairs[lo + c].key;
        }
        *pair0 = lo;
        return true;
    }
    int out_trace(int r) const override { return in.row_o[(size_t)inst * in.n + r]; }
};

struct Planner {
    const BfPlanIn &in;
    BfPlan &out;
    std::vector<float> omega;   // [nw]
    std::vector<int> cut;       // first frequency with a zero antialias weight, per key
    std::vector<Tree> trees;    // per L
    std::string err;
    const RowSource *src = nullptr;
    int inst = 0;

    Planner(const BfPlanIn &i, BfPlan &o) : in(i), out(o) {}

    // antialias_weight_iw (kirchdat2d_frequency_cs.cpp:203-233) is zero for
    // every iw >= the returned index (nw when it never is)
    int aa_cut(int k) const
    {
        if (!(in.antialias > 0.0f)) return in.nw;
        const float dtau = std::fabs(in.aarg[k]) * in.antialias;
        if (!(dtau > 1e-12f) || !std::isfinite(dtau)) return in.nw;
        const float fnyq = 0.5f / in.dt, fmax = 0.5f / dtau;
        if (!(fmax < fnyq)) return in.nw;
        if (fmax <= 0.0f) return 1;
        // the first iw >= 1 with iw / (nfft dt) >= fmax in float (monotone in
        // iw): start below the real-valued answer and step up
        const float den = (float)in.nfft * in.dt;
        int iw = std::max(1, (int)std::floor((double)fmax * (double)den) - 2);
        while (iw > 1 && (float)(iw - 1) / den >= fmax) iw--;
        while (iw < in.nw && (float)iw / den < fmax) iw++;
        return iw;
    }

    void walk(int r0, int c0, int rows, int cols, std::vector<int> &anc)
    {
        if (rows <= 0 || cols <= 0) return;
        const int leaf = in.leaf;
        const bool legal = rows == cols && rows >= 2 * leaf && rows % leaf == 0 &&
                           pow2(rows / leaf);
        bool pushed = false;
        if (legal) {
            const int n = rows;
            int L = 0;
            while ((leaf << L) < n) L++;
            // every entry has a summation pair (inside the band, weight
            // nonzero) and a positive traveltime (else H = 0 and the entry is
            // inactive at all frequencies)
            std::vector<int> keys((size_t)n * n), pair0(n);
            bool full = true;
            int hi = in.nw - 1;
            for (int r = 0; r < n && full; r++) {
                if (!src->span(r0 + r, c0, n, &keys[(size_t)r * n], &pair0[r])) {
                    full = false;
                    break;
                }
                for (int c = 0; c < n; c++) {
                    const int k = keys[(size_t)r * n + c];
                    const float t = in.tau_key[k];
                    if (!(t > 0.0f) || !std::isfinite(t)) {
                        full = false;
                        break;
                    }
                    hi = std::min(hi, cut[(size_t)k]);  // the cut itself kept as a margin
                }
            }
            if (full) {
                // residual (bf1d.cpp:816-877) as the largest |difference|
                auto T = [&](int r, int c) {
                    return (double)in.tau_key[keys[(size_t)(r - r0) * n + (c - c0)]];
                };
                const int rs[5] = {r0, r0 + n - 1, r0 + n / 2, r0 + n / 4, r0 + (3 * n) / 4};
                const int cs[5] = {c0, c0 + n - 1, c0 + n / 2, c0 + n / 4, c0 + (3 * n) / 4};
                const double t00 = T(r0, c0);
                double dmax = 0.0;
                for (int a = 0; a < 5; a++)
                    for (int b = 0; b < 5; b++) {
                        const double sep = T(rs[a], c0) + T(r0, cs[b]) - t00;
                        dmax = std::max(dmax, std::fabs(T(rs[a], cs[b]) - sep));
                    }
                int ires = -1;
                for (int iw = 0; iw < in.nw; iw++) {
                    if ((double)omega[iw] * dmax <= (double)in.phase_tol)
                        ires = iw;
                    else
                        break;
                }
                hi = std::min(hi, ires);
                if (hi >= 0) {
                    if (L > bf::MAX_L || ((1 << L) * in.p) > bf::MAX_SIGMA) {
                        char b[256];
                        snprintf(b, sizeof b,
                                 "butterfly panels of %d points with bf_p=%d exceed the GPU "
                                 "butterfly ((panel / bf_n_leaf) * bf_p <= %d, panel <= "
                                 "bf_n_leaf * %d)",
                                 n, in.p, bf::MAX_SIGMA, 1 << bf::MAX_L);
                        err = b;
                        return;
                    }
                    Node nd{};
                    nd.row0 = r0;
                    nd.col0 = c0;
                    nd.L = L;
                    nd.iw_res = hi;
                    nd.inst = inst;
                    nd.rows_off = -1;
                    if (pair0[0] >= 0) {
                        nd.rows_off = (int)out.node_pair0.size();
                        for (int r = 0; r < n; r++) {
                            out.node_pair0.push_back(pair0[r]);
                            out.node_out.push_back(src->out_trace(r0 + r));
                        }
                    }
                    nd.nanc = (int)std::min<size_t>(anc.size(), bf::MAX_L + 1);
                    for (int q = 0; q < nd.nanc; q++) nd.anc[q] = anc[anc.size() - 1 - q];
                    nd.tau_off = (int)out.tau.size();
                    samples(nd, keys);
                    anc.push_back((int)out.nodes.size());
                    out.nodes.push_back(nd);
                    pushed = true;
                }
            }
        }
        if (!err.empty()) return;
        if (!(rows <= leaf && cols <= leaf)) {
            if (rows >= cols && rows > leaf) {
                const int h = rows / 2;
                walk(r0, c0, h, cols, anc);
                walk(r0 + h, c0, rows - h, cols, anc);
            } else if (cols > leaf) {
                const int h = cols / 2;
                walk(r0, c0, rows, h, anc);
                walk(r0, c0 + h, rows, cols - h, anc);
            }
        }
        if (pushed) anc.pop_back();
    }

    // every traveltime sample of the node's factor, in TauLayout order
    void samples(const Node &nd, const std::vector<int> &keys)
    {
        const Tree &T = trees[nd.L];
        const int L = nd.L, p = in.p, leaf = in.leaf, nloc = leaf << L, nb = 1 << L,
                  mid = L / 2;
        std::vector<float> t((size_t)nloc * nloc);
        for (size_t e = 0; e < t.size(); e++) t[e] = in.tau_key[keys[e]];
        std::vector<double> &s = out.tau;
        const size_t base = s.size();
        const bf::TauLayout lay = bf::tau_layout(L, p, leaf);
        s.resize(base + lay.total);
        double *o = s.data() + base;
        const double rc = T.ctr(0, 0), sc = T.ctr(0, 0);
        for (int k = 0; k < nloc; k++) o[lay.inpt + k] = sample(t, nloc, (double)k, rc);
        for (int b = 0; b < nb; b++)
            for (int it = 0; it < p; it++)
                o[lay.innode + b * p + it] = sample(t, nloc, T.nd(L, b, it), rc);
        for (int lv = 1; lv <= mid; lv++) {
            const int nr = 1 << lv, nsc = 1 << (L - lv + 1), nsp = 1 << (L - lv);
            for (int br = 0; br < nr; br++) {
                const double r0 = T.ctr(lv, br);
                for (int cs = 0; cs < nsc; cs++)
                    for (int jt = 0; jt < p; jt++)
                        o[lay.r1in[lv] + (br * nsc + cs) * p + jt] =
                            sample(t, nloc, T.nd(L - lv + 1, cs, jt), r0);
                for (int bs = 0; bs < nsp; bs++)
                    for (int it = 0; it < p; it++)
                        o[lay.r1out[lv] + (br * nsp + bs) * p + it] =
                            sample(t, nloc, T.nd(L - lv, bs, it), r0);
            }
        }
        {
            const int nr = 1 << mid, nsp = 1 << (L - mid);
            for (int bs = 0; bs < nsp; bs++)
                for (int br = 0; br < nr; br++)
                    for (int it = 0; it < p; it++)
                        for (int jt = 0; jt < p; jt++)
                            o[lay.sw + ((bs * nr + br) * p + it) * p + jt] =
                                sample(t, nloc, T.nd(L - mid, bs, jt), T.nd(mid, br, it));
        }
        for (int lv = mid + 1; lv <= L; lv++) {
            const int nr = 1 << lv, nsc = 1 << (L - lv + 1);
            for (int br = 0; br < nr; br++)
                for (int cs = 0; cs < nsc; cs++) {
                    const double s0 = T.ctr(L - lv + 1, cs);
                    for (int jt = 0; jt < p; jt++)
                        o[lay.r2in[lv] + (br * nsc + cs) * p + jt] =
                            sample(t, nloc, s0, T.nd(lv - 1, br / 2, jt));
                    for (int it = 0; it < p; it++)
                        o[lay.r2out[lv] + (br * nsc + cs) * p + it] =
                            sample(t, nloc, s0, T.nd(lv, br, it));
                }
        }
        for (int b = 0; b < nb; b++)
            for (int it = 0; it < p; it++)
                o[lay.outnode + b * p + it] = sample(t, nloc, sc, T.nd(L, b, it));
        for (int r = 0; r < nloc; r++) o[lay.outpt + r] = sample(t, nloc, sc, (double)r);
    }

    // Lagrange values and switch samples of one class
    void tables(int L)
    {
        const Tree &T = trees[L];
        const int p = in.p, leaf = in.leaf, nloc = leaf << L, nb = 1 << L, mid = L / 2;
        const bf::LagLayout g = bf::lag_layout(L, p, leaf);
        out.lag_off[L] = (int)out.lag.size();
        out.lag.resize(out.lag.size() + g.total);
        float *o = out.lag.data() + out.lag_off[L];
        std::vector<double> b(p);
        for (int bx = 0; bx < nb; bx++) {
            const int e = T.box(L, bx);
            for (int k = 0; k < leaf; k++) {
                lagrange(p, &T.node[(size_t)e * p], (double)(T.i1[e] + k), b.data());
                for (int it = 0; it < p; it++) {
                    o[g.in + (bx * leaf + k) * p + it] = (float)b[it];
                    o[g.out + (bx * leaf + k) * p + it] = (float)b[it];
                }
            }
        }
        for (int lv = 1; lv <= mid; lv++)
            for (int cs = 0; cs < (nb >> (lv - 1)); cs++)
                for (int jt = 0; jt < p; jt++) {
                    lagrange(p, &T.node[(size_t)T.box(L - lv, cs / 2) * p],
                             T.nd(L - lv + 1, cs, jt), b.data());
                    for (int it = 0; it < p; it++)
                        o[g.r1[lv] + (cs * p + jt) * p + it] = (float)b[it];
                }
        for (int lv = mid + 1; lv <= L; lv++)
            for (int br = 0; br < (1 << lv); br++)
                for (int it = 0; it < p; it++) {
                    lagrange(p, &T.node[(size_t)T.box(lv - 1, br / 2) * p], T.nd(lv, br, it),
                             b.data());
                    for (int jt = 0; jt < p; jt++)
                        o[g.r2[lv] + (br * p + it) * p + jt] = (float)b[jt];
                }
        out.sws_off[L] = (int)out.sws.size();
        const int nr = 1 << mid, nsp = 1 << (L - mid);
        for (int bs = 0; bs < nsp; bs++)
            for (int br = 0; br < nr; br++)
                for (int it = 0; it < p; it++)
                    for (int jt = 0; jt < p; jt++)
                        out.sws.push_back(bilinear_at(nloc, T.nd(L - mid, bs, jt), T.nd(mid, br, it)));
    }
};

}  // namespace

int bf_plan(const BfPlanIn &in, BfPlan &out, std::string &err)
{
    out = BfPlan{};
    if (in.p <= 0 || in.p > bf::MAX_P || in.leaf <= 1 || in.p >= in.leaf) {
        err = "the GPU butterfly takes 0 < bf_p < bf_n_leaf, bf_p <= 32";
        return KMDAT_E_PARAM;
    }
    Planner P(in, out);
    P.omega.resize(in.nw);
    for (int iw = 0; iw < in.nw; iw++)
        P.omega[iw] = 2.0f * (float)M_PI * (float)iw / ((float)in.nfft * in.dt);
    out.omega = P.omega;
    P.cut.resize(in.nkey);
    for (int k = 0; k < in.nkey; k++) P.cut[k] = P.aa_cut(k);
    // top panels (bf1d.cpp:935-961)
    const int lev = in.levels < 0 ? 2 : in.levels;
    int pn = in.leaf;
    for (int k = 0; k < lev && pn < (1 << 29); k++) pn <<= 1;
    if (pn > in.n) {
        int m = in.leaf;
        while (m < in.n) m <<= 1;
        pn = m;
    }
    P.trees.resize(bf::MAX_L + 1);
    for (int L = 1; L <= bf::MAX_L; L++) P.trees[L].build(in.leaf << L, in.leaf, in.p);
    std::vector<int> anc;
    BandedRows banded(in);
    const int ninst = in.pairs ? in.ninst : 1;
    for (int k = 0; k < ninst && P.err.empty(); k++) {
        PairRows pr(in, k);
        P.src = in.pairs ? (const RowSource *)&pr : (const RowSource *)&banded;
        P.inst = k;
        for (int r0 = 0; r0 < in.n && P.err.empty(); r0 += pn)
            for (int c0 = 0; c0 < in.n && P.err.empty(); c0 += pn)
                P.walk(r0, c0, std::min(pn, in.n - r0), std::min(pn, in.n - c0), anc);
    }
    if (!P.err.empty()) {
        err = P.err;
        return KMDAT_E_PARAM;
    }
    int maxL = 0;
    for (const Node &nd : out.nodes) maxL = std::max(maxL, nd.L);
    out.maxL = maxL;
    for (int L = 1; L <= maxL; L++) P.tables(L);
    // items: (node, chunk) for every chunk up to the node's last frequency
    long long d = 0, sw = 0;
    for (int k = 0; k < (int)out.nodes.size(); k++) {
        const Node &nd = out.nodes[k];
        const int nloc = in.leaf << nd.L;
        for (int ch = 0; ch * 64 <= nd.iw_res; ch++) {
            bf::Item it{};
            it.node = k;
            it.chunk = ch;
            it.d_off = d;
            it.sw_off = sw;
            if (!in.pairs) {  // difference blocks and switch kernels: banded form only
                d += (long long)nloc * nloc * 64;
                sw += (long long)((1 << nd.L) * in.p * in.p) * 64;
            }
            out.items.push_back(it);
        }
    }
    out.d_elems = d;
    out.sw_elems = sw;
    // groups: per chunk, items that write a common output trace (overlapping
    // rows of one matrix, or -- pair lists -- rows of different gathers that
    // land on one trace) share a group, so that no two waves update a trace
    auto outputs = [&](const bf::Item &it, std::vector<long long> &o) {
        const Node &nd = out.nodes[it.node];
        const int nloc = in.leaf << nd.L;
        o.clear();
        for (int r = 0; r < nloc; r++)
            o.push_back(nd.rows_off >= 0 ? (long long)out.node_out[(size_t)nd.rows_off + r]
                                         : (long long)nd.row0 + r);
    };
    std::vector<int> byc(out.items.size());
    for (size_t i = 0; i < byc.size(); i++) byc[i] = (int)i;
    std::stable_sort(byc.begin(), byc.end(), [&](int a, int b) {
        return out.items[a].chunk < out.items[b].chunk;
    });
    std::vector<int> parent(out.items.size());
    for (size_t i = 0; i < parent.size(); i++) parent[i] = (int)i;
    std::function<int(int)> root = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    std::vector<long long> o;
    for (size_t b = 0; b < byc.size();) {
        size_t e = b;
        while (e < byc.size() && out.items[byc[e]].chunk == out.items[byc[b]].chunk) e++;
        std::unordered_map<long long, int> owner;
        for (size_t i = b; i < e; i++) {
            outputs(out.items[byc[i]], o);
            for (long long t : o) {
                auto f = owner.find(t);
                if (f == owner.end())
                    owner.emplace(t, byc[i]);
                else
                    parent[root(byc[i])] = root(f->second);
            }
        }
        b = e;
    }
    // a group's items in (chunk, first row) order; banded groups cover the
    // contiguous rows [grp_row0, grp_row0 + grp_rows)
    std::vector<int> order(out.items.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        const int ra = root(a), rb = root(b);
        if (ra != rb) {
            const bf::Item &x = out.items[ra], &y = out.items[rb];
            if (x.chunk != y.chunk) return x.chunk < y.chunk;
            return ra < rb;
        }
        return out.nodes[out.items[a].node].row0 < out.nodes[out.items[b].node].row0 ||
               (out.nodes[out.items[a].node].row0 == out.nodes[out.items[b].node].row0 && a < b);
    });
    out.grp_ptr.push_back(0);
    for (size_t b = 0; b < order.size();) {
        size_t e = b;
        int lo = 1 << 30, hi = -1;
        while (e < order.size() && root(order[e]) == root(order[b])) {
            const Node &nd = out.nodes[out.items[order[e]].node];
            lo = std::min(lo, nd.row0);
            hi = std::max(hi, nd.row0 + (in.leaf << nd.L));
            out.grp.push_back(order[e]);
            e++;
        }
        out.grp_row0.push_back(lo);
        out.grp_rows.push_back(hi - lo);
        out.grp_chunk.push_back(out.items[order[b]].chunk);
        out.grp_ptr.push_back((int)out.grp.size());
        b = e;
    }
    for (size_t g = 0; g + 1 < out.grp_ptr.size(); g++)
        for (int q = 0; q < (out.grp_rows[g] + 3) / 4; q++) out.quads.push_back({(int)g, q});
    return KMDAT_OK;
}

}  // namespace kmdat
