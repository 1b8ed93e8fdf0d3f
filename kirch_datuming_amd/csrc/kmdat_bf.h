// Layout of the segmented butterfly (use_bf=1) shared by the host planner
// (kmdat_bfplan.cpp) and the kernels (kmdat_bf.hip).
//
// A butterfly panel ("node") is an nloc x nloc block of one operator's
// matrix (rows = outputs, columns = summation indices), nloc = leaf 2^L, at
// (row0, col0).  Its strict factor has boxes of 2^l per level l, p Chebyshev
// nodes per box, and stages as SERECKIRCH/src/bf1d.cpp:283-473 forms them:
// leaf initialisation, source-side levels 1 .. mid (mid = L / 2), the switch
// at mid, target-side levels mid + 1 .. L, termination.  Every phase of the
// factor is exp(+-i alpha t~) with alpha = -omega and t~ a bilinear sample
// of the block's traveltimes; the samples of one node are stored in double,
// stage after stage, at the offsets TauLayout gives.  The Lagrange values of
// a class (one L) are stored in float, at the offsets LagLayout gives.
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define KMDAT_HD __host__ __device__
#else
#define KMDAT_HD
#endif

namespace kmdat {

struct PairRec;

namespace bf {

constexpr int MAX_L = 3;        // nloc up to leaf 2^3
constexpr int MAX_SIGMA = 48;   // (nloc / leaf) p: coefficients of a level per lane
                                // (2 levels + p products of 64 lanes fit 64 KiB of LDS)
constexpr int MAX_P = 32;

// Offsets (in samples) of one node's traveltime samples.
struct TauLayout {
    int inpt;        // [nloc]            column k, row = target root centre   (+alpha)
    int innode;      // [2^L][p]          source leaf nodes, target root centre (-alpha)
    int r1in[MAX_L + 1];   // level lv: [2^lv][2^(L-lv+1)][p] child nodes, target box centre (+)
    int r1out[MAX_L + 1];  // level lv: [2^lv][2^(L-lv)][p] parent nodes, target box centre (-)
    int sw;          // [2^(L-mid)][2^mid][p][p]  source node jt, target node it (+)
    int r2in[MAX_L + 1];   // level lv: [2^lv][2^(L-lv+1)][p] child centre, parent target nodes (-)
    int r2out[MAX_L + 1];  // level lv: [2^lv][2^(L-lv+1)][p] child centre, target nodes (+)
    int outnode;     // [2^L][p]          source root centre, target leaf nodes (-)
    int outpt;       // [nloc]            source root centre, row r (+)
    int total;
};

KMDAT_HD inline TauLayout tau_layout(int L, int p, int leaf)
{
    TauLayout t{};
    const int nloc = leaf << L, nb = 1 << L, mid = L / 2;
    int o = 0;
    t.inpt = o;
    o += nloc;
    t.innode = o;
    o += nb * p;
    for (int lv = 1; lv <= mid; lv++) {
        t.r1in[lv] = o;
        o += 2 * nb * p;
        t.r1out[lv] = o;
        o += nb * p;
    }
    t.sw = o;
    o += nb * p * p;
    for (int lv = mid + 1; lv <= L; lv++) {
        t.r2in[lv] = o;
        o += 2 * nb * p;
        t.r2out[lv] = o;
        o += 2 * nb * p;
    }
    t.outnode = o;
    o += nb * p;
    t.outpt = o;
    o += nloc;
    t.total = o;
    return t;
}

// Offsets (in floats) of one class's Lagrange values.
struct LagLayout {
    int in;                // [2^L][leaf][p]          leaf box points against its nodes
    int r1[MAX_L + 1];     // level lv: [2^(L-lv+1)][p][p] child node jt against parent node it
    int r2[MAX_L + 1];     // level lv: [2^lv][p][p]  target node it against parent node jt
    int out;               // [2^L][leaf][p]
    int total;
};

KMDAT_HD inline LagLayout lag_layout(int L, int p, int leaf)
{
    LagLayout g{};
    const int nb = 1 << L, mid = L / 2;
    int o = 0;
    g.in = o;
    o += nb * leaf * p;
    for (int lv = 1; lv <= mid; lv++) {
        g.r1[lv] = o;
        o += (nb >> (lv - 1)) * p * p;
    }
    for (int lv = mid + 1; lv <= L; lv++) {
        g.r2[lv] = o;
        o += (1 << lv) * p * p;
    }
    g.out = o;
    o += nb * leaf * p;
    g.total = o;
    return g;
}

// One bilinear amplitude sample of the switch kernel: four entries of the
// block (local row * nloc + column) and their weights, in the order the
// reference sums them (bf1d.cpp:235-239).
struct SwSample {
    int e[4];
    double w[4];
};

// One butterfly panel candidate.
struct Node {
    int row0, col0, L;
    int inst;        // matrix instance (pair lists: shot or receiver gather; banded: 0)
    int rows_off;    // pair lists: first of the node's rows in node_pair0 / node_out, else -1
    int tau_off;     // first traveltime sample (TauLayout of L)
    int iw_res;      // largest frequency index whose phase residual passes (-1: none)
    int nanc;        // candidate ancestors (panels containing this one)
    int anc[MAX_L + 1];
};

// One (node, 64-frequency chunk) work item.
struct Item {
    int node, chunk;
    long long d_off;   // first float2 of its difference block [nloc][nloc][64]
    long long sw_off;  // first float2 of its switch kernel [2^L p^2][64]
};

// Four output rows of one group (a chunk's items whose rows overlap).
struct Quad {
    int g, q;          // group, rows grp_row0[g] + 4 q .. + 3
};

struct Args {
    // operator (banded form)
    const int *pkey;
    const float *pw;
    const float *tau_key;   // traveltime of each table key
    const float *aarg;      // antialias argument of each key
    const float2 *ktab;     // [key][nw]
    const float *omega;     // [nw] 2 pi iw / (nfft dt) in float
    int nb, aper, nw, nfft;
    float dt, antialias, eps2;
    // plan
    const Node *nodes;
    const Item *items;
    int nitems;
    const double *tau;      // node samples
    const float *lag;       // Lagrange values
    const SwSample *sws;    // switch samples
    int lag_off[MAX_L + 1], sws_off[MAX_L + 1];
    int p, leaf, maxL;
    // work
    unsigned long long *mask;  // [item]
    float2 *sw;             // switch kernels
    float2 *d;              // difference blocks
    // correction
    const int *grp_ptr;     // [ngroup + 1] into grp
    const int *grp_row0;    // [ngroup] first output row
    const int *grp_rows;    // [ngroup] rows
    const int *grp_chunk;   // [ngroup]
    const int *grp;         // items of each group
    const Quad *quads;
    int nquads;
    const float2 *in;
    float2 *out;
    int n_b;
    size_t in_b, in_i, out_b, out_o;
    // pair-list operators (cmp=1): entry (r, c) of a node is pair
    // node_pair0[rows_off + r] + c; its column's input trace is that pair's tr
    // (spectra [trace][nw]) and row r's output trace node_out[rows_off + r]
    const PairRec *pairs;
    const int *node_pair0, *node_out;
    int ngroup;
};

}  // namespace bf
}  // namespace kmdat
