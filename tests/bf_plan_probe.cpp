// Host driver of the butterfly planner (kirch_datuming_amd/csrc/kmdat_bfplan.cpp),
// built with g++ by tests/test_bf_plan.py: reads one operator from a binary
// file, plans it and prints the plan.
//
// Input (little-endian):
//   int32  mode (0 banded, 1 pair lists), n, nb, aper, nkey, nw, nfft, p, leaf,
//          levels, ninst, col_base, col_by_div, npairs
//   float  dt, antialias, phase_tol
//   mode 0: int32 pkey[n * nb], float pw[n * nb]
//   mode 1: {int32 tr, int32 key, float w, int32 pad}[npairs],
//           int32 row_o[ninst * n], row_q0[ninst * n], row_q1[ninst * n]
//   float  tau[nkey], aarg[nkey]
// Output lines:
//   node <inst> <row0> <col0> <L> <iw_res> <nanc> <anc...>
//   out <node> <output trace of each row>          (pair lists)
//   group <chunk> <node of each item>
//   items <n> d <float2 elements> sw <float2 elements>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "kmdat_internal.h"

using namespace kmdat;

template <class T>
static void rd(FILE *f, T *p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s input\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int h[14];
    float fl[3];
    rd(f, h, 14);
    rd(f, fl, 3);
    const int mode = h[0], n = h[1], nb = h[2], nkey = h[4], ninst = h[10], npairs = h[13];
    std::vector<int> pkey, ro, rq0, rq1;
    std::vector<float> pw, tau(nkey), aarg(nkey);
    std::vector<PairRec> pairs;
    if (mode == 0) {
        pkey.resize((size_t)n * nb);
        pw.resize((size_t)n * nb);
        rd(f, pkey.data(), pkey.size());
        rd(f, pw.data(), pw.size());
    } else {
        pairs.resize(npairs);
        ro.resize((size_t)ninst * n);
        rq0.resize(ro.size());
        rq1.resize(ro.size());
        rd(f, pairs.data(), pairs.size());
        rd(f, ro.data(), ro.size());
        rd(f, rq0.data(), rq0.size());
        rd(f, rq1.data(), rq1.size());
    }
    rd(f, tau.data(), tau.size());
    rd(f, aarg.data(), aarg.size());
    fclose(f);

    BfPlanIn in{};
    in.n = n;
    in.nb = nb;
    in.aper = h[3];
    in.pkey = pkey.empty() ? nullptr : pkey.data();
    in.pw = pw.empty() ? nullptr : pw.data();
    in.tau_key = tau.data();
    in.aarg = aarg.data();
    in.nkey = nkey;
    in.nw = h[5];
    in.nfft = h[6];
    in.p = h[7];
    in.leaf = h[8];
    in.levels = h[9];
    in.dt = fl[0];
    in.antialias = fl[1];
    in.phase_tol = fl[2];
    if (mode == 1) {
        in.pairs = pairs.data();
        in.row_o = ro.data();
        in.row_q0 = rq0.data();
        in.row_q1 = rq1.data();
        in.ninst = ninst;
        in.col_base = h[11];
        in.col_by_div = h[12];
    }
    BfPlan out;
    std::string err;
    const int rc = bf_plan(in, out, err);
    if (rc) {
        printf("error %d %s\n", rc, err.c_str());
        return 0;
    }
    for (size_t k = 0; k < out.nodes.size(); k++) {
        const bf::Node &nd = out.nodes[k];
        printf("node %d %d %d %d %d %d", nd.inst, nd.row0, nd.col0, nd.L, nd.iw_res, nd.nanc);
        for (int q = 0; q < nd.nanc; q++) printf(" %d", nd.anc[q]);
        printf("\n");
        if (nd.rows_off >= 0) {
            printf("out %zu", k);
            for (int r = 0; r < (in.leaf << nd.L); r++)
                printf(" %d", out.node_out[(size_t)nd.rows_off + r]);
            printf("\n");
        }
    }
    for (size_t g = 0; g + 1 < out.grp_ptr.size(); g++) {
        printf("group %d", out.grp_chunk[g]);
        for (int e = out.grp_ptr[g]; e < out.grp_ptr[g + 1]; e++)
            printf(" %d", out.items[(size_t)out.grp[(size_t)e]].node);
        printf("\n");
    }
    printf("items %zu d %lld sw %lld\n", out.items.size(), out.d_elems, out.sw_elems);
    return 0;
}
