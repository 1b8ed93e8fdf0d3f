// kirchmig2d_cig — drop-in command line for the adj=1 path of the reference
// (and its adj=0 modelling)
// program of the same name (src/kirchmig2d_cig.cpp:494-1401): same key=value
// parameters and RSF inputs/outputs, migration on the GPU through libkmig.
//
// Inputs: seismic_data= (t,h,s), stable=/sderiv= (z,x,s), rtable=/rderiv=
// (z,x,r), optionally sgradx=/sgradz=/rgradx=/rgradz=.  Outputs: migration=
// (z,x,angle|offset|1) and the direct_mute_qc_before=/after= gathers.
// Errors print "ERROR: <the reference's message>" and exit with status 1.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "kmig.h"
#include "kmig_internal.h"
#include "rsf.h"

namespace {

[[noreturn]] void die(const std::string &msg)
{
    fprintf(stderr, "ERROR: %s\n", msg.c_str());
    exit(EXIT_FAILURE);
}

std::string need(const std::map<std::string, std::string> &kv, const char *k)
{
    auto it = kv.find(k);
    if (it == kv.end()) die(std::string("Need ") + k + "=");
    return it->second;
}

struct Vol {
    rsf::Header h;
    std::vector<float> v;
};

Vol load(const std::string &path)
{
    Vol r;
    std::string err;
    if (!rsf::read_header(path, r.h, err)) die(err);
    if (!rsf::read_floats(path, r.h, r.v, err)) die(err);
    return r;
}

}  // namespace

int main(int argc, char **argv)
{
    std::map<std::string, std::string> kv;
    char err[1024] = {0};
    if (kmig::collect_kv(kv, argc - 1, (const char *const *)(argv + 1), err, sizeof(err)))
        die(err);
    const int adj = kv.count("adj") ? atoi(kv["adj"].c_str()) : 1;
    const std::string dat_f = need(kv, "seismic_data"), mig_f = need(kv, "migration");
    const std::string stim_f = need(kv, "stable"), sder_f = need(kv, "sderiv");
    const std::string rtim_f = need(kv, "rtable"), rder_f = need(kv, "rderiv");

    Vol dat;
    int nt, nh, ns;
    float t0, dt, h0, dh, s0, ds;
    if (adj) {
        dat = load(dat_f);
        if (dat.h.ndim < 3) die("Need 3D data volume for seismic_data= (t,h/r,s)");
        nt = dat.h.n[0], nh = dat.h.n[1], ns = dat.h.n[2];
        t0 = (float)dat.h.o[0], dt = (float)dat.h.d[0];
        h0 = (float)dat.h.o[1], dh = (float)dat.h.d[1];
        s0 = (float)dat.h.o[2], ds = (float)dat.h.d[2];
    } else {
        // the data axes of the modelled volume (src/kirchmig2d_cig.cpp:662-699)
        auto geti = [&](const char *k, int def) {
            return kv.count(k) ? atoi(kv[k].c_str()) : def;
        };
        auto getf = [&](const char *k, float def) {
            return kv.count(k) ? (float)atof(kv[k].c_str()) : def;
        };
        nt = atoi(need(kv, "nt").c_str());
        nh = geti("nh", 1);
        ns = geti("ns", 1);
        t0 = getf("t0", 0.0f);
        dt = (float)atof(need(kv, "dt").c_str());
        h0 = getf("h0", 0.0f);
        dh = (float)atof(need(kv, "dh").c_str());
        s0 = getf("s0", 0.0f);
        ds = (float)atof(need(kv, "ds").c_str());
    }

    kmig_params p;
    int rc = kmig::parse_kv(&p, kv, nh, h0, dh, err, sizeof(err));
    if (rc) die(err);
    if ((rc = kmig_params_check(&p, ns, err, sizeof(err)))) die(err);
    const bool angle = p.cig && p.cig_is_angle;
    const bool qc = p.direct_mute && p.direct_mute_qc_shot >= 0;
    std::string qcb_f, qca_f;
    if (qc) {
        if (!kv.count("direct_mute_qc_before") || !kv.count("direct_mute_qc_after"))
            die("direct_mute_qc_shot requires direct_mute_qc_before= and direct_mute_qc_after=.");
        qcb_f = kv["direct_mute_qc_before"];
        qca_f = kv["direct_mute_qc_after"];
    }
    const char *gk[4] = {"sgradx", "sgradz", "rgradx", "rgradz"};
    int ngk = 0;
    for (const char *k : gk) ngk += kv.count(k) ? 1 : 0;
    const bool use_grad = angle && ngk > 0;
    if (use_grad && ngk != 4)
        die("For mode=angle, please provide all of sgradx=, sgradz=, rgradx=, and rgradz=, or "
            "provide none of them.");

    Vol stim = load(stim_f), sder = load(sder_f);
    if (stim.h.ndim < 3) die("Need 3D data volume for stable= (z,x,s)");
    if (sder.h.ndim < 3) die("Need 3D data volume for sderiv= (z,x,s)");
    const int nz = stim.h.n[0], nx = stim.h.n[1], sny = stim.h.n[2];
    if (sder.h.n[0] != nz || sder.h.n[1] != nx || sder.h.n[2] != sny) {
        snprintf(err, sizeof(err),
                 "Dimension mismatch between stable= and sderiv=. Expected sderiv n1=%d n2=%d "
                 "n3=%d",
                 nz, nx, sny);
        die(err);
    }
    Vol rtim = load(rtim_f), rder = load(rder_f);
    if (rtim.h.ndim < 3) die("Need 3D data volume for rtable= (z,x,r)");
    if (rder.h.ndim < 3) die("Need 3D data volume for rderiv= (z,x,r)");
    if (rtim.h.n[0] != nz || rtim.h.n[1] != nx) {
        snprintf(err, sizeof(err), "Dimension mismatch in rtable=. Expected n1=%d n2=%d", nz, nx);
        die(err);
    }
    const int rny = rtim.h.n[2];
    if (rder.h.n[0] != nz || rder.h.n[1] != nx || rder.h.n[2] != rny) {
        snprintf(err, sizeof(err),
                 "Dimension mismatch between rtable= and rderiv=. Expected rderiv n1=%d n2=%d "
                 "n3=%d",
                 nz, nx, rny);
        die(err);
    }
    Vol g[4];
    if (use_grad) {
        for (int k = 0; k < 4; k++) {
            g[k] = load(kv[gk[k]]);
            const int n3 = k < 2 ? sny : rny;
            if (g[k].h.ndim < 3) {
                snprintf(err, sizeof(err), "Need 3D data volume for %s= (z,x,source_or_receiver)",
                         gk[k]);
                die(err);
            }
            if (g[k].h.n[0] != nz || g[k].h.n[1] != nx || g[k].h.n[2] != n3) {
                snprintf(err, sizeof(err), "Dimension mismatch in %s=. Expected n1=%d n2=%d n3=%d",
                         gk[k], nz, nx, n3);
                die(err);
            }
        }
    }

    kmig_ctx *ctx = nullptr;
    if ((rc = kmig_create(&ctx, &p))) die(ctx ? kmig_last_error(ctx) : "kmig_create failed");
    auto chk = [&](int r) {
        if (r) die(kmig_last_error(ctx));
    };
    const float z0 = (float)stim.h.o[0], dz = (float)stim.h.d[0];
    const float x0 = (float)stim.h.o[1], dx = (float)stim.h.d[1];
    chk(kmig_set_tables(ctx, 0, stim.v.data(), sder.v.data(), use_grad ? g[0].v.data() : nullptr,
                        use_grad ? g[1].v.data() : nullptr, nz, nx, sny, z0, dz, x0, dx,
                        (float)stim.h.o[2], (float)stim.h.d[2]));
    chk(kmig_set_tables(ctx, 1, rtim.v.data(), rder.v.data(), use_grad ? g[2].v.data() : nullptr,
                        use_grad ? g[3].v.data() : nullptr, nz, nx, rny, z0, dz, x0, dx,
                        (float)rtim.h.o[2], (float)rtim.h.d[2]));
    const int ng = p.cig ? p.ng : 1;
    if (!adj) {
        // modelling: image in (migration=), data out (seismic_data=)
        Vol img = load(mig_f);
        if (img.h.n[0] != nz || img.h.n[1] != nx || img.h.n[2] != ng) {
            snprintf(err, sizeof(err),
                     "Dimension mismatch in migration=. Expected n1=%d n2=%d n3=%d", nz, nx, ng);
            die(err);
        }
        std::vector<float> data((size_t)nt * nh * ns);
        auto m0 = std::chrono::steady_clock::now();
        chk(kmig_model(ctx, img.v.data(), nt, t0, dt, nh, h0, dh, ns, s0, ds, data.data()));
        auto m1 = std::chrono::steady_clock::now();
        kmig_stats st;
        kmig_get_stats(ctx, &st);
        printf("INFO: Done. Elapsed time: %.3f s. (model kernels %.3f ms)\n",
               std::chrono::duration<double>(m1 - m0).count(), st.ms_migrate);
        rsf::Header dh_;
        dh_.ndim = 3;
        dh_.n[0] = nt;
        dh_.o[0] = t0;
        dh_.d[0] = dt;
        dh_.n[1] = nh;
        dh_.o[1] = h0;
        dh_.d[1] = nh == 1 ? 0.0f : dh;
        dh_.n[2] = ns;
        dh_.o[2] = s0;
        dh_.d[2] = ns == 1 ? 0.0f : ds;
        dh_.extra["label1"] = "Time";
        dh_.extra["unit1"] = "s";
        dh_.extra["label2"] = p.cmp ? "Offset" : "Receiver";
        dh_.extra["label3"] = "Shot";
        std::string werr;
        if (!rsf::write(dat_f, data.data(), dh_, werr)) die(werr);
        kmig_destroy(ctx);
        return 0;
    }
    auto t_start = std::chrono::steady_clock::now();
    chk(kmig_set_data(ctx, dat.v.data(), nt, t0, dt, nh, h0, dh, ns, s0, ds));
    chk(kmig_run(ctx, 0, ns));
    chk(kmig_sync(ctx));
    auto t_end = std::chrono::steady_clock::now();
    std::vector<float> out((size_t)ng * nz * nx);
    chk(kmig_get_cig(ctx, out.data()));
    kmig_stats st;
    kmig_get_stats(ctx, &st);
    const double el = std::chrono::duration<double>(t_end - t_start).count();
    printf("INFO: Done. Elapsed time: %.3f s. (migrate kernel %.3f ms, %.3e updates/s)\n", el,
           st.ms_migrate, st.ms_migrate > 0 ? st.updates / (st.ms_migrate * 1e-3) : 0.0);

    rsf::Header oh;
    oh.ndim = 3;
    oh.n[0] = nz;
    oh.o[0] = z0;
    oh.d[0] = dz;
    oh.n[1] = nx;
    oh.o[1] = x0;
    oh.d[1] = dx;
    oh.extra["label1"] = "Depth";
    oh.extra["label2"] = "Lateral";
    if (dat.h.extra.count("unit1")) oh.extra["unit1"] = dat.h.extra["unit1"];
    if (p.cig) {
        oh.n[2] = ng;
        oh.o[2] = p.g0;
        oh.d[2] = p.dg;
        if (angle) {
            oh.extra["label3"] = p.half_angle ? "Half-opening angle" : "Scattering angle";
            oh.extra["unit3"] = "degree";
        } else {
            oh.extra["label3"] = p.cmp ? "Offset" : "Receiver";
            if (dat.h.extra.count("unit1")) oh.extra["unit3"] = dat.h.extra["unit1"];
        }
    } else {
        oh.n[2] = 1;
        oh.o[2] = 0.0;
        oh.d[2] = 1.0;
    }
    std::string werr;
    if (!rsf::write(mig_f, out.data(), oh, werr)) die(werr);
    if (qc) {
        std::vector<float> b((size_t)nt * nh), a((size_t)nt * nh);
        chk(kmig_get_qc(ctx, b.data(), a.data()));
        rsf::Header qh;
        qh.ndim = 2;
        qh.n[0] = nt;
        qh.o[0] = t0;
        qh.d[0] = dt;
        qh.n[1] = nh;
        qh.o[1] = h0;
        qh.d[1] = dh;
        qh.extra["label1"] = "Time";
        qh.extra["unit1"] = "s";
        qh.extra["label2"] = p.cmp ? "Offset" : "Receiver";
        if (!rsf::write(qcb_f, b.data(), qh, werr)) die(werr);
        if (!rsf::write(qca_f, a.data(), qh, werr)) die(werr);
    }
    kmig_destroy(ctx);
    return 0;
}
