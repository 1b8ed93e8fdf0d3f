// kirchdat2d_freq — drop-in command line for the reference program of the
// same name (src/kirchdat2d_frequency_cs.cpp:1177-1413): the same key=value
// parameters and RSF files, the datuming on the GPU through libkmdat.
//
// Inputs: input_file= (t,h,s), sgreen_file= / rgreen_file= (the datum tables
// greensol_auto writes: n1 = positions, n1 x n1 samples), model_file= (its
// header's should_datum and depth axis give the datum when auto_datum=1).
// Outputs: output_file= (t,h,s) and, with interm=, the receiver-side result.
// Errors print "ERROR: <the reference's message>" and exit with status 1.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "kmig.h"
#include "kmig_dat.h"
#include "kmig_internal.h"
#include "rsf.h"

namespace {

[[noreturn]] void die(const std::string &msg)
{
    fprintf(stderr, "ERROR: %s\n", msg.c_str());
    exit(EXIT_FAILURE);
}

using KV = std::map<std::string, std::string>;

std::string need(const KV &kv, const char *k)
{
    auto it = kv.find(k);
    if (it == kv.end()) die(std::string("Need ") + k + "=");
    return it->second;
}

int geti(const KV &kv, const char *k, int def)
{
    auto it = kv.find(k);
    return it == kv.end() ? def : (int)strtol(it->second.c_str(), nullptr, 10);
}

float getf(const KV &kv, const char *k, float def)
{
    auto it = kv.find(k);
    return it == kv.end() ? def : (float)strtod(it->second.c_str(), nullptr);
}

rsf::Header header(const std::string &path)
{
    rsf::Header h;
    std::string err;
    if (!rsf::read_header(path, h, err)) die(err);
    return h;
}

std::vector<float> floats(const std::string &path, const rsf::Header &h)
{
    std::vector<float> v;
    std::string err;
    if (!rsf::read_floats(path, h, v, err)) die(err);
    return v;
}

}  // namespace

int main(int argc, char **argv)
{
    KV kv;
    char err[1024] = {0};
    if (kmig::collect_kv(kv, argc - 1, (const char *const *)(argv + 1), err, sizeof(err)))
        die(err);
    const std::string in_f = need(kv, "input_file"), out_f = need(kv, "output_file");
    const std::string sg_f = need(kv, "sgreen_file"), rg_f = need(kv, "rgreen_file");
    const std::string model_f = need(kv, "model_file");

    kmdat_config cfg;
    kmdat_config_default(&cfg);
    const int verb = geti(kv, "verb", 1);
    cfg.cmp = geti(kv, "cmp", 1);
    if (cfg.cmp != 0 && cfg.cmp != 1) die("cmp must be 0 or 1.");
    const rsf::Header mh = header(model_f);
    const int auto_datum = geti(kv, "auto_datum", 1);
    if (auto_datum != 0 && auto_datum != 1) die("auto_datum must be 0 or 1.");
    if (auto_datum) {
        auto it = mh.extra.find("should_datum");
        if (it == mh.extra.end()) die("Need should_datum in model_file header when auto_datum=1");
        const int should_datum = (int)strtol(it->second.c_str(), nullptr, 10);
        cfg.sdatum = (float)(mh.o[0] + should_datum * mh.d[0]);
        cfg.rdatum = cfg.sdatum;
    } else {
        if (!kv.count("sdatum")) die("Need sdatum= when auto_datum=0");
        if (!kv.count("rdatum")) die("Need rdatum= when auto_datum=0");
        cfg.sdatum = getf(kv, "sdatum", 0.0f);
        cfg.rdatum = getf(kv, "rdatum", 0.0f);
    }
    cfg.use_bf = geti(kv, "use_bf", 1);
    cfg.bf_p = geti(kv, "bf_p", 12);
    cfg.bf_n_leaf = geti(kv, "bf_n_leaf", 16);
    cfg.bf_panel_levels = geti(kv, "bf_panel_levels", 1);
    cfg.bf_amp_eps = getf(kv, "bf_amp_eps", 1e-20f);
    cfg.bf_phase_tol = getf(kv, "bf_phase_tol", 2.0f);
    cfg.aperture = geti(kv, "aperture", 50);
    cfg.taper = geti(kv, "taper", 10);
    cfg.data_taper = geti(kv, "data_taper", 0);
    cfg.antialias = getf(kv, "antialias", 1.0f);
    cfg.length = getf(kv, "length", 0.025f);
    const bool want_interm = kv.count("interm") > 0;
    cfg.interm = want_interm ? 1 : 0;
    if (kmdat_config_check(&cfg, err, sizeof err)) die(err);
    if (verb) {
        printf("INFO: Use butterfly: %s (the operators are applied exactly on the GPU).\n",
               cfg.use_bf ? "yes" : "no");
        printf("INFO: data_taper = %d, antialias = %g, aperture = %d, taper = %d.\n",
               cfg.data_taper, (double)cfg.antialias, cfg.aperture, cfg.taper);
    }

    const rsf::Header ih = header(in_f);
    if (ih.ndim < 3) die("Input must be 3D.");
    const int nt = ih.n[0], nh = ih.n[1], ns = ih.n[2];
    const float dt = (float)ih.d[0], h0 = (float)ih.o[1], dh = (float)ih.d[1];
    const float s0 = (float)ih.o[2], ds = (float)ih.d[2];
    // the datum tables: n1 positions, n1 x n1 samples (greensol_auto sets
    // only n1, :56-59)
    auto green = [&](const std::string &f, const char *what) {
        rsf::Header h = header(f);
        if (h.ndim < 2) die(std::string(what) + " Green's function must be 2D.");
        rsf::Header r = h;
        for (int k = 1; k < 9; k++) r.n[k] = 1;
        r.n[1] = h.n[0];
        return std::make_pair(h, floats(f, r));
    };
    auto sg = green(sg_f, "Source");
    auto rg = green(rg_f, "Receiver");
    const std::vector<float> data = floats(in_f, ih);

    kmdat_ctx *ctx = nullptr;
    int rc = kmdat_create(&ctx, &cfg);
    if (rc) die(ctx ? kmdat_last_error(ctx) : "kmdat_create failed");
    auto chk = [&](int r) {
        if (r) die(kmdat_last_error(ctx));
    };
    chk(kmdat_set_green(ctx, 0, sg.second.data(), sg.first.n[0], (float)sg.first.o[0],
                        (float)sg.first.d[0]));
    chk(kmdat_set_green(ctx, 1, rg.second.data(), rg.first.n[0], (float)rg.first.o[0],
                        (float)rg.first.d[0]));
    std::vector<float> out((size_t)nt * nh * ns), mid(want_interm ? out.size() : 0);
    const auto t0 = std::chrono::steady_clock::now();
    chk(kmdat_datum(ctx, data.data(), nt, dt, nh, h0, dh, ns, s0, ds, out.data(),
                    want_interm ? mid.data() : nullptr));
    const auto t1 = std::chrono::steady_clock::now();
    kmdat_stats st;
    kmdat_get_stats(ctx, &st);
    printf("INFO: Done. Elapsed time: %.3f s. (nfft %d, GPU %.3f ms, %.3e complex MAC/s)\n",
           std::chrono::duration<double>(t1 - t0).count(), st.nfft, st.ms_total,
           st.ms_total > 0 ? st.cmacs / (st.ms_total * 1e-3) : 0.0);
    kmdat_destroy(ctx);

    // output header as :1379-1388
    rsf::Header oh;
    oh.ndim = 3;
    oh.n[0] = nt;
    oh.n[1] = nh;
    oh.n[2] = ns;
    oh.d[0] = dt;
    oh.d[1] = dh;
    oh.d[2] = ds;
    oh.o[0] = 0.0;
    oh.o[1] = h0;
    oh.o[2] = s0;
    std::string werr;
    if (!rsf::write(out_f, out.data(), oh, werr)) die(werr);
    if (want_interm && !rsf::write(kv["interm"], mid.data(), oh, werr)) die(werr);
    return 0;
}
