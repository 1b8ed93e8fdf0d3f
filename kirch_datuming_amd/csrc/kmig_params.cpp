// Parameter defaults, key=value parsing and validation of the kirchmig2d_cig
// adj=1 path.  Host code only (no HIP calls), so it runs without a GPU.
//
// Follows src/kirchmig2d_cig.cpp:548-875 (defaults and key handling),
// parse_output_axis :323-368, and the parameter store of
// SEFILESYSTEM/src/se_par_sep.cpp:185-330 (key=value and par=file, later
// keys replace earlier ones) with its getters :432-446
// (int = (int32_t)double, float = (float)double).
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "kmig.h"
#include "kmig_internal.h"

namespace kmig {

void set_err(char *err, int errlen, const char *fmt, ...)
{
    if (!err || errlen <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
}

using KV = std::map<std::string, std::string>;

static void put_token(KV &kv, const std::string &tok)
{
    size_t eq = tok.find('=');
    if (eq == std::string::npos || eq == 0) return;
    std::string k = tok.substr(0, eq), v = tok.substr(eq + 1);
    if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') ||
                          (v.front() == '\'' && v.back() == '\'')))
        v = v.substr(1, v.size() - 2);
    kv[k] = v;
}

// par=file: whitespace-separated key=value tokens, '#' starts a comment.
static bool read_par_file(KV &kv, const std::string &path)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ss(line);
        std::string tok;
        while (ss >> tok) put_token(kv, tok);
    }
    return true;
}

int collect_kv(KV &kv, int argc, const char *const *argv, char *err, int errlen)
{
    for (int i = 0; i < argc; i++) {
        if (!argv[i]) continue;
        std::string a(argv[i]);
        if (a.rfind("par=", 0) == 0) {
            if (!read_par_file(kv, a.substr(4))) {
                set_err(err, errlen, "Cannot open parameter file %s", a.c_str() + 4);
                return KMIG_E_IO;
            }
        } else {
            put_token(kv, a);
        }
    }
    return KMIG_OK;
}

static bool has(const KV &kv, const char *k) { return kv.find(k) != kv.end(); }
static double dbl(const KV &kv, const char *k) { return strtod(kv.at(k).c_str(), nullptr); }
static int geti(const KV &kv, const char *k) { return (int)(int32_t)dbl(kv, k); }
static float getf(const KV &kv, const char *k) { return (float)dbl(kv, k); }

static bool ieq(const std::string &a, const char *b)
{
    if (a.size() != strlen(b)) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (tolower((unsigned char)a[i]) != tolower((unsigned char)b[i])) return false;
    return true;
}

// parse_output_axis, src/kirchmig2d_cig.cpp:323-368 (float arithmetic).
static int output_axis(const KV &kv, int is_angle, int nh, float h0, float dh, float *g0,
                       float *dg, int *ng, int half_angle, char *err, int errlen)
{
    float hend = h0 + (float)(nh - 1) * dh;
    float vmin = is_angle ? 0.0f : std::fmin(h0, hend);
    float vmax = is_angle ? (half_angle ? 90.0f : 180.0f) : std::fmax(h0, hend);
    float vstep = is_angle ? 1.0f : (dh != 0.0f ? fabsf(dh) : 1.0f);
    int n_given = 0;
    int nval = is_angle ? (half_angle ? 91 : 181) : nh;
    const char *pre[2] = {"cig", is_angle ? "angle" : "offset"};
    for (const char *p : pre) {
        std::string b(p);
        if (has(kv, (b + "_min").c_str())) vmin = getf(kv, (b + "_min").c_str());
        if (has(kv, (b + "_max").c_str())) vmax = getf(kv, (b + "_max").c_str());
        if (has(kv, (b + "_step").c_str())) vstep = getf(kv, (b + "_step").c_str());
        if (has(kv, (b + "_n").c_str())) {
            nval = geti(kv, (b + "_n").c_str());
            n_given = 1;
        }
    }
    if (vstep <= 0.0f) {
        set_err(err, errlen, "CIG axis step must be positive.");
        return KMIG_E_PARAM;
    }
    if (!n_given) nval = (int)floorf((vmax - vmin) / vstep + 0.5f) + 1;
    if (nval <= 0) {
        set_err(err, errlen, "CIG axis sample number must be positive.");
        return KMIG_E_PARAM;
    }
    *g0 = vmin;
    *dg = vstep;
    *ng = nval;
    return KMIG_OK;
}

int parse_kv(kmig_params *p, const KV &kv, int nh, float h0, float dh, char *err, int errlen)
{
    kmig_params_default(p);
    if (has(kv, "aperture_trace")) p->aperture_trace = geti(kv, "aperture_trace");
    const bool mode_given = has(kv, "mode") || has(kv, "cig_mode");
    std::string mode = has(kv, "mode") ? kv.at("mode")
                       : has(kv, "cig_mode") ? kv.at("cig_mode") : std::string("offset");
    if (ieq(mode, "offset")) {
        p->cig_is_angle = 0;
    } else if (ieq(mode, "angle")) {
        p->cig_is_angle = 1;
    } else {
        set_err(err, errlen, "Unknown CIG mode: %s. Supported modes are offset and angle.",
                mode.c_str());
        return KMIG_E_PARAM;
    }
    p->adj = has(kv, "adj") ? geti(kv, "adj") : 1;
    p->cig = has(kv, "cig") ? geti(kv, "cig") : (mode_given ? 1 : 0);
    p->cmp = has(kv, "cmp") ? geti(kv, "cmp") : 1;
    if (has(kv, "half_angle")) p->half_angle = geti(kv, "half_angle");
    if (p->half_angle) {
        p->oazmin = 40.0f;
        p->oazmax = 25.0f;
    }
    if (has(kv, "oazmin")) p->oazmin = getf(kv, "oazmin");
    if (has(kv, "oazmax")) p->oazmax = getf(kv, "oazmax");
    if (has(kv, "oaztaper")) p->oaztaper = getf(kv, "oaztaper");
    p->diff = has(kv, "diff") ? geti(kv, "diff") : 0;
    if (has(kv, "direct_mute")) p->direct_mute = geti(kv, "direct_mute");
    if (has(kv, "direct_mute_slope")) p->direct_mute_slope = getf(kv, "direct_mute_slope");
    if (has(kv, "direct_mute_invert")) p->direct_mute_invert = geti(kv, "direct_mute_invert");
    if (has(kv, "direct_mute_intercept"))
        p->direct_mute_intercept = getf(kv, "direct_mute_intercept");
    if (has(kv, "direct_mute_taper")) p->direct_mute_taper = getf(kv, "direct_mute_taper");
    if (has(kv, "direct_mute_qc_shot")) p->direct_mute_qc_shot = geti(kv, "direct_mute_qc_shot");
    if (has(kv, "tau")) p->tau = getf(kv, "tau");
    if (has(kv, "aperture")) p->aperture = getf(kv, "aperture");
    if (has(kv, "aperture_taper")) p->aperture_taper = getf(kv, "aperture_taper");
    if (has(kv, "aperture_rmin")) p->aperture_rmin = getf(kv, "aperture_rmin");
    if (has(kv, "aperture_rtaper")) p->aperture_rtaper = getf(kv, "aperture_rtaper");
    if (has(kv, "antialias")) p->antialias = getf(kv, "antialias");
    if (has(kv, "angle_interp")) p->angle_interp = geti(kv, "angle_interp");
    if (has(kv, "angle_taper_start")) p->angle_taper_start = getf(kv, "angle_taper_start");
    if (has(kv, "angle_taper_end")) p->angle_taper_end = getf(kv, "angle_taper_end");
    if (has(kv, "angle_grad_radius")) p->angle_grad_radius = geti(kv, "angle_grad_radius");
    if (p->angle_grad_radius < 1) p->angle_grad_radius = 1;
    if (has(kv, "angle_smooth")) p->angle_smooth = geti(kv, "angle_smooth");
    if (p->angle_smooth < 0) p->angle_smooth = 0;
    if (has(kv, "aperture_trace_taper"))
        p->aperture_trace_taper = geti(kv, "aperture_trace_taper");
    if (p->aperture_trace_taper < 0) p->aperture_trace_taper = 0;
    if (has(kv, "type")) {
        const std::string &t = kv.at("type");
        if (t.empty() || (t[0] != 'l' && t[0] != 'p' && t[0] != 'h')) {
            set_err(err, errlen,
                    "Unknown interpolation type: %s. Supported types are linear (l), partial "
                    "(p), and hermite (h)",
                    t.c_str());
            return KMIG_E_PARAM;
        }
        p->interp_type = t[0];
    }
    if (has(kv, "precision")) {
        const std::string &t = kv.at("precision");
        if (ieq(t, "fp32")) p->precision = KMIG_PREC_FP32;
        else if (ieq(t, "fp16")) p->precision = KMIG_PREC_FP16;
        else if (ieq(t, "bf16")) p->precision = KMIG_PREC_BF16;
        else if (ieq(t, "u16")) p->precision = KMIG_PREC_U16;
        else {
            set_err(err, errlen, "Unknown precision: %s. Supported are fp32, fp16, bf16, u16.",
                    t.c_str());
            return KMIG_E_PARAM;
        }
    }
    if (has(kv, "tile")) p->tile = geti(kv, "tile");
    if (has(kv, "variant")) {
        const std::string &t = kv.at("variant");
        if (ieq(t, "auto")) p->variant = KMIG_VARIANT_AUTO;
        else if (ieq(t, "stream")) p->variant = KMIG_VARIANT_STREAM;
        else if (ieq(t, "resident")) p->variant = KMIG_VARIANT_RESIDENT;
        else {
            set_err(err, errlen, "Unknown variant: %s. Supported are auto, stream, resident.",
                    t.c_str());
            return KMIG_E_PARAM;
        }
    }
    if (nh == 1) dh = 0.0f; /* :700 */
    if (p->cig) {
        int rc = output_axis(kv, p->cig_is_angle, nh, h0, dh, &p->g0, &p->dg, &p->ng,
                             p->half_angle, err, errlen);
        if (rc) return rc;
    } else {
        p->ng = 1;
        p->g0 = 0.0f;
        p->dg = 1.0f;
    }
    return KMIG_OK;
}

}  // namespace kmig

extern "C" void kmig_params_default(kmig_params *p)
{
    memset(p, 0, sizeof(*p));
    p->adj = 1;
    p->cig = 0;
    p->cig_is_angle = 0;
    p->cmp = 1;
    p->half_angle = 0;
    p->oazmin = 80.0f;
    p->oazmax = 50.0f;
    p->oaztaper = 10.0f;
    p->diff = 0;
    p->aperture_trace = -1;
    p->aperture_trace_taper = 20;
    p->direct_mute = 0;
    p->direct_mute_invert = 0;
    p->direct_mute_qc_shot = -1;
    p->direct_mute_slope = 0.0f;
    p->direct_mute_intercept = 0.0f;
    p->direct_mute_taper = 0.05f;
    p->tau = 0.0f;
    p->aperture = 90.0f;
    p->aperture_taper = 10.0f;
    p->aperture_rmin = 0.0f;
    p->aperture_rtaper = 0.05f;
    p->antialias = 1.0f;
    p->angle_interp = 1;
    p->angle_taper_start = -1.0f;
    p->angle_taper_end = -1.0f;
    p->angle_grad_radius = 2;
    p->angle_smooth = 2;
    p->interp_type = 'h';
    p->g0 = 0.0f;
    p->dg = 1.0f;
    p->ng = 1;
    p->precision = KMIG_PREC_FP32;
    p->tile = 0;
    p->variant = KMIG_VARIANT_AUTO;
}

extern "C" int kmig_params_parse(kmig_params *p, int argc, const char *const *argv, int nh,
                                 float h0, float dh, char *err, int errlen)
{
    if (!p) return KMIG_E_PARAM;
    kmig::KV kv;
    int rc = kmig::collect_kv(kv, argc, argv, err, errlen);
    if (rc) return rc;
    return kmig::parse_kv(p, kv, nh, h0, dh, err, errlen);
}

extern "C" int kmig_params_check(const kmig_params *p, int ns, char *err, int errlen)
{
    using kmig::set_err;
    if (!p) return KMIG_E_PARAM;
    if (p->adj != 0 && p->adj != 1) {
        set_err(err, errlen, "adj=%d: use adj=1 (migration) or adj=0 (modelling).", p->adj);
        return KMIG_E_PARAM;
    }
    if (p->adj == 0 && p->direct_mute) { /* src/kirchmig2d_cig.cpp:719-720 */
        set_err(err, errlen,
                "direct_mute=1 is an input-data preprocessing option and requires adj=1.");
        return KMIG_E_PARAM;
    }
    if (p->interp_type != 'l' && p->interp_type != 'p' && p->interp_type != 'h') {
        set_err(err, errlen,
                "Unknown interpolation type: %c. Supported types are linear (l), partial (p), "
                "and hermite (h)",
                p->interp_type);
        return KMIG_E_PARAM;
    }
    if (p->direct_mute) { /* src/kirchmig2d_cig.cpp:719-738 */
        if (!std::isfinite(p->direct_mute_slope) || p->direct_mute_slope < 0.0f) {
            set_err(err, errlen, "direct_mute_slope must be finite and non-negative.");
            return KMIG_E_PARAM;
        }
        if (!std::isfinite(p->direct_mute_intercept)) {
            set_err(err, errlen, "direct_mute_intercept must be finite.");
            return KMIG_E_PARAM;
        }
        if (!std::isfinite(p->direct_mute_taper) || p->direct_mute_taper <= 0.0f) {
            set_err(err, errlen,
                    "direct_mute_taper must be finite and positive to avoid a hard cutoff.");
            return KMIG_E_PARAM;
        }
        if (p->direct_mute_invert != 0 && p->direct_mute_invert != 1) {
            set_err(err, errlen, "direct_mute_invert must be 0 or 1.");
            return KMIG_E_PARAM;
        }
        if (p->direct_mute_qc_shot < -1) {
            set_err(err, errlen,
                    "direct_mute_qc_shot must be -1 (disabled) or a zero-based shot index.");
            return KMIG_E_PARAM;
        }
        if (ns > 0 && p->direct_mute_qc_shot >= ns) {
            set_err(err, errlen, "direct_mute_qc_shot=%d is outside the input shot range [0,%d].",
                    p->direct_mute_qc_shot, ns - 1);
            return KMIG_E_PARAM;
        }
    }
    if (p->cig) {
        if (!(p->dg > 0.0f)) {
            set_err(err, errlen, "CIG axis step must be positive.");
            return KMIG_E_PARAM;
        }
        if (p->ng <= 0) {
            set_err(err, errlen, "CIG axis sample number must be positive.");
            return KMIG_E_PARAM;
        }
    }
    if (p->variant < KMIG_VARIANT_AUTO || p->variant > KMIG_VARIANT_RESIDENT) {
        set_err(err, errlen, "Unknown kernel variant %d.", p->variant);
        return KMIG_E_PARAM;
    }
    if (p->precision < KMIG_PREC_FP32 || p->precision > KMIG_PREC_U16) {
        set_err(err, errlen, "Unknown precision mode %d.", p->precision);
        return KMIG_E_PARAM;
    }
    return KMIG_OK;
}
