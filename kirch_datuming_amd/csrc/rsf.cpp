// RSF float I/O (see rsf.h for the header semantics it follows).
#include "rsf.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace rsf {

size_t Header::count() const
{
    size_t c = 1;
    for (int i = 0; i < ndim; i++) c *= (size_t)n[i];
    return c;
}

static std::string dirname_of(const std::string &p)
{
    size_t s = p.rfind('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

static std::string basename_of(const std::string &p)
{
    size_t s = p.rfind('/');
    return s == std::string::npos ? p : p.substr(s + 1);
}

// split a header line into whitespace-separated fields, honouring quotes
static std::vector<std::string> fields(const std::string &line)
{
    std::vector<std::string> out;
    std::string cur;
    char q = 0;
    for (char ch : line) {
        if (q) {
            if (ch == q) q = 0;
            else cur += ch;
        } else if (ch == '"' || ch == '\'') {
            q = ch;
        } else if (isspace((unsigned char)ch)) {
            if (!cur.empty()) out.push_back(cur), cur.clear();
        } else {
            cur += ch;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

bool read_header(const std::string &path, Header &h, std::string &err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "Cannot open SEP file '" + path + "'";
        return false;
    }
    h = Header();
    h.n[0] = 0;
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t eot = text.find('\x04');
    if (eot != std::string::npos) text.resize(eot);
    std::istringstream ss(text);
    std::string line;
    while (std::getline(ss, line)) {
        size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        for (const std::string &fld : fields(line)) {
            size_t eq = fld.find('=');
            if (eq == std::string::npos || eq == 0) continue;
            std::string k = fld.substr(0, eq), v = fld.substr(eq + 1);
            if (k.size() == 2 && (k[0] == 'n' || k[0] == 'o' || k[0] == 'd') && k[1] >= '1' &&
                k[1] <= '9') {
                const int idx = k[1] - '1';
                char *end = nullptr;
                const double dv = strtod(v.c_str(), &end);
                if (end == v.c_str()) continue;
                if (k[0] == 'n') {
                    if (dv < 1.0) continue;
                    h.n[idx] = (int)dv;
                } else if (k[0] == 'o') {
                    h.o[idx] = dv;
                } else {
                    if (dv == 0.0) continue;
                    h.d[idx] = dv;
                }
                if (idx + 1 > h.ndim) h.ndim = idx + 1;
            } else if (k == "in") {
                if (!v.empty()) h.in = v;
            } else if (k == "esize") {
                const int e = atoi(v.c_str());
                if (e == 1 || e == 2 || e == 4 || e == 8 || e == 12) h.esize = e;
            } else if (k == "data_format") {
                if (v == "native_float") h.little = true;
                else if (v == "xdr_float") h.little = false;
            } else {
                h.extra[k] = v;
            }
        }
    }
    if (h.in.empty()) {
        err = "Missing in= in SEP file '" + path + "'";
        return false;
    }
    if (h.in[0] != '/') {
        std::string cand = dirname_of(path) + "/" + h.in;
        std::ifstream t(cand);
        if (t) h.in = cand;
    }
    return true;
}

bool read_floats(const std::string &path, const Header &h, std::vector<float> &out,
                 std::string &err)
{
    if (h.esize != 4) {
        err = "Only esize=4 float data is supported in '" + path + "'";
        return false;
    }
    const size_t n = h.count();
    out.resize(n);
    FILE *f = fopen(h.in.c_str(), "rb");
    if (!f) {
        err = "Cannot open binary '" + h.in + "' of '" + path + "'";
        return false;
    }
    const size_t got = fread(out.data(), sizeof(float), n, f);
    fclose(f);
    if (got != n) {
        err = "Failed to read " + std::to_string(n) + " floats from '" + h.in + "'";
        return false;
    }
    if (!h.little) {
        for (float &x : out) {
            uint32_t u;
            memcpy(&u, &x, 4);
            u = __builtin_bswap32(u);
            memcpy(&x, &u, 4);
        }
    }
    return true;
}

static std::string num(double v)
{
    char b[64];
    snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

bool write(const std::string &path, const float *data, const Header &h, std::string &err)
{
    const std::string bin = path + "@";
    FILE *f = fopen(bin.c_str(), "wb");
    if (!f) {
        err = "Cannot open '" + bin + "' for writing";
        return false;
    }
    const size_t n = h.count();
    if (fwrite(data, sizeof(float), n, f) != n) {
        fclose(f);
        err = "Failed to write '" + bin + "'";
        return false;
    }
    fclose(f);
    std::ofstream o(path);
    if (!o) {
        err = "Cannot open '" + path + "' for writing";
        return false;
    }
    o << "\tin=\"" << basename_of(bin) << "\"\n";
    o << "\tdata_format=\"native_float\"\n";
    for (int i = 0; i < h.ndim; i++) o << "\tn" << i + 1 << "=" << h.n[i] << "\n";
    for (int i = 0; i < h.ndim; i++) o << "\td" << i + 1 << "=" << num(h.d[i]) << "\n";
    for (int i = 0; i < h.ndim; i++) o << "\to" << i + 1 << "=" << num(h.o[i]) << "\n";
    o << "\tesize=4\n";
    for (auto &kv : h.extra) o << "\t" << kv.first << "=\"" << kv.second << "\"\n";
    return (bool)o;
}

}  // namespace rsf
