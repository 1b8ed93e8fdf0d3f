// Minimal RSF (Madagascar/SEP) float I/O for the kirchmig2d_cig CLI.
// Header semantics follow SEFILESYSTEM/src/se_fs_sep.cpp:440-640 (parse) and
// :780-975 (write): key=value fields, later ones win; n/o/d 1..9; esize=4;
// data_format native_float (little-endian) or xdr_float (big-endian, also the
// default when absent); in= names the binary, relative to the header's
// directory; written as <header>@.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace rsf {

struct Header {
    int n[9] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
    double o[9] = {0};
    double d[9] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
    int ndim = 1;
    int esize = 4;
    bool little = false;
    std::string in;
    std::map<std::string, std::string> extra;
    size_t count() const;
};

// Returns false and sets err on failure.
bool read_header(const std::string &path, Header &h, std::string &err);
bool read_floats(const std::string &path, const Header &h, std::vector<float> &out,
                 std::string &err);
// Writes native_float header + binary "<path>@".
bool write(const std::string &path, const float *data, const Header &h, std::string &err);

}  // namespace rsf
