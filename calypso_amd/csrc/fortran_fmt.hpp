// fortran_fmt.hpp -- the Fortran edit descriptors the reference's text files
// are written with, for byte-identical host-side writers.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace calypso {

// 1PEw.dE3: one digit before the point, d after, a signed three-digit
// exponent, right-justified in w characters
inline std::string fort_e(double v, int w, int d)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*E", d, v);
    char *e = std::strchr(buf, 'E');
    int ex = 0;
    if (e) {
        ex = std::atoi(e + 1);
        *e = 0;
    }
    char out[96];
    std::snprintf(out, sizeof out, "%sE%c%03d", buf, ex < 0 ? '-' : '+', std::abs(ex));
    std::string s(out);
    if ((int)s.size() < w) s.insert(0, (size_t)(w - (int)s.size()), ' ');
    return s;
}

// Iw
inline std::string fort_i(long long v, int w)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%*lld", w, v);
    return buf;
}

// '(8i16)' of a list: eight values a line; an empty list gives one empty line
inline std::string fort_i16_rows(const std::vector<long long> &v)
{
    std::string s;
    if (v.empty()) return "\n";
    for (size_t i = 0; i < v.size(); ++i) {
        s += fort_i(v[i], 16);
        if (i % 8 == 7 || i + 1 == v.size()) s += "\n";
    }
    return s;
}

}  // namespace calypso
