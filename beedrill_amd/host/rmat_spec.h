// rmat_spec.h -- the generator's graph names for the CLIs' --generate flag
// (generator/rmat_args.h:26-55): "A-B-C-D-E-V.rmat" (probabilities, edge and
// vertex counts with K/M/G/T suffixes) or "graph500-scaleN".
#pragma once

#include <regex>
#include <string>

inline bool parse_rmat(const std::string &s, double *abcd, long *ne, long *nv) {
    std::smatch m;
    auto suffixed = [](const std::string &t) {
        long long n = std::stoll(t);
        switch (t.back()) {
        case 'K': n <<= 10; break;
        case 'M': n <<= 20; break;
        case 'G': n <<= 30; break;
        case 'T': n <<= 40; break;
        default: break;
        }
        return (long)n;
    };
    std::regex r1(R"((\d+[.]\d+)-(\d+[.]\d+)-(\d+[.]\d+)-(\d+[.]\d+)-(\d+[KMGT]?)-(\d+[KMGT]?)\.rmat)");
    std::regex r2(R"(graph500-scale(\d+)(\.mtx)?)");
    if (std::regex_match(s, m, r1)) {
        for (int i = 0; i < 4; ++i) abcd[i] = std::stod(m[i + 1]);
        *ne = suffixed(m[5]);
        *nv = suffixed(m[6]);
        return true;
    }
    if (std::regex_match(s, m, r2)) {
        abcd[0] = 0.57; abcd[1] = 0.19; abcd[2] = 0.19; abcd[3] = 0.05;
        *nv = 1L << std::stol(m[1]);
        *ne = 16 * *nv;
        return true;
    }
    return false;
}
