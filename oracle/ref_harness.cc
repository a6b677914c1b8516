// ref_harness.cc -- drives the reference's own generator headers
// (generator/prng_engine.hpp, generator/rmat_generator.h,
// generator/edge_list_utils.h) and lcg.h, compiled where they lie under the
// reference tree, to produce golden vectors for the oracle.  TEST
// INFRASTRUCTURE ONLY: the output binary lives in oracle/_ref/.
//
//   ref_harness sitmo <seed> <n> <out.bin>          n uint32 engine outputs
//   ref_harness rmat <nv> <ne> <a> <b> <c> <d> <seed> <raw.bin> <dedup.bin>
//                                                   int64 (src,dst) pairs
//   ref_harness lcg <step> <n> <out.bin>            n uint64 lcg outputs
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "prng_engine.hpp"
#include "rmat_generator.h"
#include "edge_list_utils.h"
#include "lcg.h"

static int write_all(const char* path, const void* p, size_t bytes) {
    FILE* fp = fopen(path, "wb");
    if (!fp) return 1;
    size_t w = fwrite(p, 1, bytes, fp);
    fclose(fp);
    return w == bytes ? 0 : 1;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    if (!strcmp(argv[1], "sitmo") && argc == 5) {
        uint32_t seed = (uint32_t)strtoul(argv[2], nullptr, 10);
        long n = atol(argv[3]);
        sitmo::prng_engine eng(seed);
        std::vector<uint32_t> out(n);
        for (long i = 0; i < n; ++i) out[i] = eng();
        return write_all(argv[4], out.data(), out.size() * sizeof(uint32_t));
    }
    if (!strcmp(argv[1], "rmat") && argc == 11) {
        long nv = atol(argv[2]);
        long ne = atol(argv[3]);
        double a = atof(argv[4]), b = atof(argv[5]), c = atof(argv[6]), d = atof(argv[7]);
        uint32_t seed = (uint32_t)strtoul(argv[8], nullptr, 10);
        rmat_edge_generator gen(nv, a, b, c, d, seed);
        std::vector<edge> edges(ne);
        rmat_fill(gen, edges.begin(), edges.end());
        if (write_all(argv[9], edges.data(), edges.size() * sizeof(edge))) return 1;
        flip_edges(edges.begin(), edges.end());
        sort_edges(edges.begin(), edges.end());
        auto new_end = dedup_edges(edges.begin(), edges.end());
        edges.resize(new_end - edges.begin());
        return write_all(argv[10], edges.data(), edges.size() * sizeof(edge));
    }
    if (!strcmp(argv[1], "lcg") && argc == 5) {
        unsigned long step = strtoul(argv[2], nullptr, 10);
        long n = atol(argv[3]);
        lcg rng(step);
        std::vector<unsigned long> out(n);
        for (long i = 0; i < n; ++i) out[i] = rng();
        return write_all(argv[4], out.data(), out.size() * sizeof(unsigned long));
    }
    return 2;
}
