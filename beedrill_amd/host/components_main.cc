// components_main.cc -- command-line driver for connected components with the
// reference's flags and output lines (components_main.cc:10-215).  Additive
// flags: --generate (a seeded device-generated RMAT graph, the generator's
// name syntax), --seed and --device.
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>

#include "components.h"
#include "graph.h"
#include "rmat_spec.h"

#define LOG(...)                      \
    do {                              \
        fprintf(stdout, __VA_ARGS__); \
        fflush(stdout);               \
    } while (0)

#ifndef BD_VERSION
#define BD_VERSION "beedrill_amd-gfx950"
#endif

static const struct option long_options[] = {{"graph_filename", required_argument, 0, 0},
                                             {"distributed_load", no_argument, 0, 0},
                                             {"num_trials", required_argument, 0, 0},
                                             {"dump_edge_list", no_argument, 0, 0},
                                             {"check_graph", no_argument, 0, 0},
                                             {"dump_graph", no_argument, 0, 0},
                                             {"check_results", no_argument, 0, 0},
                                             {"dump_results", no_argument, 0, 0},
                                             {"version", no_argument, 0, 0},
                                             {"help", no_argument, 0, 0},
                                             {"generate", required_argument, 0, 0},
                                             {"seed", required_argument, 0, 0},
                                             {"device", required_argument, 0, 0},
                                             {nullptr, 0, 0, 0}};

static void print_help(const char *argv0) {
    LOG("Usage: %s [OPTIONS]\n", argv0);
    LOG("\t--graph_filename     el64 edge-list file to read (see --generate for a synthetic graph)\n");
    LOG("\t--distributed_load   Kept for scripts that pass it; the file is always read once\n");
    LOG("\t--num_trials         Number of labelling runs\n");
    LOG("\t--dump_edge_list     Write the loaded edges to stdout (large output)\n");
    LOG("\t--check_graph        Compare the built CSR with the edge list (takes a while)\n");
    LOG("\t--dump_graph         Write the built CSR to stdout (large output)\n");
    LOG("\t--check_results      Verify the component labels after the runs\n");
    LOG("\t--dump_results       Write each vertex's component label to stdout\n");
    LOG("\t--version            Show the build version and exit\n");
    LOG("\t--help               Show this list and exit\n");
    LOG("\t--generate SPEC      Build graph500-scaleN or A-B-C-D-E-V.rmat on the GPU\n");
    LOG("\t--seed N             Generator seed (default 0)\n");
    LOG("\t--device N           GPU ordinal (default 0)\n");
}

struct cc_args {
    const char *graph_filename = nullptr;
    long num_trials = 1;
    bool dump_edge_list = false, check_graph = false, dump_graph = false;
    bool check_results = false, dump_results = false;
    const char *generate = nullptr;
    unsigned seed = 0;
    int device = 0;

    static cc_args parse(int argc, char **argv) {
        cc_args a;
        int idx = 0;
        while (true) {
            int c = getopt_long(argc, argv, "", long_options, &idx);
            if (c == -1) break;
            if (c == '?') {
                LOG("Invalid arguments\n");
                print_help(argv[0]);
                exit(1);
            }
            const std::string n = long_options[idx].name;
            if (n == "graph_filename") a.graph_filename = optarg;
            else if (n == "num_trials") a.num_trials = atol(optarg);
            else if (n == "dump_edge_list") a.dump_edge_list = true;
            else if (n == "check_graph") a.check_graph = true;
            else if (n == "dump_graph") a.dump_graph = true;
            else if (n == "check_results") a.check_results = true;
            else if (n == "dump_results") a.dump_results = true;
            else if (n == "version") {
                LOG("%s\n", BD_VERSION);
                exit(0);
            } else if (n == "help") {
                print_help(argv[0]);
                exit(1);
            } else if (n == "generate") a.generate = optarg;
            else if (n == "seed") a.seed = (unsigned)strtoul(optarg, nullptr, 10);
            else if (n == "device") a.device = atoi(optarg);
        }
        if (a.graph_filename == nullptr && a.generate == nullptr) {
            LOG("Missing graph filename\n");
            exit(1);
        }
        if (a.num_trials <= 0) {
            LOG("num_trials must be > 0\n");
            exit(1);
        }
        return a;
    }
};

int main(int argc, char **argv) {
    bool success = true;
    cc_args args = cc_args::parse(argc, argv);
    try {
        auto t0 = std::chrono::steady_clock::now();
        dist_edge_list::handle el;
        std::unique_ptr<emu::repl_shallow<graph>> g;
        if (args.generate) {
            double abcd[4];
            long ne = 0, nv = 0;
            if (!parse_rmat(args.generate, abcd, &ne, &nv)) {
                LOG("Cannot parse --generate '%s'\n", args.generate);
                exit(1);
            }
            bd_handle ctx = bd_open(args.device);
            int64_t m = 0;
            bd_check(ctx.get(), bd_graph_generate(ctx.get(), BD_GEN_RMAT, nv, ne, abcd[0], abcd[1],
                                                  abcd[2], abcd[3], args.seed, &m),
                     "generate");
            g = emu::make_repl_shallow<graph>(ctx);
        } else {
            el = dist_edge_list::load_binary(args.graph_filename, args.device);
            g = create_graph_from_edge_list<graph>(*el);
        }
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                        .count();
        LOG("Loaded %li edges in %3.2f ms, %3.2f MB/s\n", g->num_edges(), ms,
            (1e-6 * g->num_edges() * 16) / (1e-3 * ms));
        if (args.dump_edge_list && el) {
            LOG("Dumping edge list...\n");
            el->dump();
        }
        LOG("Constructing graph...\n");
        g->print_distribution();
        if (args.check_graph) {
            LOG("Checking graph...");
            if (!el || g->check(*el)) {
                LOG("PASS\n");
            } else {
                LOG("FAIL\n");
                success = false;
            }
        }
        if (args.dump_graph) {
            LOG("Dumping graph...\n");
            g->dump();
        }

        LOG("Initializing data structures...\n");
        auto cc = emu::make_repl_shallow<components>(*g);
        for (long trial = 0; trial < args.num_trials; ++trial) {
            cc->clear();
            LOG("Finding connected components...\n");
            components::stats s = cc->run();
            const double time_ms = cc->last_ms();
            const double teps = s.num_iters * g->num_edges() / (1e-3 * time_ms);
            LOG("Found %li components in %li iterations (%3.2f ms, %3.2f GTEPS)\n",
                s.num_components, s.num_iters, time_ms, 1e-9 * teps);
        }
        if (args.check_results) {
            LOG("Checking results...");
            if (cc->check()) {
                LOG("PASS\n");
            } else {
                LOG("FAIL\n");
                success = false;
            }
        }
        if (args.dump_results) cc->dump();
    } catch (const std::exception &e) {
        LOG("%s\n", e.what());
        return 2;
    }
    return !success;
}
