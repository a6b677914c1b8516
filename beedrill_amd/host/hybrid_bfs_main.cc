// hybrid_bfs_main.cc -- command-line driver with the reference's flags and
// output lines (hybrid_bfs_main.cc:10-309).  Additive flags: --generate (a
// seeded device-generated RMAT graph instead of a file; the reference
// generator's argument syntax, generator/rmat_args.h:26-55), --seed,
// --policy, --mlp_weights and --device.
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <regex>
#include <string>
#include <vector>

#include "graph.h"
#include "hybrid_bfs.h"
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
                                             {"heavy_threshold", required_argument, 0, 0},
                                             {"num_trials", required_argument, 0, 0},
                                             {"source_vertex", required_argument, 0, 0},
                                             {"algorithm", required_argument, 0, 0},
                                             {"alpha", required_argument, 0, 0},
                                             {"beta", required_argument, 0, 0},
                                             {"sort_edge_blocks", no_argument, 0, 0},
                                             {"dump_edge_list", no_argument, 0, 0},
                                             {"check_graph", no_argument, 0, 0},
                                             {"dump_graph", no_argument, 0, 0},
                                             {"check_results", no_argument, 0, 0},
                                             {"version", no_argument, 0, 0},
                                             {"help", no_argument, 0, 0},
                                             {"generate", required_argument, 0, 0},
                                             {"seed", required_argument, 0, 0},
                                             {"policy", required_argument, 0, 0},
                                             {"mlp_weights", required_argument, 0, 0},
                                             {"device", required_argument, 0, 0},
                                             {nullptr, 0, 0, 0}};

static void print_help(const char *argv0) {
    LOG("Usage: %s [OPTIONS]\n", argv0);
    LOG("\t--graph_filename     el64 edge-list file to read (see --generate for a synthetic graph)\n");
    LOG("\t--distributed_load   Kept for scripts that pass it; the file is always read once\n");
    LOG("\t--heavy_threshold    Kept for scripts that pass it; has no effect\n");
    LOG("\t--num_trials         Number of searches (one root each)\n");
    LOG("\t--source_vertex      Root of every search; by default each search draws a root\n");
    LOG("\t--algorithm          beamer_hybrid | migrating_threads | remote_writes |\n");
    LOG("\t                     remote_writes_hybrid\n");
    LOG("\t--alpha              Top-down to bottom-up switch factor (default 15)\n");
    LOG("\t--beta               Bottom-up to top-down switch factor (default 18)\n");
    LOG("\t--sort_edge_blocks   Order every adjacency list by neighbour id before searching\n");
    LOG("\t--dump_edge_list     Write the loaded edges to stdout (large output)\n");
    LOG("\t--check_graph        Compare the built CSR with the edge list (takes a while)\n");
    LOG("\t--dump_graph         Write the built CSR to stdout (large output)\n");
    LOG("\t--check_results      Verify every search's parent tree after it runs\n");
    LOG("\t--version            Show the build version and exit\n");
    LOG("\t--help               Show this list and exit\n");
    LOG("\t--generate SPEC      Build graph500-scaleN or A-B-C-D-E-V.rmat on the GPU\n");
    LOG("\t--seed N             Generator seed (default 0)\n");
    LOG("\t--policy NAME        beamer (default) | learned | td_only | bu_only\n");
    LOG("\t--mlp_weights FILE   float32 parameters of the 8-32-32-2 direction network\n");
    LOG("\t--device N           GPU ordinal (default 0)\n");
}

struct bfs_args {
    const char *graph_filename = nullptr;
    long num_trials = 1;
    long source_vertex = -1;
    const char *algorithm = "beamer_hybrid";
    long alpha = 15, beta = 18;
    bool sort_edge_blocks = false, dump_edge_list = false, check_graph = false;
    bool dump_graph = false, check_results = false;
    const char *generate = nullptr;
    unsigned seed = 0;
    int policy = BD_POLICY_BEAMER;
    const char *mlp_weights = nullptr;
    int device = 0;

    static bfs_args parse(int argc, char **argv) {
        bfs_args a;
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
            else if (n == "source_vertex") a.source_vertex = atol(optarg);
            else if (n == "algorithm") a.algorithm = optarg;
            else if (n == "alpha") a.alpha = atol(optarg);
            else if (n == "beta") a.beta = atol(optarg);
            else if (n == "sort_edge_blocks") a.sort_edge_blocks = true;
            else if (n == "dump_edge_list") a.dump_edge_list = true;
            else if (n == "check_graph") a.check_graph = true;
            else if (n == "dump_graph") a.dump_graph = true;
            else if (n == "check_results") a.check_results = true;
            else if (n == "version") {
                LOG("%s\n", BD_VERSION);
                exit(0);
            } else if (n == "help") {
                print_help(argv[0]);
                exit(1);
            } else if (n == "generate") a.generate = optarg;
            else if (n == "seed") a.seed = (unsigned)strtoul(optarg, nullptr, 10);
            else if (n == "policy") {
                const std::string p = optarg;
                if (p == "beamer") a.policy = BD_POLICY_BEAMER;
                else if (p == "learned") a.policy = BD_POLICY_LEARNED;
                else if (p == "td_only") a.policy = BD_POLICY_TD_ONLY;
                else if (p == "bu_only") a.policy = BD_POLICY_BU_ONLY;
                else {
                    LOG("Unknown policy '%s'\n", optarg);
                    exit(1);
                }
            } else if (n == "mlp_weights") a.mlp_weights = optarg;
            else if (n == "device") a.device = atoi(optarg);
        }
        if (a.graph_filename == nullptr && a.generate == nullptr) {
            LOG("Missing graph filename\n");
            exit(1);
        }
        if (a.num_trials <= 0) { LOG("num_trials must be > 0\n"); exit(1); }
        if (a.alpha <= 0) { LOG("alpha must be > 0\n"); exit(1); }
        if (a.beta <= 0) { LOG("beta must be > 0\n"); exit(1); }
        return a;
    }
};

int main(int argc, char **argv) {
    bool success = true;
    bfs_args args = bfs_args::parse(argc, argv);
    const uint64_t lcg_mul = 6364136223846793005ULL;
    uint64_t lcg_state = 1;  // lcg rng(0) (hybrid_bfs_main.cc:159)

    std::unique_ptr<emu::repl_shallow<graph>> g;
    try {
        auto t0 = std::chrono::steady_clock::now();
        dist_edge_list::handle el;
        bd_handle ctx;
        if (args.generate) {
            double abcd[4];
            long ne = 0, nv = 0;
            if (!parse_rmat(args.generate, abcd, &ne, &nv)) {
                LOG("Cannot parse --generate '%s'\n", args.generate);
                exit(1);
            }
            ctx = bd_open(args.device);
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
        if (args.sort_edge_blocks) LOG("Sorting edge lists is not needed on one GPU; skipped.\n");
        g->print_distribution();
        if (args.check_graph) {
            LOG("Checking graph...");
            if (!el) {
                LOG("(generated graph: no edge list file) ");
            }
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
        if (args.source_vertex >= g->num_vertices()) {
            LOG("Source vertex %li out of range.\n", args.source_vertex);
            exit(1);
        }
        if (!strcmp(args.algorithm, "none")) return !success;

        enum { REMOTE_WRITES, MIGRATING_THREADS, REMOTE_WRITES_HYBRID, BEAMER_HYBRID } alg;
        if (!strcmp(args.algorithm, "remote_writes")) alg = REMOTE_WRITES;
        else if (!strcmp(args.algorithm, "migrating_threads")) alg = MIGRATING_THREADS;
        else if (!strcmp(args.algorithm, "remote_writes_hybrid")) alg = REMOTE_WRITES_HYBRID;
        else if (!strcmp(args.algorithm, "beamer_hybrid")) alg = BEAMER_HYBRID;
        else {
            LOG("Algorithm '%s' not implemented!\n", args.algorithm);
            exit(1);
        }
        LOG("Initializing BFS data structures...\n");
        auto bfs = emu::make_repl_shallow<hybrid_bfs>(*g);
        bfs->set_policy(args.policy);
        if (args.mlp_weights) {
            FILE *fp = fopen(args.mlp_weights, "rb");
            if (!fp) {
                LOG("Unable to open %s\n", args.mlp_weights);
                exit(1);
            }
            std::vector<float> p(32 * 8 + 32 + 32 * 32 + 32 + 2 * 32 + 2);
            size_t got = fread(p.data(), sizeof(float), p.size(), fp);
            fclose(fp);
            if (got != p.size()) {
                LOG("Short weight file %s\n", args.mlp_weights);
                exit(1);
            }
            bd_check(g->context(), bd_mlp_set(g->context(), p.data(), 32, 32), "mlp_set");
            // the override margin chosen at training time, from the JSON
            // summary written next to the weights ("margin_milli": N)
            std::string meta(args.mlp_weights);
            const size_t dot = meta.rfind('.');
            meta = (dot == std::string::npos ? meta : meta.substr(0, dot)) + ".json";
            if (FILE *mf = fopen(meta.c_str(), "rb")) {
                std::string text;
                char buf[4096];
                size_t k;
                while ((k = fread(buf, 1, sizeof(buf), mf)) > 0) text.append(buf, k);
                fclose(mf);
                std::smatch mm;
                if (std::regex_search(text, mm, std::regex("\"margin_milli\"\\s*:\\s*([0-9]+)"))) {
                    bd_tuning t;
                    memset(&t, 0, sizeof(t));
                    t.learned_margin_milli = std::stoi(mm[1].str());
                    bd_check(g->context(), bd_set_tuning(g->context(), &t), "set_tuning");
                }
            }
        }

        long edges_all = 0;
        double ms_all = 0;
        std::vector<double> teps;
        for (long s = 0; s < args.num_trials; ++s) {
            long source;
            if (args.source_vertex >= 0) {
                source = args.source_vertex;
            } else {
                // pick_random_vertex (hybrid_bfs_main.cc:136-144)
                do {
                    lcg_state = lcg_mul * lcg_state + 1ULL;
                    source = (long)(lcg_state % (uint64_t)g->num_vertices());
                } while (g->out_degree(source) == 0);
            }
            bfs->clear();
            LOG("Doing breadth-first search from vertex %li (sample %li of %li)\n", source, s + 1,
                args.num_trials);
            switch (alg) {
            case REMOTE_WRITES: bfs->run_with_remote_writes(source); break;
            case MIGRATING_THREADS: bfs->run_with_migrating_threads(source); break;
            case REMOTE_WRITES_HYBRID:
                bfs->run_with_remote_writes_hybrid(source, args.alpha, args.beta);
                break;
            case BEAMER_HYBRID: bfs->run_beamer(source, args.alpha, args.beta); break;
            }
            const double time_ms = bfs->last_stats().ms;
            if (args.check_results) {
                LOG("Checking results...\n");
                if (bfs->check(source)) {
                    LOG("PASS\n");
                } else {
                    LOG("FAIL\n");
                    success = false;
                }
            }
            const long traversed = bfs->count_num_traversed_edges();
            edges_all += traversed;
            ms_all += time_ms;
            teps.push_back(traversed / (time_ms / 1000));
            LOG("Traversed %li edges in %3.2f ms, %3.2f MTEPS \n", traversed, time_ms,
                (1e-6 * traversed) / (time_ms / 1000));
        }
        LOG("Mean performance over all trials: %3.2f MTEPS \n",
            (1e-6 * edges_all) / (ms_all / 1000));
        if (args.num_trials > 1) {
            // combined.cc:108-146, 315-327
            double rs = 0, sum = 0, mn = teps[0], mx = teps[0];
            for (double x : teps) {
                rs += 1.0 / x;
                sum += x;
                mn = x < mn ? x : mn;
                mx = x > mx ? x : mx;
            }
            const double mean = sum / teps.size(), hm = teps.size() / rs;
            double sq = 0;
            for (double x : teps) sq += (x - mean) * (x - mean);
            const double sd = sqrt(sq / (teps.size() - 1));
            LOG("    BFS: %3.2f +/- %3.2f GTEPS, min/max %3.2f/%3.2f GTEPS\n", 1e-9 * hm, 1e-9 * sd,
                1e-9 * mn, 1e-9 * mx);
        }
    } catch (const std::exception &e) {
        LOG("%s\n", e.what());
        return 1;
    }
    return !success;
}
