/*
 * oracle.h -- CPU restatement of the reference's direction-optimizing BFS path
 * (hybrid_bfs --alg beamer_hybrid) and of its input generator.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, and by nothing else.  The product (beedrill_amd/) never links
 * or calls it; the product fails loudly when its HIP library is missing.
 *
 * Every function names the reference file:line it restates.  Layout follows
 * the product's device CSR: int64 offsets[V+1], int32 neighbour ids[nnz].
 *
 * Pinning (what checks each function independently of its own formula):
 *   oracle_sitmo_*, oracle_rmat_raw, oracle_flip_sort_dedup
 *       -> byte-exact against oracle/_ref/ref_harness, built from the
 *          reference's own generator headers (tests/golden fixtures).
 *   oracle_lcg_*, oracle_pick_roots
 *       -> byte-exact against oracle/_ref/ref_harness built from lcg.h, plus
 *          the closed-form known answers in tests.
 *   oracle_build_csr     -> graph_base::check invariant + scipy COO->CSR.
 *   oracle_bfs_depth     -> scipy.sparse.csgraph / networkx shortest paths,
 *                           brute force (Floyd-Warshall) on tiny graphs.
 *   oracle_beamer        -> hand-derived direction sequences (path, star,
 *                           two components, complete graph), depth vs the
 *                           above, planted-mutation tests.
 *   oracle_check         -> planted corrupt parents must FAIL.
 *   oracle_count_traversed, oracle_aggregate -> closed forms / numpy.
 *   oracle_head_index    -> numpy lexsort over (vertex, -degree, position).
 *   oracle_vperm, oracle_shuffle_keys -> build-defined permutation (the
 *          reference's std::random_shuffle under _GLIBCXX_PARALLEL is
 *          thread-count dependent); bijection and multiset tests.
 *   oracle_mlp_*         -> finite differences and numpy.
 *   oracle_cc            -> scipy.sparse.csgraph.connected_components (the
 *                           same partition, each part labelled by its
 *                           smallest id) and closed forms on special graphs.
 *   oracle_cc_check      -> planted wrong labels must FAIL.
 */
#ifndef BEEDRILL_ORACLE_H
#define BEEDRILL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- generator (generator/prng_engine.hpp, rmat_generator.h) */

/* One sitmo block: 256-bit counter s under key k -> 256 bits of output
 * (prng_engine.hpp:305-341 encrypt_counter). */
void oracle_sitmo_block(const uint64_t k[4], const uint64_t s[4], uint64_t out[4]);

/* n-th 32-bit output of a sitmo engine seeded with `seed`
 * (prng_engine.hpp:139-150 seed(uint32), :190-211 operator()). */
uint32_t oracle_sitmo_word(uint32_t seed, uint64_t n);

/* Raw RMAT edges exactly as rmat_fill() produces them, including the serial
 * self-loop re-roll (rmat_generator.h:70-193).  nv = number of vertices
 * (SCALE = floor(log2 nv)).  Writes num_edges (src,dst) pairs. */
int oracle_rmat_raw(int64_t nv, int64_t num_edges, double a, double b, double c,
                    double d, uint32_t seed, int64_t *src, int64_t *dst);

/* flip_edges + sort_edges + dedup_edges (edge_list_utils.h:13-57).
 * Returns the deduplicated count; arrays are compacted in place. */
int64_t oracle_flip_sort_dedup(int64_t n, int64_t *src, int64_t *dst);

/* Build-defined seeded vertex relabelling, a bijection on [0, nv). */
int64_t oracle_vperm(int64_t x, int64_t nv, uint64_t seed);

/* Build-defined seeded shuffle key of deduplicated edge j. */
uint64_t oracle_shuffle_key(uint64_t j, uint64_t seed);

/* Full seeded RMAT pipeline: raw -> flip/sort/dedup -> relabel -> shuffle.
 * src/dst must hold num_edges entries; *m_out receives the final count. */
int oracle_make_rmat(int64_t nv, int64_t num_edges, double a, double b, double c,
                     double d, uint32_t seed, int64_t *src, int64_t *dst,
                     int64_t *m_out);

/* Erdos-Renyi with uniform endpoints (no reference counterpart): raw ->
 * self-loops dropped -> flip/sort/dedup -> shuffle. */
int oracle_make_er(int64_t nv, int64_t num_edges, uint32_t seed, int64_t *src,
                   int64_t *dst, int64_t *m_out);

/* rows x cols 4-neighbour lattice, vertex r*cols+c, then shuffled. */
int oracle_make_lattice(int64_t rows, int64_t cols, uint32_t seed, int64_t *src,
                        int64_t *dst, int64_t *m_out);

/* el64 file writer (generator/edge_list_utils.h:171-198 dump_bin). */
int oracle_write_el64(const char *path, int64_t nv, int64_t m, const int64_t *src,
                      const int64_t *dst);

/* ---------------- graph (graph_base.h:335-436) */

/* Symmetric CSR in edge-list insertion order.  off[V+1], adj[2M].
 * Returns 0, or -1 when an id is out of range (the reference asserts,
 * graph_base.h:362-363).  Self-loops and duplicates are stored as given. */
int oracle_build_csr(int64_t nv, int64_t m, const int64_t *src, const int64_t *dst,
                     int64_t *off, int32_t *adj);

/* graph_base::check (graph_base.h:110-134): every edge present both ways and
 * no adjacent duplicate in any (sorted) neighbour list.  1 = pass. */
int oracle_check_graph(int64_t nv, int64_t m, const int64_t *src, const int64_t *dst,
                       const int64_t *off, const int32_t *adj);

/* ---------------- BFS (hybrid_bfs.cc) */

enum { ORACLE_TD = 0, ORACLE_BU = 1 };

typedef struct {
    int64_t level;          /* depth of the frontier being expanded          */
    int64_t dir;            /* ORACLE_TD or ORACLE_BU                        */
    int64_t frontier;       /* |F| = sliding_queue window size              */
    int64_t frontier_deg;   /* sum of degrees over F                         */
    int64_t unvisited;      /* vertices with no parent before the step       */
    int64_t unvisited_deg;  /* sum of their degrees                          */
    int64_t edges_to_check; /* Beamer counter after this step's update       */
    int64_t result;         /* scout (TD) or awake (BU) returned by the step */
    int64_t next;           /* |N| newly visited                             */
    int64_t next_deg;       /* sum of degrees over N                         */
    int64_t edges_examined; /* TD: sum deg(F); BU: in-order early-exit scans */
    int64_t max_frontier_deg;
} oracle_level_t;

/* run_beamer (hybrid_bfs.cc:152-191) with top_down_step_with_migrating_threads
 * (:71-100) and bottom_up_step (:111-143).  parent[] uses the reference's
 * encoding: unreached v keeps -deg(v) (or -1 when deg 0) from clear()
 * (:433-443).  depth[] (may be NULL) is -1 for unreached.  log (may be NULL)
 * receives one record per step.  Returns the number of steps, or -1. */
int64_t oracle_beamer(int64_t nv, int64_t m, const int64_t *off, const int32_t *adj,
                      int64_t source, int64_t alpha, int64_t beta, int64_t *parent,
                      int32_t *depth, oracle_level_t *log, int64_t log_cap);

/* Bottom-up head index (no reference counterpart; a derived index of the
 * product's bottom-up step): head[v] = the neighbour of v of largest degree,
 * the first in CSR order among equals, degrees saturated at 2^31-1; -1 when
 * v has no neighbours. */
void oracle_head_index(int64_t nv, const int64_t *off, const int32_t *adj, int32_t *head);

/* With head != NULL, oracle_beamer's bottom-up edges_examined follows the
 * product's head-first scan: each unvisited vertex reads head[v] (one id),
 * and only when head[v] is not yet visited scans its list in CSR order up to
 * the first visited neighbour.  Depths, parents and every other counter are
 * unchanged.  NULL restores the in-order definition. */
void oracle_set_bu_head(const int32_t *head);

/* run_beamer on every host core (OpenMP; oracle_bfs_omp.c), the CPU
 * baseline's parallel leg.  Same rules as oracle_beamer: the direction of
 * every step equals its (dirs[i] = ORACLE_TD or ORACLE_BU, may be NULL);
 * parents are any valid BFS tree.  Returns the number of steps, or -1.
 * Pinned by tests against oracle_beamer's directions, check() and the
 * traversed-edge count. */
int64_t oracle_beamer_omp(int64_t nv, int64_t m, const int64_t *off, const int32_t *adj,
                          int64_t source, int64_t alpha, int64_t beta, int64_t *parent,
                          int8_t *dirs, int64_t dirs_cap);
/* Threads oracle_beamer_omp uses (omp_get_max_threads). */
int oracle_omp_threads(void);

/* run_with_migrating_threads (hybrid_bfs.cc:197-213): top-down only. */
int64_t oracle_td_only(int64_t nv, int64_t m, const int64_t *off, const int32_t *adj,
                       int64_t source, int64_t *parent, int32_t *depth);

/* Textbook serial BFS depth, as in check() (hybrid_bfs.cc:284-307). */
void oracle_bfs_depth(int64_t nv, const int64_t *off, const int32_t *adj,
                      int64_t source, int32_t *depth);

/* hybrid_bfs::check (hybrid_bfs.cc:281-379).  1 = PASS, 0 = FAIL.  Any
 * negative parent means "not in the tree". */
int oracle_check(int64_t nv, const int64_t *off, const int32_t *adj, int64_t source,
                 const int64_t *parent);

/* count_num_traversed_edges (hybrid_bfs.cc:401-414). */
int64_t oracle_count_traversed(int64_t nv, const int64_t *off, const int64_t *parent);

/* ---------------- driver helpers */

/* lcg (lcg.h:4-33): state after the skip-ahead constructor, and next(). */
uint64_t oracle_lcg_init(uint64_t step);
uint64_t oracle_lcg_next(uint64_t *state);

/* pick_random_vertex (hybrid_bfs_main.cc:136-144) n times from lcg(0). */
void oracle_pick_roots(int64_t nv, const int64_t *off, int64_t n, int64_t *roots);

/* aggregate_stats::compute (combined.cc:108-146):
 * out = {sum, mean, hmean, stddev, min, max}. */
void oracle_aggregate(const double *x, int64_t n, double out[6]);

/* Plant a known mistake in the BFS restatement (0 = none; ids documented in
 * oracle_bfs.c).  Tests use it to show their known answers catch each one. */
void oracle_set_mutation(int id);

/* ---------------- connected components (components.cc) */
/* components::run (components.cc:38-104): labels[nv] (each component carries
 * its smallest vertex id), *ncomp; returns the iteration count. */
int64_t oracle_cc(int64_t nv, const int64_t *off, const int32_t *adj, int64_t *label,
                  int64_t *ncomp);
/* components::check (components.cc:152-202).  1 = PASS, 0 = FAIL. */
int oracle_cc_check(int64_t nv, const int64_t *off, const int32_t *adj, const int64_t *label);

/* ---------------- direction-cost MLP (no reference counterpart)
 * 8 -> H1 -> H2 -> 2, ReLU, fp32.  Parameter layout (row-major, out x in):
 *   W1[H1*8] b1[H1] W2[H2*H1] b2[H2] W3[2*H2] b3[2]                       */
int64_t oracle_mlp_nparams(int h1, int h2);
void oracle_mlp_forward(const float *params, int h1, int h2, const float *x,
                        int64_t n, float *y);
/* Mean-squared-error loss over n samples and its gradient w.r.t. params
 * (grad sized like params).  mask[n*2] selects which outputs carry a target
 * (1) or not (0); NULL means all.  Returns the loss. */
float oracle_mlp_loss_grad(const float *params, int h1, int h2, const float *x,
                           const float *target, const float *mask, int64_t n,
                           float *grad);
/* One Adam step (lr, b1=0.9, b2=0.999, eps=1e-8), t is 1-based. */
void oracle_mlp_adam(float *params, const float *grad, float *m, float *v,
                     int64_t np, float lr, int64_t t);

#ifdef __cplusplus
}
#endif
#endif
