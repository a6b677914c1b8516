/*
 * beedrill_amd.h -- C ABI of the MI355X (gfx950) direction-optimizing BFS.
 *
 * The reference (ehein6/beedrill) exposes this path as the C++ class
 * `hybrid_bfs` (hybrid_bfs.h:11-47) over the CSR view `graph`
 * (graph_base.h:231-305), built by create_graph_from_edge_list
 * (graph_base.h:335-436) from an el64 file read by dist_edge_list::load_binary
 * (dist_edge_list.cc:50-143).  Each entry point below names the reference
 * interface it replaces.  The C++ drop-in classes in beedrill_amd/host/ wrap
 * these calls with the reference's own method names.
 *
 * Conventions
 *   - Every call returns BD_OK (0) or a negative bd_status; bd_last_error()
 *     gives the message.  No C++ exception crosses this boundary.
 *   - One bd_ctx owns one HIP stream and all device buffers.  A context is
 *     single-caller and not re-entrant, as a hybrid_bfs instance is
 *     (hybrid_bfs.h:22-23 shared scout/awake counters).
 *   - Host buffers are caller-owned; a NULL output pointer means "skip".
 *   - Vertex ids are int64 at the boundary (the reference's `long`), stored
 *     on the device as int32 neighbour ids with int64 CSR offsets.
 */
#ifndef BEEDRILL_AMD_H
#define BEEDRILL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bd_ctx bd_ctx;

typedef enum {
    BD_OK = 0,
    BD_EINVAL = -1,  /* bad argument (reference: LOG + exit(1))            */
    BD_ENOMEM = -2,  /* device allocation failed (reference: std::terminate) */
    BD_EHIP = -3,    /* HIP runtime error                                   */
    BD_ESTATE = -4,  /* call out of order (e.g. BFS before a graph)          */
    BD_EIO = -5,     /* file missing, short or malformed                    */
    BD_ERANGE = -6,  /* vertex id out of range                              */
    BD_ENOEXT = -7   /* no usable GPU                                       */
} bd_status;

/* Direction policies.  BEAMER is run_beamer (hybrid_bfs.cc:152-191);
 * TD_ONLY is run_with_migrating_threads (hybrid_bfs.cc:197-213); LEARNED
 * evaluates the MLP cost model beside the two predicates at
 * hybrid_bfs.cc:167 and :178-179 and takes its direction when it predicts a
 * gain larger than bd_tuning.learned_margin_milli; with no model loaded it
 * is Beamer; REMOTE_WRITES is run_with_remote_writes (hybrid_bfs.cc:218-236),
 * every step a top-down step with remote writes; REMOTE_WRITES_HYBRID is
 * run_with_remote_writes_hybrid (hybrid_bfs.cc:243-275), Beamer's state
 * machine with remote-write top-down steps where Beamer goes bottom-up.
 * TD_ONLY and the two REMOTE policies ignore bd_tuning.td_split_edges.
 * BU_ONLY and SCRIPT exist for testing and for timing both directions. */
typedef enum {
    BD_POLICY_BEAMER = 0,
    BD_POLICY_LEARNED = 1,
    BD_POLICY_TD_ONLY = 2,
    BD_POLICY_BU_ONLY = 3,
    BD_POLICY_SCRIPT = 4,
    BD_POLICY_REMOTE_WRITES = 5,
    BD_POLICY_REMOTE_WRITES_HYBRID = 6
} bd_policy;

enum { BD_DIR_TD = 0, BD_DIR_BU = 1 };

/* Per-step record (one per top-down or bottom-up step).  The first fields
 * mirror the reference's step-level quantities: |F| (sliding_queue
 * combined_size, sliding_queue.h:126-136), scout/awake returned by the step
 * (hybrid_bfs.cc:99, :142) and edges_to_check (:162, :182).  A top-down step
 * run with remote writes (bd_tuning.td_split_edges) takes two launches and
 * gets one record. */
typedef struct {
    int64_t level;            /* depth of the frontier expanded            */
    int64_t dir;              /* BD_DIR_TD or BD_DIR_BU                    */
    int64_t frontier;         /* |F|                                       */
    int64_t frontier_deg;     /* sum deg(F)                                */
    int64_t unvisited;        /* vertices unreached before the step        */
    int64_t unvisited_deg;    /* sum of their degrees                      */
    int64_t edges_to_check;   /* Beamer counter after this step's update   */
    int64_t result;           /* scout (TD) or awake (BU)                  */
    int64_t next;             /* |N| newly reached                         */
    int64_t next_deg;         /* sum deg(N)                                */
    int64_t edges_examined;   /* TD: sum deg(F); BU: neighbour ids read    */
    int64_t max_frontier_deg; /* max deg over F                            */
    int64_t tasks;            /* edge slices of F's high-degree vertices   */
    int64_t queued;           /* F's other vertices, read from the queue   */
    double ms;                /* device time of the step (timed runs only) */
    double ms_other;          /* profile runs: time of the other direction */
    int64_t launch;           /* index of the step's first kernel launch   */
    int64_t launches;         /* its launches: 1, or 2 for a top-down step
                                 run with remote writes                    */
} bd_level_rec;

typedef struct {
    int64_t source;
    int64_t steps;            /* TD + BU steps                             */
    int64_t td_steps;
    int64_t bu_steps;
    int64_t reached;          /* vertices with a parent                    */
    int64_t traversed_edges;  /* count_num_traversed_edges()               */
    double ms;                /* device time of the whole search           */
} bd_run_stats;

/* ---------------------------------------------------------------- context */

/* hybrid_bfs::hybrid_bfs(graph&) allocation half (hybrid_bfs.cc:10-21) and
 * ack_control_init (ack_control.h) -- on the GPU: pick the device, create
 * the stream. */
int bd_create(bd_ctx **out, int device);
int bd_destroy(bd_ctx *ctx);
const char *bd_last_error(const bd_ctx *ctx);
/* Number of visible GPUs (0 when none). */
int bd_device_count(void);

/* ----------------------------------------------------------------- graphs */

/* Upload an already built symmetric CSR (graph_base.h:30-35 view):
 * offsets[nv+1] (int64), cols[nnz] (int32), m_undirected = num_edges()
 * (graph_base.h:235-237).  Validates offsets are monotone and ids in range. */
int bd_graph_upload(bd_ctx *ctx, const int64_t *offsets, const int32_t *cols,
                    int64_t nv, int64_t nnz, int64_t m_undirected);

/* create_graph_from_edge_list<graph> (graph_base.h:335-436) on the device:
 * both directions of every edge, each adjacency list in edge-list order.
 * Out-of-range ids are rejected (the reference asserts, graph_base.h:362-363);
 * like the reference, self-loops and duplicates are stored as given. */
int bd_graph_from_edges(bd_ctx *ctx, const int64_t *src, const int64_t *dst,
                        int64_t nv, int64_t m);

/* dist_edge_list::load_binary (dist_edge_list.cc:50-143) + the above:
 * reads an el64 file (header edge_list.cc:20-89) and builds the CSR. */
int bd_graph_load_el64(bd_ctx *ctx, const char *path);

/* Seeded generators, run on the device.  kind 0 = RMAT (a,b,c,d used;
 * restates generator/rmat_generator.h and edge_list_utils.h flip/sort/dedup,
 * then a seeded relabel and shuffle), 1 = Erdos-Renyi with uniform
 * endpoints, 2 = rows x cols 4-neighbour lattice (nv = rows*cols, rows in
 * `num_edges`).  *m_out receives the deduplicated edge count. */
enum { BD_GEN_RMAT = 0, BD_GEN_ER = 1, BD_GEN_LATTICE = 2 };
int bd_graph_generate(bd_ctx *ctx, int kind, int64_t nv, int64_t num_edges,
                      double a, double b, double c, double d, uint32_t seed,
                      int64_t *m_out);

/* Size of the resident graph (num_vertices()/num_edges(), graph_base.h:231-237). */
int bd_graph_info(bd_ctx *ctx, int64_t *nv, int64_t *nnz, int64_t *m_undirected);
/* Copy the device CSR back (for checks and for the CPU baseline). */
int bd_graph_download(bd_ctx *ctx, int64_t *offsets, int32_t *cols);
/* Copy the generated/loaded edge list (el64 order) back; BD_ESTATE when the
 * graph came from bd_graph_upload. */
int bd_graph_edges_download(bd_ctx *ctx, int64_t *src, int64_t *dst);
/* dist_edge_list -> el64 file (generator/edge_list_utils.h:171-198 dump_bin). */
int bd_graph_save_el64(bd_ctx *ctx, const char *path);
/* pick_random_vertex (hybrid_bfs_main.cc:136-144) n times from lcg(0). */
int bd_pick_roots(bd_ctx *ctx, int64_t n, int64_t *roots);

/* -------------------------------------------------------------------- BFS */

/* hybrid_bfs::run_beamer(source, alpha, beta) (hybrid_bfs.cc:152-191) and the
 * other policies above, preceded by clear() (hybrid_bfs.cc:433-443), which is
 * not timed, as in hybrid_bfs_main.cc:261-282.  alpha, beta > 0
 * (hybrid_bfs_main.cc:127-130).  source in [0, nv) (hybrid_bfs.cc:155). */
int bd_bfs_run(bd_ctx *ctx, int64_t source, int64_t alpha, int64_t beta, int policy,
               bd_run_stats *out);

/* For BD_POLICY_SCRIPT: dirs[i] is the direction of step i; steps past n use
 * the last entry. */
int bd_bfs_set_script(bd_ctx *ctx, const int32_t *dirs, int64_t n);

/* Results of the last run.  depth: -1 unreached.  parent: -1 unreached, the
 * source is its own parent (reference: parent_ after run_beamer; the
 * reference leaves -deg(v) in unreached entries, normalised here to -1). */
int bd_bfs_fetch(bd_ctx *ctx, int32_t *depth, int64_t *parent);

/* Per-step records of the last run (bd_bfs_run or bd_bfs_profile). */
int bd_bfs_levels(bd_ctx *ctx, bd_level_rec *recs, int64_t cap, int64_t *n);

/* hybrid_bfs::count_num_traversed_edges (hybrid_bfs.cc:401-414). */
int bd_bfs_count_traversed(bd_ctx *ctx, int64_t *out);

/* hybrid_bfs::check (hybrid_bfs.cc:281-379) against the last run, on the
 * device: the same PASS/FAIL as the reference's textbook-BFS comparison
 * (every parent a neighbour one level up, the source its own parent,
 * reachability equal), tested as a tree whose every edge spans at most one
 * level and whose reached set is closed under adjacency.
 * *ok = 1 PASS, 0 FAIL. */
int bd_bfs_check(bd_ctx *ctx, int64_t source, int *ok);

/* hybrid_bfs::check (hybrid_bfs.cc:281-379) of a caller's parent array
 * (nv entries; negative = unreached, as the reference's -deg(v) encoding
 * after clear(), hybrid_bfs.cc:433-443) against the resident graph, on the
 * device.  Lets a caller validate a tree from any source, e.g. the
 * reference's own parent_ array, without a host copy of the graph. */
int bd_bfs_check_parents(bd_ctx *ctx, int64_t source, const int64_t *parent, int *ok);

/* Time both directions at every step from the same state (no reference
 * counterpart: the per-step hook sites are hybrid_bfs.cc:173-177, 184-188).
 * The search follows `policy`; each record gets ms (chosen direction) and
 * ms_other.  Used to log training data for the learned policy. */
int bd_bfs_profile(bd_ctx *ctx, int64_t source, int64_t alpha, int64_t beta, int policy,
                   bd_run_stats *out);

/* Kernel time of every step launch of the last search run with
 * bd_tuning.time_steps = 3 (each launch timed) or 4 (each chunk of launches
 * timed, its launches given the chunk's mean), in launch order (a level's
 * launches are its record's launch .. launch + launches - 1; launches queued
 * after the search ended return at once and are listed too).  No
 * reference counterpart: the measurement behind the step kernel's roofline. */
int bd_bfs_launch_times(bd_ctx *ctx, double *ms, int64_t cap, int64_t *n);

/* --------------------------------------------------- connected components */

/* components::stats (components.h:27-32), plus the device time of the run
 * and of its edge passes. */
typedef struct {
    int64_t num_components;   /* components found                            */
    int64_t num_iters;        /* edge passes, the last of which changed none */
    double ms;                /* device time of run(): passes and the count  */
    double pass_ms;           /* kernel time of the edge passes alone        */
    int64_t labels_read;      /* neighbour labels the edge passes read       */
    int64_t edge_bytes;       /* algorithmic bytes of the edge passes: 8 per
                                 label read (column id + label), 4 per
                                 column id of the first pass, 12 per vertex
                                 of a pull, per changed vertex of a push,
                                 and the changed-vertex bitmap              */
    int64_t pass_launches;    /* kernel launches of the edge passes (a tiled
                                 pull pass takes 8)                         */
    double prep_ms;           /* device time of the per-graph regrouping the
                                 passes use (slices of the long lists, lists
                                 regrouped by column block), built by the
                                 first run on a graph and not part of `ms` */
    int64_t prep_bytes;       /* device memory that regrouping holds       */
} bd_cc_stats;

/* components::run (components.cc:38-104) on the resident graph: label
 * propagation (each vertex takes its neighbours' smallest label) and
 * shortcutting until an edge pass changes nothing.  Each component ends
 * labelled by its smallest vertex id.  The first pass takes neighbour ids
 * for labels (each label is still its vertex's id); once few labels change,
 * a pass offers only the changed vertices' labels to their neighbours (the
 * same pairs as a pull over changed neighbours; the outcome is unchanged). */
int bd_cc_run(bd_ctx *ctx, bd_cc_stats *out);
/* components::component_ (components.h:13-14) of the last run: labels[nv]. */
int bd_cc_fetch(bd_ctx *ctx, int64_t *labels);
/* components::check (components.cc:152-202), on the host from downloaded
 * copies.  *ok = 1 PASS, 0 FAIL. */
int bd_cc_check(bd_ctx *ctx, int *ok);

/* Launch tuning (autotuner output).  Any field <= 0 keeps the default. */
typedef struct {
    int32_t block;           /* threads per workgroup of the step kernel:
                                128, 256 or 512                           */
    int32_t grid_per_cu;     /* workgroups per CU of the top-down (or
                                two-direction) kernel; the grid is
                                persistent                                */
    int32_t wave_threshold;  /* top-down: degree at which a vertex gets a
                                whole wave                                */
    int32_t task_threshold;  /* degree at which a frontier vertex is split
                                into task slices                          */
    int32_t task_edges;      /* edges per task slice                       */
    int32_t bu_words;        /* bottom-up: visited words per wave chunk
                                (1..32)                                   */
    int32_t bu_cap;          /* bottom-up: neighbours a lane scans before
                                its wave finishes the list together       */
    int32_t level_driver;    /* 1 = direction chosen on device and steps
                                enqueued without a per-step host sync;
                                2 = host reads the counters after each step */
    int32_t time_steps;      /* 1 = time each step's kernel with HIP events
                                (implies the host-driven loop); 2 = off;
                                3 = device-driven levels, every step launch
                                timed by events stamped at its dispatch;
                                4 = device-driven levels, one event pair per
                                chunk of launches (bd_bfs_launch_times)   */
    int32_t bu_head;         /* bottom-up: 1 = test each vertex's head index
                                (its highest-degree neighbour) before its
                                list; 2 = scan lists only, in CSR order   */
    int32_t step_kernels;    /* device-driven levels: 1 = one step kernel per
                                direction (both are launched per level; the
                                one not chosen on the device returns after
                                its state rebuild); 2 = one kernel carrying
                                both directions.  Host-driven levels always
                                launch the chosen direction's kernel      */
    int32_t bu_grid_per_cu;  /* workgroups per CU of the bottom-up kernel
                                (step_kernels 1; otherwise grid_per_cu)   */
    int32_t learned_margin_milli; /* LEARNED: the network overrides Beamer's
                                choice only when it predicts the other
                                direction cheaper by more than this many
                                thousandths of a log2 unit (1 = the network
                                decides alone)                            */
    int32_t ms_alpha;        /* bit-parallel search: a level pushes while the
                                frontier's degree sum * ms_alpha is below
                                the degree sum of the vertices still open,
                                else it pulls                             */
    int32_t learned_min_skew; /* LEARNED consults the network only on graphs
                                whose largest degree is at least this many
                                times the mean degree (the graphs it was
                                trained on have hubs); otherwise Beamer   */
    int32_t td_split_edges;  /* a top-down step whose frontier has at least
                                this many edges runs with remote writes
                                (hybrid_bfs.cc:41-67): one launch writes
                                each edge's source as the candidate parent
                                of its unvisited end, the next claims one
                                candidate per vertex; 2^31-1 = never       */
    int32_t vertex_order;    /* numbering of the CSR copy the searches read
                                (results are always in graph ids): 1 =
                                degree order, largest first, degree-0
                                vertices last, the 4096 largest one per
                                64-bit visited word; 2 = the graph's own
                                ids (no copy); 3 = vertices with edges
                                first, in graph order; 4 (default) =
                                degree order without the spreading        */
    int32_t bu_batch;        /* bottom-up: 1 = one round of 64 vertices per
                                wave at a time (the next round's head
                                entries prefetched); 4 = four rounds' head
                                entries loaded together, the next four in
                                flight meanwhile; 8 (default) = six rounds'
                                neighbour halves in one trip, then their
                                degree halves with the frontier words     */
    int32_t learned_near;    /* LEARNED: besides Beamer's switch points, the
                                network is consulted at a top-down step
                                Beamer keeps top-down when the frontier's
                                edges * alpha * learned_near exceed the
                                unvisited edges (an early switch to
                                bottom-up); 1 = switch points only;
                                default 4                                 */
} bd_tuning;
int bd_set_tuning(bd_ctx *ctx, const bd_tuning *t);
int bd_get_tuning(bd_ctx *ctx, bd_tuning *t);

/* A plain device copy of `bytes` bytes (read + write), `reps` times (no
 * reference counterpart: the measurement reference for the bottom-up
 * kernel's hardware traffic counters and achievable HBM bandwidth).  *gbps
 * receives the best rate, read plus written bytes per second / 1e9. */
int bd_copy_bandwidth(bd_ctx *ctx, int64_t bytes, int reps, double *gbps);

/* ------------------------------------------------------- batched searches */

/* Bit-parallel search from up to 64 roots at once on the resident graph (no
 * reference counterpart; replaces up to 64 trips round the sequential trial
 * loop, hybrid_bfs_main.cc:252-302).  Every vertex carries a 64-bit mask of
 * the roots that reached it; each level pushes frontier masks to neighbours
 * or pulls them from neighbours, chosen per level from the frontier's and
 * the open vertices' degree sums.  depths (may be NULL) receives
 * depths[r*nv + v] = depth of v from roots[r] (-1 unreached); *ms the time
 * from the first kernel to the last; *levels the number of levels.
 * bd_bfs_levels then returns one record per level: dir 0 = push, 1 = pull;
 * frontier / frontier_deg = vertices holding any root's frontier and their
 * degree sum; next = vertices gaining a root; ms = the level's time. */
int bd_bfs_multi_root(bd_ctx *ctx, const int64_t *roots, int n, int32_t *depths,
                      double *ms, int64_t *levels);

/* Many disjoint graphs searched at once (no reference counterpart).  The
 * resident graph is their union: graph g owns vertices [vbase[g],
 * vbase[g+1]) (vbase has ngraphs+1 entries, vbase[0] = 0, vbase[ngraphs] =
 * nv) and no edge joins two graphs (BD_EINVAL otherwise).  roots[g] is a
 * vertex of graph g, numbered within it.  One direction-optimizing search
 * (Beamer, alpha 15, beta 18) starts from every root; depths (may be NULL,
 * nv entries) gets each vertex's depth from its own graph's root, *ms the
 * device time of the search.  bd_bfs_fetch / bd_bfs_levels then describe
 * this search. */
int bd_bfs_multi_graph(bd_ctx *ctx, int ngraphs, const int64_t *vbase, const int64_t *roots,
                       int32_t *depths, double *ms);

/* bd_bfs_multi_graph without allocating: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * staging (the sources, the graph bounds, the cross-edge flag) lives in a
 * static device array of the library, one per GPU, sized for at most 4096
 * graphs; more are rejected with BD_ERANGE.  Same search and results as
 * bd_bfs_multi_graph, read afterwards with bd_bfs_fetch / bd_bfs_levels
 * (there is no depths argument).  Calls that overlap on one GPU share the
 * array and race. */
int bd_bfs_multi_graph_static(bd_ctx *ctx, int ngraphs, const int64_t *vbase,
                              const int64_t *roots, double *ms);

/* --------------------------------------------------- learned direction MLP */

/* 8 -> h1 -> h2 -> 2 ReLU MLP predicting log step cost of (TD, BU).
 * Parameters fp32, layout W1[h1*8] b1[h1] W2[h2*h1] b2[h2] W3[2*h2] b3[2]. */
int bd_mlp_set(bd_ctx *ctx, const float *params, int h1, int h2);
int bd_mlp_get(bd_ctx *ctx, float *params, int64_t cap, int *h1, int *h2);
/* Features of a step record, as fed to the network (8 floats). */
int bd_mlp_features(const bd_level_rec *rec, int64_t nv, int64_t m_undirected,
                    int64_t prev_dir, float *x8);
/* Device forward pass (fp32 reference kernel and bf16 MFMA single-wave
 * kernel, selected by use_mfma). */
int bd_mlp_infer(bd_ctx *ctx, const float *x, int64_t n, float *y, int use_mfma);
/* Device loss and gradient (MSE over masked outputs). */
int bd_mlp_loss_grad(bd_ctx *ctx, const float *x, const float *t, const float *mask,
                     int64_t n, float *grad, float *loss);
/* Device training: `epochs` full-batch Adam steps at learning rate lr;
 * *loss receives the final loss. */
int bd_mlp_train(bd_ctx *ctx, const float *x, const float *t, const float *mask, int64_t n,
                 int epochs, float lr, uint32_t seed, float *loss);

#ifdef __cplusplus
}
#endif
#endif
