// bd_internal.h -- context, device control block and shared helpers of the
// MI355X BFS library.  Not part of the public ABI (include/beedrill_amd.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/beedrill_amd.h"

namespace bd {

constexpr int kWave = 64;

// One task = a slice of a huge vertex's adjacency list (the GPU counterpart of
// worklist<Edge>'s 64-edge granules, worklist.h:96-117).
struct td_task {
    int64_t beg;
    int64_t end;
    int32_t src;
    int32_t pad;
};

// Hubs of the search layout: lists of at least kHubDeg entries; those of
// kHubBigDeg or more are cut into slices of kHubSlice entries.
constexpr int64_t kHubDeg = 256;
constexpr int64_t kHubBigDeg = 8192;
constexpr int64_t kHubSlice = 4096;
// A top-down step over a bitmap frontier spreads the hubs over the grid only
// when the frontier has this many edges (below it, the hubs it holds are few).
constexpr int64_t kHubPassEdges = 1 << 20;

// Frontier lists (queue of ordinary vertices, task slices of high-degree
// vertices) are segmented: workgroup b appends to segment b % kSeg, so no
// address is shared by more than a few dozen workgroups -- on MI355X a
// same-address atomic costs ~12 ns and they serialize, so one counter hit by
// 2048 workgroups costs ~25 us.  Segment kSeg is an overflow segment that
// takes what does not fit.  The same spreading applies to the step's
// counters (kSlots partial accumulators, each on its own cache line).
constexpr int kSeg = 64;
constexpr int kSlots = 64;
constexpr int kLine = 16;  // 64-bit words per 128-byte line

// Step L appends to list (L+1)&1 and counts its segments in qlen/tlen
// [(L+1)%3]; its counters go to acc[L%3].  Three rotating copies of the
// counters let the step that reads one copy clear another that nobody
// touches during that step.
struct seg_counters {
    unsigned long long qlen[3][kSeg + 1][kLine];
    unsigned long long tlen[3][kSeg + 1][kLine];
    // per-slot partial accumulators: new, new_deg, examined, max_deg
    unsigned long long acc[3][kSlots][kLine];
    unsigned long long task_overflow;
    uint64_t t_start;  // device timestamps (wall_clock64 ticks)
    uint64_t t_end;
};

// State of the search before a step: the step's direction and frontier
// description, the policy state and the totals so far.  Two copies: the
// step kernel for level L rebuilds the state before step L from copy
// (L-1)&1 and the counters of step L-1 (every workgroup does it, with the
// same result) and workgroup 0 publishes it as copy L&1.  qpref/tpref are
// the prefix sums of the segment lengths of frontier L (index kSeg+1 =
// total).
struct ctl_block {
    // run constants
    int64_t nv;
    int64_t m2;        // 2 * num_edges()   (hybrid_bfs.cc:162)
    int64_t alpha;
    int64_t beta;
    int32_t policy;
    int32_t has_mlp;
    int32_t script_len;
    int32_t log_cap;
    // Beamer state (hybrid_bfs.cc:162-190)
    int64_t edges_to_check;
    int64_t scout;
    int64_t old_awake;
    int32_t in_bu_loop;
    // current step
    int32_t level;
    int32_t dir;
    int32_t done;
    int32_t td_bitmap;   // frontier L is only in bm[L%3] (step L-1 was BU)
    int32_t prev_dir;    // -1 before the first step
    int32_t td_nolist;   // this top-down step leaves its frontier in the bitmap only
    int64_t frontier;
    int64_t frontier_deg;
    int64_t max_frontier_deg;
    int64_t unvisited;
    int64_t unvisited_deg;
    int64_t qpref[kSeg + 2];
    int64_t tpref[kSeg + 2];
    // totals
    int64_t steps;
    int64_t td_steps;
    int64_t bu_steps;
    int64_t reached;
    int64_t reached_deg;
};

struct ctx_state;

}  // namespace bd

struct bd_ctx {
    int device = 0;
    int cu_count = 256;
    int64_t wall_khz = 100000;
    hipStream_t stream = nullptr;
    std::string err;

    // graph (flat device CSR)
    int64_t nv = 0, nnz = 0, m = 0, nwords = 0;
    int64_t *d_off = nullptr;
    int32_t *d_cols = nullptr;
    uint64_t *d_dead = nullptr;   // 1 = vertex of degree 0, or padding bit
    uint64_t *d_head = nullptr;   // bottom-up head index, in search ids (built with the layout)
    int64_t max_deg = 0;          // largest degree (found with the head index)
    uint64_t *d_el = nullptr;     // el64 order, packed (src << 32) | dst
    int64_t el_m = 0;

    // search layout: the CSR the BFS kernels read.  Vertices are renumbered
    // (tune.vertex_order): search id i is graph vertex perm[i], graph vertex
    // v is search id inv[v]; every adjacency list keeps its CSR order, ids
    // mapped.  Searches run on search ids; sources are mapped on the device
    // and results mapped back on readout.  With the graph's own order the
    // pointers alias d_off / d_cols / d_dead and perm / inv are null.
    int64_t *d_soff = nullptr;
    int32_t *d_scols = nullptr;
    uint64_t *d_sdead = nullptr;
    int32_t *d_perm = nullptr;
    int32_t *d_inv = nullptr;
    int64_t nv_active = 0;        // vertices with edges
    // hubs of the search layout (lists of kHubDeg entries or more), which a
    // top-down step over a bitmap frontier spreads over the whole grid
    uint64_t *d_hub_bm = nullptr;  // bit v: v is a hub
    int32_t *d_hub_med = nullptr;  // hubs with fewer than kHubBigDeg neighbours
    bd::td_task *d_hub_big = nullptr;  // the others' lists, in slices of kHubSlice
    int64_t n_hub_med = 0, n_hub_big = 0;
    int s_order = 0;              // vertex_order the layout was built with (0: none yet)

    // search state
    int32_t *d_depth = nullptr;
    int32_t *d_parent = nullptr;
    int32_t *d_cand = nullptr;    // candidate parents of a remote-write top-down step (-1 = none)
    uint8_t *d_dirty = nullptr;   // per visited word: a candidate was written into it (0 = none)
    int32_t *d_comp = nullptr;    // connected-component labels (components_)
    int32_t *d_cc_flag = nullptr; // [0] an edge pass changed a label, [1..] counters
    bool have_cc = false;         // d_comp holds a finished run
    uint64_t *d_cc_chg = nullptr;       // two bitmaps of vertices whose label changed
    int32_t *d_cc_bcols = nullptr;      // short lists regrouped by column block
    uint64_t *d_cc_bend = nullptr;      // per vertex: end of each block's run (one byte each)
    int cc_bshift = 0;                  // column blocks of 2^cc_bshift vertices
    bd::td_task *d_cc_pieces = nullptr; // long-list pieces by (row block, column block)
    int64_t *d_cc_pbase = nullptr;      // [65] start of each bucket's pieces
    bd::td_task *d_cc_tasks = nullptr;  // slices of the long lists of the edge pass
    int64_t cc_ntasks = -1;             // -1: not built for the resident graph
    double cc_prep_ms = 0;              // device time of that one-time build
    int64_t cc_prep_bytes = 0;          // device memory it holds
    uint64_t *d_vis = nullptr;
    uint64_t *d_vis_save = nullptr;
    uint64_t *d_bm[3] = {nullptr, nullptr, nullptr};
    int32_t *d_q[2] = {nullptr, nullptr};
    bd::td_task *d_tasks[2] = {nullptr, nullptr};
    int64_t task_cap = 0;    // most task slices one frontier can need
    int64_t qseg_cap = 0;    // entries per queue segment (overflow holds nv)
    int64_t tseg_cap = 0;    // entries per task segment (overflow holds task_cap)
    bd::seg_counters *d_seg = nullptr;
    bd::ctl_block *d_ctl = nullptr;  // two published copies (see ctl_block)
    bd::ctl_block *h_ctl = nullptr;  // pinned mirror of both copies
    bd_level_rec *d_log = nullptr;
    int64_t log_cap = 0;
    int32_t *d_script = nullptr;
    int64_t script_len = 0;
    std::vector<int32_t> script;

    // results of the last run
    bool have_result = false;
    bool depth_valid = false;  // d_depth holds the last search's depths
    bool split_off = false;    // profiling: top-down steps never run with remote writes
    bool cand_clean = false;   // d_cand is all -1 (set when a search finishes)
    int chunk_hint = 0;        // first chunk of device-driven launches (0: 16)
    int64_t last_source = -1;
    int64_t last_steps = 0;
    std::vector<bd_level_rec> last_log;

    // learned policy
    float *d_mlp = nullptr;
    uint4 *d_mlp_ops = nullptr;  // the step kernels' per-lane operand table
    int mlp_h1 = 0, mlp_h2 = 0;
    bool has_mlp = false;
    std::vector<float> h_mlp;

    // multi-root state
    uint64_t *d_ms[4] = {nullptr, nullptr, nullptr, nullptr};  // F, N, seen, counters
    uint64_t *d_ms_bits[2] = {nullptr, nullptr};              // frontier bitmaps
    int32_t *d_ms_depth = nullptr;
    int64_t ms_depth_rows = 0;
    void *d_ms_slices[2] = {nullptr, nullptr};  // long frontier lists as push slices
    int64_t ms_max_slices = 0, ms_seg_cap = 0;
    int32_t *d_ms_hubs = nullptr;      // search ids of the vertices a pull cuts into slices
    void *d_ms_hub_slices = nullptr;   // their lists as slices
    int64_t ms_nhubs = 0, ms_nhub_slices = 0;
    int ms_hub_order = 0;              // the layout (s_order) they were listed in
    void *ms_graph = nullptr;      // hipGraphExec_t of the captured level chunk
    std::string ms_graph_key;      // the buffers and sizes it was captured with

    int64_t *d_static = nullptr;  // this device's address of bd_bfs.hip's `scratch`

    bd_tuning tune{};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> ev_pool;     // per-launch timing events (device-scope release)
    std::vector<double> last_launch_ms;  // kernel time of every step launch of the last search
};

namespace bd {

int set_err(bd_ctx *c, int code, const std::string &msg);
int hip_err(bd_ctx *c, hipError_t e, const char *what);
// Allocate n bytes of device memory (zero-size requests allocate 256 B).
int dev_alloc(bd_ctx *c, void **p, size_t bytes, const char *what);
void dev_free(void *p);
// Release all graph-dependent buffers.
void free_graph(bd_ctx *c);
// (Re)allocate search buffers for the resident graph (and build its head
// index if it has none).
int alloc_search(bd_ctx *c);
// Head index of the bottom-up step for the resident graph.
int build_head_index(bd_ctx *c);
// Search layout of the resident graph in the tuned vertex order (no-op when
// it is built in that order); frees it first when the order changed.
int build_search_layout(bd_ctx *c);
void free_search_layout(bd_ctx *c);
// Drop the bit-parallel search's captured graph (its buffers are going).
void ms_free_graph(bd_ctx *c);
// Look up this device's address of the allocation-free entry points'
// staging array (with the context's device current).
int static_scratch_init(bd_ctx *c);
// Rebuild the step kernels' operand table of the direction network.
int refresh_mlp_ops(bd_ctx *c);
// Build device CSR from a device edge list (packed src<<32|dst, el order).
int build_csr_from_el(bd_ctx *c, int64_t nv, int64_t m);
int64_t grid_for(const bd_ctx *c, int64_t work_items, int block);

}  // namespace bd

#define BD_HIP(c, call)                                            \
    do {                                                           \
        hipError_t e__ = (call);                                   \
        if (e__ != hipSuccess) return bd::hip_err((c), e__, #call); \
    } while (0)
