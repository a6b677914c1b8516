// bd_bfs.hip -- direction-optimizing BFS on gfx950 (hybrid_bfs::run_beamer,
// hybrid_bfs.cc:152-191).
//
// Device state per search (V vertices, W = ceil(V/64) 64-bit words):
//   parent[V]            int32     result (-1 = unreached); as in the
//                                  reference (parent_ only, hybrid_bfs.h:15)
//                                  the steps write parents alone, and depths
//                                  are derived from the parent tree on
//                                  readout (k_tree_*)
//   vis[W]               bitmap    visited; degree-0 vertices are pre-set
//   bm[3][W]             bitmaps   bm[L%3] = depth-L set (bottom-up input),
//                                  bm[(L+1)%3] = set found by step L,
//                                  bm[(L+2)%3] zeroed during step L
//   q[2]                 int32     frontier vertices of degree < task_t
//   tasks[2]                       edge slices of frontier vertices of degree
//                                  >= task_t (the reference's 64-edge
//                                  worklist granules, worklist.h:96-117)
//   seg                            segment lengths and per-slot counters
//   ctl                            step state and direction-policy state
//
// One BFS level is one launch of k_step.  The first wave of every workgroup
// rebuilds the state before the step from the previous step's counters and
// published state (the same result in every workgroup, so none waits for
// another), and workgroup 0 publishes it with the previous step's record.
//   Top-down (hybrid_bfs.cc:71-100): blocks take tiles of the queue; per
//   tile, vertices of degree >= wave_t get one wave each, the rest are
//   concatenated by a block-wide scan and every thread takes one edge.
//   Then waves take task slices.  A vertex is claimed by atomicOr on vis.
//   Bottom-up (hybrid_bfs.cc:111-143): each wave owns chunks of `bu_words`
//   visited words; lanes load the words and list the unvisited vertices in
//   LDS, and then each lane takes one unvisited vertex: its head index (the
//   neighbour of largest degree) first, then its list in CSR order, stopping
//   at the first neighbour whose bit is set in bm[L%3].  Lists longer than
//   bu_cap are finished by the whole wave.
//   Reading the depth-L bitmap, not vis, keeps the reference's two-phase rule
//   (:117-141): a vertex found in this step is never a parent in this step.
// A top-down step emits the next frontier as queue entries (staged in LDS and
// flushed once per block into the block's segment), task slices for
// high-degree vertices, and the depth-(L+1) bitmap; a bottom-up step writes
// only the bitmap, and a top-down step after it reads the bitmap.  The
// direction is
// picked on the device, so a search is a stream of launches with no host
// round trip.
#include <hip/hip_ext.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "bd_device.h"
#include "bd_internal.h"
#include "bd_mlp.h"

namespace bd {

namespace {

constexpr int kMaxLevelsPerSync = 64;
constexpr int kQCap = 2048;  // LDS-staged queue entries per workgroup
constexpr int64_t kPrecheckEdges = 1 << 16;  // top-down steps this heavy pre-check claims
// A top-down step over at least this many frontier edges builds no queue or
// task list: its frontier is large enough that the next step, bottom-up as a
// rule, reads the bitmap, and a top-down step after it ranks the bitmap.  On
// a graph whose largest degree is under wave_t no step builds lists: the
// bitmap's tiles take every vertex (there are no hubs to slice), and ranking
// the bitmap costs less than staging, flushing and reading the lists (a
// 4096 x 4096 lattice, 6278 levels: 68 ms against 92 ms).
constexpr int64_t kListEdges = 1 << 14;
// Top-down rounds of 64 edges a wave visits together (see td_visit_n).
constexpr int kTdRounds = 2;
// Largest bottom-up chunk (visited words per wave): a chunk's unvisited
// vertices are listed in LDS, 2 bytes each.
constexpr int kBuChunkWords = 32;
// Rounds of 64 bottom-up vertices whose head loads are issued together
// (bu_batch > 1).
constexpr int kBuBatch = 4;

// A top-down step run with remote writes (hybrid_bfs.cc:41-67) takes two
// launches: kDirTDWrite writes each frontier edge's source as the candidate
// parent of its unvisited end (plain stores, no claims), kDirTDClaim sweeps
// the visited words and gives every unvisited vertex with a candidate that
// parent.  Records, counts and the direction policy see one top-down step.
constexpr int32_t kDirTDWrite = 4, kDirTDClaim = 5;
// the step a direction belongs to
__host__ __device__ constexpr int32_t step_dir(int32_t d) {
    return (d == kDirTDWrite || d == kDirTDClaim) ? BD_DIR_TD : d;
}
// the kernel that runs it: the claim sweep is bitmap work, like bottom-up
__host__ __device__ constexpr int32_t kernel_dir(int32_t d) {
    return d == kDirTDWrite ? BD_DIR_TD : (d == kDirTDClaim ? BD_DIR_BU : d);
}

struct step_args {
    const int64_t *off;
    const int32_t *cols;
    const uint64_t *head;  // bottom-up head index: (degree << 32) | neighbour of largest degree
    const int32_t *inv;    // graph id -> search id (null: the search runs on graph ids)
    uint64_t *vis;
    const uint64_t *front;
    uint64_t *next_bm;
    uint64_t *clr;
    const int32_t *qin;
    int32_t *qout;
    const td_task *tin;
    td_task *tout;
    int32_t *parent;
    int32_t *cand;         // candidate parents (-1 = none)
    uint8_t *dirty;        // per visited word: a candidate was written into it
    const uint64_t *hub_bm;     // bit v: deg(v) >= kHubDeg
    const int32_t *hub_med;     // hubs under kHubBigDeg neighbours
    const td_task *hub_big;     // the other hubs' lists in slices
    int64_t n_hub_med, n_hub_big;
    ctl_block *c;
    seg_counters *seg;
    bd_level_rec *log;
    const int32_t *script;
    const float *mlp;
    const uint4 *mlp_ops;  // the network's operands in MFMA lane order
    int64_t nwords;
    int64_t live_nwords;   // words holding every vertex with edges (the sweeps cover these)
    int64_t task_cap;
    int64_t qseg_cap, tseg_cap;
    int64_t nv, m, nnz;
    int64_t alpha, beta;
    int64_t nv_beta;       // num_vertices() / beta (hybrid_bfs.cc:179)
    int64_t td_split;      // frontier edges from which top-down writes remotely
    int64_t max_deg;       // the graph's largest degree
    int32_t level;
    int32_t close_only;    // only publish the state before this step
    int32_t force_dir;     // -1: ctl->dir; 0/1: run this direction (profiling)
    int32_t wave_t, task_t, task_edges;
    int32_t bu_words, bu_cap, bu_head, bu_batch;
    int32_t policy, script_len, log_cap;
    int32_t mlp_h1, mlp_h2, has_mlp;
    float mlp_margin;      // log2-cost margin for the network to override Beamer
    int64_t mlp_near;      // also consulted at top-down steps this close to Beamer's switch
};

// ----------------------------------------------------------------- helpers


// Index of the segment holding item i: the largest s with pref[s] <= i.
__device__ __forceinline__ int seg_of(const int64_t *pref, int64_t i) {
    int s = 0;
#pragma unroll
    for (int st = 64; st > 0; st >>= 1) {
        const int mid = s + st;
        if (mid <= kSeg && pref[mid] <= i) s = mid;
    }
    return s;
}

// Per-workgroup LDS state.  The top-down staging and tile arrays and the
// bottom-up found words, pending lists and first hits share one union, sized
// by the workgroup.
template <int BLOCK>
struct step_smem {
    ctl_block st;  // the search state before this step
    long long s_new, s_new_deg;
    int s_dir;
    int s_split;  // the step closed was a top-down step with remote writes
    float x[8];
    unsigned long long red[4][BLOCK / 64];
    int32_t qn;
    int64_t qpos, qopos;
    int32_t qfit;
    union {
        struct {  // top-down
            int32_t q[kQCap];
            int64_t t_beg[BLOCK];
            int64_t t_end[BLOCK];
            int32_t t_src[BLOCK];
            int32_t t_pref[BLOCK + 1];
            int32_t t_med[BLOCK];
            int32_t t_wsum[BLOCK / 64];
            int32_t t_nmed;
            // bitmap frontier: a range's words and their rank prefix
            uint64_t w_bits[BLOCK];
            int32_t w_pref[BLOCK + 1];
            int32_t w_wsum[BLOCK / 64];
        };
        struct {  // bottom-up: found bits of each wave's chunk, the chunk's
                  // unvisited vertices, head misses, first hit per pending list
            uint64_t b_found[BLOCK / 64][64];
            uint16_t b_list[BLOCK / 64][kBuChunkWords * 64];
            int32_t b_pend[BLOCK / 64][64 * (kBuBatch + 1)];
            unsigned long long b_hit[BLOCK / 64][64];
        };
    };
};

// Reserve n slots of the output list for this workgroup's segment; returns
// how many fit there (at *pos) and where the rest goes in the overflow
// segment (at *opos).
__device__ __forceinline__ int64_t seg_reserve(unsigned long long (*len)[kLine], int64_t seg_cap,
                                               int64_t n, int64_t *pos, int64_t *opos) {
    const int sidx = blockIdx.x % kSeg;
    *pos = (int64_t)atomicAdd(&len[sidx][0], (unsigned long long)n);
    int64_t fit = seg_cap - *pos;
    fit = fit < 0 ? 0 : (fit > n ? n : fit);
    *opos = 0;
    if (fit < n) *opos = (int64_t)atomicAdd(&len[kSeg][0], (unsigned long long)(n - fit));
    return fit;
}

__device__ __forceinline__ int64_t seg_base(int64_t seg_cap) {
    return (int64_t)(blockIdx.x % kSeg) * seg_cap;
}

// Emit the task slices of the lanes' high-degree vertices (wave-collective):
// one reservation for the whole wave, then the slices are written 64 at a
// time, lane j taking slice j of the concatenation (its owner found by a
// binary search over the lanes' exclusive prefix).
__device__ __forceinline__ void emit_tasks(const step_args &a, bool huge, int32_t u, int64_t b, int64_t e) {
    const int lane = lane_id();
    const long long nt = huge ? (e - b + a.task_edges - 1) / a.task_edges : 0;
    const long long incl = wave_incl_scan(nt);
    const long long total = lane_bcast((int64_t)incl, 63);
    const long long excl = incl - nt;
    long long pos = 0, opos = 0, fit = 0;
    if (lane == 0) {
        int64_t ps, ops;
        fit = seg_reserve(a.seg->tlen[(a.level + 1) % 3], a.tseg_cap, total, &ps, &ops);
        pos = ps;
        opos = ops;
    }
    pos = lane_bcast((int64_t)pos, 0);
    opos = lane_bcast((int64_t)opos, 0);
    fit = lane_bcast((int64_t)fit, 0);
    for (long long j0 = 0; j0 < total; j0 += 64) {
        const long long j = j0 + lane;
        const long long jj = j < total ? j : total - 1;
        int o = 0;
#pragma unroll
        for (int st = 32; st > 0; st >>= 1) {
            const int mid = o + st;
            const long long em = __shfl(excl, mid < 64 ? mid : 63);
            if (mid < 64 && em <= jj) o = mid;
        }
        const long long t = jj - __shfl(excl, o);
        const int64_t bo = __shfl(b, o), eo = __shfl(e, o);
        const int32_t uo = __shfl(u, o);
        if (j < total) {
            td_task tk;
            tk.beg = bo + t * a.task_edges;
            tk.end = tk.beg + a.task_edges < eo ? tk.beg + a.task_edges : eo;
            tk.src = uo;
            tk.pad = 0;
            if (j < fit) {
                a.tout[seg_base(a.tseg_cap) + pos + j] = tk;
            } else {
                const int64_t ov = opos + (j - fit);
                if (ov < a.task_cap)
                    a.tout[kSeg * a.tseg_cap + ov] = tk;
                else
                    atomicAdd(&a.seg->task_overflow, 1ULL);
            }
        }
    }
}

// Emit vertex u of the next frontier: high-degree vertices become task
// slices, the others are staged in LDS.  Every lane of the wave must call
// this together (emit = false for idle lanes).
template <int BLOCK>
__device__ __forceinline__ void emit_next(const step_args &a, step_smem<BLOCK> &s, bool emit, int32_t u,
                                          int64_t b, int64_t e) {
    if (s.st.td_nolist) return;  // the frontier stays in the bitmap only
    const bool huge = emit && (e - b) >= a.task_t;
    if (__ballot(huge)) emit_tasks(a, huge, u, b, e);
    const bool q = emit && !huge;
    const uint64_t m = __ballot(q);
    if (m == 0) return;
    const int lane = lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int cnt = __popcll(m);
    int p = 0;
    if (lane == leader) p = atomicAdd(&s.qn, cnt);
    p = lane_bcast((int32_t)p, leader);
    const int rank = __popcll(m & ((1ULL << lane) - 1));
    // Slots below kQCap are staged in LDS (the flush writes exactly those);
    // the part of the batch past the end goes straight to HBM.
    const int in_lds = p >= kQCap ? 0 : (cnt < kQCap - p ? cnt : kQCap - p);
    long long pos = 0, opos = 0, fit = 0;
    if (in_lds < cnt) {
        if (lane == leader) {
            int64_t ps, ops;
            fit = seg_reserve(a.seg->qlen[(a.level + 1) % 3], a.qseg_cap, cnt - in_lds, &ps, &ops);
            pos = ps;
            opos = ops;
        }
        pos = lane_bcast((int64_t)pos, leader);
        opos = lane_bcast((int64_t)opos, leader);
        fit = lane_bcast((int64_t)fit, leader);
    }
    if (q) {
        if (rank < in_lds) {
            s.q[p + rank] = u;
        } else {
            const int64_t r = rank - in_lds;
            if (r < fit)
                a.qout[seg_base(a.qseg_cap) + pos + r] = u;
            else
                a.qout[kSeg * a.qseg_cap + opos + (r - fit)] = u;
        }
    }
}

// Write the staged queue entries to this workgroup's segment (whole
// workgroup).
template <int BLOCK>
__device__ __forceinline__ void flush_queue(const step_args &a, step_smem<BLOCK> &s) {
    __syncthreads();
    const int n = s.qn < kQCap ? s.qn : kQCap;
    if (n > 0) {
        if (threadIdx.x == 0) {
            int64_t pos, opos;
            s.qfit = (int32_t)seg_reserve(a.seg->qlen[(a.level + 1) % 3], a.qseg_cap, n, &pos,
                                          &opos);
            s.qpos = pos;
            s.qopos = opos;
        }
        __syncthreads();
        const int fit = s.qfit;
        const int64_t base = seg_base(a.qseg_cap) + s.qpos;
        const int64_t obase = kSeg * a.qseg_cap + s.qopos;
        for (int i = threadIdx.x; i < n; i += BLOCK) {
            if (i < fit)
                a.qout[base + i] = s.q[i];
            else
                a.qout[obase + (i - fit)] = s.q[i];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) s.qn = 0;
    __syncthreads();
}

// ---------------------------------------------------------------- top-down

// Visit edge (src -> u).  Wave-collective.  The visited word is first read
// with a device-scope load, which bypasses the CU's L1: a stale L1 copy would
// send every later edge into a popular vertex to the atomic, and atomics on
// one word serialize (~12 ns each).  With tdw (the write half of a top-down
// step with remote writes) the edge only writes a candidate parent.
// A remote-write step marks the visited words it writes candidates into,
// and its claim sweep lists only those, when its frontier has fewer than
// kDirtyWordEdges edges per live visited word.  Heavier steps write into
// nearly every word, and the marks would only add a scattered store per
// candidate.  Both launches of the step see the same frontier_deg.
constexpr int64_t kDirtyWordEdges = 8;
__device__ __forceinline__ bool use_dirty(const step_args &a, int64_t frontier_deg) {
    return frontier_deg < kDirtyWordEdges * a.live_nwords;
}

template <int BLOCK>
__device__ __forceinline__ void td_visit(const step_args &a, step_smem<BLOCK> &s, bool tdw, bool valid, int32_t u,
                                         int32_t src, unsigned long long &cnt,
                                         unsigned long long &dsum, unsigned long long &dmax) {
    if (tdw) {
        // remote write (hybrid_bfs.cc:47-54): a plain store of the source as
        // the candidate parent of an unvisited end; nothing sets visited
        // bits in this launch, so the cached word is current
        // The next-frontier bitmap, zero and otherwise unused during this
        // launch, marks the ends already written: a plain read-modify-write
        // that each L2 keeps for its own XCD, so most repeated ends cost no
        // second scattered store.  A lost or unseen mark only costs a
        // repeated store of another valid candidate; the bitmap is cleared
        // before it next holds a frontier.
        if (valid && !((a.vis[u >> 6] >> (u & 63)) & 1)) {
            uint64_t *mw = a.next_bm + (u >> 6);
            const uint64_t w = *mw;
            const uint64_t bit = 1ULL << (u & 63);
            if (!(w & bit)) {
                a.cand[u] = src;
                *mw = w | bit;
                // same value from every writer: a plain byte store
                if (use_dirty(a, s.st.frontier_deg)) a.dirty[u >> 6] = 1;
            }
        }
        return;
    }
    bool won = false;
    const uint64_t bit = 1ULL << (u & 63);
    if (valid) {
        uint64_t *wp = a.vis + (u >> 6);
        // heavy steps pre-check the word (a popular vertex's word would
        // otherwise take an atomic per edge); light steps go straight to it
        const uint64_t seen =
            s.st.frontier_deg >= kPrecheckEdges
                ? __hip_atomic_load((unsigned long long *)wp, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT)
                : 0ULL;
        if (!(seen & bit)) {
            const unsigned long long old =
                atomicOr((unsigned long long *)wp, (unsigned long long)bit);
            won = !(old & bit);
        }
    }
    int64_t b = 0, e = 0;
    if (won) {
        a.parent[u] = src;
        b = a.off[u];
        e = a.off[u + 1];
        const unsigned long long d = (unsigned long long)(e - b);
        ++cnt;
        dsum += d;
        dmax = d > dmax ? d : dmax;
        atomicOr((unsigned long long *)(a.next_bm + (u >> 6)), (unsigned long long)bit);
    }
    emit_next(a, s, won, u, b, e);
}

// kTdRounds rounds of edges (u[r] on each lane, all from `src` or each
// with its own source) visited together: each round's claim is a chain of
// dependent memory round trips (pre-check, claim, the new vertex's
// offsets), so the rounds' loads and claims are issued back to back and a
// wave waits one chain for up to 64 * kTdRounds edges.  Same effect as
// td_visit on each round in turn.
template <int BLOCK>
__device__ __forceinline__ void td_visit_n(const step_args &a, step_smem<BLOCK> &s, bool tdw,
                                           const bool (&valid)[kTdRounds],
                                           const int32_t (&u)[kTdRounds],
                                           const int32_t (&src)[kTdRounds], unsigned long long &cnt,
                                           unsigned long long &dsum, unsigned long long &dmax) {
    if (tdw) {
#pragma unroll
        for (int r = 0; r < kTdRounds; ++r)
            td_visit(a, s, true, valid[r], u[r], src[r], cnt, dsum, dmax);
        return;
    }
    // Loads and claims are issued for every round before any result is
    // used, with in-range addresses for idle lanes (their results are
    // dropped), so the compiler keeps them in flight together.
    uint64_t seen[kTdRounds];
    if (s.st.frontier_deg >= kPrecheckEdges) {
#pragma unroll
        for (int r = 0; r < kTdRounds; ++r)
            seen[r] = __hip_atomic_load((unsigned long long *)(a.vis + (u[r] >> 6)),
                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
#pragma unroll
        for (int r = 0; r < kTdRounds; ++r) seen[r] = 0;
    }
    unsigned long long old[kTdRounds];
#pragma unroll
    for (int r = 0; r < kTdRounds; ++r) {
        const uint64_t bit = 1ULL << (u[r] & 63);
        old[r] = ~0ULL;  // no claim
        if (valid[r] && !(seen[r] & bit))
            old[r] = atomicOr((unsigned long long *)(a.vis + (u[r] >> 6)),
                              (unsigned long long)bit);
    }
    bool won[kTdRounds];
    int64_t b[kTdRounds], e[kTdRounds];
#pragma unroll
    for (int r = 0; r < kTdRounds; ++r) {
        won[r] = !(old[r] & (1ULL << (u[r] & 63)));
        const int32_t w = won[r] ? u[r] : 0;
        b[r] = a.off[w];
        e[r] = a.off[w + 1];
    }
#pragma unroll
    for (int r = 0; r < kTdRounds; ++r) {
        if (won[r]) {
            a.parent[u[r]] = src[r];
            const unsigned long long d = (unsigned long long)(e[r] - b[r]);
            ++cnt;
            dsum += d;
            dmax = d > dmax ? d : dmax;
            atomicOr((unsigned long long *)(a.next_bm + (u[r] >> 6)),
                     (unsigned long long)(1ULL << (u[r] & 63)));
        } else {
            b[r] = e[r] = 0;
        }
        emit_next(a, s, won[r], u[r], b[r], e[r]);
    }
}

// One tile of frontier vertices, one per thread (v < 0: none).  Vertices of
// degree >= wave_t get a wave each; the edges of the others are concatenated
// by a block-wide scan and every thread takes one.  Whole workgroup.
template <int BLOCK>
__device__ __forceinline__ void td_tile(const step_args &a, step_smem<BLOCK> &s, bool tdw, int32_t v,
                        unsigned long long &cnt, unsigned long long &dsum,
                        unsigned long long &dmax) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) s.t_nmed = 0;
    __syncthreads();
    int64_t b = 0, e = 0;
    if (v >= 0) {
        b = a.off[v];
        e = a.off[v + 1];
    }
    const int64_t dg = e - b;
    int32_t small = 0;
    if (dg >= a.wave_t) {
        const int p = atomicAdd(&s.t_nmed, 1);
        s.t_med[p] = tid;
    } else {
        small = (int32_t)dg;
    }
    s.t_beg[tid] = b;
    s.t_end[tid] = e;
    s.t_src[tid] = v;
    const int32_t incl = wave_incl_scan32(small);
    if (lane == 63) s.t_wsum[wid] = incl;
    __syncthreads();
    int32_t wbase = 0;
    for (int w2 = 0; w2 < wid; ++w2) wbase += s.t_wsum[w2];
    s.t_pref[tid] = wbase + incl - small;
    if (tid == BLOCK - 1) s.t_pref[BLOCK] = wbase + incl;
    __syncthreads();
    const int32_t total = s.t_pref[BLOCK];
    const int nmed = s.t_nmed;
    for (int i = wid; i < nmed; i += BLOCK / 64) {
        const int t = s.t_med[i];
        const int64_t mb = s.t_beg[t], me = s.t_end[t];
        const int32_t src = s.t_src[t];
        for (int64_t k0 = mb; k0 < me; k0 += 64 * kTdRounds) {
            bool valid[kTdRounds];
            int32_t u[kTdRounds], sr[kTdRounds];
#pragma unroll
            for (int r = 0; r < kTdRounds; ++r) {
                const int64_t k = k0 + 64 * r + lane;
                valid[r] = k < me;
                u[r] = valid[r] ? __builtin_nontemporal_load(a.cols + k) : 0;
                sr[r] = src;
            }
            td_visit_n(a, s, tdw, valid, u, sr, cnt, dsum, dmax);
        }
    }
    for (int32_t i0 = 0; i0 < total; i0 += BLOCK * kTdRounds) {
        bool valid[kTdRounds];
        int32_t u[kTdRounds], sr[kTdRounds];
#pragma unroll
        for (int r = 0; r < kTdRounds; ++r) {
            const int32_t i = i0 + BLOCK * r + tid;
            valid[r] = i < total;
            u[r] = 0;
            sr[r] = 0;
            if (valid[r]) {
                int lo = 0, hi = BLOCK;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s.t_pref[mid] <= i)
                        lo = mid;
                    else
                        hi = mid;
                }
                sr[r] = s.t_src[lo];
                u[r] = a.cols[s.t_beg[lo] + (i - s.t_pref[lo])];
            }
        }
        td_visit_n(a, s, tdw, valid, u, sr, cnt, dsum, dmax);
    }
    __syncthreads();
    if (s.qn >= kQCap / 2) flush_queue<BLOCK>(a, s);
}

// Task slices of the frontier's high-degree vertices, one wave each, in
// workgroup-synchronous rounds.
template <int BLOCK>
__device__ __forceinline__ void td_tasks(const step_args &a, step_smem<BLOCK> &s, bool tdw, unsigned long long &cnt,
                         unsigned long long &dsum, unsigned long long &dmax) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    const int64_t ntasks = s.st.tpref[kSeg + 1];
    const int64_t waves = (int64_t)gridDim.x * (BLOCK / 64);
    // Slices are cut into pieces of a multiple of 64 edges, each taken by
    // its own wave, sized so the frontier's edges make about four pieces
    // per wave: a frontier of a few high-degree vertices has too few slices
    // to occupy the grid, and each slice is a chain of dependent claims.
    int64_t piece = (s.st.frontier_deg / (4 * waves) + 63) / 64 * 64;
    piece = piece < 64 ? 64 : (piece > a.task_edges ? a.task_edges : piece);
    const int64_t per = (a.task_edges + piece - 1) / piece;  // pieces per slice
    const int64_t npieces = ntasks * per;
    for (int64_t t0 = (int64_t)blockIdx.x * (BLOCK / 64); t0 < npieces; t0 += waves) {
        const int64_t t = t0 + wid;
        if (t < npieces) {
            const int64_t ti = t / per;
            const int sg = seg_of(s.st.tpref, ti);
            const td_task tk = a.tin[(int64_t)sg * a.tseg_cap + (ti - s.st.tpref[sg])];
            const int64_t beg = tk.beg + (t - ti * per) * piece;
            const int64_t end = beg + piece < tk.end ? beg + piece : tk.end;
            for (int64_t k0 = beg; k0 < end; k0 += 64 * kTdRounds) {
                bool valid[kTdRounds];
                int32_t u[kTdRounds], sr[kTdRounds];
#pragma unroll
                for (int r = 0; r < kTdRounds; ++r) {
                    const int64_t k = k0 + 64 * r + lane;
                    valid[r] = k < end;
                    u[r] = valid[r] ? __builtin_nontemporal_load(a.cols + k) : 0;
                    sr[r] = tk.src;
                }
                td_visit_n(a, s, tdw, valid, u, sr, cnt, dsum, dmax);
            }
        }
        __syncthreads();
        if (s.qn >= kQCap / 2) flush_queue<BLOCK>(a, s);
    }
}

template <int BLOCK>
__device__ __forceinline__ void td_body(const step_args &a, step_smem<BLOCK> &s, bool tdw, unsigned long long &cnt,
                        unsigned long long &dsum, unsigned long long &dmax) {
    const int64_t nitems = s.st.qpref[kSeg + 1];
    for (int64_t tile = blockIdx.x; tile * BLOCK < nitems; tile += gridDim.x) {
        const int64_t slot = tile * BLOCK + threadIdx.x;
        int32_t v = -1;
        if (slot < nitems) {
            const int sg = seg_of(s.st.qpref, slot);
            v = a.qin[(int64_t)sg * a.qseg_cap + (slot - s.st.qpref[sg])];
        }
        td_tile<BLOCK>(a, s, tdw, v, cnt, dsum, dmax);
    }
    td_tasks<BLOCK>(a, s, tdw, cnt, dsum, dmax);
}

// Top-down step whose frontier is only in the depth-L bitmap (the previous
// step was bottom-up, which does not build lists).  Each workgroup takes
// ranges of BLOCK frontier words, ranks their set bits with a block-wide
// scan, and feeds them to td_tile BLOCK vertices at a time.  When the
// frontier may hold many hubs (its largest degree is at least kHubDeg and
// it has kHubPassEdges edges or more), the ranges skip them and the hubs are
// spread over the grid: 64 medium hubs checked per wave, the big ones as
// slices, one per wave.  In degree order the hubs share the first ranges,
// whose workgroup would otherwise walk all their lists.
template <int BLOCK>
__device__ __forceinline__ void td_bitmap_body(const step_args &a, step_smem<BLOCK> &s, bool tdw, unsigned long long &cnt,
                               unsigned long long &dsum, unsigned long long &dmax) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const bool hubs = s.st.max_frontier_deg >= kHubDeg && s.st.frontier_deg >= kHubPassEdges;
    const int64_t nranges = (a.live_nwords + BLOCK - 1) / BLOCK;
    for (int64_t r = blockIdx.x; r < nranges; r += gridDim.x) {
        const int64_t wi = r * BLOCK + tid;
        uint64_t w = wi < a.live_nwords ? a.front[wi] : 0ULL;
        if (hubs && w) w &= ~a.hub_bm[wi];
        const int32_t c = __popcll(w);
        s.w_bits[tid] = w;
        const int32_t incl = wave_incl_scan32(c);
        if (lane == 63) s.w_wsum[wid] = incl;
        __syncthreads();
        int32_t wbase = 0;
        for (int w2 = 0; w2 < wid; ++w2) wbase += s.w_wsum[w2];
        s.w_pref[tid] = wbase + incl - c;
        if (tid == BLOCK - 1) s.w_pref[BLOCK] = wbase + incl;
        __syncthreads();
        const int32_t total = s.w_pref[BLOCK];
        for (int32_t t0 = 0; t0 < total; t0 += BLOCK) {
            const int32_t i = t0 + tid;
            int32_t v = -1;
            if (i < total) {
                int lo = 0, hi = BLOCK;  // largest j with w_pref[j] <= i
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s.w_pref[mid] <= i)
                        lo = mid;
                    else
                        hi = mid;
                }
                v = (int32_t)((r * BLOCK + lo) * 64 +
                              select_bit(s.w_bits[lo], i - s.w_pref[lo]));
            }
            td_tile<BLOCK>(a, s, tdw, v, cnt, dsum, dmax);
        }
        __syncthreads();
    }
    if (!hubs) return;
    const int64_t nmed = (a.n_hub_med + 63) / 64;
    const int64_t nh = nmed + (s.st.max_frontier_deg >= kHubBigDeg ? a.n_hub_big : 0);
    const int64_t waves = (int64_t)gridDim.x * (BLOCK / 64);
    for (int64_t t0 = (int64_t)blockIdx.x * (BLOCK / 64); t0 < nh; t0 += waves) {
        const int64_t t = t0 + wid;
        if (t < nmed) {
            const int64_t i = t * 64 + lane;
            const int32_t v = i < a.n_hub_med ? a.hub_med[i] : -1;
            const bool in_f = v >= 0 && ((a.front[v >> 6] >> (v & 63)) & 1);
            int64_t b = 0, e = 0;
            if (in_f) {
                b = a.off[v];
                e = a.off[v + 1];
            }
            for (uint64_t m = __ballot(in_f); m; m &= m - 1) {
                const int j = __ffsll((unsigned long long)m) - 1;
                const int64_t hb = lane_bcast(b, j), he = lane_bcast(e, j);
                const int32_t src = lane_bcast(v, j);
                for (int64_t k0 = hb; k0 < he; k0 += 64 * kTdRounds) {
                    bool valid[kTdRounds];
                    int32_t u[kTdRounds], sr[kTdRounds];
#pragma unroll
                    for (int r = 0; r < kTdRounds; ++r) {
                        const int64_t k = k0 + 64 * r + lane;
                        valid[r] = k < he;
                        u[r] = valid[r] ? __builtin_nontemporal_load(a.cols + k) : 0;
                        sr[r] = src;
                    }
                    td_visit_n(a, s, tdw, valid, u, sr, cnt, dsum, dmax);
                }
            }
        } else if (t < nh) {
            const td_task tk = a.hub_big[t - nmed];
            if ((a.front[tk.src >> 6] >> (tk.src & 63)) & 1) {
                for (int64_t k0 = tk.beg; k0 < tk.end; k0 += 64 * kTdRounds) {
                    bool valid[kTdRounds];
                    int32_t u[kTdRounds], sr[kTdRounds];
#pragma unroll
                    for (int r = 0; r < kTdRounds; ++r) {
                        const int64_t k = k0 + 64 * r + lane;
                        valid[r] = k < tk.end;
                        u[r] = valid[r] ? __builtin_nontemporal_load(a.cols + k) : 0;
                        sr[r] = tk.src;
                    }
                    td_visit_n(a, s, tdw, valid, u, sr, cnt, dsum, dmax);
                }
            }
        }
        __syncthreads();
        if (s.qn >= kQCap / 2) flush_queue<BLOCK>(a, s);
    }
}

// --------------------------------------------------------------- bottom-up

// Scan the lists of the wave's pending vertices (those whose head index is
// not in the frontier), pend[i] for i < n.  The first bu_cap neighbours of
// all n lists are laid end to end and taken 64 at a time, one per lane;
// a hit records (position << 32 | neighbour) with an LDS atomic minimum, so
// each vertex keeps its first hit in CSR order, as a sequential scan would.
// Lists that hold no hit in their first bu_cap neighbours are finished by
// the whole wave.  Found vertices get their parent and their bit in the
// wave's found words (chunk word (v >> 6) - w0).
template <int BLOCK>
__device__ __forceinline__ void bu_pending(const step_args &a, step_smem<BLOCK> &s, const int32_t *pend, int n,
                           int64_t w0, unsigned long long &cnt, unsigned long long &dsum,
                           unsigned long long &dmax, unsigned long long &exam) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    const bool valid = lane < n;
    const int64_t v = valid ? pend[lane] : 0;
    int64_t b = 0, e = 0;
    if (valid) {
        b = a.off[v];
        e = a.off[v + 1];
    }
    const int32_t len = (int32_t)(e - b < a.bu_cap ? e - b : a.bu_cap);
    int32_t excl = wave_incl_scan32(len);
    const int32_t total = lane_bcast(excl, 63);
    excl -= len;
    unsigned long long *hit = s.b_hit[wid];
    hit[lane] = ~0ULL;
    wave_lds_sync();
    for (int32_t k0 = 0; k0 < total; k0 += 64) {
        const int32_t idx = k0 + lane;
        // owner: the largest lane i with excl_i <= idx
        int i = 0;
#pragma unroll
        for (int st = 32; st > 0; st >>= 1) {
            const int mid = i + st;
            const int32_t em = __shfl(excl, mid < 64 ? mid : 63);
            if (mid < 64 && em <= idx) i = mid;
        }
        const int32_t pos = idx - __shfl(excl, i);
        const int64_t bi = __shfl(b, i);
        if (idx < total) {
            const int32_t u = __builtin_nontemporal_load(a.cols + bi + pos);
            if ((a.front[u >> 6] >> (u & 63)) & 1)
                atomicMin(&hit[i], ((unsigned long long)(uint32_t)pos << 32) | (uint32_t)u);
        }
    }
    wave_lds_sync();
    const unsigned long long hk = hit[lane];
    int32_t par = -1;
    if (valid) {
        if (hk != ~0ULL) {
            par = (int32_t)(uint32_t)hk;
            exam += (hk >> 32) + 1;
        } else {
            exam += (unsigned long long)len;
        }
    }
    // lists with no hit in their first bu_cap neighbours: the whole wave
    uint64_t lm = __ballot(valid && par < 0 && e - b > len);
    while (lm) {
        const int src_lane = __ffsll((unsigned long long)lm) - 1;
        lm &= lm - 1;
        const int64_t kb = lane_bcast(b, src_lane) + lane_bcast(len, src_lane),
                      ke = lane_bcast(e, src_lane);
        int32_t found = -1;
        for (int64_t k0 = kb; k0 < ke && found < 0; k0 += 64) {
            const int64_t kk = k0 + lane;
            const int32_t u = kk < ke ? a.cols[kk] : -1;
            const bool h = u >= 0 && ((a.front[u >> 6] >> (u & 63)) & 1);
            const uint64_t hm = __ballot(h);
            const int64_t rest = ke - k0 < 64 ? ke - k0 : 64;
            const int64_t seen = hm ? (int64_t)__ffsll((unsigned long long)hm) : rest;
            if (lane == src_lane) exam += (unsigned long long)seen;
            if (hm) found = lane_bcast(u, __ffsll((unsigned long long)hm) - 1);
        }
        if (lane == src_lane) par = found;
    }
    if (par >= 0) {
        a.parent[v] = par;
        const unsigned long long d = (unsigned long long)(e - b);
        ++cnt;
        dsum += d;
        dmax = d > dmax ? d : dmax;
        atomicOr((unsigned long long *)&s.b_found[wid][(v >> 6) - w0],
                 (unsigned long long)(1ULL << (v & 63)));
    }

    wave_lds_sync();  // hit[] is rewritten by the next batch
}

// List the unvisited vertices of a wave's chunk in LDS: lane i holds visited
// word w0 + i (i < CW), and writes the positions (chunk-relative, word * 64 +
// bit) of its word's clear bits, in order, after those of the lanes before
// it.  Returns the count (wave-uniform); `unv` receives the lane's clear
// bits.  One pass over the bits instead of a rank search per vertex.
__device__ __forceinline__ int bu_list_chunk(const step_args &a, uint16_t *list, int64_t w0,
                                             int CW, uint64_t &vw, uint64_t &unv,
                                             bool claim = false) {
    const int lane = lane_id();
    const int64_t wi = w0 + lane;
    const bool own = lane < CW && wi < a.live_nwords;
    vw = own ? a.vis[wi] : ~0ULL;
    if (own) a.clr[wi] = 0;
    unv = ~vw;
    if (claim && own) {
        // a claim sweep of a light step lists only the words the write half
        // wrote a candidate into (and resets their flags)
        if (a.dirty[wi])
            a.dirty[wi] = 0;
        else
            unv = 0;
    }
    const int cntw = __popcll(unv);
    int pos = wave_incl_scan32(cntw);
    const int total = lane_bcast(pos, 63);
    pos -= cntw;
    for (uint64_t u = unv; u; u &= u - 1)
        list[pos++] = (uint16_t)(lane * 64 + (__ffsll((unsigned long long)u) - 1));
    wave_lds_sync();
    return total;
}

// Claim half of a top-down step run with remote writes (hybrid_bfs.cc:56-66):
// every unvisited vertex with a candidate takes it as its parent and joins
// the next frontier, which is left in the bitmap only, as after a bottom-up
// step.  Waves own chunks of bu_words visited words and list their unvisited
// vertices as the bottom-up step does; each consumed candidate is reset, so
// the array is all -1 again when the step ends.
template <int BLOCK>
__device__ __forceinline__ void td_claim_body(const step_args &a, step_smem<BLOCK> &s, unsigned long long &cnt,
                              unsigned long long &dsum, unsigned long long &dmax) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    const int CW = a.bu_words;
    const int64_t nchunks = (a.live_nwords + CW - 1) / CW;
    uint16_t *list = s.b_list[wid];
    for (int64_t chunk = (int64_t)blockIdx.x * (BLOCK / 64) + wid; chunk < nchunks;
         chunk += (int64_t)gridDim.x * (BLOCK / 64)) {
        const int64_t w0 = chunk * CW;
        s.b_found[wid][lane] = 0;
        uint64_t vw, unv;
        const int total =
            bu_list_chunk(a, list, w0, CW, vw, unv, use_dirty(a, s.st.frontier_deg));
        // kClaimRounds rounds of 64 vertices at a time: their candidates
        // are loaded together, then the degrees of the claimed ones (the
        // degree half of the head entry)
        constexpr int kClaimRounds = 2 * kBuBatch;
        const uint32_t *hw = (const uint32_t *)a.head;
        for (int r0 = 0; r0 < total; r0 += 64 * kClaimRounds) {
            int idx[kClaimRounds];
            int32_t p[kClaimRounds];
#pragma unroll
            for (int k = 0; k < kClaimRounds; ++k) {
                const int g = r0 + 64 * k + lane;
                idx[k] = g < total ? list[g] : 0;
                p[k] = g < total ? __builtin_nontemporal_load(a.cand + w0 * 64 + idx[k]) : -1;
            }
            uint32_t d[kClaimRounds];
#pragma unroll
            for (int k = 0; k < kClaimRounds; ++k) {
                const int64_t v = w0 * 64 + idx[k];
                d[k] = 0;
                if (p[k] >= 0)
                    d[k] = a.bu_head ? hw[2 * v + 1] : (uint32_t)(a.off[v + 1] - a.off[v]);
            }
#pragma unroll
            for (int k = 0; k < kClaimRounds; ++k) {
                if (p[k] < 0) continue;
                const int64_t v = w0 * 64 + idx[k];
                a.cand[v] = -1;
                a.parent[v] = p[k];
                ++cnt;
                dsum += d[k];
                dmax = d[k] > dmax ? d[k] : dmax;
                atomicOr((unsigned long long *)&s.b_found[wid][idx[k] >> 6],
                         (unsigned long long)(1ULL << (idx[k] & 63)));
            }

        }
        wave_lds_sync();
        const int64_t wi = w0 + lane;
        if (lane < CW && wi < a.live_nwords) {
            const uint64_t f = s.b_found[wid][lane];
            a.next_bm[wi] = f;
            if (f) a.vis[wi] = vw | f;
        }
        wave_lds_sync();
    }
}

// Bottom-up step.  Each wave owns chunks of bu_words visited words and lists
// their unvisited vertices in LDS; per round of 64 vertices each lane reads
// its vertex's head index (the neighbour most likely to be in the frontier,
// read from a dense array in vertex order) and tests it.  Vertices whose
// head misses are queued in LDS and their lists scanned 64 at a time, so the
// rare long scans do not hold up every round.  Found bits are gathered in
// LDS and written by the word owner: no global atomics.
template <int BLOCK>
__device__ __forceinline__ void bu_body(const step_args &a, step_smem<BLOCK> &s, unsigned long long &cnt,
                        unsigned long long &dsum, unsigned long long &dmax,
                        unsigned long long &exam) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    const int CW = a.bu_words;
    const int64_t nchunks = (a.live_nwords + CW - 1) / CW;
    const uint64_t below = (1ULL << lane) - 1;
    int32_t *pend = s.b_pend[wid];
    uint16_t *list = s.b_list[wid];
    for (int64_t chunk = (int64_t)blockIdx.x * (BLOCK / 64) + wid; chunk < nchunks;
         chunk += (int64_t)gridDim.x * (BLOCK / 64)) {
        const int64_t w0 = chunk * CW;
        s.b_found[wid][lane] = 0;
        uint64_t vw, unv;
        const int total = bu_list_chunk(a, list, w0, CW, vw, unv);
        int np = 0;  // pending vertices (wave-uniform)
        // One vertex per lane: test its head entry hd against frontier word
        // fw; on a miss queue it for the list scan.  Wave-collective.
        auto visit = [&](bool valid, int idx, uint64_t hd, uint64_t fw) {
            const int64_t v = w0 * 64 + idx;
            const int32_t h = (int32_t)(uint32_t)hd;
            bool miss = valid;
            if (valid && a.bu_head && h >= 0) {
                ++exam;
                if ((hd >> 32) == 1 && !((fw >> (h & 63)) & 1)) {
                    // the only neighbour is the head, which missed: the list
                    // scan would read it again and find nothing
                    miss = false;
                    ++exam;
                } else if ((fw >> (h & 63)) & 1) {
                    miss = false;
                    a.parent[v] = h;
                    const unsigned long long d = hd >> 32;
                    ++cnt;
                    dsum += d;
                    dmax = d > dmax ? d : dmax;
                    atomicOr((unsigned long long *)&s.b_found[wid][idx >> 6],
                             (unsigned long long)(1ULL << (idx & 63)));

                }
            }
            const uint64_t mm = __ballot(miss);
            if (miss) pend[np + __popcll(mm & below)] = (int32_t)v;
            np += __popcll(mm);
        };
        // Scan the queued lists 64 at a time while 64 or more are queued
        // (the queue holds up to 64 * (kBuBatch + 1): fewer than 64 left
        // over, plus one batch of rounds).
        auto drain = [&]() {
            while (np >= 64) {
                wave_lds_sync();
                bu_pending(a, s, pend, 64, w0, cnt, dsum, dmax, exam);
                const int rest = np - 64;
                // move the rest down, 64 at a time in increasing order: each
                // block is read before the block written over it
                for (int i0 = 0; i0 < rest; i0 += 64) {
                    const int32_t t = i0 + lane < rest ? pend[64 + i0 + lane] : 0;
                    wave_lds_sync();
                    if (i0 + lane < rest) pend[i0 + lane] = t;
                    wave_lds_sync();
                }
                np = rest;
            }
        };
        auto head_of = [&](bool valid, int idx) -> uint64_t {
            return (valid && a.bu_head) ? __builtin_nontemporal_load(a.head + w0 * 64 + idx)
                                        : 0ULL;
        };
        auto front_of = [&](bool valid, uint64_t hd) -> uint64_t {
            const int32_t h = (int32_t)(uint32_t)hd;
            return (valid && a.bu_head && h >= 0) ? a.front[h >> 6] : 0ULL;
        };
        // R rounds of 64 vertices at a time: every round's head entry, then
        // every round's frontier word, are loaded before any is used, so the
        // wave waits one chain of two loads per 64 * R vertices; the next
        // batch's head entries are issued before this batch is visited
        auto batched = [&](auto rounds) {
            constexpr int R = decltype(rounds)::value;
            bool valid[R];
            int idx[R];
            uint64_t hd[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int g = 64 * k + lane;
                valid[k] = g < total;
                idx[k] = valid[k] ? list[g] : 0;
                hd[k] = head_of(valid[k], idx[k]);
            }
            for (int r0 = 0; r0 < total; r0 += 64 * R) {
                uint64_t fw[R];
#pragma unroll
                for (int k = 0; k < R; ++k) fw[k] = front_of(valid[k], hd[k]);
                bool validn[R];
                int idxn[R];
                uint64_t hdn[R];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int g = r0 + 64 * (R + k) + lane;
                    validn[k] = g < total;
                    idxn[k] = validn[k] ? list[g] : 0;
                    hdn[k] = head_of(validn[k], idxn[k]);
                }
#pragma unroll
                for (int k = 0; k < R; ++k) visit(valid[k], idx[k], hd[k], fw[k]);
                drain();
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    valid[k] = validn[k];
                    idx[k] = idxn[k];
                    hd[k] = hdn[k];
                }
            }
        };
        if (a.bu_batch >= 8) {
            // kBuBatch + 2 rounds: their neighbour halves of the head entries
            // (4 bytes each) in one trip, then the degree halves (the same
            // lines) with the frontier words; the queue is drained after
            // kBuBatch rounds, so it never holds more than 64 * (kBuBatch + 1)
            // (six rounds fit in the registers of five waves per SIMD;
            // eight, with the list entries read again, ran slower)
            constexpr int R = kBuBatch + 2;
            const uint32_t *hw = (const uint32_t *)(a.head + w0 * 64);
            const uint32_t *fh = (const uint32_t *)a.front;  // half words of the frontier
            for (int r0 = 0; r0 < total; r0 += 64 * R) {
                bool valid[R];
                int idx[R];
                uint32_t lo[R], hi[R], fw[R];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int g = r0 + 64 * k + lane;
                    valid[k] = g < total;
                    idx[k] = valid[k] ? list[g] : 0;
                    lo[k] = (valid[k] && a.bu_head) ? hw[2 * idx[k]] : 0u;
                }
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const bool hv = valid[k] && a.bu_head && (int32_t)lo[k] >= 0;
                    hi[k] = (valid[k] && a.bu_head) ? hw[2 * idx[k] + 1] : 0u;
                    fw[k] = hv ? fh[(int32_t)lo[k] >> 5] : 0u;
                }
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    // the half word back in its place in the 64-bit word
                    const uint64_t f64 = (lo[k] & 32) ? ((uint64_t)fw[k] << 32) : (uint64_t)fw[k];
                    visit(valid[k], idx[k], ((uint64_t)hi[k] << 32) | lo[k], f64);
                    if (k == kBuBatch - 1) drain();
                }
                drain();
            }
        } else if (a.bu_batch > 1) {
            batched(std::integral_constant<int, kBuBatch>{});
        } else {
            // rounds of 64 vertices, software-pipelined: the next round's
            // head entry is loaded while this round's frontier word is in
            // flight
            bool valid = lane < total;
            int idx = valid ? list[lane] : 0;
            uint64_t hd = head_of(valid, idx);
            for (int r = 0; r < total; r += 64) {
                const uint64_t fw = front_of(valid, hd);
                const bool validn = r + 64 + lane < total;
                const int idxn = validn ? list[r + 64 + lane] : 0;
                const uint64_t hdn = head_of(validn, idxn);
                visit(valid, idx, hd, fw);
                drain();
                valid = validn;
                idx = idxn;
                hd = hdn;
            }
        }
        if (np > 0) {
            wave_lds_sync();
            bu_pending(a, s, pend, np, w0, cnt, dsum, dmax, exam);
        }
        wave_lds_sync();
        const int64_t wi = w0 + lane;
        if (lane < CW && wi < a.live_nwords) {
            const uint64_t f = s.b_found[wid][lane];
            a.next_bm[wi] = f;
            if (f) a.vis[wi] = vw | f;
        }
        wave_lds_sync();
    }
}

// --------------------------------------------------------- direction policy

// hybrid_bfs.cc:162-190 as a per-step state machine.  After a TD step the
// scout count is sum deg(N) (the -deg encoding of clear(), :93, :438); after
// a BU step the do-while test (:178-179) decides whether to stay bottom-up;
// leaving sets scout = 1 (:180).  Then the outer test (:167) picks the
// direction, charging scout to edges_to_check before a TD step (:182).
// s > e / alpha with C division (alpha > 0), without a 64-bit divide when
// s, e >= 0: floor(e / alpha) < s exactly when e < s * alpha.
__device__ __forceinline__ bool gt_quotient(int64_t s, int64_t e, int64_t alpha) {
    if (s < 0 || e < 0) return s > e / alpha;
    long long p;
    if (__builtin_mul_overflow((long long)s, (long long)alpha, &p)) return true;
    return p > e;
}

__device__ __forceinline__ void beamer_next(const step_args &a, ctl_block *c, int32_t last_dir, int64_t nnew,
                            int64_t nnew_deg) {
    if (last_dir == BD_DIR_TD) {
        c->scout = nnew_deg;
    } else if (c->in_bu_loop) {
        if (nnew >= c->old_awake || nnew > a.nv_beta) {
            c->old_awake = nnew;
            c->dir = BD_DIR_BU;
            return;
        }
        c->in_bu_loop = 0;
        c->scout = 1;
    }
    if (gt_quotient(c->scout, c->edges_to_check, c->alpha)) {
        c->in_bu_loop = 1;
        c->old_awake = c->frontier;
        c->dir = BD_DIR_BU;
    } else {
        c->edges_to_check -= c->scout;
        c->dir = BD_DIR_TD;
    }
}

// Choose the direction of the step described by *c on one wave (no
// workgroup barrier): lane 0 updates the state; the learned policy evaluates
// its network on the wave.
template <int BLOCK>
__device__ __forceinline__ void choose_dir(const step_args &a, step_smem<BLOCK> &s, ctl_block *c, int32_t last_dir,
                           int64_t nnew, int64_t nnew_deg) {
    if (lane_id() == 0) {
        switch (a.policy) {
        case BD_POLICY_TD_ONLY:
            c->dir = BD_DIR_TD;
            break;
        case BD_POLICY_BU_ONLY:
            c->dir = BD_DIR_BU;
            break;
        case BD_POLICY_REMOTE_WRITES:
            c->dir = kDirTDWrite;
            break;
        case BD_POLICY_REMOTE_WRITES_HYBRID:
            // Beamer's state machine; its bottom-up phase writes remotely
            beamer_next(a, c, last_dir, nnew, nnew_deg);
            if (c->dir == BD_DIR_BU) c->dir = kDirTDWrite;
            break;
        case BD_POLICY_SCRIPT: {
            const int32_t i = c->level < a.script_len ? c->level : a.script_len - 1;
            c->dir = (a.script_len > 0 && a.script[i] == BD_DIR_BU) ? BD_DIR_BU : BD_DIR_TD;
            break;
        }
        case BD_POLICY_LEARNED:
            // Beamer's choice and state, which the network may override
            beamer_next(a, c, last_dir, nnew, nnew_deg);
            break;
        default:
            beamer_next(a, c, last_dir, nnew, nnew_deg);
        }
    }
    // the network runs only for LEARNED with a network loaded (kernel
    // arguments, the same in every lane), and only where Beamer's rules
    // switch direction (hybrid_bfs.cc:167 turning bottom-up, :178-179
    // leaving it): it decides whether to switch now or to stay; and at a
    // top-down step Beamer keeps top-down whose frontier edges come within
    // a factor mlp_near of the switch (scout * alpha * near > the unvisited
    // edges before this step's charge): whether to switch early.  Elsewhere
    // Beamer's choice stands without a network evaluation, whose latency
    // every step would otherwise pay.
    if (a.policy == BD_POLICY_LEARNED && a.has_mlp) wave_lds_sync();
    if (a.policy == BD_POLICY_LEARNED && a.has_mlp && last_dir >= 0 &&
        (c->dir != last_dir ||
         (a.mlp_near > 1 && last_dir == BD_DIR_TD && c->dir == BD_DIR_TD &&
          gt_quotient(c->scout, c->edges_to_check + c->scout,
                      c->alpha > (1LL << 46) ? (1LL << 62) : c->alpha * a.mlp_near)))) {
        // the operands are loaded only here: every workgroup loading the
        // 6 KB table at every step cost more than a consultation
        mlp_wave_regs w;
        mlp_load(a.mlp_ops, w);
        const int l = lane_id();
        if (l < kMlpIn)
            s.x[l] = mlp_feature_k(l, c->frontier, c->frontier_deg, c->max_frontier_deg,
                                   c->unvisited, c->unvisited_deg, c->level, c->nv, a.m,
                                   last_dir);
        wave_lds_sync();
        float y[2];
        mlp_eval(w, s.x, y);
        // y = predicted log2 cost of (top-down, bottom-up).  The network
        // overrides Beamer only when it predicts the other direction cheaper
        // by more than the margin; a NaN prediction keeps Beamer's choice.
        if (lane_id() == 0) {
            const float m = a.mlp_margin;
            int want = c->dir;
            if (c->dir == BD_DIR_TD && y[1] + m < y[0]) want = BD_DIR_BU;
            if (c->dir == BD_DIR_BU && y[0] + m < y[1]) want = BD_DIR_TD;
            if (want != c->dir) {
                // leave Beamer's state as if it had chosen `want`
                if (want == BD_DIR_BU) {
                    c->edges_to_check += c->scout;  // undo the top-down charge
                    c->in_bu_loop = 1;
                    c->old_awake = c->frontier;
                } else {
                    c->in_bu_loop = 0;
                    c->edges_to_check -= c->scout;
                }
                c->dir = want;
            }
        }
    }
    if (lane_id() == 0) {
        if (c->dir == BD_DIR_TD && a.policy != BD_POLICY_TD_ONLY &&
            a.policy != BD_POLICY_REMOTE_WRITES_HYBRID && a.td_split > 0 &&
            c->frontier_deg >= a.td_split)
            c->dir = kDirTDWrite;
        c->td_nolist =
            (c->dir == BD_DIR_TD && (c->frontier_deg >= kListEdges || a.max_deg < a.wave_t)) ? 1
                                                                                             : 0;
    }
    wave_lds_sync();
}

// Prefix sums of the segment lengths of a list into pref[0..kSeg+1] on one
// wave: lane s holds the length of segment s (capped at seg_cap), `ovf` the
// overflow segment's (capped at ovf_cap; read by every lane).
__device__ __forceinline__ void seg_prefix(int64_t n, int64_t ovf, int64_t seg_cap, int64_t ovf_cap,
                           int64_t *pref) {
    const int l = lane_id();
    n = n < seg_cap ? n : seg_cap;
    ovf = ovf < ovf_cap ? ovf : ovf_cap;
    const int64_t incl = wave_incl_scan(n);
    pref[l + 1] = incl;
    if (l == 0) pref[0] = 0;
    if (l == 63) pref[kSeg + 1] = incl + ovf;
}

constexpr int kCtlWords = (int)(sizeof(ctl_block) / 8);
static_assert(kCtlWords <= 3 * 64, "ctl_block is copied as three words per lane");

// Rebuild, in s.st, the state before step L = a.level from the published
// state before step L-1 and the counters of step L-1: close step L-1
// (record, totals), describe frontier L (segment prefix sums) and choose
// its direction.  Every workgroup of the step runs this on its first wave
// and gets the same result, so no workgroup waits for another; `publish`
// (workgroup 0) also writes the step record.  All global loads are issued
// before any is used, so the rebuild costs one memory round trip.  Returns
// true, having changed nothing, when the search ended before this level.
template <int BLOCK>
__device__ __forceinline__ bool rebuild_state(const step_args &a, step_smem<BLOCK> &s, bool publish) {
    const int L = a.level;
    const int l = lane_id();
    seg_counters *g = a.seg;
    const uint64_t *src = (const uint64_t *)&a.c[(L - 1) & 1];
    const int pa = (L - 1) % 3, pq = L % 3;
    uint64_t cw[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) cw[i] = l + 64 * i < kCtlWords ? src[l + 64 * i] : 0;
    unsigned long long ac[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) ac[k] = g->acc[pa][l][k];
    const int64_t ql = (int64_t)g->qlen[pq][l][0], qo = (int64_t)g->qlen[pq][kSeg][0];
    const int64_t tl = (int64_t)g->tlen[pq][l][0], to = (int64_t)g->tlen[pq][kSeg][0];
    {
        uint64_t *dst = (uint64_t *)&s.st;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if (l + 64 * i < kCtlWords) dst[l + 64 * i] = cw[i];
    }
    wave_lds_sync();
    // the search ended before this level: leave everything as it is (the
    // launch only has to see the flag, which is part of the state it loaded)
    if (s.st.done) return true;
    const unsigned long long nn = wave_sum(ac[0]);
    const unsigned long long nd = wave_sum(ac[1]);
    const unsigned long long ex = wave_sum(ac[2]);
    const unsigned long long mx = wave_max(ac[3]);
    wave_lds_sync();
    ctl_block *c = &s.st;
    if (l == 0) {
        const int32_t dir = c->dir;
        s.s_dir = dir;
        s.s_new = 0;
        s.s_split = dir == kDirTDClaim;
        if (dir == kDirTDWrite) {
            // the write half found nothing by itself: its claim half runs
            // next, at the same level and from the same state (the frontier
            // lists stay as they were, for the step's record)
            c->dir = kDirTDClaim;
        } else {
            const int64_t nnew = (int64_t)nn;
            const int64_t nnew_deg = (int64_t)nd;
            const int32_t sd = step_dir(dir);
            const int64_t idx = c->steps;  // records are per level
            if (publish && idx < a.log_cap) {
                bd_level_rec r;
                r.level = idx;
                r.dir = sd;
                r.frontier = c->frontier;
                r.frontier_deg = c->frontier_deg;
                r.unvisited = c->unvisited;
                r.unvisited_deg = c->unvisited_deg;
                r.edges_to_check = c->edges_to_check;
                r.result = sd == BD_DIR_TD ? nnew_deg : nnew;
                r.next = nnew;
                r.next_deg = nnew_deg;
                r.edges_examined = sd == BD_DIR_TD ? c->frontier_deg : (int64_t)ex;
                r.max_frontier_deg = c->max_frontier_deg;
                r.tasks = c->tpref[kSeg + 1];
                r.queued = c->qpref[kSeg + 1];
                r.ms = -1.0;
                r.ms_other = -1.0;
                r.launches = dir == kDirTDClaim ? 2 : 1;
                r.launch = L - r.launches;
                a.log[idx] = r;
            }
            c->steps += 1;
            if (sd == BD_DIR_TD)
                c->td_steps += 1;
            else
                c->bu_steps += 1;
            c->reached += nnew;
            c->reached_deg += nnew_deg;
            c->unvisited -= nnew;
            c->unvisited_deg -= nnew_deg;
            c->frontier = nnew;
            c->frontier_deg = nnew_deg;
            c->max_frontier_deg = (int64_t)mx;
            c->prev_dir = sd;
            // after a bottom-up step or a claim sweep the frontier is only
            // in the bitmap
            c->td_bitmap = (kernel_dir(dir) == BD_DIR_BU || c->td_nolist) ? 1 : 0;
            c->level = (int32_t)c->steps;
            if (nnew == 0) c->done = 1;
            s.s_new = nnew;
            s.s_new_deg = nnew_deg;
            s.s_dir = sd;
        }
    }
    wave_lds_sync();
    if (s.s_dir != kDirTDWrite) {
        seg_prefix(ql, qo, a.qseg_cap, a.nv, c->qpref);
        seg_prefix(tl, to, a.tseg_cap, a.task_cap, c->tpref);
        wave_lds_sync();
        // the remote-write steps of REMOTE_WRITES_HYBRID are its bottom-up
        // phase (hybrid_bfs.cc:262-268)
        const int32_t last = (a.policy == BD_POLICY_REMOTE_WRITES_HYBRID && s.s_split)
                                 ? BD_DIR_BU
                                 : s.s_dir;
        if (s.s_new > 0) choose_dir(a, s, c, last, s.s_new, s.s_new_deg);
    }
    return false;
}

// Workgroup 0, first wave: publish the state before step L as copy L&1 (both
// copies once the search is done, so later launches see it), and clear the
// counters the next step fills: acc[(L+1)%3] and the segment lengths
// [(L+2)%3], neither of which anyone reads or writes during step L.
template <int BLOCK>
__device__ __forceinline__ void publish_state(const step_args &a, step_smem<BLOCK> &s) {
    const int L = a.level;
    const int l = lane_id();
    seg_counters *g = a.seg;
    const bool done = s.st.done != 0;
    {
        const uint64_t *src = (const uint64_t *)&s.st;
        uint64_t *d0 = (uint64_t *)&a.c[L & 1];
        uint64_t *d1 = (uint64_t *)&a.c[(L + 1) & 1];
        for (int i = l; i < (int)(sizeof(ctl_block) / 8); i += 64) {
            d0[i] = src[i];
            if (done) d1[i] = src[i];
        }
    }
    for (int k = 0; k < 4; ++k) g->acc[(L + 1) % 3][l][k] = 0;
    g->qlen[(L + 2) % 3][l][0] = 0;
    g->tlen[(L + 2) % 3][l][0] = 0;
    if (l == 0) {
        g->qlen[(L + 2) % 3][kSeg][0] = 0;
        g->tlen[(L + 2) % 3][kSeg][0] = 0;
        // the first launch to publish "done" stamps the end (the step
        // kernel of the other direction republishes the same state)
        if (done && g->t_end == 0) g->t_end = wall_clock64();
    }
}

// ---------------------------------------------------------------- kernels

// Sum the block's counters and add them to its slot (spread over kSlots
// cache lines so no line sees more than a few dozen atomics per step).
template <int BLOCK>
__device__ __forceinline__ void block_flush_counters(const step_args &a, step_smem<BLOCK> &s, unsigned long long cnt,
                                     unsigned long long dsum, unsigned long long exam,
                                     unsigned long long dmax) {
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    cnt = wave_sum(cnt);
    dsum = wave_sum(dsum);
    exam = wave_sum(exam);
    dmax = wave_max(dmax);
    if (lane == 0) {
        s.red[0][wid] = cnt;
        s.red[1][wid] = dsum;
        s.red[2][wid] = exam;
        s.red[3][wid] = dmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; ++w) {
            cnt += s.red[0][w];
            dsum += s.red[1][w];
            exam += s.red[2][w];
            dmax = s.red[3][w] > dmax ? s.red[3][w] : dmax;
        }
        unsigned long long *slot = a.seg->acc[a.level % 3][blockIdx.x % kSlots];
        if (cnt) atomicAdd(&slot[0], cnt);
        if (dsum) atomicAdd(&slot[1], dsum);
        if (exam) atomicAdd(&slot[2], exam);
        if (dmax) atomicMax(&slot[3], dmax);
    }
}

// Which directions a step kernel carries: one each (the host launches both
// for a level, and the one that does not match the device's choice returns
// after the state rebuild), or both in one kernel.
constexpr int kStepTD = BD_DIR_TD, kStepBU = BD_DIR_BU, kStepBoth = 2;

// One BFS level.  The first wave of every workgroup rebuilds the state
// before the step (closing the previous step and choosing this step's
// direction); then the workgroup runs its share of the step.  close_only
// launches (one workgroup) only publish the state.  The rebuild is
// idempotent, so a level may be launched more than once.
template <int BLOCK, int DIRS>
__global__ __launch_bounds__(BLOCK) void k_step(step_args a) {
    const int L = a.level;
    __shared__ step_smem<BLOCK> s;
    if (threadIdx.x < 64) {
        if (L == 0) {
            const uint64_t *src = (const uint64_t *)&a.c[0];
            uint64_t *dst = (uint64_t *)&s.st;
            for (int i = threadIdx.x; i < (int)(sizeof(ctl_block) / 8); i += 64) dst[i] = src[i];
        } else {
            // one round trip: the previous state (done flag included) and the
            // counters are loaded together; a finished search stops there
            const bool was_done = rebuild_state(a, s, blockIdx.x == 0);
            if (blockIdx.x == 0 && !was_done) publish_state(a, s);
        }
    }
    if (threadIdx.x == 0) s.qn = 0;
    __syncthreads();
    if (s.st.done || a.close_only) return;
    const int dir = a.force_dir >= 0 ? a.force_dir : s.st.dir;
    const int kdir = kernel_dir(dir);
    if (DIRS != kStepBoth && kdir != DIRS) return;
    unsigned long long cnt = 0, dsum = 0, exam = 0, dmax = 0;
    if (DIRS != kStepBU && kdir == BD_DIR_TD) {
        const bool tdw = dir == kDirTDWrite;
        for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < a.live_nwords;
             i += (int64_t)gridDim.x * BLOCK)
            a.clr[i] = 0;
        if (s.st.td_bitmap)
            td_bitmap_body<BLOCK>(a, s, tdw, cnt, dsum, dmax);
        else
            td_body<BLOCK>(a, s, tdw, cnt, dsum, dmax);
        flush_queue<BLOCK>(a, s);
    } else if (DIRS != kStepTD) {
        if (dir == kDirTDClaim)
            td_claim_body<BLOCK>(a, s, cnt, dsum, dmax);
        else
            bu_body<BLOCK>(a, s, cnt, dsum, dmax, exam);
    }
    block_flush_counters<BLOCK>(a, s, cnt, dsum, exam, dmax);
}

// Start a search: clear() (hybrid_bfs.cc:433-443) has been done by k_clear;
// here the counters are zeroed, the sources are pushed and marked
// (hybrid_bfs.cc:157-163), and the first direction chosen.  One source is
// the reference's search; several (srcs[0..nsrc), one wave claiming them 64
// at a time, duplicates once) start every search of a batch of disjoint
// graphs at once.  Frontier 0 is written to segment 0 of list 0 and its
// overflow segment.
__global__ __launch_bounds__(64) void k_init(step_args a, int64_t source,
                                             const int64_t *__restrict__ srcs, int64_t nsrc) {
    seg_counters *g = a.seg;
    __shared__ step_smem<64> s;
    const int l = lane_id();
    // the timed region starts here: pushing the sources and the first
    // direction choice are inside run_beamer's timed call
    // (hybrid_bfs.cc:157-167, hybrid_bfs_main.cc:267-282)
    if (l == 0) g->t_start = wall_clock64();
    const uint64_t below = (1ULL << l) - 1;
    for (int w = 0; w < 3; ++w) {
        for (int i = l; i <= kSeg; i += 64) {
            g->qlen[w][i][0] = 0;
            g->tlen[w][i][0] = 0;
        }
        for (int k = 0; k < 4; ++k) g->acc[w][l][k] = 0;
    }
    // the state is built in LDS and published with one coalesced copy
    ctl_block *c = &s.st;
    for (int i = l; i < kCtlWords; i += 64) ((uint64_t *)&s.st)[i] = 0;
    if (l == 0) {
        g->task_overflow = 0;
        g->t_end = 0;
    }
    wave_lds_sync();
    int64_t nq = 0, nt = 0;  // wave-uniform list lengths
    unsigned long long cnt = 0, dsum = 0, dmax = 0, lost = 0;
    for (int64_t i0 = 0; i0 < nsrc; i0 += 64) {
        const int64_t i = i0 + l;
        const bool valid = i < nsrc;
        int64_t v = valid ? (srcs ? srcs[i] : source) : 0;
        if (a.inv) v = a.inv[v];  // sources are given in graph ids
        bool won = false;
        int64_t b = 0, e = 0;
        if (valid) {
            // the list bounds are loaded beside the claim, not after it
            const int64_t lb = a.off[v], le = a.off[v + 1];
            // claims go through bm[0] (zeroed by k_clear): vis already holds
            // the degree-0 vertices, which can still be sources
            const uint64_t bit = 1ULL << (v & 63);
            const unsigned long long old =
                atomicOr((unsigned long long *)&a.next_bm[v >> 6], (unsigned long long)bit);
            won = !(old & bit);
            if (won) {
                atomicOr((unsigned long long *)&a.vis[v >> 6], (unsigned long long)bit);
                a.parent[v] = (int32_t)v;
                b = lb;
                e = le;
                const unsigned long long d = (unsigned long long)(e - b);
                ++cnt;
                dsum += d;
                dmax = d > dmax ? d : dmax;
            }
        }
        const bool huge = won && (e - b) >= a.task_t;
        const bool q = won && !huge;
        const uint64_t qm = __ballot(q);
        if (q) {
            const int64_t p = nq + __popcll(qm & below);
            a.qout[p < a.qseg_cap ? p : kSeg * a.qseg_cap + (p - a.qseg_cap)] = (int32_t)v;
        }
        nq += __popcll(qm);
        const long long ntk = huge ? (e - b + a.task_edges - 1) / a.task_edges : 0;
        const long long incl = wave_incl_scan(ntk);
        for (long long t = 0; t < ntk; ++t) {
            td_task tk;
            tk.beg = b + t * a.task_edges;
            tk.end = tk.beg + a.task_edges < e ? tk.beg + a.task_edges : e;
            tk.src = (int32_t)v;
            tk.pad = 0;
            const int64_t p = nt + incl - ntk + t;
            if (p < a.tseg_cap)
                a.tout[p] = tk;
            else if (p - a.tseg_cap < a.task_cap)
                a.tout[kSeg * a.tseg_cap + (p - a.tseg_cap)] = tk;
            else
                ++lost;
        }
        nt += lane_bcast((int64_t)incl, 63);
    }
    cnt = wave_sum(cnt);
    dsum = wave_sum(dsum);
    dmax = wave_max(dmax);
    lost = wave_sum(lost);
    if (l == 0) {
        if (lost) g->task_overflow = lost;
        c->nv = a.nv;
        c->m2 = 2 * a.m;
        c->alpha = a.alpha;
        c->beta = a.beta;
        c->policy = a.policy;
        c->edges_to_check = 2 * a.m;
        c->scout = (int64_t)dsum;
        c->prev_dir = -1;
        c->frontier = (int64_t)cnt;
        c->frontier_deg = (int64_t)dsum;
        c->max_frontier_deg = (int64_t)dmax;
        c->unvisited = a.nv - (int64_t)cnt;
        c->unvisited_deg = a.nnz - (int64_t)dsum;
        c->reached = (int64_t)cnt;
        c->reached_deg = (int64_t)dsum;
    }
    {
        const int64_t q0 = nq < a.qseg_cap ? nq : a.qseg_cap;
        const int64_t t0 = nt < a.tseg_cap ? nt : a.tseg_cap;
        const int64_t tall = nt < t0 + a.task_cap ? nt : t0 + a.task_cap;
        for (int k = l; k < kSeg + 2; k += 64) {
            c->qpref[k] = k == 0 ? 0 : (k == kSeg + 1 ? nq : q0);
            c->tpref[k] = k == 0 ? 0 : (k == kSeg + 1 ? tall : t0);
        }
    }
    wave_lds_sync();
    // level 0: with no previous step Beamer applies only the outer test to
    // scout = deg(sources) (hybrid_bfs.cc:163-167); the network is consulted
    // only after a step (last_dir >= 0), so its operands are not loaded
    choose_dir(a, s, c, -1, 0, 0);
    wave_lds_sync();
    // publish as the state before step 0; the other copy starts zeroed
    uint64_t *d0 = (uint64_t *)&a.c[0];
    uint64_t *d1 = (uint64_t *)&a.c[1];
    for (int i = l; i < kCtlWords; i += 64) {
        d0[i] = ((const uint64_t *)&s.st)[i];
        d1[i] = 0;
    }
}

// clear() (hybrid_bfs.cc:433-443) for the device layout: visited = the
// degree-0 vertices, frontier bitmaps zero.
__global__ void k_clear(step_args a, int32_t *__restrict__ parent, int64_t nv,
                        const uint64_t *__restrict__ dead, uint64_t *__restrict__ b0,
                        uint64_t *__restrict__ b1, uint64_t *__restrict__ b2, int64_t nwords) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride)
        parent[i] = -1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) {
        a.vis[i] = dead[i];
        b0[i] = 0;
        b1[i] = 0;
        b2[i] = 0;
    }
}

// Candidate parents to -1.  A search leaves them so (each claim sweep resets
// what its write half wrote), so this runs only after a search that did not
// finish, and after allocation: clearing them with every search would push
// another V-sized array through the caches just before the timed steps.
__global__ void k_fill_cand(int32_t *__restrict__ cand, int64_t nv, uint8_t *__restrict__ dirty,
                            int64_t nwords) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * blockDim.x)
        cand[i] = -1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
         i += (int64_t)gridDim.x * blockDim.x)
        dirty[i] = 0;
}

// Profiling: undo the first direction's step at level L (its counters and
// the lengths of the list it filled) before timing the other direction.
__global__ __launch_bounds__(64) void k_reset_step(seg_counters *g, int32_t level) {
    const int l = threadIdx.x;
    for (int k = 0; k < 4; ++k) g->acc[level % 3][l][k] = 0;
    g->qlen[(level + 1) % 3][l][0] = 0;
    g->tlen[(level + 1) % 3][l][0] = 0;
    if (l == 0) {
        g->qlen[(level + 1) % 3][kSeg][0] = 0;
        g->tlen[(level + 1) % 3][kSeg][0] = 0;
    }
}

// count_num_traversed_edges (hybrid_bfs.cc:401-414): sum of degrees over
// reached vertices (one atomic per wave, spread over kSlots lines).
__global__ void k_count_traversed(const int32_t *__restrict__ parent,
                                  const int64_t *__restrict__ off, int64_t nv,
                                  unsigned long long *__restrict__ out) {
    unsigned long long s = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        if (parent[v] >= 0) s += (unsigned long long)(off[v + 1] - off[v]);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&out[(blockIdx.x % kSlots) * kLine], s);
}

// A search's result in graph ids: vertex v is search id inv[v], and a parent
// p (a search id) is graph vertex perm[p].  Either output may be null.
__global__ void k_to_graph_ids(const int32_t *__restrict__ inv, const int32_t *__restrict__ perm,
                               const int32_t *__restrict__ sparent,
                               const int32_t *__restrict__ sdepth, int64_t nv,
                               int32_t *__restrict__ parent, int32_t *__restrict__ depth) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t i = inv[v];
        if (parent) {
            const int32_t p = sparent[i];
            parent[v] = p < 0 ? -1 : perm[p];
        }
        if (depth) depth[v] = sdepth[i];
    }
}

// Depths from the parent tree.  The steps write parents only, as
// hybrid_bfs::run_beamer does; a vertex's depth is its distance to the
// source its parent chain ends at (a source is its own parent).
// t[v] = (k << 32) | x names an ancestor x of v, k levels up, starting from
// the parent (k = 1, or k = 0 for a source).  Each round of pointer jumping
// replaces (x, k) by (y, k + j) where t[x] = (j << 32) | y, until every x is
// a source.  Entries are single 8-byte words, read and written whole, so a
// jump that reads a word another thread has already advanced in this round
// still gets an ancestor and its distance; such reads only shorten the
// rounds, which are no more than with separate buffers (log2 of the depth).
// err[1] is set when a parent is itself unreached, err[2] when a chain
// closes on a vertex that is not a source.
constexpr uint64_t kNoTree = ~0ULL;

__global__ void k_tree_init(const int32_t *__restrict__ parent, uint64_t *__restrict__ t,
                            int64_t nv) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = parent[v];
        t[v] = p < 0 ? kNoTree : ((uint64_t)(p == v ? 0u : 1u) << 32) | (uint32_t)p;
    }
}

__global__ void k_tree_jump(uint64_t *t, int64_t nv, int32_t *__restrict__ err) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = __hip_atomic_load(t + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (x == kNoTree) continue;
        const uint32_t a = (uint32_t)x;
        if ((int64_t)a >= nv) {
            err[1] = 1;
            continue;
        }
        const uint64_t y = __hip_atomic_load(t + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (y == kNoTree) {
            err[1] = 1;
            continue;
        }
        if ((uint32_t)y == a) continue;  // a is a source (or closes a cycle)
        __hip_atomic_store(t + v, (((x >> 32) + (y >> 32)) << 32) | (uint32_t)y, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
        err[0] = 1;  // a jump was made: another round
    }
}

__global__ void k_tree_depth(const uint64_t *__restrict__ t, int32_t *__restrict__ depth,
                             int64_t nv, int32_t *__restrict__ err) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = t[v];
        if (x == kNoTree) {
            depth[v] = -1;
            continue;
        }
        if (t[(uint32_t)x] != (uint64_t)(uint32_t)x) err[2] = 1;  // ends at a non-source
        depth[v] = (int32_t)(x >> 32);
    }
}

// hybrid_bfs::check (hybrid_bfs.cc:281-379) on the device, given the depths
// of the parent tree (k_tree_*, which already failed any chain that does not
// end at a self-parent vertex).  The reference runs a textbook BFS and
// compares; here every reached vertex v is tested instead:
//   - only `source` is its own parent (the reference fails any other
//     self-parent: its parent is not one level up), and the source is reached;
//   - v's parent is in v's list (the tree's depths make it one level up);
//   - every neighbour u of v is reached and |depth u - depth v| <= 1.
// The tests pass exactly when the reference's do: the reached set then
// contains the source and is closed under adjacency, so it is the source's
// component, and along a shortest path depths rise by at most one per edge
// while the tree path to the source is depth v long, so depth v is the
// distance.  Conversely a tree the reference accepts has BFS depths, which
// pass every test.  err[0]: a parent is not a neighbour, err[1]: an edge
// spans two levels or leaves the reached set, err[2]: a second root, or the
// source unreached or not its own parent.  One wave per 64 vertices; lists
// of 64 or more are scanned by the whole wave.
__global__ __launch_bounds__(256) void k_check_tree(const int64_t *__restrict__ off,
                                                    const int32_t *__restrict__ cols,
                                                    const int32_t *__restrict__ parent,
                                                    const int32_t *__restrict__ depth, int64_t nv,
                                                    int64_t source, int32_t *__restrict__ err) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t base = wave * 64; base < nv; base += nwaves * 64) {
        const int64_t v = base + lane;
        const bool valid = v < nv;
        const int32_t d = valid ? depth[v] : -1;
        const int32_t p = valid ? parent[v] : -1;
        bool bad_root = false, bad_par = false, bad_edge = false;
        if (valid && v == source && (d != 0 || p != (int32_t)v)) bad_root = true;
        if (d >= 0 && p == (int32_t)v && v != source) bad_root = true;
        int64_t b = 0, e = 0;
        const bool scan = d >= 0;
        if (scan) {
            b = off[v];
            e = off[v + 1];
        }
        const bool alone = scan && e - b < 64;
        bool found = !scan || v == source;
        if (alone) {
            for (int64_t k = b; k < e; ++k) {
                const int32_t u = cols[k];
                const int32_t du = depth[u];
                found |= u == p;
                if (du < 0 || du > d + 1 || du < d - 1) bad_edge = true;
            }
        }
        uint64_t lm = __ballot(scan && !alone);
        while (lm) {
            const int src = __ffsll((unsigned long long)lm) - 1;
            lm &= lm - 1;
            const int64_t kb = lane_bcast(b, src), ke = lane_bcast(e, src);
            const int32_t ds = lane_bcast(d, src), ps = lane_bcast(p, src);
            bool f = false, be = false;
            for (int64_t k = kb + lane; k < ke; k += 64) {
                const int32_t u = cols[k];
                const int32_t du = depth[u];
                f |= u == ps;
                if (du < 0 || du > ds + 1 || du < ds - 1) be = true;
            }
            const bool fw = __ballot(f) != 0, bw = __ballot(be) != 0;
            if (lane == src) {
                found |= fw;
                bad_edge |= bw;
            }
        }
        bad_par = scan && !found;
        if (__ballot(bad_par) && lane == 0) err[0] = 1;
        if (__ballot(bad_edge) && lane == 0) err[1] = 1;
        if (__ballot(bad_root) && lane == 0) err[2] = 1;
    }
}

// Head index of the bottom-up step: head[v] = (deg(v) << 32) | the
// neighbour of v of largest degree, the first in CSR order among equals
// (degrees saturated at 2^31-1), or 0xffffffff when v has none.  One 8-byte
// load gives the bottom-up step both the likeliest parent and the degree.  A vertex's highest-degree neighbour is the
// one most likely to be reached first, so it is the likeliest parent when v
// is scanned bottom-up; a dense array of them is read in vertex order, where
// the first entry of each adjacency list costs a cache line per vertex.
__device__ __forceinline__ unsigned long long head_key(const int64_t *off, int32_t u,
                                                       int64_t pos) {
    int64_t d = off[u + 1] - off[u];
    d = d < 0x7fffffff ? d : 0x7fffffff;
    return ((unsigned long long)d << 32) | (0xffffffffULL - (unsigned long long)pos);
}

// One wave per 64 vertices: a lane does its own vertex when the list is
// shorter than 64, longer lists are done by the whole wave.
__global__ __launch_bounds__(256) void k_head_index(const int64_t *__restrict__ off,
                                                    const int32_t *__restrict__ cols, int64_t nv,
                                                    uint64_t *__restrict__ head,
                                                    unsigned long long *__restrict__ max_deg) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long dmax = 0;
    for (int64_t base = wave * 64; base < nv; base += nwaves * 64) {
        const int64_t v = base + lane;
        const bool valid = v < nv;
        int64_t b = 0, e = 0;
        if (valid) {
            b = off[v];
            e = off[v + 1];
            dmax = (unsigned long long)(e - b) > dmax ? (unsigned long long)(e - b) : dmax;
        }
        const bool alone = valid && e - b < 64;
        if (alone) {
            unsigned long long best = 0;
            for (int64_t k = b; k < e; ++k) {
                const unsigned long long key = head_key(off, cols[k], k - b);
                best = key > best ? key : best;
            }
            const uint32_t id =
                best ? (uint32_t)cols[b + (int64_t)(0xffffffffULL - (best & 0xffffffffULL))]
                     : 0xffffffffu;
            head[v] = ((uint64_t)(e - b) << 32) | id;
        }
        // lists of kHubBigDeg or more: k_head_big and k_head_big_done
        const bool big = valid && e - b >= kHubBigDeg;
        if (big) head[v] = 0;
        uint64_t lm = __ballot(valid && !alone && !big);
        while (lm) {
            const int src = __ffsll((unsigned long long)lm) - 1;
            lm &= lm - 1;
            const int64_t kb = lane_bcast(b, src), ke = lane_bcast(e, src);
            unsigned long long best = 0;
            for (int64_t k = kb + lane; k < ke; k += 64) {
                const unsigned long long key = head_key(off, cols[k], k - kb);
                best = key > best ? key : best;
            }
            best = wave_max(best);
            if (lane == src)
                head[v] = ((uint64_t)(ke - kb) << 32) |
                          (uint32_t)cols[kb + (int64_t)(0xffffffffULL - (best & 0xffffffffULL))];
        }
    }
    dmax = wave_max(dmax);
    if (lane == 0 && dmax) atomicMax(max_deg, dmax);
}

// The lists of kHubBigDeg entries or more, one hub slice per wave: each
// slice's best key (degree, then first position) is combined into head[v]
// (cleared by k_head_index) with a 64-bit atomic maximum.
__global__ __launch_bounds__(256) void k_head_big(const int64_t *__restrict__ off,
                                                  const int32_t *__restrict__ cols,
                                                  const td_task *__restrict__ big, int64_t nbig,
                                                  uint64_t *__restrict__ head) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t t = wave; t < nbig; t += nwaves) {
        const td_task tk = big[t];
        const int64_t b = off[tk.src];
        unsigned long long best = 0;
        for (int64_t k = tk.beg + lane; k < tk.end; k += 64) {
            const unsigned long long key = head_key(off, cols[k], k - b);
            best = key > best ? key : best;
        }
        best = wave_max(best);
        if (lane == 0) atomicMax((unsigned long long *)&head[tk.src], best);
    }
}

// Then each such vertex's key becomes its entry: degree and neighbour id
// (the slice that starts its list does it).
__global__ void k_head_big_done(const int64_t *__restrict__ off, const int32_t *__restrict__ cols,
                                const td_task *__restrict__ big, int64_t nbig,
                                uint64_t *__restrict__ head) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nbig;
         t += (int64_t)gridDim.x * blockDim.x) {
        const td_task tk = big[t];
        const int64_t b = off[tk.src], e = off[tk.src + 1];
        if (tk.beg != b) continue;
        const unsigned long long best = head[tk.src];
        head[tk.src] = ((uint64_t)(e - b) << 32) |
                       (uint32_t)cols[b + (int64_t)(0xffffffffULL - (best & 0xffffffffULL))];
    }
}

// Streams a buffer larger than the last-level cache so a timed step starts
// from cold caches.
__global__ void k_scrub(const uint4 *__restrict__ p, int64_t n, uint4 *__restrict__ sink) {
    uint4 acc = {0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint4 x = p[i];
        acc.x ^= x.x;
        acc.y ^= x.y;
    }
    if (acc.x == 0x9E3779B9u && acc.y == 0x7F4A7C15u) sink[0] = acc;
}

int norm_block(int b) { return (b == 128 || b == 512) ? b : 256; }

step_args make_args(bd_ctx *c, int32_t level, int64_t alpha, int64_t beta, int policy) {
    step_args a;
    a.off = c->d_soff;
    a.cols = c->d_scols;
    a.head = c->d_head;
    a.inv = c->d_inv;
    a.vis = c->d_vis;
    a.front = c->d_bm[level % 3];
    a.next_bm = c->d_bm[(level + 1) % 3];
    a.clr = c->d_bm[(level + 2) % 3];
    a.qin = c->d_q[level & 1];
    a.qout = c->d_q[(level + 1) & 1];
    a.tin = c->d_tasks[level & 1];
    a.tout = c->d_tasks[(level + 1) & 1];
    a.parent = c->d_parent;
    a.cand = c->d_cand;
    a.dirty = c->d_dirty;
    a.hub_bm = c->d_hub_bm;
    a.hub_med = c->d_hub_med;
    a.hub_big = c->d_hub_big;
    a.n_hub_med = c->n_hub_med;
    a.n_hub_big = c->n_hub_big;
    a.c = c->d_ctl;
    a.seg = c->d_seg;
    a.log = c->d_log;
    a.script = c->d_script;
    a.mlp = c->d_mlp;
    a.mlp_ops = c->d_mlp_ops;
    a.nwords = c->nwords;
    // vertices past nv_active have no edges when the layout is renumbered:
    // clear() sets their visited bits, no step finds them, and only an
    // isolated source (whose search ends after one step) is ever in a
    // frontier, so no sweep reads or clears their words
    a.live_nwords = c->d_perm ? std::max<int64_t>(1, (c->nv_active + 63) / 64) : c->nwords;
    a.task_cap = c->task_cap;
    a.qseg_cap = c->qseg_cap;
    a.tseg_cap = c->tseg_cap;
    a.nv = c->nv;
    a.m = c->m;
    a.nnz = c->nnz;
    a.alpha = alpha;
    a.beta = beta;
    a.nv_beta = c->nv / beta;
    a.td_split = c->split_off ? 0 : (int64_t)c->tune.td_split_edges;
    a.max_deg = c->max_deg;
    a.level = level;
    a.close_only = 0;
    a.force_dir = -1;
    a.wave_t = std::max(1, c->tune.wave_threshold);
    a.task_t = std::max(16, c->tune.task_threshold);
    a.task_edges = std::max(64, c->tune.task_edges);
    a.bu_words = c->tune.bu_words > 0 ? std::min(kBuChunkWords, c->tune.bu_words) : kBuChunkWords;
    a.bu_cap = c->tune.bu_cap > 0 ? std::min(c->tune.bu_cap, 1 << 16) : 32;
    a.bu_head = c->tune.bu_head == 2 ? 0 : 1;
    a.bu_batch = c->tune.bu_batch;
    a.policy = policy;
    a.script_len = (int32_t)c->script_len;
    a.log_cap = (int32_t)c->log_cap;
    a.mlp_h1 = c->mlp_h1;
    a.mlp_h2 = c->mlp_h2;
    // The network was trained on graphs with hubs; on a graph whose largest
    // degree is under learned_min_skew times the mean, LEARNED runs Beamer.
    const bool skewed = (double)c->max_deg * (double)c->nv >=
                        (double)std::max(1, c->tune.learned_min_skew) * (double)c->nnz;
    a.has_mlp = (c->has_mlp && skewed) ? 1 : 0;
    a.mlp_margin = (float)c->tune.learned_margin_milli * 1e-3f;
    a.mlp_near = std::max(1, c->tune.learned_near);
    return a;
}

int read_ctl(bd_ctx *c);

template <int B, int D>
void launch_one(bd_ctx *c, const step_args &a, int64_t grid, hipEvent_t e0, hipEvent_t e1) {
    if (e0 || e1)  // the events are stamped by the dispatch itself: kernel time only
        hipExtLaunchKernelGGL((k_step<B, D>), dim3(grid), dim3(B), 0, c->stream, e0, e1, 0, a);
    else
        hipLaunchKernelGGL((k_step<B, D>), dim3(grid), dim3(B), 0, c->stream, a);
}

template <int B>
void launch_step_block(bd_ctx *c, const step_args &a, int dirs, int64_t grid, hipEvent_t e0,
                       hipEvent_t e1) {
    if (dirs == kStepTD)
        launch_one<B, kStepTD>(c, a, grid, e0, e1);
    else if (dirs == kStepBU)
        launch_one<B, kStepBU>(c, a, grid, e0, e1);
    else
        launch_one<B, kStepBoth>(c, a, grid, e0, e1);
}

// Launch the step kernel carrying `dirs` (kStepTD, kStepBU or kStepBoth);
// grid <= 0 takes the tuned workgroups per CU of that kernel.  With events,
// the launch stamps e0 at the kernel's start and e1 at its end.
int launch_step(bd_ctx *c, const step_args &a, int dirs, int64_t grid = 0,
                hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const int blk = norm_block(c->tune.block);
    if (grid <= 0) {
        // the bottom-up grid is its own only when it has its own kernel
        const int per_cu = (dirs == kStepBU && c->tune.step_kernels == 1) ? c->tune.bu_grid_per_cu
                                                                         : c->tune.grid_per_cu;
        grid = (int64_t)c->cu_count * std::max(1, per_cu);
    }
    switch (blk) {
    case 128:
        launch_step_block<128>(c, a, dirs, grid, e0, e1);
        break;
    case 512:
        launch_step_block<512>(c, a, dirs, grid, e0, e1);
        break;
    default:
        launch_step_block<256>(c, a, dirs, grid, e0, e1);
    }
    BD_HIP(c, hipGetLastError());
    return BD_OK;
}

// All work of level L in device-driven order: the kernel of each direction
// (the one the device did not choose returns after its state rebuild), or
// the kernel carrying both.
int launch_level(bd_ctx *c, const step_args &a, hipEvent_t e0 = nullptr,
                 hipEvent_t e1 = nullptr) {
    int rc;
    if (c->tune.step_kernels == 2) return launch_step(c, a, kStepBoth, 0, e0, e1);
    if ((rc = launch_step(c, a, kStepBU, 0, e0, nullptr))) return rc;
    return launch_step(c, a, kStepTD, 0, nullptr, e1);
}

// Timing events for the launches of one chunk of device-driven levels.
int ensure_event_pool(bd_ctx *c, size_t n) {
    while (c->ev_pool.size() < n) {
        hipEvent_t e = nullptr;
        BD_HIP(c, hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
        c->ev_pool.push_back(e);
    }
    return BD_OK;
}

// Host-driven levels: publish the state before level L with one workgroup
// (level 0 was published by k_init) and read it back.
int close_level(bd_ctx *c, const step_args &a) {
    int rc;
    if (a.level > 0) {
        step_args ca = a;
        ca.close_only = 1;
        if ((rc = launch_step(c, ca, kStepBoth, 1))) return rc;
    }
    return read_ctl(c);
}

int launch_init(bd_ctx *c, const step_args &a, int64_t source, const int64_t *srcs,
                int64_t nsrc, hipEvent_t stop = nullptr) {
    if (stop)  // stamped when k_init ends, i.e. when the first step starts
        hipExtLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, c->stream, nullptr, stop, 0, a, source,
                              srcs, nsrc);
    else
        hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, c->stream, a, source, srcs, nsrc);
    BD_HIP(c, hipGetLastError());
    return BD_OK;
}

int check_run_args(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy) {
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph");
    if (source < 0 || source >= c->nv) return set_err(c, BD_ERANGE, "Source vertex out of range.");
    if (alpha <= 0) return set_err(c, BD_EINVAL, "alpha must be > 0");
    if (beta <= 0) return set_err(c, BD_EINVAL, "beta must be > 0");
    if (policy < BD_POLICY_BEAMER || policy > BD_POLICY_REMOTE_WRITES_HYBRID)
        return set_err(c, BD_EINVAL, "unknown policy");
    if (policy == BD_POLICY_SCRIPT && c->script_len <= 0)
        return set_err(c, BD_ESTATE, "no script set");
    return BD_OK;
}

int start_search(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy,
                 const int64_t *srcs = nullptr, int64_t nsrc = 1, hipEvent_t init_end = nullptr) {
    const int64_t n = std::max(c->nv, c->nwords);
    const step_args ca = make_args(c, 0, alpha, beta, policy);
    hipLaunchKernelGGL(k_clear, dim3(grid_for(c, n, 256)), dim3(256), 0, c->stream, ca,
                       c->d_parent, c->nv, c->d_sdead, c->d_bm[0], c->d_bm[1], c->d_bm[2],
                       c->nwords);
    if (!c->cand_clean)
        hipLaunchKernelGGL(k_fill_cand, dim3(grid_for(c, c->nv, 256)), dim3(256), 0, c->stream,
                           c->d_cand, c->nv, c->d_dirty, c->nwords);
    c->cand_clean = false;  // until this search finishes
    c->depth_valid = false;
    BD_HIP(c, hipGetLastError());
    // frontier 0: bitmap bm[0], queue q[0], tasks tasks[0]
    step_args a = make_args(c, 0, alpha, beta, policy);
    a.next_bm = c->d_bm[0];
    a.qout = c->d_q[0];
    a.tout = c->d_tasks[0];
    return launch_init(c, a, source, srcs, nsrc, init_end);
}

// Copy both published states to the pinned mirror and wait.
int read_ctl(bd_ctx *c) {
    BD_HIP(c, hipMemcpyAsync(c->h_ctl, c->d_ctl, 2 * sizeof(ctl_block), hipMemcpyDeviceToHost,
                             c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    return BD_OK;
}

// The final state: once the search is done both copies hold it.
const ctl_block *final_ctl(const bd_ctx *c) {
    return c->h_ctl[0].done ? &c->h_ctl[0] : (c->h_ctl[1].done ? &c->h_ctl[1] : nullptr);
}

int finish_search(bd_ctx *c, int64_t source, bd_run_stats *out, double event_ms) {
    const ctl_block *hp = final_ctl(c);
    if (!hp) return set_err(c, BD_EHIP, "search state not final");
    const ctl_block &h = *hp;
    struct {
        unsigned long long task_overflow;
        uint64_t t_start, t_end;
    } tail;
    BD_HIP(c, hipMemcpy(&tail, &c->d_seg->task_overflow, sizeof(tail), hipMemcpyDeviceToHost));
    if (tail.task_overflow) return set_err(c, BD_EHIP, "task list overflow");
    c->last_steps = h.steps;
    c->last_source = source;
    c->have_result = true;
    c->cand_clean = true;  // every write half was followed by its claim sweep
    const int64_t n = std::min<int64_t>(h.steps, c->log_cap);
    c->last_log.resize(n);
    if (n)
        BD_HIP(c, hipMemcpy(c->last_log.data(), c->d_log, sizeof(bd_level_rec) * n,
                            hipMemcpyDeviceToHost));
    if (out) {
        out->source = source;
        out->steps = h.steps;
        out->td_steps = h.td_steps;
        out->bu_steps = h.bu_steps;
        out->reached = h.reached;
        out->traversed_edges = h.reached_deg / 2;
        const double dev_ms = (double)(tail.t_end - tail.t_start) / (double)c->wall_khz;
        out->ms = (tail.t_end >= tail.t_start && dev_ms > 0) ? dev_ms : event_ms;
    }
    return BD_OK;
}

}  // namespace

int build_head_index(bd_ctx *c) {
    if (c->d_head || c->nv == 0) return BD_OK;
    int rc = dev_alloc(c, (void **)&c->d_head, sizeof(uint64_t) * c->nv, "head index");
    if (rc) return rc;
    unsigned long long *d_max = nullptr;
    if ((rc = dev_alloc(c, (void **)&d_max, sizeof(unsigned long long), "max degree"))) return rc;
    struct guard {
        void *p;
        ~guard() { dev_free(p); }
    } g{d_max};
    BD_HIP(c, hipMemsetAsync(d_max, 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_head_index, dim3(grid_for(c, c->nv, 256)), dim3(256), 0, c->stream,
                       c->d_soff, c->d_scols, c->nv, c->d_head, d_max);
    if (c->n_hub_big) {
        hipLaunchKernelGGL(k_head_big, dim3(grid_for(c, c->n_hub_big * 64, 256)), dim3(256), 0,
                           c->stream, c->d_soff, c->d_scols, c->d_hub_big, c->n_hub_big,
                           c->d_head);
        hipLaunchKernelGGL(k_head_big_done, dim3(grid_for(c, c->n_hub_big, 256)), dim3(256), 0,
                           c->stream, c->d_soff, c->d_scols, c->d_hub_big, c->n_hub_big,
                           c->d_head);
    }
    BD_HIP(c, hipGetLastError());
    unsigned long long h = 0;
    BD_HIP(c, hipMemcpyAsync(&h, d_max, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    c->max_deg = (int64_t)h;
    return BD_OK;
}

// A search is k_init, then one k_step launch per level.  Step L publishes
// the state before it (closing step L-1), so the launch that finds the
// frontier empty marks the search done and later launches return at once.
// With srcs (device, nsrc entries) the search starts from all of them.
int run_search(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy,
               bd_run_stats *out, const int64_t *srcs = nullptr, int64_t nsrc = 1) {
    int rc;
    if ((rc = check_run_args(c, source, alpha, beta, policy))) return rc;
    BD_HIP(c, hipSetDevice(c->device));
    const bool per_chunk = c->tune.time_steps == 4;
    if (per_chunk && (rc = ensure_event_pool(c, 2 * kMaxLevelsPerSync))) return rc;
    // time_steps 4: the first chunk's span starts where k_init ends
    if ((rc = start_search(c, source, alpha, beta, policy, srcs, nsrc,
                           per_chunk ? c->ev_pool[0] : nullptr)))
        return rc;
    BD_HIP(c, hipEventRecord(c->ev[0], c->stream));
    // launches: a level takes one, or two when it writes remotely
    const int64_t max_levels = 2 * c->nv + 4;
    int32_t L = 0;
    const bool timed = c->tune.time_steps == 1;
    std::vector<double> step_ms;
    if (c->tune.level_driver == 2 || timed) {
        for (;; ++L) {
            if (L > max_levels) return set_err(c, BD_EHIP, "search did not terminate");
            const step_args a = make_args(c, L, alpha, beta, policy);
            if ((rc = close_level(c, a))) return rc;
            if (c->h_ctl[L & 1].done) break;
            // the host knows the direction: launch the kernel that runs it
            const int dirs = kernel_dir(c->h_ctl[L & 1].dir);
            if ((rc = launch_step(c, a, dirs, 0, timed ? c->ev[2] : nullptr,
                                  timed ? c->ev[3] : nullptr)))
                return rc;
            if (timed) {
                BD_HIP(c, hipEventSynchronize(c->ev[3]));
                float ms = 0;
                BD_HIP(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
                step_ms.push_back(ms);
            }
        }
    } else {
        // time_steps 3: every launch carries kernel-stamped events;
        // time_steps 4: one event pair spans each chunk of launches, and each
        // launch of the chunk is given the chunk's mean
        const bool per_launch = c->tune.time_steps == 3;
        if (per_launch && (rc = ensure_event_pool(c, 2 * kMaxLevelsPerSync))) return rc;
        c->last_launch_ms.clear();
        // the first chunk: as many launches as the last search on this
        // graph took, and a few more (16 for the first search); launches
        // past the end return at once but still occupy the device
        int chunk = c->chunk_hint > 0 ? c->chunk_hint : 16;
        for (;;) {
            for (int k = 0; k < chunk; ++k, ++L) {
                const step_args a = make_args(c, L, alpha, beta, policy);
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (per_launch) {
                    e0 = c->ev_pool[2 * k];
                    e1 = c->ev_pool[2 * k + 1];
                } else if (per_chunk) {
                    // later chunks start after a host wait: their first
                    // launch stamps the start
                    e0 = (k == 0 && L > 0) ? c->ev_pool[0] : nullptr;
                    e1 = k == chunk - 1 ? c->ev_pool[1] : nullptr;
                }
                if ((rc = launch_level(c, a, e0, e1))) return rc;
            }
            if ((rc = read_ctl(c))) return rc;
            for (int k = 0; per_launch && k < chunk; ++k) {
                float ms = 0;
                BD_HIP(c, hipEventElapsedTime(&ms, c->ev_pool[2 * k], c->ev_pool[2 * k + 1]));
                c->last_launch_ms.push_back(ms);
            }
            if (per_chunk) {
                float ms = 0;
                BD_HIP(c, hipEventElapsedTime(&ms, c->ev_pool[0], c->ev_pool[1]));
                for (int k = 0; k < chunk; ++k) c->last_launch_ms.push_back(ms / chunk);
            }
            if (final_ctl(c)) break;
            if (L > max_levels) return set_err(c, BD_EHIP, "search did not terminate");
            chunk = std::min(chunk * 2, kMaxLevelsPerSync);
        }
    }
    BD_HIP(c, hipEventRecord(c->ev[1], c->stream));
    BD_HIP(c, hipEventSynchronize(c->ev[1]));
    float ev_ms = 0;
    BD_HIP(c, hipEventElapsedTime(&ev_ms, c->ev[0], c->ev[1]));
    if ((rc = read_ctl(c))) return rc;
    if ((rc = finish_search(c, source, out, ev_ms))) return rc;
    if (!c->last_log.empty() && (int64_t)c->last_log.size() == c->last_steps) {
        const bd_level_rec &r = c->last_log.back();
        // its level launches and the launch that found the frontier empty
        const int64_t used = r.launch + r.launches + 1;
        c->chunk_hint = (int)std::min<int64_t>(std::max<int64_t>(used + 3, 4), kMaxLevelsPerSync);
    }
    if (c->tune.time_steps == 3) step_ms = c->last_launch_ms;  // per-launch times only
    for (auto &r : c->last_log) {
        if (r.launch < 0 || r.launch + r.launches > (int64_t)step_ms.size()) continue;
        r.ms = 0;
        for (int64_t k = r.launch; k < r.launch + r.launches; ++k) r.ms += step_ms[k];
    }
    return BD_OK;
}

}  // namespace bd

using namespace bd;

extern "C" int bd_bfs_run(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy,
                          bd_run_stats *out) {
    if (!c) return BD_EINVAL;
    return run_search(c, source, alpha, beta, policy, out);
}

extern "C" int bd_bfs_set_script(bd_ctx *c, const int32_t *dirs, int64_t n) {
    if (!c || n < 0 || (n > 0 && !dirs)) return BD_EINVAL;
    BD_HIP(c, hipSetDevice(c->device));
    dev_free(c->d_script);
    c->d_script = nullptr;
    c->script_len = 0;
    if (n == 0) return BD_OK;
    int rc = dev_alloc(c, (void **)&c->d_script, sizeof(int32_t) * n, "script");
    if (rc) return rc;
    BD_HIP(c, hipMemcpy(c->d_script, dirs, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    c->script.assign(dirs, dirs + n);
    c->script_len = n;
    return BD_OK;
}

// Depths of the parent tree `parent` (nv entries, -1 unreached) into
// `depth` (k_tree_*).  *why is null when the tree is well formed, else the
// reason it is not: a reached vertex with an unreached or out-of-range
// parent, a chain that never ends (a cycle), or one that ends at a vertex
// that is not its own parent.  Returns a status only for runtime errors.
int tree_depths(bd_ctx *c, const int32_t *parent, int32_t *depth, const char **why) {
    *why = nullptr;
    const int64_t nv = c->nv;
    uint64_t *t = nullptr;
    // nv tree words, then err[4]
    int rc = dev_alloc(c, (void **)&t, sizeof(uint64_t) * (nv + 2), "parent tree");
    if (rc) return rc;
    struct guard {
        uint64_t *p;
        ~guard() { dev_free(p); }
    } g{t};
    int32_t *err = (int32_t *)(t + nv);
    const dim3 grid(grid_for(c, nv, 256)), block(256);
    hipLaunchKernelGGL(k_tree_init, grid, block, 0, c->stream, parent, t, nv);
    BD_HIP(c, hipGetLastError());
    int32_t h[4] = {1, 0, 0, 0};
    // a chain of length D takes ceil(log2 D) + 1 rounds; ids are 31 bits
    for (int r = 0; r < 40 && h[0]; ++r) {
        BD_HIP(c, hipMemsetAsync(err, 0, sizeof(h), c->stream));
        hipLaunchKernelGGL(k_tree_jump, grid, block, 0, c->stream, t, nv, err);
        BD_HIP(c, hipGetLastError());
        BD_HIP(c, hipMemcpyAsync(h, err, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        if (h[1]) {
            *why = "parent tree: a reached vertex has an unreached parent";
            return BD_OK;
        }
    }
    if (h[0]) {
        *why = "parent tree: chains do not end";
        return BD_OK;
    }
    hipLaunchKernelGGL(k_tree_depth, grid, block, 0, c->stream, t, depth, nv, err);
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipMemcpyAsync(h, err, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    if (h[2]) *why = "parent tree: a chain closes on a non-source";
    return BD_OK;
}

// Depths of the last search into d_depth, from its parent tree.
int derive_depth(bd_ctx *c) {
    if (c->depth_valid) return BD_OK;
    const char *why = nullptr;
    int rc = tree_depths(c, c->d_parent, c->d_depth, &why);
    if (rc) return rc;
    if (why) return set_err(c, BD_EHIP, why);
    c->depth_valid = true;
    return BD_OK;
}

// check() of the parent tree `parent` from `source` on the device; `depth`
// holds its depths when have_depth, else receives them (scratch of nv
// entries when the tree is not the last search's).  *ok = 1 PASS, 0 FAIL.
int check_tree(bd_ctx *c, const int32_t *parent, int32_t *depth, bool have_depth,
               int64_t source, int *ok) {
    int rc;
    if (!have_depth) {
        const char *why = nullptr;
        if ((rc = tree_depths(c, parent, depth, &why))) return rc;
        if (why) {
            *ok = 0;
            return BD_OK;
        }
    }
    int32_t *err = nullptr;
    if ((rc = dev_alloc(c, (void **)&err, sizeof(int32_t) * 4, "check flags"))) return rc;
    struct guard {
        int32_t *p;
        ~guard() { dev_free(p); }
    } g{err};
    int32_t h[4] = {0, 0, 0, 0};
    BD_HIP(c, hipMemsetAsync(err, 0, sizeof(h), c->stream));
    hipLaunchKernelGGL(k_check_tree, dim3(grid_for(c, c->nv, 256)), dim3(256), 0, c->stream,
                       c->d_off, c->d_cols, parent, depth, c->nv, source, err);
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipMemcpyAsync(h, err, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    *ok = (h[0] | h[1] | h[2]) ? 0 : 1;
    return BD_OK;
}

// The last search's parents (and, with want_depth, depths) in graph ids on
// the device: the search's own arrays when it ran on graph ids, else mapped
// into `buf` (allocated here, 2 nv entries; the caller frees it).
int graph_id_result(bd_ctx *c, bool want_depth, int32_t **buf, const int32_t **parent,
                    const int32_t **depth) {
    int rc;
    *buf = nullptr;
    if (want_depth && (rc = derive_depth(c))) return rc;
    if (!c->d_perm) {
        *parent = c->d_parent;
        *depth = c->d_depth;
        return BD_OK;
    }
    if ((rc = dev_alloc(c, (void **)buf, sizeof(int32_t) * 2 * c->nv, "results in graph ids")))
        return rc;
    hipLaunchKernelGGL(k_to_graph_ids, dim3(grid_for(c, c->nv, 256)), dim3(256), 0, c->stream,
                       c->d_inv, c->d_perm, c->d_parent, c->d_depth, c->nv, *buf,
                       want_depth ? *buf + c->nv : nullptr);
    BD_HIP(c, hipGetLastError());
    *parent = *buf;
    *depth = *buf + c->nv;
    return BD_OK;
}

extern "C" int bd_bfs_fetch(bd_ctx *c, int32_t *depth, int64_t *parent) {
    if (!c) return BD_EINVAL;
    if (!c->have_result) return set_err(c, BD_ESTATE, "no search has run");
    if (!depth && !parent) return BD_OK;
    BD_HIP(c, hipSetDevice(c->device));
    int32_t *buf = nullptr;
    const int32_t *dp = nullptr, *dd = nullptr;
    int rc = graph_id_result(c, depth != nullptr, &buf, &dp, &dd);
    struct guard {
        int32_t *p;
        ~guard() { dev_free(p); }
    } g{buf};
    if (rc) return rc;
    if (depth)
        BD_HIP(c, hipMemcpyAsync(depth, dd, sizeof(int32_t) * c->nv, hipMemcpyDeviceToHost,
                                 c->stream));
    std::vector<int32_t> p;
    if (parent) {
        p.resize(c->nv);
        BD_HIP(c, hipMemcpyAsync(p.data(), dp, sizeof(int32_t) * c->nv, hipMemcpyDeviceToHost,
                                 c->stream));
    }
    BD_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t v = 0; parent && v < c->nv; ++v) parent[v] = p[v];
    return BD_OK;
}

extern "C" int bd_bfs_levels(bd_ctx *c, bd_level_rec *recs, int64_t cap, int64_t *n) {
    if (!c || !n || cap < 0 || (cap > 0 && !recs)) return BD_EINVAL;
    const int64_t k = std::min<int64_t>(cap, (int64_t)c->last_log.size());
    for (int64_t i = 0; i < k; ++i) recs[i] = c->last_log[i];
    *n = (int64_t)c->last_log.size();
    return BD_OK;
}

extern "C" int bd_bfs_count_traversed(bd_ctx *c, int64_t *out) {
    if (!c || !out) return BD_EINVAL;
    if (!c->have_result) return set_err(c, BD_ESTATE, "no search has run");
    BD_HIP(c, hipSetDevice(c->device));
    unsigned long long *d = nullptr;
    const size_t bytes = sizeof(unsigned long long) * kSlots * kLine;
    int rc = dev_alloc(c, (void **)&d, bytes, "count");
    if (rc) return rc;
    std::vector<unsigned long long> h(kSlots * kLine, 0);
    hipError_t e = hipMemsetAsync(d, 0, bytes, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_count_traversed, dim3(grid_for(c, c->nv, 256)), dim3(256), 0,
                           c->stream, c->d_parent, c->d_soff, c->nv, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(d);
    if (e != hipSuccess) return hip_err(c, e, "count_traversed");
    unsigned long long sum = 0;
    for (int s = 0; s < kSlots; ++s) sum += h[s * kLine];
    *out = (int64_t)(sum / 2);
    return BD_OK;
}

// hybrid_bfs::check (hybrid_bfs.cc:281-379) of the last search, on the
// device (k_check_tree: the same PASS/FAIL as the reference's textbook-BFS
// comparison, without a host copy of the graph).
extern "C" int bd_bfs_check(bd_ctx *c, int64_t source, int *ok) {
    if (!c || !ok) return BD_EINVAL;
    if (!c->have_result) return set_err(c, BD_ESTATE, "no search has run");
    if (source < 0 || source >= c->nv) return set_err(c, BD_ERANGE, "source out of range");
    BD_HIP(c, hipSetDevice(c->device));
    if (c->d_perm) {
        // the tree as the caller reads it (graph ids), checked on the graph
        int32_t *buf = nullptr;
        const int32_t *dp = nullptr, *dd = nullptr;
        int rc = graph_id_result(c, false, &buf, &dp, &dd);
        struct guard {
            int32_t *p;
            ~guard() { dev_free(p); }
        } g{buf};
        if (rc) return rc;
        return check_tree(c, dp, buf + c->nv, false, source, ok);
    }
    if (!c->depth_valid) {
        // a malformed tree is a FAIL, not an error
        const char *why = nullptr;
        int rc = tree_depths(c, c->d_parent, c->d_depth, &why);
        if (rc) return rc;
        if (why) {
            *ok = 0;
            return BD_OK;
        }
        c->depth_valid = true;
    }
    return check_tree(c, c->d_parent, c->d_depth, true, source, ok);
}

// hybrid_bfs::check of a caller's parent array (nv entries, negative =
// unreached) against the resident graph, on the device.
extern "C" int bd_bfs_check_parents(bd_ctx *c, int64_t source, const int64_t *parent, int *ok) {
    if (!c || !ok || !parent) return BD_EINVAL;
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph");
    if (source < 0 || source >= c->nv) return set_err(c, BD_ERANGE, "source out of range");
    BD_HIP(c, hipSetDevice(c->device));
    const int64_t nv = c->nv;
    // out-of-range parents become nv, which the tree kernels reject
    std::vector<int32_t> p(nv);
    for (int64_t v = 0; v < nv; ++v)
        p[v] = parent[v] < 0 ? -1 : (parent[v] >= nv ? (int32_t)nv : (int32_t)parent[v]);
    int32_t *d = nullptr;
    int rc = dev_alloc(c, (void **)&d, sizeof(int32_t) * 2 * nv, "checked tree");
    if (rc) return rc;
    struct guard {
        int32_t *p;
        ~guard() { dev_free(p); }
    } g{d};
    BD_HIP(c, hipMemcpy(d, p.data(), sizeof(int32_t) * nv, hipMemcpyHostToDevice));
    return check_tree(c, d, d + nv, false, source, ok);
}

// Time both directions at every step from the same state; the search follows
// `policy`.  The direction not taken runs first; then the state it changed
// (vis, next-frontier counts) is restored and the chosen direction runs, so
// the search ends exactly as a normal run would.  Both directions must find
// the same number of vertices.
extern "C" int bd_bfs_profile(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy,
                              bd_run_stats *out) {
    if (!c) return BD_EINVAL;
    int rc;
    if ((rc = check_run_args(c, source, alpha, beta, policy))) return rc;
    if (policy == BD_POLICY_REMOTE_WRITES || policy == BD_POLICY_REMOTE_WRITES_HYBRID)
        return set_err(c, BD_EINVAL, "profile times one-launch steps: not for remote-write policies");
    BD_HIP(c, hipSetDevice(c->device));
    // each level is timed in both directions as one launch each: top-down
    // steps run without remote writes here
    c->split_off = true;
    struct split_reset {
        bd_ctx *c;
        ~split_reset() { c->split_off = false; }
    } sr{c};
    const int64_t scrub_n = ((int64_t)512 << 20) / 16;
    uint4 *scrub = nullptr;
    if ((rc = dev_alloc(c, (void **)&scrub, sizeof(uint4) * (scrub_n + 1), "scrub"))) return rc;
    struct guard {
        uint4 *p;
        ~guard() { dev_free(p); }
    } g{scrub};
    BD_HIP(c, hipMemsetAsync(scrub, 0, sizeof(uint4) * (scrub_n + 1), c->stream));
    if ((rc = start_search(c, source, alpha, beta, policy))) return rc;
    if ((rc = read_ctl(c))) return rc;
    std::vector<std::pair<double, double>> times;
    std::vector<unsigned long long> acc(kSlots * kLine);
    double total = 0;
    for (int32_t L = 0;; ++L) {
        if (L > c->nv + 2) return set_err(c, BD_EHIP, "search did not terminate");
        if (L > 0) {
            // one workgroup closes step L-1 and publishes the state before L
            if ((rc = close_level(c, make_args(c, L, alpha, beta, policy)))) return rc;
            if (c->h_ctl[L & 1].done) break;
        }
        const int32_t chosen = c->h_ctl[L & 1].dir;
        const int32_t other = 1 - chosen;
        float ms[2] = {0, 0};
        unsigned long long found[2] = {0, 0};
        for (int pass = 0; pass < 2; ++pass) {
            step_args a = make_args(c, L, alpha, beta, policy);
            a.force_dir = pass == 0 ? other : chosen;
            if (pass == 1) {
                BD_HIP(c, hipMemcpyAsync(c->d_vis, c->d_vis_save, sizeof(uint64_t) * c->nwords,
                                         hipMemcpyDeviceToDevice, c->stream));
                hipLaunchKernelGGL(k_reset_step, dim3(1), dim3(64), 0, c->stream, c->d_seg, L);
                BD_HIP(c, hipGetLastError());
            } else {
                BD_HIP(c, hipMemcpyAsync(c->d_vis_save, c->d_vis, sizeof(uint64_t) * c->nwords,
                                         hipMemcpyDeviceToDevice, c->stream));
            }
            hipLaunchKernelGGL(k_scrub, dim3(c->cu_count * 4), dim3(256), 0, c->stream, scrub,
                               scrub_n, scrub + scrub_n);
            if ((rc = launch_step(c, a, a.force_dir, 0, c->ev[2], c->ev[3]))) return rc;
            BD_HIP(c, hipEventSynchronize(c->ev[3]));
            BD_HIP(c, hipEventElapsedTime(&ms[pass], c->ev[2], c->ev[3]));
            BD_HIP(c, hipMemcpy(acc.data(), c->d_seg->acc[L % 3],
                                sizeof(unsigned long long) * acc.size(), hipMemcpyDeviceToHost));
            for (int s = 0; s < kSlots; ++s) found[pass] += acc[s * kLine];
        }
        if (found[0] != found[1])
            return set_err(c, BD_EHIP, "profile: directions disagree on the next frontier");
        times.push_back({ms[1], ms[0]});
        total += ms[1];
    }
    if ((rc = finish_search(c, source, out, total))) return rc;
    for (size_t i = 0; i < times.size() && i < c->last_log.size(); ++i) {
        c->last_log[i].ms = times[i].first;
        c->last_log[i].ms_other = times[i].second;
    }
    if (out) out->ms = total;
    return BD_OK;
}

// ------------------------------------------------------- many graphs at once

namespace bd {
namespace {

// Flags an edge whose ends lie in different graphs of the batch (graph g
// owns vertices [vbase[g], vbase[g+1])).
__global__ void k_check_parts(const int64_t *__restrict__ off, const int32_t *__restrict__ cols,
                              int64_t nv, const int64_t *__restrict__ vbase, int ng,
                              int *__restrict__ bad) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = ng;  // largest g with vbase[g] <= v
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (vbase[mid] <= v)
                lo = mid;
            else
                hi = mid;
        }
        const int64_t gb = vbase[lo], ge = vbase[lo + 1];
        for (int64_t k = off[v]; k < off[v + 1]; ++k) {
            const int64_t u = cols[k];
            if (u < gb || u >= ge) {
                *bad = 1;
                break;
            }
        }
    }
}

}  // namespace
}  // namespace bd

// Many disjoint graphs (the resident graph is their union) searched from
// their roots at once: one direction-optimizing search whose frontier 0
// holds every root.  As no edge joins two graphs, each vertex's depth is its
// distance from its own graph's root.  Beamer's rules (alpha 15, beta 18)
// run on the union.
extern "C" int bd_bfs_multi_graph(bd_ctx *c, int ngraphs, const int64_t *vbase,
                                  const int64_t *roots, int32_t *depths, double *ms) {
    if (!c || ngraphs < 1 || !vbase || !roots) return BD_EINVAL;
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph");
    if (vbase[0] != 0 || vbase[ngraphs] != c->nv)
        return set_err(c, BD_EINVAL, "vbase must run from 0 to the number of vertices");
    std::vector<int64_t> src(ngraphs);
    for (int g = 0; g < ngraphs; ++g) {
        if (vbase[g + 1] <= vbase[g]) return set_err(c, BD_EINVAL, "empty or unordered graph");
        if (roots[g] < 0 || roots[g] >= vbase[g + 1] - vbase[g])
            return set_err(c, BD_ERANGE, "root out of range of its graph");
        src[g] = vbase[g] + roots[g];
    }
    BD_HIP(c, hipSetDevice(c->device));
    int rc;
    int64_t *d = nullptr;
    if ((rc = dev_alloc(c, (void **)&d, sizeof(int64_t) * (2 * ngraphs + 1) + sizeof(int),
                        "batch sources")))
        return rc;
    struct guard {
        void *p;
        ~guard() { dev_free(p); }
    } gd{d};
    int64_t *d_src = d, *d_vbase = d + ngraphs;
    int *d_bad = (int *)(d + 2 * ngraphs + 1);
    BD_HIP(c, hipMemcpy(d_src, src.data(), sizeof(int64_t) * ngraphs, hipMemcpyHostToDevice));
    BD_HIP(c, hipMemcpy(d_vbase, vbase, sizeof(int64_t) * (ngraphs + 1), hipMemcpyHostToDevice));
    BD_HIP(c, hipMemset(d_bad, 0, sizeof(int)));
    hipLaunchKernelGGL(k_check_parts, dim3(grid_for(c, c->nv, 256)), dim3(256), 0, c->stream,
                       c->d_off, c->d_cols, c->nv, d_vbase, ngraphs, d_bad);
    BD_HIP(c, hipGetLastError());
    int bad = 0;
    BD_HIP(c, hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    if (bad) return set_err(c, BD_EINVAL, "an edge joins two graphs of the batch");
    bd_run_stats st;
    if ((rc = run_search(c, src[0], 15, 18, BD_POLICY_BEAMER, &st, d_src, ngraphs))) return rc;
    if (ms) *ms = st.ms;
    if (depths) return bd_bfs_fetch(c, depths, nullptr);
    return BD_OK;
}

// ---------------------------------------- many graphs at once, no allocation

// Staging of bd_bfs_multi_graph_static, in a file-scope device array instead
// of memory allocated per call: [0, n) the search's sources, [n, 2n] the
// graph bounds vbase, [2n + 1] the flag k_check_parts raises.  Every GPU the
// library runs on holds its own copy (kStaticGraphs * 16 + 16 bytes, 64 KB)
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives for that device (bd_ctx::d_static).
constexpr int kStaticGraphs = 4096;
static __device__ __align__(256) int64_t scratch[2 * kStaticGraphs + 2];

namespace bd {
namespace {

// sources = vbase[g] + roots[g] (roots were copied to [0, n)); clears the flag.
__global__ void k_static_sources(int64_t *st, int n) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x)
        st[g] += st[n + g];
    if (blockIdx.x == 0 && threadIdx.x == 0) st[2 * n + 1] = 0;
}

}  // namespace

// The device's address of `scratch`, with the context's device current.
int static_scratch_init(bd_ctx *c) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return hip_err(c, e, "hipGetSymbolAddress(scratch)");
    }
    c->d_static = (int64_t *)p;
    return BD_OK;
}

}  // namespace bd

// bd_bfs_multi_graph without its per-call device allocation (the staging of
// sources, bounds and the check flag lives in `scratch`, at most
// kStaticGraphs graphs) and without the depths output.  No counterpart in the
// reference; for callers that cannot allocate.  Calls overlapping on one GPU
// share the array and race.
extern "C" int bd_bfs_multi_graph_static(bd_ctx *c, int ngraphs, const int64_t *vbase,
                                         const int64_t *roots, double *ms) {
    if (!c || ngraphs < 1 || !vbase || !roots) return BD_EINVAL;
    if (ngraphs > kStaticGraphs)
        return set_err(c, BD_ERANGE,
                       "bd_bfs_multi_graph_static: at most 4096 graphs (its static staging)");
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph");
    if (!c->d_static) return set_err(c, BD_ESTATE, "no static staging on this device");
    if (vbase[0] != 0 || vbase[ngraphs] != c->nv)
        return set_err(c, BD_EINVAL, "vbase must run from 0 to the number of vertices");
    for (int g = 0; g < ngraphs; ++g) {
        if (vbase[g + 1] <= vbase[g]) return set_err(c, BD_EINVAL, "empty or unordered graph");
        if (roots[g] < 0 || roots[g] >= vbase[g + 1] - vbase[g])
            return set_err(c, BD_ERANGE, "root out of range of its graph");
    }
    BD_HIP(c, hipSetDevice(c->device));
    int64_t *st = c->d_static;
    BD_HIP(c, hipMemcpyAsync(st, roots, sizeof(int64_t) * ngraphs, hipMemcpyHostToDevice,
                             c->stream));
    BD_HIP(c, hipMemcpyAsync(st + ngraphs, vbase, sizeof(int64_t) * (ngraphs + 1),
                             hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_static_sources, dim3((ngraphs + 255) / 256), dim3(256), 0, c->stream, st,
                       ngraphs);
    int *d_bad = (int *)(st + 2 * ngraphs + 1);
    hipLaunchKernelGGL(k_check_parts, dim3(grid_for(c, c->nv, 256)), dim3(256), 0, c->stream,
                       c->d_off, c->d_cols, c->nv, st + ngraphs, ngraphs, d_bad);
    BD_HIP(c, hipGetLastError());
    int bad = 0;
    BD_HIP(c, hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    if (bad) return set_err(c, BD_EINVAL, "an edge joins two graphs of the batch");
    bd_run_stats rs;
    const int rc = run_search(c, roots[0], 15, 18, BD_POLICY_BEAMER, &rs, st, ngraphs);
    if (rc) return rc;
    if (ms) *ms = rs.ms;
    return BD_OK;
}

// Kernel time of every step launch of the last search run with time_steps 3
// (device-driven, one event pair per launch) or 4 (one pair per chunk of
// launches, each launch given the chunk's mean), including the launches
// queued after the search ended, which return at once.
extern "C" int bd_bfs_launch_times(bd_ctx *c, double *ms, int64_t cap, int64_t *n) {
    if (!c || !n || cap < 0 || (cap > 0 && !ms)) return BD_EINVAL;
    const int64_t k = std::min<int64_t>(cap, (int64_t)c->last_launch_ms.size());
    for (int64_t i = 0; i < k; ++i) ms[i] = c->last_launch_ms[i];
    *n = (int64_t)c->last_launch_ms.size();
    return BD_OK;
}

