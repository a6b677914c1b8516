// bd_bfs.hip -- direction-optimizing BFS on gfx950 (hybrid_bfs::run_beamer,
// hybrid_bfs.cc:152-191).
//
// Device state per search (V vertices, W = ceil(V/64) 64-bit words):
//   depth[V], parent[V]  int32     result arrays (-1 = unreached)
//   vis[W]               bitmap    visited; degree-0 vertices pre-set
//   bm[3][W]             bitmaps   frontier rotation: bm[L%3] = depth-L set,
//                                  bm[(L+1)%3] = next, bm[(L+2)%3] is zeroed
//                                  during step L for step L+1's use
//   q[2][V]              int32     frontier queues (top-down input/output)
//   tasks                          slices of huge adjacency lists
//   ctl                            control block (counters, policy state)
//
// One step = one BFS level.  Top-down: k_top_down expands F (from the queue,
// or from the bitmap after a bottom-up step), claims vertices with atomicOr
// on vis, and appends them to the next queue with one atomic per wave;
// k_top_down_tasks expands the slices of vertices whose degree exceeds the
// task threshold.  Bottom-up: k_bottom_up gives each wave one 64-vertex word
// of vis, so the wave owns that word and writes it (and the next-frontier
// word) with plain stores; each unvisited lane scans its neighbours in CSR
// order and stops at the first one whose bit is set in bm[L%3].  Reading the
// depth-L bitmap rather than vis keeps the reference's two-phase rule
// (hybrid_bfs.cc:117-141): a vertex found in this step is never seen as a
// parent in the same step.  k_decide closes the step, writes its record and
// picks the next direction on the device.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <queue>
#include <vector>

#include "bd_internal.h"
#include "bd_mlp.h"

namespace bd {

namespace {

constexpr int kMaxLevelsPerSync = 64;

struct td_args {
    const int64_t *off;
    const int32_t *cols;
    uint64_t *vis;
    const int32_t *qin;
    const uint64_t *front;
    uint64_t *next_bm;
    uint64_t *clr;
    int32_t *qout;
    int32_t *depth;
    int32_t *parent;
    td_task *tasks;
    ctl_block *c;
    int64_t nwords;
    int64_t task_cap;
    int32_t level;
    int32_t gate;
    int32_t wave_t;
    int32_t task_t;
    int32_t task_edges;
    int32_t pad;
};

struct bu_args {
    const int64_t *off;
    const int32_t *cols;
    uint64_t *vis;
    const uint64_t *front;
    uint64_t *next;
    uint64_t *clr;
    int32_t *depth;
    int32_t *parent;
    ctl_block *c;
    int64_t nwords;
    int32_t level;
    int32_t gate;
};

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += (unsigned long long)__shfl_xor((long long)x, o);
    return x;
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long y = (unsigned long long)__shfl_xor((long long)x, o);
        x = y > x ? y : x;
    }
    return x;
}

// Reduce four per-thread counters over the block; one atomic each per block.
template <int BLOCK>
__device__ void block_flush(ctl_block *c, unsigned long long cnt, unsigned long long dsum,
                            unsigned long long exam, unsigned long long dmax) {
    __shared__ unsigned long long red[4][BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    cnt = wave_sum(cnt);
    dsum = wave_sum(dsum);
    exam = wave_sum(exam);
    dmax = wave_max(dmax);
    if (lane == 0) {
        red[0][wid] = cnt;
        red[1][wid] = dsum;
        red[2][wid] = exam;
        red[3][wid] = dmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; ++w) {
            cnt += red[0][w];
            dsum += red[1][w];
            exam += red[2][w];
            dmax = red[3][w] > dmax ? red[3][w] : dmax;
        }
        if (cnt) atomicAdd(&c->acc_new, cnt);
        if (dsum) atomicAdd(&c->acc_new_deg, dsum);
        if (exam) atomicAdd(&c->acc_examined, exam);
        if (dmax) atomicMax(&c->acc_max_deg, dmax);
    }
}

__device__ __forceinline__ bool step_gated_out(const ctl_block *c, int32_t level, int32_t dir) {
    return c->done || c->dir != dir || c->level != level;
}

// Visit edge (src -> u) for one lane.  Every lane of the wave must call this
// together (valid = false for idle lanes): the queue append is one atomic per
// wave.
__device__ __forceinline__ void td_visit(const td_args &a, bool valid, int32_t u, int32_t src,
                                         unsigned long long &cnt, unsigned long long &dsum,
                                         unsigned long long &dmax,
                                         unsigned long long *qlen_out) {
    bool won = false;
    const uint64_t bit = 1ULL << (u & 63);
    if (valid) {
        uint64_t *wp = a.vis + (u >> 6);
        if (!(*wp & bit)) {
            const unsigned long long old = atomicOr((unsigned long long *)wp, (unsigned long long)bit);
            won = !(old & bit);
        }
    }
    if (won) {
        a.depth[u] = a.level + 1;
        a.parent[u] = src;
        const unsigned long long d = (unsigned long long)(a.off[u + 1] - a.off[u]);
        ++cnt;
        dsum += d;
        dmax = d > dmax ? d : dmax;
        atomicOr((unsigned long long *)(a.next_bm + (u >> 6)), (unsigned long long)bit);
    }
    const uint64_t m = __ballot(won);
    if (m) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((unsigned long long)m) - 1;
        long long base = 0;
        if (lane == leader) base = (long long)atomicAdd(qlen_out, (unsigned long long)__popcll(m));
        base = __shfl(base, leader);
        if (won) a.qout[base + __popcll(m & ((1ULL << lane) - 1))] = u;
    }
}

// Top-down step (hybrid_bfs.cc:71-100).  Each block takes tiles of BLOCK
// frontier slots.  Per tile: degree >= task_t -> split into tasks for
// k_top_down_tasks; degree >= wave_t -> one wave strides the list;
// otherwise the tile's short lists are concatenated by a block-wide scan and
// every thread takes one edge (owner found by binary search in LDS).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_top_down(td_args a) {
    ctl_block *c = a.c;
    if (a.gate && step_gated_out(c, a.level, BD_DIR_TD)) return;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < a.nwords;
         i += (int64_t)gridDim.x * BLOCK)
        a.clr[i] = 0;
    const bool bitmap = c->td_bitmap != 0;
    const int64_t nitems = bitmap ? a.nwords * 64 : (int64_t)c->qlen[a.level & 1];
    unsigned long long *qlen_out = &c->qlen[(a.level + 1) & 1];

    __shared__ int64_t s_beg[BLOCK];
    __shared__ int64_t s_end[BLOCK];
    __shared__ int32_t s_src[BLOCK];
    __shared__ int32_t s_pref[BLOCK + 1];
    __shared__ int32_t s_med[BLOCK];
    __shared__ int32_t s_wsum[BLOCK / 64];
    __shared__ int32_t s_nmed;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    unsigned long long cnt = 0, dsum = 0, dmax = 0;

    for (int64_t tile = blockIdx.x; tile * BLOCK < nitems; tile += gridDim.x) {
        if (tid == 0) s_nmed = 0;
        __syncthreads();
        const int64_t slot = tile * BLOCK + tid;
        int32_t v = -1;
        if (slot < nitems) {
            if (bitmap) {
                const uint64_t fw = a.front[slot >> 6];
                if ((fw >> (slot & 63)) & 1) v = (int32_t)slot;
            } else {
                v = a.qin[slot];
            }
        }
        int64_t b = 0, e = 0;
        if (v >= 0) {
            b = a.off[v];
            e = a.off[v + 1];
        }
        const int64_t dg = e - b;
        int32_t small = 0;
        if (dg >= a.task_t) {
            const int64_t nt = (dg + a.task_edges - 1) / a.task_edges;
            const int64_t base = (int64_t)atomicAdd(&c->ntasks, (unsigned long long)nt);
            for (int64_t t = 0; t < nt; ++t) {
                const int64_t idx = base + t;
                if (idx < a.task_cap) {
                    td_task tk;
                    tk.beg = b + t * a.task_edges;
                    tk.end = tk.beg + a.task_edges < e ? tk.beg + a.task_edges : e;
                    tk.src = v;
                    tk.pad = 0;
                    a.tasks[idx] = tk;
                } else {
                    atomicAdd(&c->task_overflow, 1ULL);
                }
            }
        } else if (dg >= a.wave_t) {
            const int p = atomicAdd(&s_nmed, 1);
            s_med[p] = tid;
        } else {
            small = (int32_t)dg;
        }
        s_beg[tid] = b;
        s_end[tid] = e;
        s_src[tid] = v;
        int32_t incl = small;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == 63) s_wsum[wid] = incl;
        __syncthreads();
        int32_t wbase = 0;
        for (int w2 = 0; w2 < wid; ++w2) wbase += s_wsum[w2];
        s_pref[tid] = wbase + incl - small;
        if (tid == BLOCK - 1) s_pref[BLOCK] = wbase + incl;
        __syncthreads();
        const int32_t total = s_pref[BLOCK];
        const int nmed = s_nmed;

        for (int i = wid; i < nmed; i += BLOCK / 64) {
            const int t = s_med[i];
            const int64_t mb = s_beg[t], me = s_end[t];
            const int32_t src = s_src[t];
            for (int64_t k0 = mb; k0 < me; k0 += 64) {
                const int64_t k = k0 + lane;
                const bool valid = k < me;
                const int32_t u = valid ? a.cols[k] : 0;
                td_visit(a, valid, u, src, cnt, dsum, dmax, qlen_out);
            }
        }
        for (int32_t i0 = 0; i0 < total; i0 += BLOCK) {
            const int32_t i = i0 + tid;
            const bool valid = i < total;
            int32_t u = 0, src = 0;
            if (valid) {
                int lo = 0, hi = BLOCK;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_pref[mid] <= i)
                        lo = mid;
                    else
                        hi = mid;
                }
                src = s_src[lo];
                u = a.cols[s_beg[lo] + (i - s_pref[lo])];
            }
            td_visit(a, valid, u, src, cnt, dsum, dmax, qlen_out);
        }
        __syncthreads();
    }
    block_flush<BLOCK>(c, cnt, dsum, 0, dmax);
}

// Slices of huge adjacency lists, one wave per slice (the reference's
// 64-edge granules, worklist.h:96-117, with 64 lanes instead of 64 threads).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_top_down_tasks(td_args a) {
    ctl_block *c = a.c;
    if (a.gate && step_gated_out(c, a.level, BD_DIR_TD)) return;
    const int64_t nt_all = (int64_t)c->ntasks;
    const int64_t ntasks = nt_all < a.task_cap ? nt_all : a.task_cap;
    if (ntasks == 0) return;
    unsigned long long *qlen_out = &c->qlen[(a.level + 1) & 1];
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * (BLOCK / 64);
    unsigned long long cnt = 0, dsum = 0, dmax = 0;
    for (int64_t t = wave0; t < ntasks; t += nwaves) {
        const td_task tk = a.tasks[t];
        for (int64_t k0 = tk.beg; k0 < tk.end; k0 += 64) {
            const int64_t k = k0 + lane;
            const bool valid = k < tk.end;
            const int32_t u = valid ? a.cols[k] : 0;
            td_visit(a, valid, u, tk.src, cnt, dsum, dmax, qlen_out);
        }
    }
    block_flush<BLOCK>(c, cnt, dsum, 0, dmax);
}

// Bottom-up step (hybrid_bfs.cc:111-143).  One wave per 64-vertex word.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bottom_up(bu_args a) {
    ctl_block *c = a.c;
    if (a.gate && step_gated_out(c, a.level, BD_DIR_BU)) return;
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * (BLOCK / 64);
    unsigned long long cnt = 0, dsum = 0, exam = 0, dmax = 0;
    for (int64_t w = wave0; w < a.nwords; w += nwaves) {
        const uint64_t vw = a.vis[w];
        const uint64_t unv = ~vw;
        uint64_t found = 0;
        if (unv != 0) {
            const int64_t v = w * 64 + lane;
            int32_t par = -1;
            int64_t dg = 0;
            if ((unv >> lane) & 1) {
                const int64_t b = a.off[v], e = a.off[v + 1];
                dg = e - b;
                for (int64_t k = b; k < e; ++k) {
                    const int32_t u = a.cols[k];
                    ++exam;
                    if ((a.front[u >> 6] >> (u & 63)) & 1) {
                        par = u;
                        break;
                    }
                }
            }
            found = __ballot(par >= 0);
            if (par >= 0) {
                a.depth[v] = a.level + 1;
                a.parent[v] = par;
                ++cnt;
                dsum += (unsigned long long)dg;
                dmax = (unsigned long long)dg > dmax ? (unsigned long long)dg : dmax;
            }
        }
        if (lane == 0) {
            a.next[w] = found;
            if (found) a.vis[w] = vw | found;
            a.clr[w] = 0;
        }
    }
    block_flush<BLOCK>(c, cnt, dsum, exam, dmax);
}

// hybrid_bfs.cc:162-190 as a per-step state machine.  After a TD step the
// scout count is sum deg(N) (the -deg encoding of clear(), :93, :438); after
// a BU step the do-while test (:178-179) decides whether to stay bottom-up;
// leaving sets scout = 1 (:180).  Then the outer test (:167) picks the
// direction, charging scout to edges_to_check before a TD step (:182).
__device__ void beamer_next(ctl_block *c, int32_t last_dir, int64_t nnew, int64_t nnew_deg) {
    if (last_dir == BD_DIR_TD) {
        c->scout = nnew_deg;
    } else if (c->in_bu_loop) {
        if (nnew >= c->old_awake || nnew > c->nv / c->beta) {
            c->old_awake = nnew;
            c->dir = BD_DIR_BU;
            return;
        }
        c->in_bu_loop = 0;
        c->scout = 1;
    }
    if (c->scout > c->edges_to_check / c->alpha) {
        c->in_bu_loop = 1;
        c->old_awake = c->frontier;
        c->dir = BD_DIR_BU;
    } else {
        c->edges_to_check -= c->scout;
        c->dir = BD_DIR_TD;
    }
}

struct decide_args {
    ctl_block *c;
    bd_level_rec *log;
    const int32_t *script;
    const float *mlp;
    uint64_t *vis;
    uint64_t *bm0;
    int32_t *q0;
    int32_t *depth;
    int32_t *parent;
    const int64_t *off;
    int64_t source;      // >= 0: initialise a search from this vertex
    int64_t nv, m, nnz;
    int64_t alpha, beta;
    int32_t level;
    int32_t gate;
    int32_t policy;
    int32_t script_len;
    int32_t log_cap;
    int32_t mlp_h1, mlp_h2;
    int32_t has_mlp;
};

// Choose the direction of step c->level.  Runs on one wave: lane 0 owns the
// control block; the learned policy evaluates its network on the wave.
__device__ void choose_dir(const decide_args &a, int32_t last_dir, int64_t nnew,
                           int64_t nnew_deg) {
    ctl_block *c = a.c;
    __shared__ float s_x[8];
    __shared__ int s_use_mlp;
    if (threadIdx.x == 0) {
        s_use_mlp = 0;
        switch (a.policy) {
        case BD_POLICY_TD_ONLY:
            c->dir = BD_DIR_TD;
            break;
        case BD_POLICY_BU_ONLY:
            c->dir = BD_DIR_BU;
            break;
        case BD_POLICY_SCRIPT: {
            int32_t i = c->level < a.script_len ? c->level : a.script_len - 1;
            c->dir = (a.script_len > 0 && a.script[i] == BD_DIR_BU) ? BD_DIR_BU : BD_DIR_TD;
            break;
        }
        case BD_POLICY_LEARNED:
            if (a.has_mlp) {
                s_use_mlp = 1;
                mlp_features_dev(c->frontier, c->frontier_deg, c->max_frontier_deg,
                                 c->unvisited, c->unvisited_deg, c->level, c->nv, a.m,
                                 last_dir, s_x);
                break;
            }
            // no network loaded: Beamer is the fallback
            beamer_next(c, last_dir, nnew, nnew_deg);
            break;
        default:
            beamer_next(c, last_dir, nnew, nnew_deg);
        }
    }
    __syncthreads();
    if (s_use_mlp) {
        float y[2];
        mlp_forward_wave(a.mlp, a.mlp_h1, a.mlp_h2, s_x, y);
        if (threadIdx.x == 0) {
            // predicted log-costs; NaN falls back to top-down
            c->dir = (y[1] < y[0]) ? BD_DIR_BU : BD_DIR_TD;
        }
    }
}

__global__ __launch_bounds__(64) void k_decide(decide_args a) {
    ctl_block *c = a.c;
    if (a.source >= 0) {
        // clear() (hybrid_bfs.cc:433-443) is done by k_clear; here the source
        // is pushed and marked (hybrid_bfs.cc:157-163).
        const int64_t s = a.source;
        if (threadIdx.x == 0) {
            const int64_t dsrc = a.off[s + 1] - a.off[s];
            memset(c, 0, sizeof(ctl_block));
            c->nv = a.nv;
            c->m2 = 2 * a.m;
            c->alpha = a.alpha;
            c->beta = a.beta;
            c->policy = a.policy;
            c->edges_to_check = 2 * a.m;
            c->scout = dsrc;
            c->prev_dir = -1;
            c->frontier = 1;
            c->frontier_deg = dsrc;
            c->max_frontier_deg = dsrc;
            c->unvisited = a.nv - 1;
            c->unvisited_deg = a.nnz - dsrc;
            c->qlen[0] = 1;
            c->reached = 1;
            c->reached_deg = dsrc;
            a.vis[s >> 6] |= 1ULL << (s & 63);
            a.bm0[s >> 6] = 1ULL << (s & 63);
            a.q0[0] = (int32_t)s;
            a.depth[s] = 0;
            a.parent[s] = (int32_t)s;
        }
        __syncthreads();
        // level 0: with no previous step, Beamer applies only the outer test
        // to scout = deg(source) (hybrid_bfs.cc:163-167)
        choose_dir(a, -1, 0, 0);
        if (threadIdx.x == 0) {
            __threadfence();
            c->t_start = wall_clock64();
        }
        return;
    }
    if (a.gate && (c->done || c->level != a.level)) return;
    __shared__ long long s_new, s_new_deg;
    __shared__ int s_dir;
    if (threadIdx.x == 0) {
        const int64_t nnew = (int64_t)c->acc_new;
        const int64_t nnew_deg = (int64_t)c->acc_new_deg;
        const int32_t dir = c->dir;
        const int32_t L = c->level;
        bd_level_rec r;
        r.level = L;
        r.dir = dir;
        r.frontier = c->frontier;
        r.frontier_deg = c->frontier_deg;
        r.unvisited = c->unvisited;
        r.unvisited_deg = c->unvisited_deg;
        r.edges_to_check = c->edges_to_check;
        r.result = dir == BD_DIR_TD ? nnew_deg : nnew;
        r.next = nnew;
        r.next_deg = nnew_deg;
        r.edges_examined = dir == BD_DIR_TD ? c->frontier_deg : (int64_t)c->acc_examined;
        r.max_frontier_deg = c->max_frontier_deg;
        r.td_bitmap_input = (dir == BD_DIR_TD && c->td_bitmap) ? 1 : 0;
        r.reserved = (int64_t)c->task_overflow;
        r.ms = -1.0;
        r.ms_other = -1.0;
        if (L < a.log_cap) a.log[L] = r;
        c->steps += 1;
        if (dir == BD_DIR_TD)
            c->td_steps += 1;
        else
            c->bu_steps += 1;
        c->reached += nnew;
        c->reached_deg += nnew_deg;
        c->unvisited -= nnew;
        c->unvisited_deg -= nnew_deg;
        c->frontier = nnew;
        c->frontier_deg = nnew_deg;
        c->max_frontier_deg = (int64_t)c->acc_max_deg;
        c->acc_new = 0;
        c->acc_new_deg = 0;
        c->acc_examined = 0;
        c->acc_max_deg = 0;
        c->ntasks = 0;
        c->qlen[L & 1] = 0;
        c->td_bitmap = dir == BD_DIR_BU ? 1 : 0;
        c->prev_dir = dir;
        c->level = L + 1;
        if (nnew == 0) c->done = 1;
        s_new = nnew;
        s_new_deg = nnew_deg;
        s_dir = dir;
    }
    __syncthreads();
    if (s_new > 0) choose_dir(a, s_dir, s_new, s_new_deg);
    if (threadIdx.x == 0) {
        __threadfence();
        c->t_end = wall_clock64();
    }
}

// clear() (hybrid_bfs.cc:433-443) for the device layout.
__global__ void k_clear(int32_t *__restrict__ depth, int32_t *__restrict__ parent, int64_t nv,
                        uint64_t *__restrict__ vis, const uint64_t *__restrict__ dead,
                        uint64_t *__restrict__ b0, uint64_t *__restrict__ b1,
                        uint64_t *__restrict__ b2, int64_t nwords) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        depth[i] = -1;
        parent[i] = -1;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) {
        vis[i] = dead[i];
        b0[i] = 0;
        b1[i] = 0;
        b2[i] = 0;
    }
}

// Reset the per-step accumulators (profile runs time both directions).
__global__ void k_reset_acc(ctl_block *c) {
    c->acc_new = 0;
    c->acc_new_deg = 0;
    c->acc_examined = 0;
    c->acc_max_deg = 0;
}

// count_num_traversed_edges (hybrid_bfs.cc:401-414): sum of degrees over
// reached vertices.
__global__ void k_count_traversed(const int32_t *__restrict__ depth, const int64_t *__restrict__ off,
                                  int64_t nv, unsigned long long *__restrict__ out) {
    unsigned long long s = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        if (depth[v] >= 0) s += (unsigned long long)(off[v + 1] - off[v]);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
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

// Launch kernel template K<block> for block in {128, 256, 512}.
#define BD_LAUNCH_BLOCKED(K, block, grid, stream, arg)                                 \
    do {                                                                               \
        switch (block) {                                                               \
        case 128:                                                                      \
            hipLaunchKernelGGL(K<128>, dim3(grid), dim3(128), 0, stream, arg);         \
            break;                                                                     \
        case 512:                                                                      \
            hipLaunchKernelGGL(K<512>, dim3(grid), dim3(512), 0, stream, arg);         \
            break;                                                                     \
        default:                                                                       \
            hipLaunchKernelGGL(K<256>, dim3(grid), dim3(256), 0, stream, arg);         \
        }                                                                              \
    } while (0)

int norm_block(int b) { return (b == 128 || b == 512) ? b : 256; }

struct level_ptrs {
    td_args td;
    bu_args bu;
};

level_ptrs make_level(bd_ctx *c, int32_t level, int32_t gate) {
    level_ptrs p;
    td_args &t = p.td;
    t.off = c->d_off;
    t.cols = c->d_cols;
    t.vis = c->d_vis;
    t.qin = c->d_q[level & 1];
    t.front = c->d_bm[level % 3];
    t.next_bm = c->d_bm[(level + 1) % 3];
    t.clr = c->d_bm[(level + 2) % 3];
    t.qout = c->d_q[(level + 1) & 1];
    t.depth = c->d_depth;
    t.parent = c->d_parent;
    t.tasks = c->d_tasks;
    t.c = c->d_ctl;
    t.nwords = c->nwords;
    t.task_cap = c->task_cap;
    t.level = level;
    t.gate = gate;
    t.wave_t = std::max(1, c->tune.wave_threshold);
    t.task_t = std::max(64, c->tune.task_threshold);
    t.task_edges = std::max(64, c->tune.task_edges);
    t.pad = 0;
    bu_args &b = p.bu;
    b.off = c->d_off;
    b.cols = c->d_cols;
    b.vis = c->d_vis;
    b.front = c->d_bm[level % 3];
    b.next = c->d_bm[(level + 1) % 3];
    b.clr = c->d_bm[(level + 2) % 3];
    b.depth = c->d_depth;
    b.parent = c->d_parent;
    b.c = c->d_ctl;
    b.nwords = c->nwords;
    b.level = level;
    b.gate = gate;
    return p;
}

int launch_td(bd_ctx *c, const td_args &t) {
    const int blk = norm_block(c->tune.td_block);
    const int64_t grid = (int64_t)c->cu_count * std::max(1, c->tune.grid_per_cu);
    BD_LAUNCH_BLOCKED(k_top_down, blk, grid, c->stream, t);
    BD_LAUNCH_BLOCKED(k_top_down_tasks, blk, grid, c->stream, t);
    BD_HIP(c, hipGetLastError());
    return BD_OK;
}

int launch_bu(bd_ctx *c, const bu_args &b) {
    const int blk = norm_block(c->tune.bu_block);
    const int64_t grid = grid_for(c, c->nwords * 64, blk);
    BD_LAUNCH_BLOCKED(k_bottom_up, blk, grid, c->stream, b);
    BD_HIP(c, hipGetLastError());
    return BD_OK;
}

decide_args make_decide(bd_ctx *c, int32_t level, int64_t source, int64_t alpha, int64_t beta,
                        int policy) {
    decide_args d;
    d.c = c->d_ctl;
    d.log = c->d_log;
    d.script = c->d_script;
    d.mlp = c->d_mlp;
    d.vis = c->d_vis;
    d.bm0 = c->d_bm[0];
    d.q0 = c->d_q[0];
    d.depth = c->d_depth;
    d.parent = c->d_parent;
    d.off = c->d_off;
    d.source = source;
    d.nv = c->nv;
    d.m = c->m;
    d.nnz = c->nnz;
    d.alpha = alpha;
    d.beta = beta;
    d.level = level;
    d.gate = 1;
    d.policy = policy;
    d.script_len = (int32_t)c->script_len;
    d.log_cap = (int32_t)c->log_cap;
    d.mlp_h1 = c->mlp_h1;
    d.mlp_h2 = c->mlp_h2;
    d.has_mlp = c->has_mlp ? 1 : 0;
    return d;
}

int check_run_args(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy) {
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph");
    if (source < 0 || source >= c->nv) return set_err(c, BD_ERANGE, "Source vertex out of range.");
    if (alpha <= 0) return set_err(c, BD_EINVAL, "alpha must be > 0");
    if (beta <= 0) return set_err(c, BD_EINVAL, "beta must be > 0");
    if (policy < BD_POLICY_BEAMER || policy > BD_POLICY_SCRIPT)
        return set_err(c, BD_EINVAL, "unknown policy");
    if (policy == BD_POLICY_SCRIPT && c->script_len <= 0)
        return set_err(c, BD_ESTATE, "no script set");
    return BD_OK;
}

int start_search(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy) {
    const int64_t n = std::max(c->nv, c->nwords);
    hipLaunchKernelGGL(k_clear, dim3(grid_for(c, n, 256)), dim3(256), 0, c->stream, c->d_depth,
                       c->d_parent, c->nv, c->d_vis, c->d_dead, c->d_bm[0], c->d_bm[1],
                       c->d_bm[2], c->nwords);
    decide_args d = make_decide(c, 0, source, alpha, beta, policy);
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, c->stream, d);
    BD_HIP(c, hipGetLastError());
    return BD_OK;
}

int read_ctl(bd_ctx *c) {
    BD_HIP(c, hipMemcpyAsync(c->h_ctl, c->d_ctl, sizeof(ctl_block), hipMemcpyDeviceToHost,
                             c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    return BD_OK;
}

int finish_search(bd_ctx *c, int64_t source, bd_run_stats *out, double event_ms) {
    const ctl_block &h = *c->h_ctl;
    if (h.task_overflow) return set_err(c, BD_EHIP, "top-down task list overflow");
    c->last_steps = h.steps;
    c->last_source = source;
    c->have_result = true;
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
        const double dev_ms = (double)(h.t_end - h.t_start) / (double)c->wall_khz;
        out->ms = (h.t_end >= h.t_start && dev_ms > 0) ? dev_ms : event_ms;
    }
    return BD_OK;
}

}  // namespace

int run_search(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy,
               bd_run_stats *out) {
    int rc;
    if ((rc = check_run_args(c, source, alpha, beta, policy))) return rc;
    BD_HIP(c, hipSetDevice(c->device));
    if ((rc = start_search(c, source, alpha, beta, policy))) return rc;
    BD_HIP(c, hipEventRecord(c->ev[0], c->stream));
    const int64_t max_levels = c->nv + 2;
    int32_t L = 0;
    const bool timed = c->tune.time_steps == 1;
    std::vector<double> step_ms;
    if (c->tune.level_driver == 2 || timed) {
        if ((rc = read_ctl(c))) return rc;
        while (!c->h_ctl->done) {
            level_ptrs p = make_level(c, L, 1);
            if (timed) BD_HIP(c, hipEventRecord(c->ev[2], c->stream));
            if (c->h_ctl->dir == BD_DIR_TD)
                rc = launch_td(c, p.td);
            else
                rc = launch_bu(c, p.bu);
            if (rc) return rc;
            if (timed) BD_HIP(c, hipEventRecord(c->ev[3], c->stream));
            hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, c->stream,
                               make_decide(c, L, -1, alpha, beta, policy));
            BD_HIP(c, hipGetLastError());
            if ((rc = read_ctl(c))) return rc;
            if (timed) {
                float ms = 0;
                BD_HIP(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
                step_ms.push_back(ms);
            }
            if (++L > max_levels) return set_err(c, BD_EHIP, "search did not terminate");
        }
    } else {
        int chunk = 8;
        for (;;) {
            for (int k = 0; k < chunk; ++k, ++L) {
                level_ptrs p = make_level(c, L, 1);
                if ((rc = launch_td(c, p.td))) return rc;
                if ((rc = launch_bu(c, p.bu))) return rc;
                hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, c->stream,
                                   make_decide(c, L, -1, alpha, beta, policy));
            }
            BD_HIP(c, hipGetLastError());
            if ((rc = read_ctl(c))) return rc;
            if (c->h_ctl->done) break;
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
    for (size_t i = 0; i < step_ms.size() && i < c->last_log.size(); ++i)
        c->last_log[i].ms = step_ms[i];
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

extern "C" int bd_bfs_fetch(bd_ctx *c, int32_t *depth, int64_t *parent) {
    if (!c) return BD_EINVAL;
    if (!c->have_result) return set_err(c, BD_ESTATE, "no search has run");
    BD_HIP(c, hipSetDevice(c->device));
    if (depth)
        BD_HIP(c, hipMemcpy(depth, c->d_depth, sizeof(int32_t) * c->nv, hipMemcpyDeviceToHost));
    if (parent) {
        std::vector<int32_t> p(c->nv);
        BD_HIP(c, hipMemcpy(p.data(), c->d_parent, sizeof(int32_t) * c->nv, hipMemcpyDeviceToHost));
        for (int64_t v = 0; v < c->nv; ++v) parent[v] = p[v];
    }
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
    int rc = dev_alloc(c, (void **)&d, sizeof(unsigned long long), "count");
    if (rc) return rc;
    unsigned long long h = 0;
    hipError_t e = hipMemsetAsync(d, 0, sizeof(h), c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_count_traversed, dim3(grid_for(c, c->nv, 256)), dim3(256), 0,
                           c->stream, c->d_depth, c->d_off, c->nv, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(d);
    if (e != hipSuccess) return hip_err(c, e, "count_traversed");
    *out = (int64_t)(h / 2);
    return BD_OK;
}

// hybrid_bfs::check (hybrid_bfs.cc:281-379): a fresh serial BFS gives the
// depths; every reached vertex's parent must be a neighbour one level up, the
// source its own parent, and reachability must agree.
extern "C" int bd_bfs_check(bd_ctx *c, int64_t source, int *ok) {
    if (!c || !ok) return BD_EINVAL;
    if (!c->have_result) return set_err(c, BD_ESTATE, "no search has run");
    if (source < 0 || source >= c->nv) return set_err(c, BD_ERANGE, "source out of range");
    const int64_t nv = c->nv;
    std::vector<int64_t> off(nv + 1);
    std::vector<int32_t> cols(c->nnz);
    std::vector<int64_t> parent(nv);
    int rc = bd_graph_download(c, off.data(), cols.data());
    if (!rc) rc = bd_bfs_fetch(c, nullptr, parent.data());
    if (rc) return rc;
    std::vector<int64_t> depth(nv, -1);
    std::vector<int64_t> q;
    q.reserve(nv);
    q.push_back(source);
    depth[source] = 0;
    for (size_t h = 0; h < q.size(); ++h) {
        const int64_t u = q[h];
        for (int64_t e = off[u]; e < off[u + 1]; ++e) {
            const int64_t v = cols[e];
            if (depth[v] == -1) {
                depth[v] = depth[u] + 1;
                q.push_back(v);
            }
        }
    }
    bool correct = true;
    for (int64_t u = 0; u < nv && correct; ++u) {
        if (depth[u] >= 0 && parent[u] >= 0) {
            if (u == source) {
                if (!(parent[u] == u && depth[u] == 0)) correct = false;
                continue;
            }
            bool found = false;
            for (int64_t e = off[u]; e < off[u + 1]; ++e) {
                if (cols[e] == parent[u]) {
                    if (depth[cols[e]] != depth[u] - 1) break;
                    found = true;
                    break;
                }
            }
            if (!found) correct = false;
        } else if ((depth[u] < 0) != (parent[u] < 0)) {
            correct = false;
        }
    }
    *ok = correct ? 1 : 0;
    return BD_OK;
}

// Time both directions at every step from the same state; the search follows
// `policy`.  The direction not taken runs first, then the state it changed is
// restored (vis) and the chosen direction runs, so the search ends exactly as
// a normal run would.  Both steps must find the same number of vertices.
extern "C" int bd_bfs_profile(bd_ctx *c, int64_t source, int64_t alpha, int64_t beta, int policy,
                              bd_run_stats *out) {
    if (!c) return BD_EINVAL;
    int rc;
    if ((rc = check_run_args(c, source, alpha, beta, policy))) return rc;
    BD_HIP(c, hipSetDevice(c->device));
    // scrub buffer: 512 MiB, larger than the last-level cache
    const int64_t scrub_n = ((int64_t)512 << 20) / 16;
    uint4 *scrub = nullptr;
    if ((rc = dev_alloc(c, (void **)&scrub, sizeof(uint4) * (scrub_n + 1), "scrub"))) return rc;
    struct guard {
        uint4 *p;
        ~guard() { dev_free(p); }
    } g{scrub};
    BD_HIP(c, hipMemsetAsync(scrub, 0, sizeof(uint4) * (scrub_n + 1), c->stream));
    auto do_scrub = [&]() {
        hipLaunchKernelGGL(k_scrub, dim3(c->cu_count * 4), dim3(256), 0, c->stream, scrub,
                           scrub_n, scrub + scrub_n);
    };
    if ((rc = start_search(c, source, alpha, beta, policy))) return rc;
    if ((rc = read_ctl(c))) return rc;
    std::vector<std::pair<double, double>> times;
    double total = 0;
    int32_t L = 0;
    while (!c->h_ctl->done) {
        const int32_t chosen = c->h_ctl->dir;
        const int32_t other = 1 - chosen;
        level_ptrs p = make_level(c, L, 0);
        float ms[2] = {0, 0};
        unsigned long long found[2] = {0, 0};
        for (int pass = 0; pass < 2; ++pass) {
            const int32_t d = pass == 0 ? other : chosen;
            if (pass == 1) {
                BD_HIP(c, hipMemcpyAsync(c->d_vis, c->d_vis_save, sizeof(uint64_t) * c->nwords,
                                         hipMemcpyDeviceToDevice, c->stream));
                hipLaunchKernelGGL(k_reset_acc, dim3(1), dim3(1), 0, c->stream, c->d_ctl);
            } else {
                BD_HIP(c, hipMemcpyAsync(c->d_vis_save, c->d_vis, sizeof(uint64_t) * c->nwords,
                                         hipMemcpyDeviceToDevice, c->stream));
            }
            do_scrub();
            BD_HIP(c, hipEventRecord(c->ev[2], c->stream));
            rc = d == BD_DIR_TD ? launch_td(c, p.td) : launch_bu(c, p.bu);
            if (rc) return rc;
            BD_HIP(c, hipEventRecord(c->ev[3], c->stream));
            BD_HIP(c, hipEventSynchronize(c->ev[3]));
            BD_HIP(c, hipEventElapsedTime(&ms[pass], c->ev[2], c->ev[3]));
            if ((rc = read_ctl(c))) return rc;
            found[pass] = c->h_ctl->acc_new;
        }
        if (found[0] != found[1])
            return set_err(c, BD_EHIP, "profile: directions disagree on the next frontier");
        times.push_back({ms[1], ms[0]});
        total += ms[1];
        hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, c->stream,
                           make_decide(c, L, -1, alpha, beta, policy));
        BD_HIP(c, hipGetLastError());
        if ((rc = read_ctl(c))) return rc;
        if (++L > c->nv + 2) return set_err(c, BD_EHIP, "search did not terminate");
    }
    if ((rc = finish_search(c, source, out, total))) return rc;
    for (size_t i = 0; i < times.size() && i < c->last_log.size(); ++i) {
        c->last_log[i].ms = times[i].first;
        c->last_log[i].ms_other = times[i].second;
    }
    if (out) out->ms = total;
    return BD_OK;
}
