// bd_batch.hip -- batched searches with no reference counterpart: up to 64
// roots in one bit-parallel search, and many disjoint graphs searched from
// all their roots at once (bd_bfs.hip, multi-source start).  Both replace
// trips round the reference's strictly sequential trial loop
// (hybrid_bfs_main.cc:252-302).
//
// Bit-parallel search (multi-source BFS over 64-bit root masks).  Device
// state per vertex v (8 bytes each):
//   seen[v]   roots that have reached v
//   F[v]      roots whose depth-L frontier holds v
//   N[v]      roots whose depth-(L+1) frontier holds v (filled by step L)
//   Fb, Nb    bitmaps of the vertices with F[v] != 0, N[v] != 0 (V/8 bytes,
//             so they stay in L2; a wave writes its 64 vertices' word with a
//             ballot)
// A level is either
//   push: every v with F[v] != 0 ORs F[v] & ~seen[u] into N[u] of each
//         neighbour u (device atomics), then a commit pass keeps the bits
//         not yet seen and clears F for reuse; or
//   pull: every v not yet seen by all roots ORs F[u] over its neighbours
//         that are in Fb, stopping as soon as every root it still waits for
//         is covered.
// Each level is two launches, k_ms_level (push or pull) and k_ms_post
// (commit, or clearing the old frontier masks after a pull).  Both pick the
// mode on the device from the counters of the frontier -- its degree sum
// against the degree sum of vertices still open, as Beamer's rule picks a
// direction -- so a search is a fixed sequence of launches with no host
// round trip: the host captures a chunk of levels in a HIP graph once and
// replays it; launches past the last level return at once.  The search runs
// on the degree-ordered layout the single searches read (bd_graph.hip);
// depths are written root-major in its ids (depth[r * nv + i]) when found, 64
// masked stores per wave, one per root, each covering 64 consecutive
// vertices of that root's row, and mapped to graph ids on readout.
#include <string.h>

#include <algorithm>
#include <vector>

#include "bd_device.h"
#include "bd_internal.h"

namespace bd {

namespace {

constexpr int kMsBlock = 256;
constexpr int kMsWaveDeg = 64;  // lists at least this long are scanned by a whole wave
constexpr int kMsChunk = 16;    // levels per captured graph
// Frontier lists of kMsWaveDeg entries or more are pushed as slices of up
// to kMsSlice entries, one wave each, taken by all waves of the grid: the
// step that finds a frontier vertex files its slices (a vertex group's own
// wave would otherwise walk every long list of its 64 vertices, and in the
// degree-ordered layout the high-degree vertices share the first groups).
constexpr int kMsSlice = 1024;
// A pull gives each wave 64 consecutive vertices, so lists of kMsHubDeg
// entries or more are taken out of it: they are cut into slices once per
// graph (k_ms_hubs), every wave of the grid takes slices of the vertices
// still open and ORs what it finds into N[v], and the level's second launch
// finishes those vertices.  Otherwise the hubs that share a wave's 64
// vertices (all of them at the front in plain degree order) are walked one
// after another by that wave.
constexpr int kMsHubDeg = 512;
// Slice counts of frontier L sit in its counter copy: line s holds, in
// word 3, the slices filed in segment s (by waves w with w % kSlots == s),
// and line 0, word 4, those filed in the overflow segment.
constexpr int kMsSegCount = 3, kMsOvfCount = 4;

struct ms_slice {
    int64_t beg;
    int32_t len;
    int32_t v;
};

// Device control of a bit-parallel search: done once a frontier is empty,
// `levels` the number of non-empty frontiers.
struct ms_ctl {
    int32_t done;
    int32_t levels;
};

struct ms_args {
    const int64_t *off;
    const int32_t *cols;
    const int32_t *inv;  // graph id -> search id (null: the search runs on graph ids)
    uint64_t *F;
    uint64_t *N;
    uint64_t *seen;
    uint64_t *Fb;    // bit v: F[v] != 0 (frontier membership, V/64 words)
    uint64_t *Nb;    // bit v: N[v] != 0, written by the commit or pull pass
    int32_t *depth;  // n rows of nv, or null
    // counters of frontier L in copy L % 3, kSlots lines each: frontier
    // size, its degree sum, the degree sum of vertices still open
    unsigned long long *cnt;
    ms_ctl *ctl;
    ms_slice *slices[2];  // frontier L's slices in slices[L & 1]
    const int32_t *hubs;         // vertices with kMsHubDeg or more neighbours
    const ms_slice *hub_slices;  // their lists in slices of up to kMsSlice entries
    int64_t nhubs, nhub_slices;
    int64_t seg_cap;      // slices per segment; the overflow segment holds max_slices
    int64_t max_slices;   // most slices one frontier can have
    bd_level_rec *log;
    int64_t log_cap;
    int64_t nv;
    uint64_t full;   // bits of the roots in this search
    uint64_t alpha;  // push while frontier degree sum * alpha < open degree sum
    int32_t nroots;
    int32_t level;   // the step expands depth `level` into depth `level + 1`
};

__device__ __forceinline__ unsigned long long *ms_copy(const ms_args &a, int level) {
    return a.cnt + (int64_t)(level % 3) * kSlots * kLine;
}

// Record the roots in `nf` as reaching v at depth d: root-major rows, one
// masked store per root, so each store covers the wave's 64 vertices.
__device__ __forceinline__ void ms_record(const ms_args &a, int64_t v, bool valid, uint64_t nf,
                                          int32_t d) {
    if (!a.depth) return;
    // only the roots that some lane of the wave gained
    for (uint64_t rem = wave_or(valid ? nf : 0ULL); rem; rem &= rem - 1) {
        const int r = __ffsll((unsigned long long)rem) - 1;
        if (valid && ((nf >> r) & 1)) a.depth[(int64_t)r * a.nv + v] = d;
    }
}

// Neighbour ids are read once per level: non-temporal loads, so the stream
// does not push the bitmaps and masks out of L2.
__device__ __forceinline__ int32_t ms_col(const ms_args &a, int64_t k) {
    return __builtin_nontemporal_load(a.cols + k);
}

// Roots whose frontier holds u: the membership bitmap (small enough to stay
// in L2) first, the 8-byte mask only for members.
__device__ __forceinline__ uint64_t ms_frontier(const ms_args &a, int32_t u) {
    return ((a.Fb[u >> 6] >> (u & 63)) & 1) ? a.F[u] : 0ULL;
}

// OR the bits of frontier vertex mask f that u has not seen into N[u].
__device__ __forceinline__ void ms_push_to(const ms_args &a, int32_t u, uint64_t f) {
    const uint64_t m = f & ~a.seen[u];
    if (m) atomicOr((unsigned long long *)&a.N[u], (unsigned long long)m);
}

// Per-block counters of the next frontier, added to slot blockIdx % kSlots
// of copy (level + 1) % 3.
__device__ __forceinline__ void ms_count(const ms_args &a, unsigned long long c0,
                                         unsigned long long c1, unsigned long long c2) {
    __shared__ unsigned long long red[3][kMsBlock / 64];
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    c0 = wave_sum(c0);
    c1 = wave_sum(c1);
    c2 = wave_sum(c2);
    if (lane == 0) {
        red[0][wid] = c0;
        red[1][wid] = c1;
        red[2][wid] = c2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kMsBlock / 64; ++w) {
            c0 += red[0][w];
            c1 += red[1][w];
            c2 += red[2][w];
        }
        unsigned long long *slot = ms_copy(a, a.level + 1) + (blockIdx.x % kSlots) * kLine;
        if (c0) atomicAdd(&slot[0], c0);
        if (c1) atomicAdd(&slot[1], c1);
        if (c2) atomicAdd(&slot[2], c2);
    }
}

// File the slices of the lanes' new frontier vertices whose lists have
// kMsWaveDeg entries or more (frontier `level`, so slices[level & 1] and
// counter copy level % 3).  Wave-collective; `wave` numbers the caller's wave.
__device__ __forceinline__ void ms_file_slices(const ms_args &a, int level, int64_t wave,
                                               bool is_new, int64_t v, int64_t b, int64_t e) {
    const int lane = lane_id();
    const int64_t deg = e - b;
    const int32_t ns = (is_new && deg >= kMsWaveDeg) ? (int32_t)((deg + kMsSlice - 1) / kMsSlice) : 0;
    if (!__ballot(ns > 0)) return;
    int32_t excl = wave_incl_scan32(ns);
    const int32_t total = lane_bcast(excl, 63);
    excl -= ns;
    const int seg = (int)(wave % kSlots);
    unsigned long long *copy = ms_copy(a, level);
    int64_t base = 0;
    if (lane == 0) {
        const unsigned long long p =
            atomicAdd(&copy[seg * kLine + kMsSegCount], (unsigned long long)total);
        if ((int64_t)p + total <= a.seg_cap) {
            base = seg * a.seg_cap + (int64_t)p;
        } else {
            // the segment is full: leave its count as it was and use the
            // overflow segment (no slot of the segment is written past its end)
            atomicAdd(&copy[seg * kLine + kMsSegCount], (unsigned long long)(-(int64_t)total));
            const unsigned long long q =
                atomicAdd(&copy[kMsOvfCount], (unsigned long long)total);
            base = kSlots * a.seg_cap + (int64_t)q;
        }
    }
    base = lane_bcast(base, 0);
    ms_slice *out = a.slices[level & 1];
    for (int32_t k = 0; k < ns; ++k) {
        ms_slice sl;
        sl.beg = b + (int64_t)k * kMsSlice;
        sl.len = (int32_t)(e - sl.beg < kMsSlice ? e - sl.beg : kMsSlice);
        sl.v = (int32_t)v;
        out[base + excl + k] = sl;
    }
}

// The state of level `level` as every workgroup sees it: the counters of
// its frontier (summed over the slots by the first wave) and the search's
// done flag, loaded together.  push: the mode Beamer's rule picks.
struct ms_state {
    unsigned long long front, fdeg, open;
    int32_t done;
    int32_t push;
};

__device__ __forceinline__ void ms_load_state(const ms_args &a, ms_state &st) {
    __shared__ ms_state sh;
    if (threadIdx.x < 64) {
        const int l = lane_id();
        const unsigned long long *c = ms_copy(a, a.level) + l * kLine;
        const unsigned long long f = c[0], d = c[1], o = c[2];
        const int32_t done = __hip_atomic_load(&a.ctl->done, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long fs = wave_sum(f), ds = wave_sum(d), os = wave_sum(o);
        if (l == 0) {
            sh.front = fs;
            sh.fdeg = ds;
            sh.open = os;
            sh.done = done || fs == 0;
            sh.push = ds * a.alpha < os;
        }
    }
    __syncthreads();
    st = sh;
}

// Roots start their own frontier: bit i on roots[i] (graph ids, mapped to
// search ids).  Counts frontier 0 into copy 0.
__global__ void k_ms_init(ms_args a, const int64_t *__restrict__ roots, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    if (i < n) {
        int64_t v = roots[i];
        if (a.inv) v = a.inv[v];
        const uint64_t bit = 1ULL << i;
        const unsigned long long old = atomicOr((unsigned long long *)&a.F[v], bit);
        atomicOr((unsigned long long *)&a.seen[v], bit);
        atomicOr((unsigned long long *)&a.Fb[v >> 6], 1ULL << (v & 63));
        if (a.depth) a.depth[(int64_t)i * a.nv + v] = 0;
        if (old == 0) {  // first root on this vertex
            c0 = 1;
            c1 = (unsigned long long)(a.off[v + 1] - a.off[v]);
        }
    }
    {
        int64_t v = 0, b = 0, e = 0;
        if (c0) {
            v = roots[i];
            if (a.inv) v = a.inv[v];
            b = a.off[v];
            e = a.off[v + 1];
        }
        ms_file_slices(a, 0, blockIdx.x, c0 != 0, v, b, e);
    }
    // open degree: every vertex is open before the search
    if (i == 0) c2 = (unsigned long long)a.off[a.nv];
    c0 = wave_sum(c0);
    c1 = wave_sum(c1);
    c2 = wave_sum(c2);
    if (lane_id() == 0) {
        if (c0) atomicAdd(&a.cnt[0], c0);
        if (c1) atomicAdd(&a.cnt[1], c1);
        if (c2) atomicAdd(&a.cnt[2], c2);
    }
}

// Push: expand the frontier into N with device atomics.  Lists shorter
// than a wave are laid end to end, per group of 64 vertices, and walked 64
// entries at a time, one per lane (the owner found by a search over the
// lanes' prefix); longer lists were filed as slices when their vertex
// joined the frontier, and every wave of the grid takes slices in turn.
__device__ __forceinline__ void ms_push(const ms_args &a) {
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kMsBlock + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kMsBlock) >> 6;
    for (int64_t base = wave * 64; base < a.nv; base += nwaves * 64) {
        const uint64_t fw = a.Fb[base >> 6];
        if (fw == 0) continue;  // no frontier vertex among these 64
        const int64_t v = base + lane;
        const uint64_t f = ((fw >> lane) & 1) ? a.F[v] : 0ULL;
        int64_t b = 0, e = 0;
        if (f) {
            b = a.off[v];
            e = a.off[v + 1];
        }
        const int32_t len = (int32_t)(f && e - b < kMsWaveDeg ? e - b : 0);
        int32_t excl = wave_incl_scan32(len);
        const int32_t total = lane_bcast(excl, 63);
        excl -= len;
        for (int32_t k0 = 0; k0 < total; k0 += 64) {
            const int32_t idx = k0 + lane;
            int i = 0;
#pragma unroll
            for (int st = 32; st > 0; st >>= 1) {
                const int mid = i + st;
                const int32_t em = __shfl(excl, mid < 64 ? mid : 63);
                if (mid < 64 && em <= idx) i = mid;
            }
            // every lane takes part in the shuffles (a lane past the end
            // still reads for the others)
            const int64_t bi = __shfl(b, i);
            const uint64_t fi = (uint64_t)__shfl((long long)f, i);
            const int32_t pos = idx - __shfl(excl, i);
            if (idx < total) ms_push_to(a, ms_col(a, bi + pos), fi);
        }
    }
    // the slices: segment counts and their prefix, read by every lane
    const unsigned long long *copy = ms_copy(a, a.level);
    const int64_t cs = (int64_t)copy[lane * kLine + kMsSegCount];
    const int64_t co = (int64_t)copy[kMsOvfCount];
    const int64_t incl = (int64_t)wave_scan64((uint64_t)cs, [](uint64_t x, uint64_t y) { return x + y; });
    const int64_t nseg = lane_bcast(incl, 63);
    const int64_t nslices = nseg + co;
    const ms_slice *in = a.slices[a.level & 1];
    for (int64_t t = wave; t < nslices; t += nwaves) {
        int64_t at;
        if (t < nseg) {
            // segment: the lane whose inclusive prefix first exceeds t
            const int sg = __ffsll((unsigned long long)__ballot(incl > t)) - 1;
            at = (int64_t)sg * a.seg_cap + (t - (lane_bcast(incl, sg) - lane_bcast(cs, sg)));
        } else {
            at = kSlots * a.seg_cap + (t - nseg);
        }
        const ms_slice sl = in[at];
        const uint64_t f = a.F[sl.v];
        for (int32_t k = lane; k < sl.len; k += 64) ms_push_to(a, ms_col(a, sl.beg + k), f);
    }
}

// Commit after a push: keep the unseen bits of N, mark them seen, record
// depths, clear F (it becomes the next push's target), count.
__device__ __forceinline__ void ms_commit(const ms_args &a) {
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    const int64_t stride = (int64_t)gridDim.x * kMsBlock;
    const int64_t nround = (a.nv + 63) / 64 * 64;
    for (int64_t v = (int64_t)blockIdx.x * kMsBlock + threadIdx.x; v < nround; v += stride) {
        const bool valid = v < a.nv;
        uint64_t nf = 0;
        if (valid) {
            const uint64_t n = a.N[v];
            const uint64_t s = a.seen[v];
            if (a.F[v]) a.F[v] = 0;
            nf = n & ~s;
            if (n != nf) a.N[v] = nf;
            if (nf) a.seen[v] = s | nf;
            const uint64_t open = ~(s | nf) & a.full;
            if (nf || open) {
                const unsigned long long d = (unsigned long long)(a.off[v + 1] - a.off[v]);
                if (nf) {
                    c0 += 1;
                    c1 += d;
                }
                if (open) c2 += d;
            }
        }
        const uint64_t word = __ballot(nf != 0);
        if (lane_id() == 0 && v < a.nv) a.Nb[v >> 6] = word;
        ms_record(a, v, valid, nf, a.level + 1);
        int64_t b = 0, e = 0;
        if (nf) {
            b = a.off[v];
            e = a.off[v + 1];
        }
        ms_file_slices(a, a.level + 1, v >> 6, nf != 0, v, b, e);
    }
    ms_count(a, c0, c1, c2);
}

// Pull: every open vertex ORs its neighbours' frontier masks, stopping once
// all the roots it still waits for are covered; writes N for every vertex.
__device__ __forceinline__ void ms_pull(const ms_args &a) {
    const int lane = lane_id();
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    const int64_t wave = ((int64_t)blockIdx.x * kMsBlock + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kMsBlock) >> 6;
    // slices of the open hubs; seen[v] of a hub does not change during the
    // launch, N[v] does (other slices OR into it), so lane 0's copy is used
    for (int64_t i = wave; i < a.nhub_slices; i += nwaves) {
        const ms_slice sl = a.hub_slices[i];
        const uint64_t want = ~a.seen[sl.v] & a.full;
        if (!want) continue;
        uint64_t got = (uint64_t)lane_bcast((int64_t)a.N[sl.v], 0);
        const int64_t end = sl.beg + sl.len;
        for (int64_t k0 = sl.beg; k0 < end && (got & want) != want; k0 += 64) {
            const int64_t k = k0 + lane;
            got |= wave_or(k < end ? ms_frontier(a, ms_col(a, k)) : 0ULL);
        }
        if (lane == 0 && (got & want))
            atomicOr((unsigned long long *)&a.N[sl.v], (unsigned long long)(got & want));
    }
    for (int64_t base = wave * 64; base < a.nv; base += nwaves * 64) {
        const int64_t v = base + lane;
        const bool valid = v < a.nv;
        uint64_t s = 0, want = 0;
        int64_t b = 0, e = 0;
        if (valid) {
            s = a.seen[v];
            want = ~s & a.full;
            if (want) {
                b = a.off[v];
                e = a.off[v + 1];
            }
        }
        if (e - b >= kMsHubDeg) want = 0;  // a hub: its slices, then k_ms_post
        uint64_t acc = 0;
        const bool alone = want && e - b < kMsWaveDeg;
        if (alone) {
            int64_t k = b;
            for (; k + 4 <= e && (acc & want) != want; k += 4) {
                const int32_t u0 = ms_col(a, k), u1 = ms_col(a, k + 1);
                const int32_t u2 = ms_col(a, k + 2), u3 = ms_col(a, k + 3);
                acc |= ms_frontier(a, u0) | ms_frontier(a, u1) | ms_frontier(a, u2) |
                       ms_frontier(a, u3);
            }
            for (; k < e && (acc & want) != want; ++k) acc |= ms_frontier(a, ms_col(a, k));
        }
        uint64_t lm = __ballot(want && !alone);
        while (lm) {
            const int src = __ffsll((unsigned long long)lm) - 1;
            lm &= lm - 1;
            const int64_t kb = lane_bcast(b, src), ke = lane_bcast(e, src);
            const uint64_t ws = (uint64_t)lane_bcast((int64_t)want, src);
            uint64_t got = 0;
            for (int64_t k0 = kb; k0 < ke && (got & ws) != ws; k0 += 64) {
                const int64_t k = k0 + lane;
                got |= wave_or(k < ke ? ms_frontier(a, ms_col(a, k)) : 0ULL);
            }
            if (lane == src) acc = got;
        }
        const uint64_t nf = acc & want;
        const uint64_t word = __ballot(nf != 0);
        if (lane == 0 && base < a.nv) a.Nb[base >> 6] = word;
        ms_file_slices(a, a.level + 1, wave, nf != 0, v, b, e);
        if (valid && e - b < kMsHubDeg) {
            a.N[v] = nf;
            if (nf) {
                a.seen[v] = s | nf;
                c0 += 1;
                c1 += (unsigned long long)(e - b);
            }
            if (want & ~nf) c2 += (unsigned long long)(e - b);
        }
        ms_record(a, v, valid, nf, a.level + 1);
    }
    ms_count(a, c0, c1, c2);
}

// After a pull: the hubs, whose slices ORed the roots found into N[v] (only
// bits v still waited for), join the next frontier as the pull's own
// vertices do -- seen, the Nb bit (the pull wrote Nb's words whole, so the
// bit is ORed in now), the counters, push slices and depths.
__device__ __forceinline__ void ms_finish_hubs(const ms_args &a) {
    const int lane = lane_id();
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    const int64_t wave = ((int64_t)blockIdx.x * kMsBlock + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kMsBlock) >> 6;
    for (int64_t base = wave * 64; base < a.nhubs; base += nwaves * 64) {
        const bool valid = base + lane < a.nhubs;
        int64_t v = 0, b = 0, e = 0;
        uint64_t want = 0, nf = 0;
        if (valid) {
            v = a.hubs[base + lane];
            const uint64_t s = a.seen[v];
            want = ~s & a.full;
            if (want) {
                b = a.off[v];
                e = a.off[v + 1];
                nf = a.N[v];
                if (nf) {
                    a.seen[v] = s | nf;
                    atomicOr((unsigned long long *)&a.Nb[v >> 6], 1ULL << (v & 63));
                    c0 += 1;
                    c1 += (unsigned long long)(e - b);
                }
                if (want & ~nf) c2 += (unsigned long long)(e - b);
            }
        }
        ms_file_slices(a, a.level + 1, wave, nf != 0, v, b, e);
        ms_record(a, v, valid, nf, a.level + 1);
    }
    ms_count(a, c0, c1, c2);
}

// Level L, first launch: the state, the record of levels L-1 and L
// (workgroup 0), then the push or the pull.
__global__ __launch_bounds__(kMsBlock) void k_ms_level(ms_args a) {
    ms_state st;
    ms_load_state(a, st);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const bool was_done = a.ctl->done != 0;
        if (!was_done && a.level > 0 && a.level - 1 < a.log_cap) {
            a.log[a.level - 1].next = (int64_t)st.front;
            a.log[a.level - 1].next_deg = (int64_t)st.fdeg;
        }
        if (!was_done && st.front == 0) {
            a.ctl->done = 1;
            a.ctl->levels = a.level;
        }
        if (!st.done && a.level < a.log_cap) {
            bd_level_rec r;
            memset(&r, 0, sizeof(r));
            r.level = a.level;
            r.dir = st.push ? 0 : 1;
            r.frontier = (int64_t)st.front;
            r.frontier_deg = (int64_t)st.fdeg;
            r.unvisited_deg = (int64_t)st.open;
            r.ms = -1.0;
            r.ms_other = -1.0;
            r.launch = 2 * a.level;
            r.launches = 2;
            a.log[a.level] = r;
        }
    }
    if (st.done) return;
    if (st.push)
        ms_push(a);
    else
        ms_pull(a);
}

// Level L, second launch: the commit after a push; after a pull, the old
// frontier's masks are cleared (they become the next level's N).  Workgroup
// 0 clears the counters copy that level L+1 fills.
__global__ __launch_bounds__(kMsBlock) void k_ms_post(ms_args a) {
    ms_state st;
    ms_load_state(a, st);
    if (st.done) return;
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        unsigned long long *c = ms_copy(a, a.level + 2) + threadIdx.x * kLine;
        c[0] = 0;
        c[1] = 0;
        c[2] = 0;
        c[kMsSegCount] = 0;
        c[kMsOvfCount] = 0;
    }
    if (st.push) {
        ms_commit(a);
    } else {
        const int64_t nw = (a.nv + 63) / 64;
        for (int64_t w = (int64_t)blockIdx.x * kMsBlock + threadIdx.x; w < nw;
             w += (int64_t)gridDim.x * kMsBlock)
            for (uint64_t m = a.Fb[w]; m; m &= m - 1) a.F[w * 64 + __ffsll((unsigned long long)m) - 1] = 0;
        ms_finish_hubs(a);
    }
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = x;
}

__global__ void k_fill_u64(uint64_t *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// The hubs (lists of kMsHubDeg entries or more) and their lists as slices,
// in the order atomics hand out places: a pull ORs what the slices find, so
// the order changes no result.
__global__ void k_ms_hubs(const int64_t *__restrict__ off, int64_t nv, int32_t *__restrict__ hubs,
                          ms_slice *__restrict__ sl, unsigned long long *__restrict__ cnt) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = off[v], e = off[v + 1];
        if (e - b < kMsHubDeg) continue;
        hubs[atomicAdd(&cnt[0], 1ULL)] = (int32_t)v;
        const int64_t ns = (e - b + kMsSlice - 1) / kMsSlice;
        const int64_t p = (int64_t)atomicAdd(&cnt[1], (unsigned long long)ns);
        for (int64_t k = 0; k < ns; ++k) {
            ms_slice x;
            x.beg = b + k * kMsSlice;
            x.len = (int32_t)std::min<int64_t>(kMsSlice, e - x.beg);
            x.v = (int32_t)v;
            sl[p + k] = x;
        }
    }
}

// Row r of the depths in graph ids: out[v] = depth[r * nv + inv[v]].
__global__ void k_ms_row(const int32_t *__restrict__ depth, const int32_t *__restrict__ inv,
                         int64_t nv, int32_t *__restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        out[v] = depth[inv[v]];
}

// Masks, bitmaps, three counter copies with the control block, and (nrows
// > 0) the depth rows.
int ms_buffers(bd_ctx *c, int nrows) {
    int rc;
    for (int i = 0; i < 3; ++i)
        if (!c->d_ms[i] &&
            (rc = dev_alloc(c, (void **)&c->d_ms[i], sizeof(uint64_t) * c->nv, "root masks")))
            return rc;
    if (!c->d_ms[3] &&
        (rc = dev_alloc(c, (void **)&c->d_ms[3],
                        sizeof(unsigned long long) * 3 * kSlots * kLine + sizeof(ms_ctl),
                        "mask counters")))
        return rc;
    for (int i = 0; i < 2; ++i)
        if (!c->d_ms_bits[i] && (rc = dev_alloc(c, (void **)&c->d_ms_bits[i],
                                                sizeof(uint64_t) * c->nwords, "frontier bits")))
            return rc;
    if (!c->d_ms_slices[0]) {
        // slices of one frontier: ceil(deg / kMsSlice) per vertex of degree
        // >= kMsWaveDeg, so at most nnz / kMsSlice + nnz / kMsWaveDeg
        c->ms_max_slices = c->nnz / kMsSlice + c->nnz / kMsWaveDeg + 1;
        c->ms_seg_cap = 2 * ((c->ms_max_slices + kSlots - 1) / kSlots) + 4096;
        for (int i = 0; i < 2; ++i)
            if ((rc = dev_alloc(c, (void **)&c->d_ms_slices[i],
                                sizeof(ms_slice) * (kSlots * c->ms_seg_cap + c->ms_max_slices),
                                "push slices")))
                return rc;
    }
    if (!c->d_ms_hubs || c->ms_hub_order != c->s_order) {
        // the hubs of the search layout, once per graph and layout
        dev_free(c->d_ms_hubs);
        dev_free(c->d_ms_hub_slices);
        c->d_ms_hubs = nullptr;
        c->d_ms_hub_slices = nullptr;
        const int64_t max_hubs = c->nnz / kMsHubDeg + 1;
        const int64_t max_sl = c->nnz / kMsSlice + max_hubs;
        if ((rc = dev_alloc(c, (void **)&c->d_ms_hubs, sizeof(int32_t) * max_hubs, "hubs")))
            return rc;
        if ((rc = dev_alloc(c, &c->d_ms_hub_slices, sizeof(ms_slice) * max_sl, "hub slices")))
            return rc;
        unsigned long long *cnt = (unsigned long long *)c->d_ms[3];
        BD_HIP(c, hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), c->stream));
        hipLaunchKernelGGL(k_ms_hubs, dim3(grid_for(c, c->nv, 256)), dim3(256), 0, c->stream,
                           c->d_soff, c->nv, c->d_ms_hubs, (ms_slice *)c->d_ms_hub_slices, cnt);
        BD_HIP(c, hipGetLastError());
        unsigned long long h[2] = {0, 0};
        BD_HIP(c, hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        c->ms_nhubs = (int64_t)h[0];
        c->ms_nhub_slices = (int64_t)h[1];
        c->ms_hub_order = c->s_order;
    }
    if (nrows > c->ms_depth_rows) {
        dev_free(c->d_ms_depth);
        c->d_ms_depth = nullptr;
        c->ms_depth_rows = 0;
        if ((rc = dev_alloc(c, (void **)&c->d_ms_depth, sizeof(int32_t) * c->nv * nrows,
                            "multi-root depths")))
            return rc;
        c->ms_depth_rows = nrows;
    }
    return BD_OK;
}

// The two launches of level L (F and N swap roles with the level's parity).
int ms_enqueue_level(bd_ctx *c, ms_args a, int64_t L, int64_t grid) {
    a.level = (int32_t)L;
    uint64_t *m[2] = {c->d_ms[0], c->d_ms[1]};
    uint64_t *bits[2] = {c->d_ms_bits[0], c->d_ms_bits[1]};
    a.F = m[L & 1];
    a.N = m[(L + 1) & 1];
    a.Fb = bits[L & 1];
    a.Nb = bits[(L + 1) & 1];
    hipLaunchKernelGGL(k_ms_level, dim3(grid), dim3(kMsBlock), 0, c->stream, a);
    hipLaunchKernelGGL(k_ms_post, dim3(grid), dim3(kMsBlock), 0, c->stream, a);
    BD_HIP(c, hipGetLastError());
    return BD_OK;
}

}  // namespace

void ms_free_graph(bd_ctx *c) {
    if (c->ms_graph) (void)hipGraphExecDestroy((hipGraphExec_t)c->ms_graph);
    c->ms_graph = nullptr;
    c->ms_graph_key.clear();
}

int run_multi_root(bd_ctx *c, const int64_t *roots, int n, int32_t *depths, double *ms,
                   int64_t *levels) {
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph");
    if (n < 1 || n > 64) return set_err(c, BD_EINVAL, "1 to 64 roots per bit-parallel search");
    if (!roots) return set_err(c, BD_EINVAL, "no roots");
    for (int i = 0; i < n; ++i)
        if (roots[i] < 0 || roots[i] >= c->nv)
            return set_err(c, BD_ERANGE, "Source vertex out of range.");
    BD_HIP(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ms_buffers(c, depths ? n : 0))) return rc;
    int64_t *d_roots = nullptr;
    if ((rc = dev_alloc(c, (void **)&d_roots, sizeof(int64_t) * n, "roots"))) return rc;
    struct guard {
        void *p;
        ~guard() { dev_free(p); }
    } g{d_roots};
    const int64_t nv = c->nv;
    ms_args a;
    a.off = c->d_soff;
    a.cols = c->d_scols;
    a.inv = c->d_inv;
    a.F = c->d_ms[0];
    a.N = c->d_ms[1];
    a.seen = c->d_ms[2];
    a.Fb = c->d_ms_bits[0];
    a.Nb = c->d_ms_bits[1];
    a.depth = depths ? c->d_ms_depth : nullptr;
    a.cnt = (unsigned long long *)c->d_ms[3];
    a.ctl = (ms_ctl *)(a.cnt + 3 * kSlots * kLine);
    a.log = c->d_log;
    a.log_cap = c->log_cap;
    a.slices[0] = (ms_slice *)c->d_ms_slices[0];
    a.slices[1] = (ms_slice *)c->d_ms_slices[1];
    a.seg_cap = c->ms_seg_cap;
    a.max_slices = c->ms_max_slices;
    a.hubs = c->d_ms_hubs;
    a.hub_slices = (const ms_slice *)c->d_ms_hub_slices;
    a.nhubs = c->ms_nhubs;
    a.nhub_slices = c->ms_nhub_slices;
    a.nv = nv;
    a.full = n == 64 ? ~0ULL : ((1ULL << n) - 1);
    a.alpha = (uint64_t)std::max(1, c->tune.ms_alpha);
    a.nroots = n;
    a.level = 0;
    const int64_t grid = (int64_t)c->cu_count * std::max(1, c->tune.grid_per_cu);
    // clear(): untimed, as in the reference's trial loop; kernels only
    BD_HIP(c, hipMemcpyAsync(d_roots, roots, sizeof(int64_t) * n, hipMemcpyHostToDevice,
                             c->stream));
    for (int i = 0; i < 3; ++i)
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(c, nv, 256)), dim3(256), 0, c->stream,
                           c->d_ms[i], nv);
    for (int i = 0; i < 2; ++i)
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(c, c->nwords, 256)), dim3(256), 0,
                           c->stream, c->d_ms_bits[i], c->nwords);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, c->stream, (uint64_t *)a.cnt,
                       (int64_t)(3 * kSlots * kLine + sizeof(ms_ctl) / 8));
    if (a.depth) {
        const int64_t cells = (int64_t)n * nv;
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c, cells, 256)), dim3(256), 0, c->stream,
                           a.depth, cells, -1);
    }
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipStreamSynchronize(c->stream));

    // Level launches: a captured chunk of kMsChunk levels, replayed (and kept
    // for the next search with the same buffers); host-driven (level_driver
    // 2): the same launches, enqueued and waited for level by level.
    const bool host_driven = c->tune.level_driver == 2;
    char key[160];
    snprintf(key, sizeof(key), "%p %p %p %p %ld %ld %d %lu %ld", (void *)a.off, (void *)a.depth,
             (void *)a.cnt, (const void *)a.hub_slices, (long)a.nhub_slices, (long)nv, n,
             (unsigned long)a.alpha, (long)grid);
    if (!host_driven && (!c->ms_graph || c->ms_graph_key != key)) {
        ms_free_graph(c);
        hipGraph_t graph = nullptr;
        BD_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int erc = BD_OK;
        for (int64_t L = 0; L < kMsChunk && !erc; ++L) erc = ms_enqueue_level(c, a, L, grid);
        const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        if (erc) {
            if (graph) (void)hipGraphDestroy(graph);
            return erc;
        }
        BD_HIP(c, ce);
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        BD_HIP(c, ie);
        c->ms_graph = (void *)exec;
        c->ms_graph_key = key;
    }
    BD_HIP(c, hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(k_ms_init, dim3((n + 63) / 64), dim3(64), 0, c->stream, a, d_roots, n);
    BD_HIP(c, hipGetLastError());
    ms_ctl h{0, 0};
    int64_t L = 0;
    std::vector<float> level_ms;  // host-driven: each level's two launches, timed by events
    for (;;) {
        if (L > nv + 2) return set_err(c, BD_EHIP, "multi-root search did not terminate");
        if (host_driven) {
            BD_HIP(c, hipEventRecord(c->ev[2], c->stream));
            if ((rc = ms_enqueue_level(c, a, L, grid))) return rc;
            BD_HIP(c, hipEventRecord(c->ev[3], c->stream));
            BD_HIP(c, hipEventSynchronize(c->ev[3]));
            float lm = 0;
            BD_HIP(c, hipEventElapsedTime(&lm, c->ev[2], c->ev[3]));
            level_ms.push_back(lm);
            L += 1;
        } else if (L == 0) {
            BD_HIP(c, hipGraphLaunch((hipGraphExec_t)c->ms_graph, c->stream));
            L = kMsChunk;
        } else {
            // past the captured chunk (deep searches): plain launches
            for (int k = 0; k < kMsChunk; ++k, ++L)
                if ((rc = ms_enqueue_level(c, a, L, grid))) return rc;
        }
        BD_HIP(c, hipMemcpyAsync(&h, a.ctl, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        if (h.done) break;
    }
    BD_HIP(c, hipEventRecord(c->ev[1], c->stream));
    BD_HIP(c, hipEventSynchronize(c->ev[1]));
    float t = 0;
    BD_HIP(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    if (ms) *ms = t;
    if (levels) *levels = h.levels;
    const int64_t nrec = std::min<int64_t>(h.levels, c->log_cap);
    c->last_log.resize(nrec);
    if (nrec)
        BD_HIP(c, hipMemcpy(c->last_log.data(), c->d_log, sizeof(bd_level_rec) * nrec,
                            hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nrec && i < (int64_t)level_ms.size(); ++i) c->last_log[i].ms = level_ms[i];
    c->have_result = false;  // bd_bfs_levels now describes this search, not the last single one
    if (depths) {
        if (!c->d_inv) {
            BD_HIP(c, hipMemcpy(depths, a.depth, sizeof(int32_t) * n * nv, hipMemcpyDeviceToHost));
        } else {
            int32_t *row = nullptr;
            if ((rc = dev_alloc(c, (void **)&row, sizeof(int32_t) * nv, "depth row"))) return rc;
            struct rguard {
                void *p;
                ~rguard() { dev_free(p); }
            } rg{row};
            for (int r = 0; r < n; ++r) {
                hipLaunchKernelGGL(k_ms_row, dim3(grid_for(c, nv, 256)), dim3(256), 0, c->stream,
                                   a.depth + (int64_t)r * nv, c->d_inv, nv, row);
                BD_HIP(c, hipGetLastError());
                BD_HIP(c, hipMemcpyAsync(depths + (int64_t)r * nv, row, sizeof(int32_t) * nv,
                                         hipMemcpyDeviceToHost, c->stream));
            }
            BD_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    return BD_OK;
}

}  // namespace bd

using namespace bd;

extern "C" int bd_bfs_multi_root(bd_ctx *c, const int64_t *roots, int n, int32_t *depths,
                                 double *ms, int64_t *levels) {
    if (!c) return BD_EINVAL;
    return run_multi_root(c, roots, n, depths, ms, levels);
}
