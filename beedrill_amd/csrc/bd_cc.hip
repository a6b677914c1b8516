// bd_cc.hip -- connected components on gfx950 (components::run,
// components.cc:38-104, and components::check, :152-202).
//
// Labels are int32[V]; every vertex starts as its own component.  An
// iteration is an edge pass and a shortcut pass, as in the reference:
//   edge pass (:56-66): each vertex takes the smallest label among its
//     neighbours when that is smaller than its own.  The CSR is symmetric,
//     so every directed edge (src -> dst) lowers label[src] in src's own
//     pull: a short list's label is written only by its vertex's lane, with
//     no atomics; long lists are cut into slices merged by atomicMin.
//   shortcut pass (:74-78): label[v] = label[label[v]] until it is a root.
// The run ends after the first edge pass that changes nothing (:70); the
// iteration count includes that pass.  Labels only decrease and stay ids of
// the vertex's own component, so each component ends labelled by its
// smallest vertex id, whatever order the passes read the labels in.
//
// Once few labels change, a pass only needs the labels of neighbours that
// changed in the previous pass or shortcut (a bitmap of changed vertices):
// an unchanged label was read by the vertex before, when it last changed,
// and the vertex's own label has only gone down since, so it cannot lower
// it.  Such a pass runs as a push: every changed vertex offers its label to
// its neighbours (atomicMin).  The CSR is symmetric, so these are the same
// (vertex, neighbour) pairs; the pass's outcome is the same, its cost is
// the changed vertices' edges instead of all of them.  Passes that follow
// many changes (the edges of changed vertices at least half of all edges)
// pull, with no atomics.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <vector>

#include "bd_device.h"
#include "bd_internal.h"

namespace bd {
namespace {

// Every vertex its own component; both changed-vertex bitmaps empty.
__global__ void k_cc_init(int32_t *__restrict__ comp, int64_t nv, uint64_t *__restrict__ chg0,
                          uint64_t *__restrict__ chg1) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        comp[v] = (int32_t)v;
        if ((v & 63) == 0) {
            chg0[v >> 6] = 0;
            chg1[v >> 6] = 0;
        }
    }
}

// Lists of kCcHeavy edges or more are cut into slices of kCcSlice edges,
// one wave each, merged into the label with atomicMin; shorter lists are
// walked 64 edges at a time by the wave that owns their vertex.
constexpr int kCcHeavy = 256;
constexpr int kCcSlice = 2048;

__device__ __forceinline__ int32_t shfl(int32_t x, int src) {
    return __builtin_amdgcn_ds_bpermute(src << 2, x);
}
__device__ __forceinline__ int64_t shfl(int64_t x, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(uint64_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Slices of the heavy lists; built once per graph.  *count ends as the
// number of slices the graph needs, even past cap (then only cap are kept).
__global__ void k_cc_tasks(const int64_t *__restrict__ off, int64_t nv, td_task *__restrict__ tasks,
                           unsigned long long *__restrict__ count, int64_t cap) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = off[v], e = off[v + 1];
        if (e - b < kCcHeavy) continue;
        const int64_t n = (e - b + kCcSlice - 1) / kCcSlice;
        const int64_t at = (int64_t)atomicAdd(count, (unsigned long long)n);
        for (int64_t i = 0; i < n && at + i < cap; ++i) {
            td_task t;
            t.beg = b + i * kCcSlice;
            t.end = t.beg + kCcSlice < e ? t.beg + kCcSlice : e;
            t.src = (int32_t)v;
            t.pad = 0;
            tasks[at + i] = t;
        }
    }
}

// Pull edge pass.  Units [0, ntasks) are heavy slices; unit ntasks + g is
// the vertex group [64g, 64g + 64), whose light lists are consecutive in
// cols (heavy lists between them are skipped).  The group's light edges are
// numbered q = 0..T-1 (light vertex i owns q in [s_i, s_i + d_i)) and taken
// 64 at a time, one per lane, so every lane loads whatever the degrees:
//   owner   o(q) = the last i with s_i <= q   (binary search over lanes)
//   edge    k    = off[base + o] + q - s_o
//   label   l    = comp[cols[k]]
// Owners are non-decreasing across the lanes of a chunk, so a suffix min
// restricted to equal owners leaves the chunk's min for owner i in the
// lane where i's run starts, max(s_i - t, 0), which lane i reads.
// Every vertex reads every neighbour's label.  In the first pass
// (first != 0) the neighbour's id stands in for its label: every label is
// still its vertex's id when the pass starts, and a pass may read a
// neighbour's label from before that neighbour's own update in the same
// pass, so this is one of the pass's own outcomes, with no label gathered.
// A vertex whose label goes down is set in chg_out and its degree (a
// slice's length for a long list) added to chg_deg; changed[0] is set if
// any label went down; labels_read counts the labels gathered.
__global__ __launch_bounds__(256) void k_cc_hook(const int64_t *__restrict__ off,
                                                 const int32_t *__restrict__ cols, int32_t *comp,
                                                 int64_t nv, const td_task *__restrict__ tasks,
                                                 int64_t ntasks, int first,
                                                 uint64_t *__restrict__ chg_out,
                                                 int32_t *__restrict__ changed,
                                                 unsigned long long *__restrict__ labels_read,
                                                 unsigned long long *__restrict__ chg_deg,
                                                 unsigned long long *__restrict__ edge_bytes) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t nunits = ntasks + (nv + 63) / 64;
    bool ch = false;
    // nbytes: algorithmic bytes besides the 8 B (column id + label) of each
    // label read
    unsigned long long nread = 0, ndeg = 0, nbytes = 0;
    for (int64_t u = wave; u < nunits; u += nwaves) {
        if (u < ntasks) {
            const int64_t beg = tasks[u].beg, end = tasks[u].end;
            const int32_t src = tasks[u].src;
            uint32_t m = 0xffffffffu;
            for (int64_t k0 = beg + lane; k0 < end; k0 += 4 * 64) {
                int32_t c[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t k = k0 + r * 64;
                    c[r] = __builtin_nontemporal_load(cols + (k < end ? k : end - 1));
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (k0 + r * 64 < end) {
                        if (first) {
                            const uint32_t l = (uint32_t)c[r];
                            m = l < m ? l : m;
                            nbytes += 4;
                        } else {
                            const uint32_t l = (uint32_t)comp[c[r]];
                            m = l < m ? l : m;
                            ++nread;
                        }
                    }
                }
            }
            m = (uint32_t)wave_min(m);
            // a slice is never empty, so m is a label; checked all the same
            if (lane == 0 && m != 0xffffffffu && (int32_t)m < comp[src]) {
                if ((int32_t)m < atomicMin(comp + src, (int32_t)m)) {
                    ch = true;
                    ndeg += (unsigned long long)(end - beg);
                    atomicOr((unsigned long long *)chg_out + (src >> 6), 1ull << (src & 63));
                }
            }
            continue;
        }
        const int64_t v = (u - ntasks) * 64 + lane;
        int64_t b = 0, e = 0;
        int32_t mine = 0x7fffffff;
        if (v < nv) {
            b = off[v];
            e = off[v + 1];
            mine = comp[v];
            nbytes += 12;
        }
        const int32_t d = e - b < kCcHeavy ? (int32_t)(e - b) : 0;
        const int32_t s_end = wave_incl_scan32(d);
        const int32_t s = s_end - d;
        const int32_t total = __builtin_amdgcn_readlane(s_end, 63);
        const int64_t base_k = b - s;  // k = base_k of the owner + q
        int32_t acc = mine;
        for (int32_t t = 0; t < total; t += 64) {
            const int32_t q = t + lane;
            const int32_t qa = q < total ? q : total - 1;
            int o = 0;
#pragma unroll
            for (int step = 32; step; step >>= 1)
                if (shfl(s, o + step) <= qa) o += step;
            const int64_t k = shfl(base_k, o) + qa;
            // one past the owner's last lane in this chunk
            const int32_t run_end = min(shfl(s_end, o) - t, 64);
            const int32_t c = __builtin_nontemporal_load(cols + k);
            const bool live = q < total;
            int32_t m = 0x7fffffff;
            if (live) {
                if (first) {
                    m = c;
                    nbytes += 4;
                } else {
                    m = comp[c];
                    ++nread;
                }
            }
#pragma unroll
            for (int step = 1; step < 64; step <<= 1) {
                const int32_t other = shfl(m, (lane + step) & 63);
                if (lane + step < run_end && other < m) m = other;
            }
            const bool mine_here = d > 0 && s < t + 64 && s_end > t;
            const int32_t got = shfl(m, mine_here ? (s > t ? s - t : 0) : lane);
            if (mine_here && got < acc) acc = got;
        }
        const bool down = acc < mine;
        if (down) {
            comp[v] = acc;
            ndeg += (unsigned long long)d;
        }
        const uint64_t downs = __ballot(down);
        if (downs && lane == 0) {
            atomicOr((unsigned long long *)chg_out + (u - ntasks), (unsigned long long)downs);
            ch = true;
        }
    }
    if (__ballot(ch) && lane == 0) changed[0] = 1;
    nbytes = wave_sum(nbytes + 8 * nread);
    nread = wave_sum(nread);
    ndeg = wave_sum(ndeg);
    const int slot = (blockIdx.x % kSlots) * kLine;
    if (lane == 0 && nread) atomicAdd(&labels_read[slot], nread);
    if (lane == 0 && ndeg) atomicAdd(&chg_deg[slot], ndeg);
    if (lane == 0 && nbytes) atomicAdd(&edge_bytes[slot], nbytes);
}

// Pull passes tiled by XCD.  Vertex ids are cut into 8 blocks of 2^bshift
// (the last may be short).  Workgroups go to the 8 XCDs round-robin, so the
// workgroups with blockIdx % 8 == x share XCD x's L2: in phase p they pull
// row block x (their vertices) from column block (x + p) % 8 only, and the
// labels they read sit in one block (8 MB at 2^24 vertices) instead of all
// of the label array.  Eight phases make one pull pass; a vertex's label is
// written in each phase by the one wave that owns its group.  Short lists
// are regrouped once per graph so that each column block's neighbours are
// contiguous: bcols holds the regrouped list in the vertex's own CSR range,
// and byte c of bend[v] is where block c's run ends (0 for empty and long
// lists, whose slices pull as before).
__global__ void k_cc_blocks(const int64_t *__restrict__ off, const int32_t *__restrict__ cols,
                            int64_t nv, int bshift, int32_t *__restrict__ bcols,
                            uint64_t *__restrict__ bend) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = off[v], e = off[v + 1];
        if (e == b || e - b >= kCcHeavy) {
            bend[v] = 0;
            continue;
        }
        // per-block counts, one byte each (a short list has < 256 entries)
        uint64_t cnt = 0;
        for (int64_t k = b; k < e; ++k) cnt += 1ull << (8 * (cols[k] >> bshift));
        uint64_t ends = cnt;  // bytewise inclusive prefix
        ends += ends << 8;
        ends += ends << 16;
        ends += ends << 32;
        bend[v] = ends;
        uint64_t cur = ends - cnt;  // bytewise starts
        for (int64_t k = b; k < e; ++k) {
            const int32_t u = cols[k];
            const int sh = 8 * (u >> bshift);
            bcols[b + (int64_t)((cur >> sh) & 255)] = u;
            cur += 1ull << sh;
        }
    }
}

// Long lists for the tiled pulls: each is regrouped by column block like
// the short lists (a wave per list, found through its first slice), and
// each block's run is cut into pieces of kCcSlice edges, filed under bucket
// (row block of the vertex) * 8 + column block.  mode 0 counts the pieces
// of each bucket into cursor[]; mode 1 (cursor = the buckets' starts)
// regroups and files them.
__global__ __launch_bounds__(256) void k_cc_heavy_tiles(
    const int64_t *__restrict__ off, const int32_t *__restrict__ cols, int bshift,
    const td_task *__restrict__ tasks, int64_t ntasks, int32_t *__restrict__ bcols,
    unsigned long long *__restrict__ cursor, td_task *__restrict__ pieces, int64_t cap, int mode) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t u = wave; u < ntasks; u += nwaves) {
        const int32_t src = tasks[u].src;
        const int64_t b = off[src], e = off[(int64_t)src + 1];
        if (tasks[u].beg != b) continue;  // not the list's first slice
        int64_t cnt[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) cnt[c] = 0;
        for (int64_t k0 = b; k0 < e; k0 += 64) {
            const int64_t k = k0 + lane;
            const int blk = k < e ? cols[k] >> bshift : 8;
#pragma unroll
            for (int c = 0; c < 8; ++c) cnt[c] += __popcll(__ballot(blk == c));
        }
        const int row = src >> bshift;
        if (mode == 0) {
            if (lane == 0) {
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (cnt[c])
                        atomicAdd(cursor + row * 8 + c,
                                  (unsigned long long)((cnt[c] + kCcSlice - 1) / kCcSlice));
            }
            continue;
        }
        int64_t st[8], cur[8];
        int64_t acc = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            st[c] = acc;
            acc += cnt[c];
            cur[c] = 0;
        }
        for (int64_t k0 = b; k0 < e; k0 += 64) {
            const int64_t k = k0 + lane;
            const int32_t w = k < e ? cols[k] : 0;
            const int blk = k < e ? w >> bshift : 8;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const uint64_t mask = __ballot(blk == c);
                if (blk == c) bcols[b + st[c] + cur[c] + __popcll(mask & below)] = w;
                cur[c] += __popcll(mask);
            }
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (!cnt[c]) continue;
            const int64_t n = (cnt[c] + kCcSlice - 1) / kCcSlice;
            unsigned long long at = 0;
            if (lane == 0) at = atomicAdd(cursor + row * 8 + c, (unsigned long long)n);
            at = (unsigned long long)lane_bcast((int64_t)at, 0);
            for (int64_t i = lane; i < n; i += 64) {
                if ((int64_t)at + i >= cap) break;
                td_task t;
                t.beg = b + st[c] + i * kCcSlice;
                t.end = t.beg + kCcSlice < b + st[c] + cnt[c] ? t.beg + kCcSlice : b + st[c] + cnt[c];
                t.src = src;
                t.pad = 0;
                pieces[at + i] = t;
            }
        }
    }
}

// One phase of a tiled pull pass (see above).  Same work per group as the
// pull pass's short lists, restricted to one column block.
__global__ __launch_bounds__(256) void k_cc_tile(const int64_t *__restrict__ off,
                                                 const int32_t *__restrict__ bcols,
                                                 const uint64_t *__restrict__ bend, int32_t *comp,
                                                 int64_t nv, int bshift, int phase,
                                                 const td_task *__restrict__ pieces,
                                                 const int64_t *__restrict__ piece_base,
                                                 uint64_t *__restrict__ chg_out,
                                                 int32_t *__restrict__ changed,
                                                 unsigned long long *__restrict__ labels_read,
                                                 unsigned long long *__restrict__ chg_deg,
                                                 unsigned long long *__restrict__ edge_bytes) {
    const int lane = threadIdx.x & 63;
    const int xcd = blockIdx.x & 7;
    const int col = (xcd + phase) & 7;
    const int64_t wave = ((int64_t)(blockIdx.x >> 3) * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)(gridDim.x >> 3) * blockDim.x) >> 6;
    const int64_t r0 = (int64_t)xcd << bshift;
    const int64_t r1 = r0 + (1ll << bshift) < nv ? r0 + (1ll << bshift) : nv;
    bool ch = false;
    unsigned long long nread = 0, ndeg = 0, nbytes = 0;
    // long lists: this bucket's pieces, a wave each, merged by atomicMin
    const int bucket = xcd * 8 + col;
    for (int64_t u = piece_base[bucket] + wave; u < piece_base[bucket + 1]; u += nwaves) {
        const int64_t beg = pieces[u].beg, end = pieces[u].end;
        const int32_t src = pieces[u].src;
        uint32_t m = 0xffffffffu;
        for (int64_t k0 = beg + lane; k0 < end; k0 += 4 * 64) {
            int32_t cc[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t k = k0 + r * 64;
                cc[r] = __builtin_nontemporal_load(bcols + (k < end ? k : end - 1));
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (k0 + r * 64 < end) {
                    const uint32_t l = (uint32_t)comp[cc[r]];
                    m = l < m ? l : m;
                    ++nread;
                }
            }
        }
        m = (uint32_t)wave_min(m);
        if (lane == 0 && m != 0xffffffffu && (int32_t)m < comp[src]) {
            if ((int32_t)m < atomicMin(comp + src, (int32_t)m)) {
                ch = true;
                ndeg += (unsigned long long)(end - beg);
                atomicOr((unsigned long long *)chg_out + (src >> 6), 1ull << (src & 63));
            }
        }
    }
    for (int64_t g0 = r0 + wave * 64; g0 < r1; g0 += nwaves * 64) {
        const int64_t v = g0 + lane;
        int64_t b = 0;
        int32_t d = 0, deg = 0;
        int32_t mine = 0x7fffffff;
        if (v < r1) {
            const uint64_t e8 = bend[v];
            if (e8) {
                const int32_t st = col ? (int32_t)((e8 >> (8 * (col - 1))) & 255) : 0;
                d = (int32_t)((e8 >> (8 * col)) & 255) - st;
                deg = (int32_t)(e8 >> 56);
                b = off[v] + st;
            }
            mine = comp[v];
            // a vertex's offset and own label count once per pass, as in an
            // untiled pull; the phases' re-reads are the tiling's own cost
            if (phase == 0) nbytes += 12;
        }
        const int32_t s_end = wave_incl_scan32(d);
        const int32_t s = s_end - d;
        const int32_t total = __builtin_amdgcn_readlane(s_end, 63);
        const int64_t base_k = b - s;
        int32_t acc = mine;
        for (int32_t t = 0; t < total; t += 64) {
            const int32_t q = t + lane;
            const int32_t qa = q < total ? q : total - 1;
            int o = 0;
#pragma unroll
            for (int step = 32; step; step >>= 1)
                if (shfl(s, o + step) <= qa) o += step;
            const int64_t k = shfl(base_k, o) + qa;
            const int32_t run_end = min(shfl(s_end, o) - t, 64);
            const int32_t c = __builtin_nontemporal_load(bcols + k);
            int32_t m = 0x7fffffff;
            if (q < total) {
                m = comp[c];
                ++nread;
            }
#pragma unroll
            for (int step = 1; step < 64; step <<= 1) {
                const int32_t other = shfl(m, (lane + step) & 63);
                if (lane + step < run_end && other < m) m = other;
            }
            const bool mine_here = d > 0 && s < t + 64 && s_end > t;
            const int32_t got = shfl(m, mine_here ? (s > t ? s - t : 0) : lane);
            if (mine_here && got < acc) acc = got;
        }
        const bool down = acc < mine;
        if (down) {
            comp[v] = acc;
            ndeg += (unsigned long long)deg;
        }
        const uint64_t downs = __ballot(down);
        if (downs && lane == 0) {
            atomicOr((unsigned long long *)chg_out + (g0 >> 6), (unsigned long long)downs);
            ch = true;
        }
    }
    if (__ballot(ch) && lane == 0) changed[0] = 1;
    nbytes = wave_sum(nbytes + 8 * nread);
    nread = wave_sum(nread);
    ndeg = wave_sum(ndeg);
    const int slot = (blockIdx.x % kSlots) * kLine;
    if (lane == 0 && nread) atomicAdd(&labels_read[slot], nread);
    if (lane == 0 && ndeg) atomicAdd(&chg_deg[slot], ndeg);
    if (lane == 0 && nbytes) atomicAdd(&edge_bytes[slot], nbytes);
}

// Push pass: every vertex set in chg_in offers its label to its neighbours;
// a neighbour whose label goes down (atomicMin) is set in chg_out and its
// degree added to chg_deg.  Long lists go by slices as in the pull pass;
// the short lists of the changed vertices of a 64-vertex group are taken 64
// edges at a time, one per lane, the lane finding its source by a binary
// search over the group's degree prefix.  labels_read counts the
// neighbours' labels read.
__global__ __launch_bounds__(256) void k_cc_push(const int64_t *__restrict__ off,
                                                 const int32_t *__restrict__ cols, int32_t *comp,
                                                 int64_t nv, const td_task *__restrict__ tasks,
                                                 int64_t ntasks, const uint64_t *__restrict__ chg_in,
                                                 uint64_t *__restrict__ chg_out,
                                                 int32_t *__restrict__ changed,
                                                 unsigned long long *__restrict__ labels_read,
                                                 unsigned long long *__restrict__ chg_deg,
                                                 unsigned long long *__restrict__ edge_bytes) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t nunits = ntasks + (nv + 63) / 64;
    bool ch = false;
    // nbytes: algorithmic bytes besides the 8 B (column id + label) of each
    // label read
    unsigned long long nread = 0, ndeg = 0, nbytes = 0;
    // offer label l to vertex w
    auto offer = [&](int32_t w, int32_t l) {
        ++nread;
        if (l < comp[w] && l < atomicMin(comp + w, l)) {
            ch = true;
            atomicOr((unsigned long long *)chg_out + (w >> 6), 1ull << (w & 63));
            ndeg += (unsigned long long)(off[w + 1] - off[w]);
        }
    };
    for (int64_t u = wave; u < nunits; u += nwaves) {
        if (u < ntasks) {
            const int32_t src = tasks[u].src;
            if (lane == 0) nbytes += 8;
            if (!((chg_in[src >> 6] >> (src & 63)) & 1)) continue;
            if (lane == 0) nbytes += 4;
            const int64_t beg = tasks[u].beg, end = tasks[u].end;
            const int32_t l = comp[src];
            for (int64_t k0 = beg + lane; k0 < end; k0 += 4 * 64) {
                int32_t c[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t k = k0 + r * 64;
                    c[r] = __builtin_nontemporal_load(cols + (k < end ? k : end - 1));
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (k0 + r * 64 < end) offer(c[r], l);
            }
            continue;
        }
        const int64_t g = u - ntasks;
        const uint64_t bits = chg_in[g];
        if (lane == 0) nbytes += 8;
        if (!bits) continue;
        const int64_t v = g * 64 + lane;
        int64_t b = 0, e = 0;
        if (v < nv && ((bits >> lane) & 1)) {
            b = off[v];
            e = off[v + 1];
            nbytes += 12;
        }
        const int32_t d = e - b < kCcHeavy ? (int32_t)(e - b) : 0;
        const int32_t lv = d > 0 ? comp[v] : 0;
        const int32_t s_end = wave_incl_scan32(d);
        const int32_t s = s_end - d;
        const int32_t total = __builtin_amdgcn_readlane(s_end, 63);
        const int64_t base_k = b - s;  // k = base_k of the source + q
        for (int32_t t = 0; t < total; t += 64) {
            const int32_t q = t + lane;
            const int32_t qa = q < total ? q : total - 1;
            int o = 0;
#pragma unroll
            for (int step = 32; step; step >>= 1)
                if (shfl(s, o + step) <= qa) o += step;
            const int64_t k = shfl(base_k, o) + qa;
            const int32_t l = shfl(lv, o);
            const int32_t w = __builtin_nontemporal_load(cols + k);
            if (q < total) offer(w, l);
        }
    }
    if (__ballot(ch) && lane == 0) changed[0] = 1;
    nbytes = wave_sum(nbytes + 8 * nread);
    nread = wave_sum(nread);
    ndeg = wave_sum(ndeg);
    const int slot = (blockIdx.x % kSlots) * kLine;
    if (lane == 0 && nread) atomicAdd(&labels_read[slot], nread);
    if (lane == 0 && ndeg) atomicAdd(&chg_deg[slot], ndeg);
    if (lane == 0 && nbytes) atomicAdd(&edge_bytes[slot], nbytes);
}

// Shortcut (:74-78): each label becomes its root.  Vertices whose label
// moves are added to chg (the next pass's input), and chg_done (this pass's
// input, read by nobody any more) is cleared for the pass after.  Waves
// cover 64 aligned vertices, one bitmap word.  A label is never above its
// vertex's id, so chains strictly descend to a root; a label outside
// [0, v] stops the chase and sets *bad instead of being followed.
__global__ __launch_bounds__(256) void k_cc_shortcut(const int64_t *__restrict__ off, int32_t *comp,
                                                     int64_t nv, uint64_t *__restrict__ chg,
                                                     uint64_t *__restrict__ chg_done,
                                                     int32_t *__restrict__ bad,
                                                     unsigned long long *__restrict__ chg_deg) {
    unsigned long long ndeg = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c0 = comp[v];
        int32_t c = c0;
        bool ok = c >= 0 && c <= v;
        while (ok) {
            const int32_t cc = comp[c];
            if (cc == c) break;
            ok = cc >= 0 && cc < c;
            if (ok) c = cc;
        }
        if (!ok) {
            *bad = 1;
            c = c0;
        }
        if (c != c0) {
            comp[v] = c;
            ndeg += (unsigned long long)(off[v + 1] - off[v]);
        }
        const uint64_t moved = __ballot(c != c0);
        if ((v & 63) == 0) {
            if (moved) chg[v >> 6] |= moved;
            chg_done[v >> 6] = 0;
        }
    }
    ndeg = wave_sum(ndeg);
    if ((threadIdx.x & 63) == 0 && ndeg) atomicAdd(&chg_deg[(blockIdx.x % kSlots) * kLine], ndeg);
}

// Roots (label[v] == v) are the components, one per label (:84-101).
__global__ void k_cc_count(const int32_t *__restrict__ comp, int64_t nv,
                           unsigned long long *__restrict__ out) {
    unsigned long long n = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        n += comp[v] == (int32_t)v;
    n = wave_sum(n);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&out[(blockIdx.x % kSlots) * kLine], n);
}

}  // namespace
}  // namespace bd

using namespace bd;

extern "C" int bd_cc_run(bd_ctx *c, bd_cc_stats *out) {
    if (!c) return BD_EINVAL;
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph loaded");
    BD_HIP(c, hipSetDevice(c->device));
    const int64_t nv = c->nv, nw = (nv + 63) / 64;
    int rc;
    if (!c->d_comp && (rc = dev_alloc(c, (void **)&c->d_comp, sizeof(int32_t) * (nv + 1), "labels")))
        return rc;
    if (!c->d_cc_chg &&
        (rc = dev_alloc(c, (void **)&c->d_cc_chg, sizeof(uint64_t) * 2 * nw, "changed-label bitmaps")))
        return rc;
    // one line of flags (int32: [0] a label went down, [1] a label out of
    // range), then four kSlots x kLine counter blocks: components, labels
    // read, algorithmic bytes of the edge passes, degrees of changed vertices
    const size_t flag_bytes = sizeof(unsigned long long) * (kLine + 4 * kSlots * kLine);
    if (!c->d_cc_flag && (rc = dev_alloc(c, (void **)&c->d_cc_flag, flag_bytes, "label flags")))
        return rc;
    c->have_cc = false;
    int32_t *changed = c->d_cc_flag;
    unsigned long long *count = (unsigned long long *)c->d_cc_flag + kLine;
    unsigned long long *nread = count + kSlots * kLine;
    unsigned long long *ebytes = nread + kSlots * kLine;
    unsigned long long *chg_deg = ebytes + kSlots * kLine;
    std::vector<unsigned long long> h_deg(kSlots * kLine);
    uint64_t *chg[2] = {c->d_cc_chg, c->d_cc_chg + nw};
    const dim3 grid(grid_for(c, nv, 256)), block(256);
    if (c->cc_ntasks < 0) {
        // slices of the long lists, once per graph (like the graph's own
        // indexes, outside the timed run; its time and memory are reported)
        BD_HIP(c, hipEventRecord(c->ev[0], c->stream));
        const int64_t cap = c->nnz / kCcSlice + c->nnz / kCcHeavy + 1;
        dev_free(c->d_cc_tasks);  // left by a build that failed part-way
        c->d_cc_tasks = nullptr;
        if ((rc = dev_alloc(c, (void **)&c->d_cc_tasks, sizeof(td_task) * cap, "component slices")))
            return rc;
        BD_HIP(c, hipMemsetAsync(count, 0, sizeof(unsigned long long), c->stream));
        hipLaunchKernelGGL(k_cc_tasks, grid, block, 0, c->stream, c->d_off, nv, c->d_cc_tasks, count,
                           cap);
        BD_HIP(c, hipGetLastError());
        unsigned long long n = 0;
        BD_HIP(c, hipMemcpyAsync(&n, count, sizeof(n), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        if ((int64_t)n > cap) return set_err(c, BD_EHIP, "component slices overflow");
        c->cc_ntasks = (int64_t)n;
        // short lists regrouped by column block for the tiled pull passes
        int bshift = 6;
        while ((8ll << bshift) < nv) ++bshift;
        dev_free(c->d_cc_bcols);
        dev_free(c->d_cc_bend);
        c->d_cc_bcols = nullptr;
        c->d_cc_bend = nullptr;
        c->cc_ntasks = -1;  // until the regrouping is built too
        if ((rc = dev_alloc(c, (void **)&c->d_cc_bcols, sizeof(int32_t) * c->nnz, "regrouped lists")) ||
            (rc = dev_alloc(c, (void **)&c->d_cc_bend, sizeof(uint64_t) * nv, "block ends")))
            return rc;
        hipLaunchKernelGGL(k_cc_blocks, grid, block, 0, c->stream, (const int64_t *)c->d_off,
                           (const int32_t *)c->d_cols, nv, bshift, c->d_cc_bcols, c->d_cc_bend);
        BD_HIP(c, hipGetLastError());
        // long lists: regrouped too, their pieces filed by (row, column) block
        const dim3 slice_grid(grid_for(c, std::max<int64_t>((int64_t)n, 1) * 64, 256));
        unsigned long long pcnt[64], pbase[65];
        BD_HIP(c, hipMemsetAsync(count, 0, sizeof(pcnt), c->stream));
        hipLaunchKernelGGL(k_cc_heavy_tiles, slice_grid, block, 0, c->stream,
                           (const int64_t *)c->d_off, (const int32_t *)c->d_cols, bshift,
                           (const td_task *)c->d_cc_tasks, (int64_t)n, c->d_cc_bcols, count,
                           (td_task *)nullptr, (int64_t)0, 0);
        BD_HIP(c, hipGetLastError());
        BD_HIP(c, hipMemcpyAsync(pcnt, count, sizeof(pcnt), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        pbase[0] = 0;
        for (int i = 0; i < 64; ++i) pbase[i + 1] = pbase[i] + pcnt[i];
        dev_free(c->d_cc_pieces);
        dev_free(c->d_cc_pbase);
        c->d_cc_pieces = nullptr;
        c->d_cc_pbase = nullptr;
        if ((rc = dev_alloc(c, (void **)&c->d_cc_pieces,
                            sizeof(td_task) * std::max<unsigned long long>(pbase[64], 1),
                            "long-list pieces")) ||
            (rc = dev_alloc(c, (void **)&c->d_cc_pbase, sizeof(pbase), "piece buckets")))
            return rc;
        BD_HIP(c, hipMemcpyAsync(count, pbase, sizeof(pcnt), hipMemcpyHostToDevice, c->stream));
        BD_HIP(c, hipMemcpyAsync(c->d_cc_pbase, pbase, sizeof(pbase), hipMemcpyHostToDevice,
                                 c->stream));
        hipLaunchKernelGGL(k_cc_heavy_tiles, slice_grid, block, 0, c->stream,
                           (const int64_t *)c->d_off, (const int32_t *)c->d_cols, bshift,
                           (const td_task *)c->d_cc_tasks, (int64_t)n, c->d_cc_bcols, count,
                           c->d_cc_pieces, (int64_t)pbase[64], 1);
        BD_HIP(c, hipGetLastError());
        BD_HIP(c, hipMemcpyAsync(pcnt, count, sizeof(pcnt), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < 64; ++i)
            if (pcnt[i] != pbase[i + 1]) return set_err(c, BD_EHIP, "long-list pieces miscounted");
        BD_HIP(c, hipEventRecord(c->ev[1], c->stream));
        BD_HIP(c, hipEventSynchronize(c->ev[1]));
        float pms = 0;
        BD_HIP(c, hipEventElapsedTime(&pms, c->ev[0], c->ev[1]));
        c->cc_prep_ms = pms;
        c->cc_prep_bytes = (int64_t)(sizeof(td_task) * cap + sizeof(int32_t) * c->nnz +
                                     sizeof(uint64_t) * nv +
                                     sizeof(td_task) * std::max<unsigned long long>(pbase[64], 1) +
                                     sizeof(pbase));
        c->cc_bshift = bshift;
        c->cc_ntasks = (int64_t)n;
    }
    const int64_t nunits = c->cc_ntasks + nw;
    const dim3 hook_grid(grid_for(c, nunits * 64, 256));
    // a whole number of workgroups per XCD
    const dim3 tile_grid((unsigned)std::max<int64_t>(8, grid_for(c, nv, 256) & ~7ll));
    BD_HIP(c, hipEventRecord(c->ev[0], c->stream));
    BD_HIP(c, hipMemsetAsync(changed, 0, sizeof(int32_t) * 2, c->stream));
    BD_HIP(c, hipMemsetAsync(count, 0, sizeof(unsigned long long) * 3 * kSlots * kLine, c->stream));
    // pass i reads chg[i & 1] and sets chg[(i + 1) & 1]
    hipLaunchKernelGGL(k_cc_init, grid, block, 0, c->stream, c->d_comp, nv, chg[0], chg[1]);
    BD_HIP(c, hipGetLastError());
    int64_t iters = 0, launches = 0;
    double pass_ms = 0;
    bool all = true;  // pull every label (first pass, or after many changes), else push
    for (;;) {
        ++iters;
        if (iters > nv + 2) return set_err(c, BD_EHIP, "components did not converge");
        uint64_t *in = chg[iters & 1], *next = chg[(iters + 1) & 1];
        BD_HIP(c, hipMemsetAsync(changed, 0, sizeof(int32_t), c->stream));  // [1] stays
        BD_HIP(c, hipMemsetAsync(chg_deg, 0, sizeof(unsigned long long) * h_deg.size(), c->stream));
        if (all && iters > 1) {
            // tiled pull: 8 phases, each XCD pulling from one column block
            for (int phase = 0; phase < 8; ++phase)
                hipExtLaunchKernelGGL(k_cc_tile, tile_grid, block, 0, c->stream,
                                      phase == 0 ? c->ev[2] : nullptr,
                                      phase == 7 ? c->ev[3] : nullptr, 0,
                                      (const int64_t *)c->d_off, (const int32_t *)c->d_cc_bcols,
                                      (const uint64_t *)c->d_cc_bend, c->d_comp, nv, c->cc_bshift,
                                      phase, (const td_task *)c->d_cc_pieces,
                                      (const int64_t *)c->d_cc_pbase, next, changed, nread,
                                      chg_deg, ebytes);
            launches += 8;
        } else if (all)
            hipExtLaunchKernelGGL(k_cc_hook, hook_grid, block, 0, c->stream, c->ev[2], c->ev[3], 0,
                                  (const int64_t *)c->d_off, (const int32_t *)c->d_cols, c->d_comp,
                                  nv, (const td_task *)c->d_cc_tasks, c->cc_ntasks,
                                  (int)(iters == 1), next, changed, nread, chg_deg, ebytes);
        else
            hipExtLaunchKernelGGL(k_cc_push, hook_grid, block, 0, c->stream, c->ev[2], c->ev[3], 0,
                                  (const int64_t *)c->d_off, (const int32_t *)c->d_cols, c->d_comp,
                                  nv, (const td_task *)c->d_cc_tasks, c->cc_ntasks,
                                  (const uint64_t *)in, next, changed, nread, chg_deg, ebytes);
        if (!(all && iters > 1)) ++launches;
        BD_HIP(c, hipGetLastError());
        int32_t h[2] = {0, 0};
        BD_HIP(c, hipMemcpyAsync(h, changed, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        if (h[1]) return set_err(c, BD_EHIP, "component label out of range");
        float kms = 0;
        BD_HIP(c, hipEventElapsedTime(&kms, c->ev[2], c->ev[3]));
        pass_ms += kms;
        if (!h[0]) break;
        hipLaunchKernelGGL(k_cc_shortcut, grid, block, 0, c->stream, (const int64_t *)c->d_off,
                           c->d_comp, nv, next, in, changed + 1, chg_deg);
        BD_HIP(c, hipGetLastError());
        BD_HIP(c, hipMemcpyAsync(h_deg.data(), chg_deg, sizeof(unsigned long long) * h_deg.size(),
                                 hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        unsigned long long deg = 0;
        for (int s = 0; s < kSlots; ++s) deg += h_deg[s * kLine];
        all = 2 * (int64_t)deg >= c->nnz;
    }
    hipLaunchKernelGGL(k_cc_count, grid, block, 0, c->stream, c->d_comp, nv, count);
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipEventRecord(c->ev[1], c->stream));
    std::vector<unsigned long long> h(3 * kSlots * kLine);
    BD_HIP(c, hipMemcpyAsync(h.data(), count, sizeof(unsigned long long) * h.size(),
                             hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    BD_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    unsigned long long ncomp = 0, nlab = 0, nbytes = 0;
    for (int s = 0; s < kSlots; ++s) {
        ncomp += h[s * kLine];
        nlab += h[(kSlots + s) * kLine];
        nbytes += h[(2 * kSlots + s) * kLine];
    }
    c->have_cc = true;
    if (out) {
        out->num_components = (int64_t)ncomp;
        out->num_iters = iters;
        out->ms = ms;
        out->pass_ms = pass_ms;
        out->labels_read = (int64_t)nlab;
        out->edge_bytes = (int64_t)nbytes;
        out->pass_launches = launches;
        out->prep_ms = c->cc_prep_ms;
        out->prep_bytes = c->cc_prep_bytes;
    }
    return BD_OK;
}

extern "C" int bd_cc_fetch(bd_ctx *c, int64_t *labels) {
    if (!c || !labels) return BD_EINVAL;
    if (!c->have_cc) return set_err(c, BD_ESTATE, "components have not run");
    BD_HIP(c, hipSetDevice(c->device));
    std::vector<int32_t> l(c->nv);
    BD_HIP(c, hipMemcpy(l.data(), c->d_comp, sizeof(int32_t) * c->nv, hipMemcpyDeviceToHost));
    for (int64_t v = 0; v < c->nv; ++v) labels[v] = l[v];
    return BD_OK;
}

// components::check (components.cc:152-202) on the host, from downloaded
// copies: one source per label (the last vertex carrying it); a BFS from it
// must meet only that label, and every vertex must be reached.
extern "C" int bd_cc_check(bd_ctx *c, int *ok) {
    if (!c || !ok) return BD_EINVAL;
    if (!c->have_cc) return set_err(c, BD_ESTATE, "components have not run");
    const int64_t nv = c->nv;
    std::vector<int64_t> off(nv + 1), label(nv);
    std::vector<int32_t> cols(c->nnz);
    int rc = bd_graph_download(c, off.data(), cols.data());
    if (!rc) rc = bd_cc_fetch(c, label.data());
    if (rc) return rc;
    std::vector<int64_t> source(nv, -1), queue;
    std::vector<char> visited(nv, 0);
    bool good = true;
    for (int64_t v = 0; v < nv && good; ++v) {
        if (label[v] < 0 || label[v] >= nv)
            good = false;
        else
            source[label[v]] = v;
    }
    queue.reserve(nv);
    for (int64_t l = 0; l < nv && good; ++l) {
        if (source[l] < 0) continue;
        queue.clear();
        visited[source[l]] = 1;
        queue.push_back(source[l]);
        for (size_t h = 0; h < queue.size() && good; ++h) {
            const int64_t u = queue[h];
            for (int64_t e = off[u]; e < off[u + 1]; ++e) {
                const int64_t w = cols[e];
                if (label[w] != l) {
                    good = false;
                    break;
                }
                if (!visited[w]) {
                    visited[w] = 1;
                    queue.push_back(w);
                }
            }
        }
    }
    for (int64_t v = 0; v < nv && good; ++v)
        if (!visited[v]) good = false;
    *ok = good ? 1 : 0;
    return BD_OK;
}
