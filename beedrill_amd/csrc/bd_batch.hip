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
// The host picks the mode per level from the previous level's counters (the
// frontier's degree sum against the degree sum of vertices still open), as
// Beamer's rule picks a direction.  Depths are written root-major
// (depth[r * nv + v]) when found: 64 masked stores per wave, one per root,
// each covering 64 consecutive vertices of that root's row.
#include <string.h>

#include <algorithm>
#include <vector>

#include "bd_device.h"
#include "bd_internal.h"

namespace bd {

namespace {

constexpr int kMsBlock = 256;
constexpr int kMsWaveDeg = 64;  // lists at least this long are scanned by a whole wave

struct ms_args {
    const int64_t *off;
    const int32_t *cols;
    uint64_t *F;
    uint64_t *N;
    uint64_t *seen;
    uint64_t *Fb;    // bit v: F[v] != 0 (frontier membership, V/64 words)
    uint64_t *Nb;    // bit v: N[v] != 0, written by the commit or pull pass
    int32_t *depth;  // n rows of nv, or null
    unsigned long long *cnt;  // [kSlots][kLine]: frontier size, degree sum, open degree sum
    int64_t nv;
    uint64_t full;   // bits of the roots in this search
    int32_t nroots;
    int32_t level;   // the step expands depth `level` into depth `level + 1`
};

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

// Per-block counters added to slot blockIdx % kSlots.
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
        unsigned long long *slot = a.cnt + (blockIdx.x % kSlots) * kLine;
        if (c0) atomicAdd(&slot[0], c0);
        if (c1) atomicAdd(&slot[1], c1);
        if (c2) atomicAdd(&slot[2], c2);
    }
}

// Roots start their own frontier: bit i on roots[i].
__global__ void k_ms_init(ms_args a, const int64_t *__restrict__ roots, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    if (i < n) {
        const int64_t v = roots[i];
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

// Push: expand the frontier into N with device atomics.
__global__ __launch_bounds__(kMsBlock) void k_ms_push(ms_args a) {
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
        const bool alone = f && e - b < kMsWaveDeg;
        if (alone)
            for (int64_t k = b; k < e; ++k) ms_push_to(a, ms_col(a, k), f);
        uint64_t lm = __ballot(f && !alone);
        while (lm) {
            const int src = __ffsll((unsigned long long)lm) - 1;
            lm &= lm - 1;
            const int64_t kb = lane_bcast(b, src), ke = lane_bcast(e, src);
            const uint64_t fs = (uint64_t)lane_bcast((int64_t)f, src);
            for (int64_t k = kb + lane; k < ke; k += 64) ms_push_to(a, ms_col(a, k), fs);
        }
    }
}

// Commit after a push: keep the unseen bits of N, mark them seen, record
// depths, clear F (it becomes the next push's target), count.
__global__ __launch_bounds__(kMsBlock) void k_ms_commit(ms_args a) {
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
    }
    ms_count(a, c0, c1, c2);
}

// Pull: every open vertex ORs its neighbours' frontier masks, stopping once
// all the roots it still waits for are covered; writes N for every vertex.
__global__ __launch_bounds__(kMsBlock) void k_ms_pull(ms_args a) {
    const int lane = lane_id();
    unsigned long long c0 = 0, c1 = 0, c2 = 0;
    const int64_t wave = ((int64_t)blockIdx.x * kMsBlock + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kMsBlock) >> 6;
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
        if (valid) {
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

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = x;
}

int ms_buffers(bd_ctx *c, int nrows) {
    int rc;
    for (int i = 0; i < 3; ++i)
        if (!c->d_ms[i] &&
            (rc = dev_alloc(c, (void **)&c->d_ms[i], sizeof(uint64_t) * c->nv, "root masks")))
            return rc;
    if (!c->d_ms[3] &&
        (rc = dev_alloc(c, (void **)&c->d_ms[3], sizeof(unsigned long long) * kSlots * kLine,
                        "mask counters")))
        return rc;
    for (int i = 0; i < 2; ++i)
        if (!c->d_ms_bits[i] && (rc = dev_alloc(c, (void **)&c->d_ms_bits[i],
                                                sizeof(uint64_t) * c->nwords, "frontier bits")))
            return rc;
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

int ms_read_counts(bd_ctx *c, unsigned long long out[3]) {
    std::vector<unsigned long long> h(kSlots * kLine);
    BD_HIP(c, hipMemcpyAsync(h.data(), c->d_ms[3], sizeof(unsigned long long) * h.size(),
                             hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    out[0] = out[1] = out[2] = 0;
    for (int s = 0; s < kSlots; ++s)
        for (int k = 0; k < 3; ++k) out[k] += h[s * kLine + k];
    return BD_OK;
}

}  // namespace

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
    a.off = c->d_off;
    a.cols = c->d_cols;
    a.F = c->d_ms[0];
    a.N = c->d_ms[1];
    a.seen = c->d_ms[2];
    a.Fb = c->d_ms_bits[0];
    a.Nb = c->d_ms_bits[1];
    a.depth = depths ? c->d_ms_depth : nullptr;
    a.cnt = (unsigned long long *)c->d_ms[3];
    a.nv = nv;
    a.full = n == 64 ? ~0ULL : ((1ULL << n) - 1);
    a.nroots = n;
    a.level = 0;
    const size_t mbytes = sizeof(uint64_t) * nv;
    const size_t cbytes = sizeof(unsigned long long) * kSlots * kLine;
    // clear(): untimed, as in the reference's trial loop
    BD_HIP(c, hipMemcpyAsync(d_roots, roots, sizeof(int64_t) * n, hipMemcpyHostToDevice,
                             c->stream));
    for (int i = 0; i < 3; ++i) BD_HIP(c, hipMemsetAsync(c->d_ms[i], 0, mbytes, c->stream));
    for (int i = 0; i < 2; ++i)
        BD_HIP(c, hipMemsetAsync(c->d_ms_bits[i], 0, sizeof(uint64_t) * c->nwords, c->stream));
    if (a.depth) {
        const int64_t cells = (int64_t)n * nv;
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c, cells, 256)), dim3(256), 0, c->stream,
                           a.depth, cells, -1);
        BD_HIP(c, hipGetLastError());
    }
    BD_HIP(c, hipMemsetAsync(a.cnt, 0, cbytes, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));

    const int64_t grid = (int64_t)c->cu_count * std::max(1, c->tune.grid_per_cu);
    BD_HIP(c, hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(k_ms_init, dim3((n + 63) / 64), dim3(64), 0, c->stream, a, d_roots, n);
    BD_HIP(c, hipGetLastError());
    unsigned long long cnt[3];
    if ((rc = ms_read_counts(c, cnt))) return rc;
    bool n_clean = true;  // N holds zeros (ready for a push)
    int64_t L = 0;
    // push while the frontier's edges * alpha are below the open edges
    const unsigned long long alpha = (unsigned long long)std::max(1, c->tune.ms_alpha);
    std::vector<bd_level_rec> log;
    while (cnt[0] > 0) {
        if (L > nv + 1) return set_err(c, BD_EHIP, "multi-root search did not terminate");
        a.level = (int32_t)L;
        const bool push = cnt[1] * alpha < cnt[2];
        bd_level_rec r;
        memset(&r, 0, sizeof(r));
        r.level = L;
        r.dir = push ? 0 : 1;
        r.frontier = (int64_t)cnt[0];
        r.frontier_deg = (int64_t)cnt[1];
        r.unvisited_deg = (int64_t)cnt[2];
        r.ms_other = -1.0;
        BD_HIP(c, hipEventRecord(c->ev[2], c->stream));
        BD_HIP(c, hipMemsetAsync(a.cnt, 0, cbytes, c->stream));
        if (push) {
            if (!n_clean) BD_HIP(c, hipMemsetAsync(a.N, 0, mbytes, c->stream));
            hipLaunchKernelGGL(k_ms_push, dim3(grid), dim3(kMsBlock), 0, c->stream, a);
            hipLaunchKernelGGL(k_ms_commit, dim3(grid), dim3(kMsBlock), 0, c->stream, a);
            n_clean = true;  // commit cleared F, the next level's N
        } else {
            hipLaunchKernelGGL(k_ms_pull, dim3(grid), dim3(kMsBlock), 0, c->stream, a);
            n_clean = false;  // F still holds depth L
        }
        BD_HIP(c, hipGetLastError());
        BD_HIP(c, hipEventRecord(c->ev[3], c->stream));
        if ((rc = ms_read_counts(c, cnt))) return rc;
        float lms = 0;
        BD_HIP(c, hipEventElapsedTime(&lms, c->ev[2], c->ev[3]));
        r.ms = lms;
        r.next = (int64_t)cnt[0];
        r.next_deg = (int64_t)cnt[1];
        log.push_back(r);
        std::swap(a.F, a.N);
        std::swap(a.Fb, a.Nb);
        ++L;
    }
    BD_HIP(c, hipEventRecord(c->ev[1], c->stream));
    BD_HIP(c, hipEventSynchronize(c->ev[1]));
    float t = 0;
    BD_HIP(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    if (ms) *ms = t;
    if (levels) *levels = L;
    c->last_log = log;
    c->have_result = false;  // bd_bfs_levels now describes this search, not the last single one
    if (depths)
        BD_HIP(c, hipMemcpy(depths, a.depth, sizeof(int32_t) * n * nv, hipMemcpyDeviceToHost));
    return BD_OK;
}

}  // namespace bd

using namespace bd;

extern "C" int bd_bfs_multi_root(bd_ctx *c, const int64_t *roots, int n, int32_t *depths,
                                 double *ms, int64_t *levels) {
    if (!c) return BD_EINVAL;
    return run_multi_root(c, roots, n, depths, ms, levels);
}
