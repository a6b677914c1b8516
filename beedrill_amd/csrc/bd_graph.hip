// bd_graph.hip -- device graph construction: el64 loading, seeded generators
// and the symmetric CSR build (create_graph_from_edge_list,
// graph_base.h:335-436).
//
// The CSR build is a stable radix sort of the 2M directed copies of the edge
// list, laid out as [e0 fwd, e0 rev, e1 fwd, e1 rev, ...] and keyed by source.
// Stability makes every adjacency list come out in edge-list order, exactly
// the order the reference's serial insert_edge loop produces
// (graph_base.h:423-427), without per-vertex atomics.
//
// This file is compiled with -ffp-contract=off: the RMAT generator's double
// arithmetic must round exactly as the reference's x86 build does.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "bd_device.h"
#include "bd_gen_common.h"
#include "bd_internal.h"

namespace bd {

namespace {

constexpr int kBlock = 256;

int bits_for(int64_t nv) {
    int s = 0;
    while (((int64_t)1 << s) < nv) ++s;
    return std::max(s, 1);
}

__global__ void k_pack_pairs(const int64_t *__restrict__ src, const int64_t *__restrict__ dst,
                             int64_t m, int64_t nv, uint64_t *__restrict__ el,
                             int *__restrict__ err) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t s = src[i], d = dst[i];
        if (s < 0 || s >= nv || d < 0 || d >= nv) {
            atomicOr(err, 1);
            s = 0;
            d = 0;
        }
        el[i] = ((uint64_t)(uint32_t)s << 32) | (uint32_t)d;
    }
}

__global__ void k_expand_directed(const uint64_t *__restrict__ el, int64_t m,
                                  uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t e = el[i];
        const uint32_t s = (uint32_t)(e >> 32), d = (uint32_t)e;
        keys[2 * i] = s;
        vals[2 * i] = d;
        keys[2 * i + 1] = d;
        vals[2 * i + 1] = s;
    }
}

// off[v] = index of the first sorted key >= v, for v in [0, nv].
__global__ void k_offsets_from_sorted(const uint32_t *__restrict__ keys, int64_t n, int64_t nv,
                                      int64_t *__restrict__ off) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t kprev = (i == 0) ? -1 : (int64_t)keys[i - 1];
        const int64_t kcur = (i == n) ? nv : (int64_t)keys[i];
        for (int64_t v = kprev + 1; v <= kcur; ++v) off[v] = i;
    }
}

// Bit set = vertex never reachable by a search (degree 0) or past nv.
__global__ void k_dead_bitmap(const int64_t *__restrict__ off, int64_t nv, int64_t nwords,
                              uint64_t *__restrict__ dead) {
    const int64_t total = nwords * 64;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < total;
         v += (int64_t)gridDim.x * blockDim.x) {
        const bool d = (v >= nv) || (off[v + 1] == off[v]);
        const uint64_t mask = __ballot(d);
        if ((threadIdx.x & 63) == 0) dead[v >> 6] = mask;
    }
}

// The hubs of a layout: bit v of bm set when deg(v) >= kHubDeg; those under
// kHubBigDeg listed in med, the others' lists cut into slices of kHubSlice
// in big (places handed out by atomics: the order is not a result).
__global__ void k_hub_lists(const int64_t *__restrict__ off, int64_t nv, int64_t nwords,
                            uint64_t *__restrict__ bm, int32_t *__restrict__ med,
                            td_task *__restrict__ big, unsigned long long *__restrict__ cnt) {
    const int64_t total = nwords * 64;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < total;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = v < nv ? off[v] : 0, e = v < nv ? off[v + 1] : 0;
        const bool hub = e - b >= kHubDeg;
        const uint64_t mask = __ballot(hub);
        if ((threadIdx.x & 63) == 0) bm[v >> 6] = mask;
        if (!hub) continue;
        if (e - b < kHubBigDeg) {
            med[atomicAdd(&cnt[0], 1ULL)] = (int32_t)v;
        } else {
            const int64_t ns = (e - b + kHubSlice - 1) / kHubSlice;
            const int64_t p = (int64_t)atomicAdd(&cnt[1], (unsigned long long)ns);
            for (int64_t k = 0; k < ns; ++k) {
                td_task t;
                t.beg = b + k * kHubSlice;
                t.end = std::min<int64_t>(t.beg + kHubSlice, e);
                t.src = (int32_t)v;
                t.pad = 0;
                big[p + k] = t;
            }
        }
    }
}

__global__ void k_validate_csr(const int64_t *__restrict__ off, const int32_t *__restrict__ cols,
                               int64_t nv, int64_t nnz, int *__restrict__ err) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (off[i + 1] < off[i]) atomicOr(err, 2);
    }
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = cols[i];
        if (u < 0 || (int64_t)u >= nv) atomicOr(err, 1);
    }
}

// Raw RMAT edge i from its own slice of the sitmo stream
// (rmat_generator.h:143-175); self-loops are listed for the serial re-roll.
__global__ void k_rmat_raw(int64_t ne, int scale, double a, double b, double c, double d,
                           uint64_t seed, uint64_t *__restrict__ raw,
                           unsigned long long *__restrict__ nself,
                           int64_t *__restrict__ self_idx, int64_t self_cap) {
    const uint64_t draws = rmat_draws_per_edge(scale);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ne;
         i += (int64_t)gridDim.x * blockDim.x) {
        sitmo_reader rd;
        rd.init(seed, (uint64_t)i * draws);
        int64_t s, t;
        rmat_edge(rd, scale, a, b, c, d, &s, &t);
        raw[i] = ((uint64_t)s << 32) | (uint64_t)t;
        if (s == t) {
            const unsigned long long p = atomicAdd(nself, 1ULL);
            if ((int64_t)p < self_cap) self_idx[p] = i;
        }
    }
}

__global__ void k_er_raw(int64_t ne, int64_t nv, uint64_t seed, uint64_t *__restrict__ raw) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ne;
         i += (int64_t)gridDim.x * blockDim.x) {
        sitmo_reader rd;
        rd.init(seed, (uint64_t)i * 4);
        const uint64_t w0 = rd.next(), w1 = rd.next(), w2 = rd.next(), w3 = rd.next();
        const int64_t s = mulhi_range(w0 | (w1 << 32), nv);
        const int64_t t = mulhi_range(w2 | (w3 << 32), nv);
        raw[i] = ((uint64_t)s << 32) | (uint64_t)t;
    }
}

__global__ void k_patch(uint64_t *__restrict__ raw, const int64_t *__restrict__ idx,
                        const uint64_t *__restrict__ val, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        raw[idx[i]] = val[i];
}

// flip_edges (edge_list_utils.h:13-25) fused with the sort key: (min << s) |
// max.  Self-loops (ER only; RMAT re-rolls them) map to the all-ones key,
// which sorts last and is dropped after dedup.
__global__ void k_flip_key(uint64_t *__restrict__ kv, int64_t n, int s) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t e = kv[i];
        const uint64_t a = e >> 32, b = e & 0xFFFFFFFFULL;
        kv[i] = (a == b) ? ~0ULL : ((std::min(a, b) << s) | std::max(a, b));
    }
}

__global__ void k_relabel_shuffle(const uint64_t *__restrict__ ukeys, int64_t m, int s,
                                  vperm_keys vp, int relabel, uint64_t seed,
                                  uint64_t *__restrict__ skeys, uint64_t *__restrict__ packed) {
    const uint64_t lo_mask = ((uint64_t)1 << s) - 1;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = ukeys[j];
        int64_t a = (int64_t)(k >> s), b = (int64_t)(k & lo_mask);
        if (relabel) {
            a = vp.apply(a);
            b = vp.apply(b);
        }
        packed[j] = ((uint64_t)a << 32) | (uint64_t)b;
        skeys[j] = shuffle_key((uint64_t)j, seed);
    }
}

// rows x cols lattice in (vertex, right, down) order.
__global__ void k_lattice(int64_t rows, int64_t cols, uint64_t *__restrict__ el) {
    const int64_t nv = rows * cols;
    const int64_t per_row = 2 * cols - 1;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = v / cols, c = v % cols;
        if (r + 1 < rows) {
            const int64_t base = r * per_row + 2 * c;
            if (c + 1 < cols) {
                el[base] = ((uint64_t)v << 32) | (uint64_t)(v + 1);
                el[base + 1] = ((uint64_t)v << 32) | (uint64_t)(v + cols);
            } else {
                el[base] = ((uint64_t)v << 32) | (uint64_t)(v + cols);
            }
        } else if (c + 1 < cols) {
            el[r * per_row + c] = ((uint64_t)v << 32) | (uint64_t)(v + 1);
        }
    }
}

__global__ void k_unpack(const uint64_t *__restrict__ el, int64_t m, int64_t *__restrict__ src,
                         int64_t *__restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        src[i] = (int64_t)(el[i] >> 32);
        dst[i] = (int64_t)(uint32_t)el[i];
    }
}

struct dev_buf {
    void *p = nullptr;
    ~dev_buf() { dev_free(p); }
    template <class T>
    T *as() const {
        return (T *)p;
    }
};

int64_t grid1d(const bd_ctx *c, int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    return std::max<int64_t>(1, std::min<int64_t>(g, (int64_t)c->cu_count * 16));
}

int sort_keys_u64(bd_ctx *c, const uint64_t *in, uint64_t *out, int64_t n, int bits) {
    size_t tmp = 0;
    BD_HIP(c, rocprim::radix_sort_keys(nullptr, tmp, in, out, (size_t)n, 0, (unsigned)bits,
                                       c->stream));
    dev_buf t;
    int rc = dev_alloc(c, &t.p, tmp, "sort temp");
    if (rc) return rc;
    BD_HIP(c, rocprim::radix_sort_keys(t.p, tmp, in, out, (size_t)n, 0, (unsigned)bits, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    return BD_OK;
}

template <class K, class V>
int sort_pairs(bd_ctx *c, const K *kin, K *kout, const V *vin, V *vout, int64_t n, int bits) {
    size_t tmp = 0;
    BD_HIP(c, rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, (size_t)n, 0,
                                        (unsigned)bits, c->stream));
    dev_buf t;
    int rc = dev_alloc(c, &t.p, tmp, "sort temp");
    if (rc) return rc;
    BD_HIP(c, rocprim::radix_sort_pairs(t.p, tmp, kin, kout, vin, vout, (size_t)n, 0,
                                        (unsigned)bits, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    return BD_OK;
}

int unique_u64(bd_ctx *c, const uint64_t *in, uint64_t *out, int64_t n, int64_t *m) {
    dev_buf cnt;
    int rc = dev_alloc(c, &cnt.p, sizeof(size_t), "unique count");
    if (rc) return rc;
    size_t tmp = 0;
    BD_HIP(c, rocprim::unique(nullptr, tmp, in, out, cnt.as<size_t>(), (size_t)n,
                              rocprim::equal_to<uint64_t>(), c->stream));
    dev_buf t;
    if ((rc = dev_alloc(c, &t.p, tmp, "unique temp"))) return rc;
    BD_HIP(c, rocprim::unique(t.p, tmp, in, out, cnt.as<size_t>(), (size_t)n,
                              rocprim::equal_to<uint64_t>(), c->stream));
    size_t h = 0;
    BD_HIP(c, hipMemcpyAsync(&h, cnt.p, sizeof(size_t), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    *m = (int64_t)h;
    return BD_OK;
}

__global__ void k_shuffle_keys(int64_t m, uint64_t seed, uint64_t *__restrict__ skeys) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x)
        skeys[j] = shuffle_key((uint64_t)j, seed);
}

// Order the m edges by shuffle_key(j) of their index j into c->d_el.
// `in` holds either sorted dedup keys ((min << s) | max, relabelled on the
// way) or, when packed_input, edges already packed as (src << 32) | dst.  The
// buffer is caller-owned and is reused as scratch.
int finish_edges(bd_ctx *c, uint64_t *in, int64_t m, int s, int64_t nv, bool relabel,
                 uint64_t seed, bool packed_input) {
    dev_buf skeys, packed;
    int rc;
    if ((rc = dev_alloc(c, &skeys.p, sizeof(uint64_t) * m, "shuffle keys"))) return rc;
    if ((rc = dev_alloc(c, &packed.p, sizeof(uint64_t) * m, "packed edges"))) return rc;
    if (packed_input) {
        hipLaunchKernelGGL(k_shuffle_keys, dim3(grid1d(c, m)), dim3(kBlock), 0, c->stream, m,
                           seed, skeys.as<uint64_t>());
        BD_HIP(c, hipMemcpyAsync(packed.p, in, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice,
                                 c->stream));
    } else {
        vperm_keys vp;
        vp.init(nv, seed);
        hipLaunchKernelGGL(k_relabel_shuffle, dim3(grid1d(c, m)), dim3(kBlock), 0, c->stream, in,
                           m, s, vp, relabel ? 1 : 0, seed, skeys.as<uint64_t>(),
                           packed.as<uint64_t>());
    }
    BD_HIP(c, hipGetLastError());
    dev_free(c->d_el);
    c->d_el = nullptr;
    if ((rc = dev_alloc(c, (void **)&c->d_el, sizeof(uint64_t) * m, "edge list"))) return rc;
    // `in` is no longer needed: it receives the sorted shuffle keys
    return sort_pairs<uint64_t, uint64_t>(c, skeys.as<uint64_t>(), in, packed.as<uint64_t>(),
                                          c->d_el, m, 64);
}

}  // namespace

int build_csr_from_el(bd_ctx *c, int64_t nv, int64_t m) {
    // c->d_el holds m packed edges in el64 order.
    const int64_t n2 = 2 * m;
    dev_buf keys, vals, skeys;
    int rc;
    free_search_layout(c);
    dev_free(c->d_off);
    dev_free(c->d_cols);
    dev_free(c->d_dead);
    dev_free(c->d_head);
    c->d_off = nullptr;
    c->d_cols = nullptr;
    c->d_dead = nullptr;
    c->d_head = nullptr;
    if ((rc = dev_alloc(c, &keys.p, sizeof(uint32_t) * n2, "directed keys"))) return rc;
    if ((rc = dev_alloc(c, &vals.p, sizeof(uint32_t) * n2, "directed values"))) return rc;
    if ((rc = dev_alloc(c, &skeys.p, sizeof(uint32_t) * n2, "sorted keys"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_cols, sizeof(int32_t) * n2, "cols"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_off, sizeof(int64_t) * (nv + 1), "offsets"))) return rc;
    const int64_t nwords = (nv + 63) / 64;
    if ((rc = dev_alloc(c, (void **)&c->d_dead, sizeof(uint64_t) * nwords, "dead bitmap")))
        return rc;
    hipLaunchKernelGGL(k_expand_directed, dim3(grid1d(c, m)), dim3(kBlock), 0, c->stream,
                       c->d_el, m, keys.as<uint32_t>(), vals.as<uint32_t>());
    BD_HIP(c, hipGetLastError());
    if ((rc = sort_pairs<uint32_t, uint32_t>(c, keys.as<uint32_t>(), skeys.as<uint32_t>(),
                                             vals.as<uint32_t>(), (uint32_t *)c->d_cols, n2,
                                             bits_for(nv))))
        return rc;
    hipLaunchKernelGGL(k_offsets_from_sorted, dim3(grid1d(c, n2 + 1)), dim3(kBlock), 0,
                       c->stream, skeys.as<uint32_t>(), n2, nv, c->d_off);
    hipLaunchKernelGGL(k_dead_bitmap, dim3(grid1d(c, nwords * 64)), dim3(kBlock), 0, c->stream,
                       c->d_off, nv, nwords, c->d_dead);
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipStreamSynchronize(c->stream));
    c->nv = nv;
    c->nnz = n2;
    c->m = m;
    c->nwords = nwords;
    c->el_m = m;
    return alloc_search(c);
}

// ------------------------------------------------------------ search layout
//
// The BFS kernels read a renumbered copy of the CSR.  In degree order
// (vertex_order 1) vertices are ranked by degree, largest first, ties broken
// by graph id, and the degree-0 vertices come last; the 4096 largest get a
// visited word each (spread_id), the rest follow in rank order.  Order 4 is
// the plain degree order (no spreading); with vertex_order 3 the vertices
// with edges come first, otherwise in graph order.  Either way the
// arrays indexed by vertex (offsets, head index, parents, visited and
// frontier bits) hold no isolated vertex before search id nv_active, so a
// bottom-up step reads them densely; in degree order the high-degree
// vertices, which most vertices find their parent through, also share a few
// cache lines of every frontier bitmap.  Each adjacency list keeps its CSR
// order (only the ids are mapped), so every count a step records
// (frontier sizes and degrees, edges examined before the first hit) is what
// it is on the graph's own numbering.

namespace {

constexpr int64_t kSpreadHubs = 4096;

// keys[v] = (rank class << s) | v, v < 2^s: sorting their low 32 + s bits
// gives the search order.
__global__ void k_order_keys(const int64_t *__restrict__ off, int64_t nv, int order, int s,
                             uint64_t *__restrict__ keys, unsigned long long *__restrict__ active) {
    unsigned long long n = 0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = off[v + 1] - off[v];
        uint64_t cls;
        if (order == 1 || order == 4)
            cls = 0xffffffffULL - (uint64_t)(d < 0xffffffffLL ? d : 0xffffffffLL);
        else
            cls = d == 0 ? 1 : 0;
        keys[v] = (cls << s) | (uint64_t)v;
        n += d > 0;
    }
    n = wave_sum(n);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(active, n);
}

// Search id of the vertex of rank r (position in the sorted keys).  The
// `spread` highest ranks get a 64-bit word each (search id 64 r), the next
// 63 * spread ranks fill the other bits of those words in rank order, and
// every later rank keeps its position.  Top-down claims are atomic ORs on
// the visited word: with the hubs packed 64 to a word, every edge into any
// of them would hit the same few words.
__device__ __forceinline__ int64_t spread_id(int64_t r, int64_t spread) {
    if (r < spread) return 64 * r;
    const int64_t q = r - spread;
    if (q < 63 * spread) return 64 * (q / 63) + 1 + q % 63;
    return r;
}

// perm[i] = graph vertex of search id i, inv its inverse, sdeg[i] its degree.
__global__ void k_order_perm(const uint64_t *__restrict__ skeys, const int64_t *__restrict__ off,
                             int64_t nv, int s, int64_t spread, int32_t *__restrict__ perm,
                             int32_t *__restrict__ inv, int64_t *__restrict__ sdeg) {
    const uint64_t vmask = (1ULL << s) - 1;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nv;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = (int32_t)(skeys[r] & vmask);
        const int64_t i = spread_id(r, spread);
        perm[i] = v;
        inv[v] = (int32_t)i;
        sdeg[i] = off[v + 1] - off[v];
    }
}

// The renumbered lists: one wave per 64 search ids, a lane copying its own
// list when it is shorter than 64, the whole wave otherwise.
__global__ __launch_bounds__(256) void k_order_cols(
    const int64_t *__restrict__ off, const int32_t *__restrict__ cols,
    const int32_t *__restrict__ perm, const int32_t *__restrict__ inv,
    const int64_t *__restrict__ soff, int64_t nv, int32_t *__restrict__ scols) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t base = wave * 64; base < nv; base += nwaves * 64) {
        const int64_t i = base + lane;
        const bool valid = i < nv;
        int64_t b = 0, e = 0, sb = 0;
        if (valid) {
            const int32_t v = perm[i];
            b = off[v];
            e = off[v + 1];
            sb = soff[i];
        }
        if (valid && e - b < 64)
            for (int64_t k = b; k < e; ++k) scols[sb + (k - b)] = inv[cols[k]];
        // lists of kHubBigDeg or more: k_order_cols_big
        uint64_t lm = __ballot(valid && e - b >= 64 && e - b < kHubBigDeg);
        while (lm) {
            const int src = __ffsll((unsigned long long)lm) - 1;
            lm &= lm - 1;
            const int64_t kb = lane_bcast(b, src), ke = lane_bcast(e, src),
                          so = lane_bcast(sb, src);
            for (int64_t k = kb + lane; k < ke; k += 64) scols[so + (k - kb)] = inv[cols[k]];
        }
    }
}

// The lists of kHubBigDeg entries or more, one slice of the layout's hub
// slices per wave (in degree order they would share a few waves above).
__global__ __launch_bounds__(256) void k_order_cols_big(
    const int64_t *__restrict__ off, const int32_t *__restrict__ cols,
    const int32_t *__restrict__ perm, const int32_t *__restrict__ inv,
    const int64_t *__restrict__ soff, const td_task *__restrict__ big, int64_t nbig,
    int32_t *__restrict__ scols) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t t = wave; t < nbig; t += nwaves) {
        const td_task tk = big[t];
        const int64_t ob = off[perm[tk.src]] - soff[tk.src];  // old minus new list start
        for (int64_t k = tk.beg + lane; k < tk.end; k += 64) scols[k] = inv[cols[ob + k]];
    }
}

}  // namespace

void free_search_layout(bd_ctx *c) {
    if (c->d_perm) {  // renumbered: the layout owns its arrays
        dev_free(c->d_soff);
        dev_free(c->d_scols);
        dev_free(c->d_sdead);
    }
    dev_free(c->d_perm);
    dev_free(c->d_inv);
    dev_free(c->d_head);
    dev_free(c->d_hub_bm);
    dev_free(c->d_hub_med);
    dev_free(c->d_hub_big);
    c->d_hub_bm = nullptr;
    c->d_hub_med = nullptr;
    c->d_hub_big = nullptr;
    c->n_hub_med = c->n_hub_big = 0;
    c->d_soff = nullptr;
    c->d_scols = nullptr;
    c->d_sdead = nullptr;
    c->d_perm = nullptr;
    c->d_inv = nullptr;
    c->d_head = nullptr;
    c->nv_active = 0;
    c->s_order = 0;
}

// The hub lists of the layout just built (d_soff).
static int build_hubs(bd_ctx *c) {
    const int64_t nnz = c->nnz;
    int rc;
    dev_buf cnt;
    if ((rc = dev_alloc(c, &cnt.p, 2 * sizeof(unsigned long long), "hub counts"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_hub_bm, sizeof(uint64_t) * c->nwords, "hub bitmap")))
        return rc;
    const int64_t max_med = nnz / kHubDeg + 1;
    const int64_t max_big = nnz / kHubSlice + nnz / kHubBigDeg + 1;
    if ((rc = dev_alloc(c, (void **)&c->d_hub_med, sizeof(int32_t) * max_med, "hubs"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_hub_big, sizeof(td_task) * max_big, "hub slices")))
        return rc;
    BD_HIP(c, hipMemsetAsync(cnt.p, 0, 2 * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_hub_lists, dim3(grid1d(c, c->nwords * 64)), dim3(kBlock), 0, c->stream,
                       c->d_soff, c->nv, c->nwords, c->d_hub_bm, c->d_hub_med, c->d_hub_big,
                       cnt.as<unsigned long long>());
    BD_HIP(c, hipGetLastError());
    unsigned long long h[2] = {0, 0};
    BD_HIP(c, hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    c->n_hub_med = (int64_t)h[0];
    c->n_hub_big = (int64_t)h[1];
    return BD_OK;
}

int build_search_layout(bd_ctx *c) {
    const int order = c->tune.vertex_order;
    if (c->s_order == order && c->d_soff) return BD_OK;
    free_search_layout(c);
    const int64_t nv = c->nv, nnz = c->nnz;
    int rc;
    dev_buf keys, skeys, sdeg, cnt, tmp;
    if ((rc = dev_alloc(c, &cnt.p, sizeof(unsigned long long), "active count"))) return rc;
    if ((rc = dev_alloc(c, &keys.p, sizeof(uint64_t) * nv, "order keys"))) return rc;
    BD_HIP(c, hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), c->stream));
    const int sh = bits_for(nv);
    hipLaunchKernelGGL(k_order_keys, dim3(grid1d(c, nv)), dim3(kBlock), 0, c->stream, c->d_off,
                       nv, order == 2 ? 3 : order, sh, keys.as<uint64_t>(),
                       cnt.as<unsigned long long>());
    BD_HIP(c, hipGetLastError());
    unsigned long long active = 0;
    BD_HIP(c, hipMemcpyAsync(&active, cnt.p, sizeof(active), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    c->nv_active = (int64_t)active;
    if (order == 2) {  // the graph's own order
        c->d_soff = c->d_off;
        c->d_scols = c->d_cols;
        c->d_sdead = c->d_dead;
        c->s_order = order;
        return build_hubs(c);
    }
    if ((rc = dev_alloc(c, &skeys.p, sizeof(uint64_t) * nv, "sorted order keys"))) return rc;
    if ((rc = sort_keys_u64(c, keys.as<uint64_t>(), skeys.as<uint64_t>(), nv, 32 + sh))) return rc;
    if ((rc = dev_alloc(c, &sdeg.p, sizeof(int64_t) * nv, "search degrees"))) return rc;
    // from here the layout owns what it allocates (freed by free_search_layout)
    if ((rc = dev_alloc(c, (void **)&c->d_perm, sizeof(int32_t) * nv, "search order"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_inv, sizeof(int32_t) * nv, "search ids"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_soff, sizeof(int64_t) * (nv + 1), "search offsets")))
        return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_scols, sizeof(int32_t) * nnz, "search cols")))
        return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_sdead, sizeof(uint64_t) * c->nwords,
                        "search dead bitmap")))
        return rc;
    // degree order spreads the 4096 largest degrees one to a visited word
    const int64_t spread = order == 1 ? std::min<int64_t>(kSpreadHubs, c->nv_active / 64) : 0;
    hipLaunchKernelGGL(k_order_perm, dim3(grid1d(c, nv)), dim3(kBlock), 0, c->stream,
                       skeys.as<uint64_t>(), c->d_off, nv, sh, spread, c->d_perm, c->d_inv,
                       sdeg.as<int64_t>());
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipMemsetAsync(c->d_soff, 0, sizeof(int64_t), c->stream));
    size_t tb = 0;
    BD_HIP(c, rocprim::inclusive_scan(nullptr, tb, sdeg.as<int64_t>(), c->d_soff + 1, (size_t)nv,
                                      rocprim::plus<int64_t>(), c->stream));
    if ((rc = dev_alloc(c, &tmp.p, tb, "scan temp"))) return rc;
    BD_HIP(c, rocprim::inclusive_scan(tmp.p, tb, sdeg.as<int64_t>(), c->d_soff + 1, (size_t)nv,
                                      rocprim::plus<int64_t>(), c->stream));
    if ((rc = build_hubs(c))) return rc;
    hipLaunchKernelGGL(k_order_cols, dim3(grid1d(c, nv)), dim3(kBlock), 0, c->stream, c->d_off,
                       c->d_cols, c->d_perm, c->d_inv, c->d_soff, nv, c->d_scols);
    if (c->n_hub_big)
        hipLaunchKernelGGL(k_order_cols_big, dim3(grid1d(c, c->n_hub_big * 64)), dim3(kBlock), 0,
                           c->stream, c->d_off, c->d_cols, c->d_perm, c->d_inv, c->d_soff,
                           c->d_hub_big, c->n_hub_big, c->d_scols);
    hipLaunchKernelGGL(k_dead_bitmap, dim3(grid1d(c, c->nwords * 64)), dim3(kBlock), 0, c->stream,
                       c->d_soff, nv, c->nwords, c->d_sdead);
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipStreamSynchronize(c->stream));
    c->s_order = order;
    return BD_OK;
}

}  // namespace bd

using namespace bd;

static int check_nv(bd_ctx *c, int64_t nv) {
    if (nv <= 0 || nv > ((int64_t)1 << 31) - 1)
        return set_err(c, BD_EINVAL, "num_vertices must be in [1, 2^31)");
    return BD_OK;
}

extern "C" int bd_graph_upload(bd_ctx *c, const int64_t *off, const int32_t *cols, int64_t nv,
                               int64_t nnz, int64_t m) {
    if (!c || !off || (nnz > 0 && !cols) || nnz < 0 || m < 0) return BD_EINVAL;
    int rc = check_nv(c, nv);
    if (rc) return rc;
    if (off[0] != 0 || off[nv] != nnz) return set_err(c, BD_EINVAL, "offsets[0] or offsets[nv]");
    BD_HIP(c, hipSetDevice(c->device));
    free_graph(c);
    const int64_t nwords = (nv + 63) / 64;
    if ((rc = dev_alloc(c, (void **)&c->d_off, sizeof(int64_t) * (nv + 1), "offsets"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_cols, sizeof(int32_t) * nnz, "cols"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_dead, sizeof(uint64_t) * nwords, "dead bitmap")))
        return rc;
    BD_HIP(c, hipMemcpy(c->d_off, off, sizeof(int64_t) * (nv + 1), hipMemcpyHostToDevice));
    if (nnz) BD_HIP(c, hipMemcpy(c->d_cols, cols, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    dev_buf err;
    if ((rc = dev_alloc(c, &err.p, sizeof(int), "flag"))) return rc;
    BD_HIP(c, hipMemsetAsync(err.p, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_validate_csr, dim3(grid1d(c, std::max(nv, nnz))), dim3(kBlock), 0,
                       c->stream, c->d_off, c->d_cols, nv, nnz, err.as<int>());
    hipLaunchKernelGGL(k_dead_bitmap, dim3(grid1d(c, nwords * 64)), dim3(kBlock), 0, c->stream,
                       c->d_off, nv, nwords, c->d_dead);
    BD_HIP(c, hipGetLastError());
    int h = 0;
    BD_HIP(c, hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    if (h) {
        free_graph(c);
        return set_err(c, h & 1 ? BD_ERANGE : BD_EINVAL,
                       h & 1 ? "neighbour id out of range" : "offsets not monotone");
    }
    c->nv = nv;
    c->nnz = nnz;
    c->m = m;
    c->nwords = nwords;
    return alloc_search(c);
}

// Pack host (src,dst) pairs into c->d_el at position `at`, validating ids.
static int upload_pairs(bd_ctx *c, const int64_t *src, const int64_t *dst, int64_t n, int64_t at,
                        int64_t nv, int64_t *dstage, int *derr) {
    BD_HIP(c, hipMemcpyAsync(dstage, src, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->stream));
    BD_HIP(c, hipMemcpyAsync(dstage + n, dst, sizeof(int64_t) * n, hipMemcpyHostToDevice,
                             c->stream));
    hipLaunchKernelGGL(k_pack_pairs, dim3(grid1d(c, n)), dim3(kBlock), 0, c->stream, dstage,
                       dstage + n, n, nv, c->d_el + at, derr);
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipStreamSynchronize(c->stream));
    return BD_OK;
}

extern "C" int bd_graph_from_edges(bd_ctx *c, const int64_t *src, const int64_t *dst, int64_t nv,
                                   int64_t m) {
    if (!c || m < 0 || (m > 0 && (!src || !dst))) return BD_EINVAL;
    int rc = check_nv(c, nv);
    if (rc) return rc;
    BD_HIP(c, hipSetDevice(c->device));
    free_graph(c);
    if ((rc = dev_alloc(c, (void **)&c->d_el, sizeof(uint64_t) * m, "edge list"))) return rc;
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(m, 1), (int64_t)1 << 24);
    dev_buf stage, err;
    if ((rc = dev_alloc(c, &stage.p, sizeof(int64_t) * 2 * chunk, "staging"))) return rc;
    if ((rc = dev_alloc(c, &err.p, sizeof(int), "flag"))) return rc;
    BD_HIP(c, hipMemsetAsync(err.p, 0, sizeof(int), c->stream));
    for (int64_t at = 0; at < m; at += chunk) {
        const int64_t n = std::min(chunk, m - at);
        if ((rc = upload_pairs(c, src + at, dst + at, n, at, nv, stage.as<int64_t>(),
                               err.as<int>())))
            return rc;
    }
    int h = 0;
    BD_HIP(c, hipMemcpy(&h, err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (h) {
        free_graph(c);
        return set_err(c, BD_ERANGE, "vertex id out of range");
    }
    return build_csr_from_el(c, nv, m);
}

namespace {

// edge_list.cc:20-89: the header is one text line parsed like a getopt_long
// option string (long options, unambiguous prefixes, --opt value or
// --opt=value).  Unknown options are an error; non-option words are ignored.
struct el_header {
    int64_t nv = -1, m = -1;
    bool deduped = false, sorted = false;
    std::string format;
    bool has_format = false;
    size_t length = 0;
};

int parse_header(FILE *fp, el_header *h, std::string *why) {
    char line[256];
    if (!fgets(line, sizeof(line), fp)) {
        *why = "Failed to read edge list header";
        return BD_EIO;
    }
    size_t len = strlen(line);
    h->length = len;
    if (len == 0 || line[len - 1] != '\n') {
        *why = "Invalid edge list file header";
        return BD_EIO;
    }
    line[--len] = '\0';
    std::vector<std::string> tok;
    for (char *t = strtok(line, " "); t; t = strtok(nullptr, " ")) tok.push_back(t);
    struct opt {
        const char *name;
        bool arg;
    };
    static const opt opts[] = {{"num_vertices", true}, {"num_edges", true},  {"is_sorted", false},
                               {"is_deduped", false},  {"is_permuted", false}, {"is_directed", false},
                               {"is_undirected", false}, {"format", true}};
    for (size_t i = 0; i < tok.size(); ++i) {
        const std::string &t = tok[i];
        if (t.size() < 2 || t[0] != '-') continue;  // non-option word
        if (t == "--") break;
        if (t[1] != '-') {
            *why = "Invalid field edge list header";
            return BD_EIO;
        }
        std::string name = t.substr(2), val;
        bool has_val = false;
        size_t eq = name.find('=');
        if (eq != std::string::npos) {
            val = name.substr(eq + 1);
            name = name.substr(0, eq);
            has_val = true;
        }
        const opt *hit = nullptr;
        int nhit = 0;
        for (const opt &o : opts) {
            if (name == o.name) {
                hit = &o;
                nhit = 1;
                break;
            }
            if (!name.empty() && strncmp(o.name, name.c_str(), name.size()) == 0) {
                hit = &o;
                ++nhit;
            }
        }
        if (nhit != 1) {
            *why = "Invalid field edge list header";
            return BD_EIO;
        }
        if (hit->arg && !has_val) {
            if (i + 1 >= tok.size()) {
                *why = "Invalid field edge list header";
                return BD_EIO;
            }
            val = tok[++i];
        } else if (!hit->arg && has_val) {
            *why = "Invalid field edge list header";
            return BD_EIO;
        }
        const std::string n = hit->name;
        if (n == "num_vertices") h->nv = atol(val.c_str());
        else if (n == "num_edges") h->m = atol(val.c_str());
        else if (n == "is_sorted") h->sorted = true;
        else if (n == "is_deduped") h->deduped = true;
        else if (n == "format") {
            h->format = val;
            h->has_format = true;
        }
    }
    return BD_OK;
}

}  // namespace

// dist_edge_list::load_binary (dist_edge_list.cc:50-143) with its checks
// (:55-76), then the device CSR build.
extern "C" int bd_graph_load_el64(bd_ctx *c, const char *path) {
    if (!c || !path) return BD_EINVAL;
    FILE *fp = fopen(path, "rb");
    if (!fp) return set_err(c, BD_EIO, std::string("Unable to open ") + path);
    el_header h;
    std::string why;
    int rc = parse_header(fp, &h, &why);
    if (rc) {
        fclose(fp);
        return set_err(c, rc, why);
    }
    if (h.nv <= 0 || h.m <= 0) {
        fclose(fp);
        return set_err(c, BD_EIO, "Invalid graph size in header");
    }
    if (!h.has_format || h.format != "el64") {
        fclose(fp);
        return set_err(c, BD_EIO, "Unsuppported edge list format " + h.format);
    }
    if (!h.deduped) {
        fclose(fp);
        return set_err(c, BD_EIO, "Edge list must be sorted and deduped.");
    }
    if ((rc = check_nv(c, h.nv))) {
        fclose(fp);
        return rc;
    }
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) {
        fclose(fp);
        return hip_err(c, e, "hipSetDevice");
    }
    free_graph(c);
    if ((rc = dev_alloc(c, (void **)&c->d_el, sizeof(uint64_t) * h.m, "edge list"))) {
        fclose(fp);
        return rc;
    }
    const int64_t chunk = std::min<int64_t>(h.m, (int64_t)1 << 22);
    std::vector<int64_t> buf(2 * chunk), s(chunk), d(chunk);
    dev_buf stage, err;
    if ((rc = dev_alloc(c, &stage.p, sizeof(int64_t) * 2 * chunk, "staging")) ||
        (rc = dev_alloc(c, &err.p, sizeof(int), "flag"))) {
        fclose(fp);
        return rc;
    }
    e = hipMemsetAsync(err.p, 0, sizeof(int), c->stream);
    for (int64_t at = 0; at < h.m && e == hipSuccess; at += chunk) {
        const int64_t n = std::min(chunk, h.m - at);
        size_t got = fread(buf.data(), sizeof(int64_t) * 2, (size_t)n, fp);
        if (got != (size_t)n) {
            fclose(fp);
            free_graph(c);
            return set_err(c, BD_EIO, std::string("Failed to load edge list from ") + path +
                                          (feof(fp) ? " unexpected EOF" : " read error"));
        }
        for (int64_t i = 0; i < n; ++i) {
            s[i] = buf[2 * i];
            d[i] = buf[2 * i + 1];
        }
        if ((rc = upload_pairs(c, s.data(), d.data(), n, at, h.nv, stage.as<int64_t>(),
                               err.as<int>()))) {
            fclose(fp);
            return rc;
        }
    }
    fclose(fp);
    if (e != hipSuccess) return hip_err(c, e, "hipMemsetAsync");
    int herr = 0;
    BD_HIP(c, hipMemcpy(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        free_graph(c);
        return set_err(c, BD_ERANGE, "vertex id out of range");
    }
    return build_csr_from_el(c, h.nv, h.m);
}

extern "C" int bd_graph_generate(bd_ctx *c, int kind, int64_t nv, int64_t num_edges, double a,
                                 double b, double cc, double d, uint32_t seed, int64_t *m_out) {
    if (!c) return BD_EINVAL;
    int rc = check_nv(c, nv);
    if (rc) return rc;
    if (num_edges < 0) return set_err(c, BD_EINVAL, "num_edges < 0");
    BD_HIP(c, hipSetDevice(c->device));
    free_graph(c);
    const int s = bits_for(nv);
    if (kind == BD_GEN_LATTICE) {
        const int64_t cols = num_edges;
        if (cols <= 0 || nv % cols) return set_err(c, BD_EINVAL, "lattice: nv must be rows*cols");
        const int64_t rows = nv / cols;
        const int64_t m = rows * (cols - 1) + (rows - 1) * cols;
        dev_buf el;
        if ((rc = dev_alloc(c, &el.p, sizeof(uint64_t) * std::max<int64_t>(m, 1), "lattice")))
            return rc;
        hipLaunchKernelGGL(k_lattice, dim3(grid1d(c, nv)), dim3(kBlock), 0, c->stream, rows, cols,
                           el.as<uint64_t>());
        BD_HIP(c, hipGetLastError());
        if ((rc = finish_edges(c, el.as<uint64_t>(), m, s, nv, false, seed, true))) return rc;
        if (m_out) *m_out = m;
        return build_csr_from_el(c, nv, m);
    }
    if (kind != BD_GEN_RMAT && kind != BD_GEN_ER) return set_err(c, BD_EINVAL, "unknown kind");
    if (kind == BD_GEN_RMAT && (nv & (nv - 1)))
        return set_err(c, BD_EINVAL, "RMAT needs a power-of-two vertex count");
    const int64_t ne = num_edges;
    dev_buf A, B;
    if ((rc = dev_alloc(c, &A.p, sizeof(uint64_t) * std::max<int64_t>(ne, 1), "raw edges")))
        return rc;
    if ((rc = dev_alloc(c, &B.p, sizeof(uint64_t) * std::max<int64_t>(ne, 1), "sorted edges")))
        return rc;
    if (kind == BD_GEN_RMAT) {
        const int scale = s;
        if (scale < 1) return set_err(c, BD_EINVAL, "RMAT needs nv >= 2");
        const int64_t self_cap = (int64_t)1 << 20;
        dev_buf nself, sidx;
        if ((rc = dev_alloc(c, &nself.p, sizeof(unsigned long long), "self count")) ||
            (rc = dev_alloc(c, &sidx.p, sizeof(int64_t) * self_cap, "self index")))
            return rc;
        BD_HIP(c, hipMemsetAsync(nself.p, 0, sizeof(unsigned long long), c->stream));
        hipLaunchKernelGGL(k_rmat_raw, dim3(grid1d(c, ne)), dim3(kBlock), 0, c->stream, ne, scale,
                           a, b, cc, d, (uint64_t)seed, A.as<uint64_t>(),
                           nself.as<unsigned long long>(), sidx.as<int64_t>(), self_cap);
        BD_HIP(c, hipGetLastError());
        unsigned long long hn = 0;
        BD_HIP(c, hipMemcpyAsync(&hn, nself.p, sizeof(hn), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        if ((int64_t)hn > self_cap) return set_err(c, BD_ENOMEM, "too many self-loops to re-roll");
        if (hn) {
            // rmat_generator.h:178-189: serial re-roll from the stream position
            // that follows the last generated edge.
            std::vector<int64_t> idx(hn);
            BD_HIP(c, hipMemcpy(idx.data(), sidx.p, sizeof(int64_t) * hn, hipMemcpyDeviceToHost));
            std::sort(idx.begin(), idx.end());
            std::vector<uint64_t> val(hn);
            sitmo_reader rd;
            rd.init(seed, (uint64_t)ne * rmat_draws_per_edge(scale));
            for (size_t k = 0; k < idx.size(); ++k) {
                int64_t x = 0, y = 0;
                do {
                    rmat_edge(rd, scale, a, b, cc, d, &x, &y);
                } while (x == y);
                val[k] = ((uint64_t)x << 32) | (uint64_t)y;
            }
            dev_buf pv;
            if ((rc = dev_alloc(c, &pv.p, sizeof(uint64_t) * hn, "patch"))) return rc;
            BD_HIP(c, hipMemcpy(sidx.p, idx.data(), sizeof(int64_t) * hn, hipMemcpyHostToDevice));
            BD_HIP(c, hipMemcpy(pv.p, val.data(), sizeof(uint64_t) * hn, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_patch, dim3(grid1d(c, hn)), dim3(kBlock), 0, c->stream,
                               A.as<uint64_t>(), sidx.as<int64_t>(), pv.as<uint64_t>(),
                               (int64_t)hn);
            BD_HIP(c, hipGetLastError());
        }
    } else {
        hipLaunchKernelGGL(k_er_raw, dim3(grid1d(c, ne)), dim3(kBlock), 0, c->stream, ne, nv,
                           (uint64_t)seed, A.as<uint64_t>());
        BD_HIP(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_flip_key, dim3(grid1d(c, ne)), dim3(kBlock), 0, c->stream,
                       A.as<uint64_t>(), ne, s);
    BD_HIP(c, hipGetLastError());
    if ((rc = sort_keys_u64(c, A.as<uint64_t>(), B.as<uint64_t>(), ne, 64))) return rc;
    int64_t m = 0;
    if ((rc = unique_u64(c, B.as<uint64_t>(), A.as<uint64_t>(), ne, &m))) return rc;
    if (m > 0) {
        uint64_t last = 0;
        BD_HIP(c, hipMemcpy(&last, A.as<uint64_t>() + (m - 1), sizeof(uint64_t),
                            hipMemcpyDeviceToHost));
        if (last == ~0ULL) --m;  // the self-loop sentinel (ER)
    }
    if ((rc = finish_edges(c, A.as<uint64_t>(), m, s, nv, kind == BD_GEN_RMAT, seed, false)))
        return rc;
    if (m_out) *m_out = m;
    return build_csr_from_el(c, nv, m);
}

extern "C" int bd_graph_edges_download(bd_ctx *c, int64_t *src, int64_t *dst) {
    if (!c) return BD_EINVAL;
    if (!c->d_el) return set_err(c, BD_ESTATE, "no edge list resident (graph was uploaded as CSR)");
    BD_HIP(c, hipSetDevice(c->device));
    const int64_t m = c->el_m;
    dev_buf s, d;
    int rc;
    if ((rc = dev_alloc(c, &s.p, sizeof(int64_t) * m, "unpack")) ||
        (rc = dev_alloc(c, &d.p, sizeof(int64_t) * m, "unpack")))
        return rc;
    hipLaunchKernelGGL(k_unpack, dim3(grid1d(c, m)), dim3(kBlock), 0, c->stream, c->d_el, m,
                       s.as<int64_t>(), d.as<int64_t>());
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipStreamSynchronize(c->stream));
    if (src) BD_HIP(c, hipMemcpy(src, s.p, sizeof(int64_t) * m, hipMemcpyDeviceToHost));
    if (dst) BD_HIP(c, hipMemcpy(dst, d.p, sizeof(int64_t) * m, hipMemcpyDeviceToHost));
    return BD_OK;
}

// generator/edge_list_utils.h:171-198 dump_bin.
extern "C" int bd_graph_save_el64(bd_ctx *c, const char *path) {
    if (!c || !path) return BD_EINVAL;
    if (!c->d_el) return set_err(c, BD_ESTATE, "no edge list resident");
    std::vector<int64_t> s(c->el_m), d(c->el_m);
    int rc = bd_graph_edges_download(c, s.data(), d.data());
    if (rc) return rc;
    FILE *fp = fopen(path, "wb");
    if (!fp) return set_err(c, BD_EIO, std::string("Cannot open ") + path);
    fprintf(fp,
            " --format el64 --num_edges %lld --num_vertices %lld --is_undirected --is_deduped"
            " --is_permuted\n",
            (long long)c->el_m, (long long)c->nv);
    std::vector<int64_t> buf;
    const int64_t chunk = 1 << 20;
    buf.resize(2 * chunk);
    for (int64_t at = 0; at < c->el_m; at += chunk) {
        const int64_t n = std::min(chunk, c->el_m - at);
        for (int64_t i = 0; i < n; ++i) {
            buf[2 * i] = s[at + i];
            buf[2 * i + 1] = d[at + i];
        }
        if (fwrite(buf.data(), sizeof(int64_t) * 2, (size_t)n, fp) != (size_t)n) {
            fclose(fp);
            return set_err(c, BD_EIO, "short write");
        }
    }
    return fclose(fp) == 0 ? BD_OK : set_err(c, BD_EIO, "close failed");
}
