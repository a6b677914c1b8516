// bd_cc.hip -- connected components on gfx950 (components::run,
// components.cc:38-104, and components::check, :152-202).
//
// Labels are int32[V]; every vertex starts as its own component.  An
// iteration is an edge pass and a shortcut pass, as in the reference:
//   edge pass (:56-66): each vertex takes the smallest label among its
//     neighbours when that is smaller than its own.  The CSR is symmetric,
//     so every directed edge (src -> dst) lowers label[src] in src's own
//     pass: a label is written only by its vertex's lane, with no atomics.
//     Lists shorter than 64 are done by one lane, longer ones by the wave.
//   shortcut pass (:74-78): label[v] = label[label[v]] until it is a root.
// The run ends after the first edge pass that changes nothing (:70); the
// iteration count includes that pass.  Labels only decrease and stay ids of
// the vertex's own component, so each component ends labelled by its
// smallest vertex id, whatever order the passes read the labels in.
#include <hip/hip_runtime.h>

#include <vector>

#include "bd_device.h"
#include "bd_internal.h"

namespace bd {
namespace {

__global__ void k_cc_init(int32_t *__restrict__ comp, int64_t nv) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        comp[v] = (int32_t)v;
}

__global__ __launch_bounds__(256) void k_cc_hook(const int64_t *__restrict__ off,
                                                 const int32_t *__restrict__ cols, int32_t *comp,
                                                 int64_t nv, int32_t *__restrict__ changed) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    bool ch = false;
    for (int64_t base = wave * 64; base < nv; base += nwaves * 64) {
        const int64_t v = base + lane;
        const bool valid = v < nv;
        int64_t b = 0, e = 0;
        int32_t mine = 0;
        if (valid) {
            b = off[v];
            e = off[v + 1];
            mine = comp[v];
        }
        const bool alone = valid && e - b < 64;
        if (alone) {
            int32_t m = mine;
            for (int64_t k = b; k < e; ++k) {
                const int32_t l = comp[__builtin_nontemporal_load(cols + k)];
                m = l < m ? l : m;
            }
            if (m < mine) {
                comp[v] = m;
                ch = true;
            }
        }
        uint64_t lm = __ballot(valid && !alone);
        while (lm) {
            const int src = __ffsll((unsigned long long)lm) - 1;
            lm &= lm - 1;
            const int64_t kb = lane_bcast(b, src), ke = lane_bcast(e, src);
            unsigned long long m = 0xffffffffull;
            for (int64_t k = kb + lane; k < ke; k += 64) {
                const unsigned long long l =
                    (uint32_t)comp[__builtin_nontemporal_load(cols + k)];
                m = l < m ? l : m;
            }
            m = wave_min(m);
            if (lane == src && (int32_t)m < mine) {
                comp[v] = (int32_t)m;
                ch = true;
            }
        }
    }
    if (__ballot(ch) && lane == 0) changed[0] = 1;
}

__global__ void k_cc_shortcut(int32_t *comp, int64_t nv) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        int32_t c = comp[v];
        for (int32_t cc = comp[c]; cc != c; cc = comp[c]) c = cc;
        comp[v] = c;
    }
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
    const int64_t nv = c->nv;
    int rc;
    if (!c->d_comp && (rc = dev_alloc(c, (void **)&c->d_comp, sizeof(int32_t) * (nv + 1), "labels")))
        return rc;
    const size_t flag_bytes = sizeof(unsigned long long) * (kSlots * kLine + 1);
    if (!c->d_cc_flag && (rc = dev_alloc(c, (void **)&c->d_cc_flag, flag_bytes, "label flags")))
        return rc;
    c->have_cc = false;
    int32_t *changed = c->d_cc_flag;
    unsigned long long *count = (unsigned long long *)(c->d_cc_flag + 2);
    const dim3 grid(grid_for(c, nv, 256)), block(256);
    BD_HIP(c, hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(k_cc_init, grid, block, 0, c->stream, c->d_comp, nv);
    BD_HIP(c, hipGetLastError());
    int64_t iters = 0;
    for (;;) {
        ++iters;
        if (iters > nv + 2) return set_err(c, BD_EHIP, "components did not converge");
        BD_HIP(c, hipMemsetAsync(changed, 0, sizeof(int32_t), c->stream));
        hipLaunchKernelGGL(k_cc_hook, grid, block, 0, c->stream, c->d_off, c->d_cols, c->d_comp, nv,
                           changed);
        BD_HIP(c, hipGetLastError());
        int32_t h = 0;
        BD_HIP(c, hipMemcpyAsync(&h, changed, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        BD_HIP(c, hipStreamSynchronize(c->stream));
        if (!h) break;
        hipLaunchKernelGGL(k_cc_shortcut, grid, block, 0, c->stream, c->d_comp, nv);
        BD_HIP(c, hipGetLastError());
    }
    BD_HIP(c, hipMemsetAsync(count, 0, sizeof(unsigned long long) * kSlots * kLine, c->stream));
    hipLaunchKernelGGL(k_cc_count, grid, block, 0, c->stream, c->d_comp, nv, count);
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipEventRecord(c->ev[1], c->stream));
    std::vector<unsigned long long> h(kSlots * kLine);
    BD_HIP(c, hipMemcpyAsync(h.data(), count, sizeof(unsigned long long) * h.size(),
                             hipMemcpyDeviceToHost, c->stream));
    BD_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    BD_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    unsigned long long ncomp = 0;
    for (int s = 0; s < kSlots; ++s) ncomp += h[s * kLine];
    c->have_cc = true;
    if (out) {
        out->num_components = (int64_t)ncomp;
        out->num_iters = iters;
        out->ms = ms;
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
