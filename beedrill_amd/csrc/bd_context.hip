// bd_context.hip -- context lifetime, errors, buffers and small queries.
#include <hip/hip_ext.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "bd_internal.h"

namespace bd {

int set_err(bd_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int hip_err(bd_ctx *c, hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_err(c, e == hipErrorOutOfMemory ? BD_ENOMEM : BD_EHIP, m);
}

int dev_alloc(bd_ctx *c, void **p, size_t bytes, const char *what) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        char buf[160];
        snprintf(buf, sizeof(buf), "hipMalloc(%s, %zu bytes) failed: %s", what, bytes,
                 hipGetErrorString(e));
        return set_err(c, e == hipErrorOutOfMemory ? BD_ENOMEM : BD_EHIP, buf);
    }
    return BD_OK;
}

void dev_free(void *p) {
    if (p) (void)hipFree(p);
}

static void free_search(bd_ctx *c) {
    dev_free(c->d_depth);
    dev_free(c->d_parent);
    dev_free(c->d_cand);
    dev_free(c->d_dirty);
    c->d_dirty = nullptr;
    dev_free(c->d_comp);
    dev_free(c->d_cc_flag);
    dev_free(c->d_cc_tasks);
    dev_free(c->d_cc_chg);
    dev_free(c->d_cc_bcols);
    dev_free(c->d_cc_bend);
    dev_free(c->d_cc_pieces);
    dev_free(c->d_cc_pbase);
    c->d_cc_pieces = nullptr;
    c->d_cc_pbase = nullptr;
    c->d_cc_chg = nullptr;
    c->d_cc_bcols = nullptr;
    c->d_cc_bend = nullptr;
    c->cc_bshift = 0;
    c->d_comp = nullptr;
    c->d_cc_flag = nullptr;
    c->d_cc_tasks = nullptr;
    c->cc_ntasks = -1;
    c->cc_prep_ms = 0;
    c->cc_prep_bytes = 0;
    c->have_cc = false;
    dev_free(c->d_vis);
    dev_free(c->d_vis_save);
    for (auto &b : c->d_bm) dev_free(b);
    for (auto &q : c->d_q) dev_free(q);
    for (auto &t : c->d_tasks) dev_free(t);
    ms_free_graph(c);
    for (auto &b : c->d_ms_slices) dev_free(b);
    for (auto &b : c->d_ms_slices) b = nullptr;
    c->ms_max_slices = c->ms_seg_cap = 0;
    dev_free(c->d_ms_hubs);
    dev_free(c->d_ms_hub_slices);
    c->d_ms_hubs = nullptr;
    c->d_ms_hub_slices = nullptr;
    c->ms_nhubs = c->ms_nhub_slices = 0;
    c->ms_hub_order = 0;
    for (auto &b : c->d_ms) dev_free(b);
    for (auto &b : c->d_ms_bits) dev_free(b);
    dev_free(c->d_ms_depth);
    c->d_depth = c->d_parent = c->d_cand = nullptr;
    c->d_vis = c->d_vis_save = nullptr;
    for (auto &b : c->d_bm) b = nullptr;
    for (auto &q : c->d_q) q = nullptr;
    for (auto &t : c->d_tasks) t = nullptr;
    for (auto &b : c->d_ms) b = nullptr;
    for (auto &b : c->d_ms_bits) b = nullptr;
    c->d_ms_depth = nullptr;
    c->ms_depth_rows = 0;
    c->task_cap = 0;
    c->have_result = false;
    c->chunk_hint = 0;
}

void free_graph(bd_ctx *c) {
    free_search(c);
    free_search_layout(c);
    dev_free(c->d_off);
    dev_free(c->d_cols);
    dev_free(c->d_dead);
    dev_free(c->d_head);
    dev_free(c->d_el);
    c->d_head = nullptr;
    c->max_deg = 0;
    c->d_off = nullptr;
    c->d_cols = nullptr;
    c->d_dead = nullptr;
    c->d_el = nullptr;
    c->el_m = 0;
    c->nv = c->nnz = c->m = c->nwords = 0;
}

int64_t grid_for(const bd_ctx *c, int64_t work_items, int block) {
    int64_t per_cu = c->tune.grid_per_cu > 0 ? c->tune.grid_per_cu : 8;
    int64_t cap = (int64_t)c->cu_count * per_cu;
    int64_t need = (work_items + block - 1) / block;
    return std::max<int64_t>(1, std::min(cap, need));
}

int alloc_search(bd_ctx *c) {
    free_search(c);
    c->cand_clean = false;  // the first search fills the candidate array
    int rc;
    const int64_t nv = c->nv, nw = c->nwords;
    if ((rc = dev_alloc(c, (void **)&c->d_depth, sizeof(int32_t) * nv, "depth"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_parent, sizeof(int32_t) * nv, "parent"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_cand, sizeof(int32_t) * nv, "candidate parents")))
        return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_dirty, nw, "candidate words"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_vis, sizeof(uint64_t) * nw, "visited"))) return rc;
    if ((rc = dev_alloc(c, (void **)&c->d_vis_save, sizeof(uint64_t) * nw, "visited copy")))
        return rc;
    for (auto &b : c->d_bm)
        if ((rc = dev_alloc(c, (void **)&b, sizeof(uint64_t) * nw, "frontier bitmap"))) return rc;
    // Segmented lists: kSeg segments of twice the even share, then an
    // overflow segment as large as the whole list can be.
    c->qseg_cap = 2 * ((nv + kSeg - 1) / kSeg) + 4096;
    for (auto &q : c->d_q)
        if ((rc = dev_alloc(c, (void **)&q, sizeof(int32_t) * (kSeg * c->qseg_cap + nv),
                            "frontier queue")))
            return rc;
    // Each vertex of degree >= task_threshold yields <= deg/task_edges + 1
    // tasks and appears in at most one frontier, so this bound is exact.
    const int64_t te = std::max(64, c->tune.task_edges);
    const int64_t tt = std::max(16, c->tune.task_threshold);
    c->task_cap = c->nnz / te + c->nnz / tt + 16;
    c->tseg_cap = 2 * ((c->task_cap + kSeg - 1) / kSeg) + 1024;
    for (auto &t : c->d_tasks)
        if ((rc = dev_alloc(c, (void **)&t, sizeof(td_task) * (kSeg * c->tseg_cap + c->task_cap),
                            "tasks")))
            return rc;
    if ((rc = build_search_layout(c))) return rc;
    return build_head_index(c);
}

}  // namespace bd

using namespace bd;

extern "C" int bd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

extern "C" int bd_create(bd_ctx **out, int device) {
    if (!out) return BD_EINVAL;
    *out = nullptr;
    int n = bd_device_count();
    if (n <= 0) return BD_ENOEXT;
    if (device < 0 || device >= n) return BD_EINVAL;
    bd_ctx *c = new bd_ctx();
    c->device = device;
    // launch shape and thresholds: beedrill_amd/autotune.py's best on MI355X
    // (RMAT scale 24, generator seed 1, models/tuning_rmat_s24.json), except
    // bu_cap (its 64 gained 0.8% there and cost an Erdos-Renyi graph 6%) and
    // bu_batch 8 (six rounds per trip, added after that run)
    c->tune.block = 128;
    c->tune.grid_per_cu = 10;
    c->tune.bu_words = 16;
    c->tune.bu_cap = 32;
    c->tune.wave_threshold = 256;
    c->tune.task_threshold = 16;
    c->tune.task_edges = 1536;
    c->tune.level_driver = 1;
    c->tune.time_steps = 2;
    c->tune.bu_head = 1;
    c->tune.step_kernels = 2;
    c->tune.bu_grid_per_cu = 4;
    c->tune.learned_margin_milli = 250;
    c->tune.ms_alpha = 4;
    c->tune.learned_min_skew = 8;
    c->tune.td_split_edges = 1 << 16;
    c->tune.vertex_order = 4;
    c->tune.bu_batch = 8;
    c->tune.learned_near = 4;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, device);
        if (e == hipSuccess) c->cu_count = prop.multiProcessorCount;
    }
    int khz = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device);
    if (e == hipSuccess && khz > 0) c->wall_khz = khz;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // ev[0..1] bracket whole searches; ev[2..3] time single kernels and use
    // a device-scope release, so a kernel's time does not include writing
    // the L2 back for the host (results are read with hipMemcpy, which
    // orders itself)
    for (int i = 0; i < 4 && e == hipSuccess; ++i)
        e = i < 2 ? hipEventCreate(&c->ev[i])
                  : hipEventCreateWithFlags(&c->ev[i], hipEventReleaseToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_ctl, 2 * sizeof(ctl_block));
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_seg, sizeof(seg_counters));
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_ctl, 2 * sizeof(ctl_block), hipHostMallocDefault);
    if (e == hipSuccess) {
        c->log_cap = 1 << 16;
        e = hipMalloc((void **)&c->d_log, sizeof(bd_level_rec) * c->log_cap);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        bd_destroy(c);
        return BD_EHIP;
    }
    if (static_scratch_init(c) != BD_OK) {
        bd_destroy(c);
        return BD_EHIP;
    }
    *out = c;
    return BD_OK;
}

extern "C" int bd_destroy(bd_ctx *c) {
    if (!c) return BD_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_graph(c);
    dev_free(c->d_ctl);
    dev_free(c->d_seg);
    dev_free(c->d_log);
    dev_free(c->d_script);
    dev_free(c->d_mlp);
    dev_free(c->d_mlp_ops);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->ev_pool)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BD_OK;
}

extern "C" const char *bd_last_error(const bd_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int bd_graph_info(bd_ctx *c, int64_t *nv, int64_t *nnz, int64_t *m) {
    if (!c) return BD_EINVAL;
    if (nv) *nv = c->nv;
    if (nnz) *nnz = c->nnz;
    if (m) *m = c->m;
    return BD_OK;
}

extern "C" int bd_graph_download(bd_ctx *c, int64_t *off, int32_t *cols) {
    if (!c) return BD_EINVAL;
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph");
    BD_HIP(c, hipSetDevice(c->device));
    if (off)
        BD_HIP(c, hipMemcpy(off, c->d_off, sizeof(int64_t) * (c->nv + 1), hipMemcpyDeviceToHost));
    if (cols)
        BD_HIP(c, hipMemcpy(cols, c->d_cols, sizeof(int32_t) * c->nnz, hipMemcpyDeviceToHost));
    return BD_OK;
}

// hybrid_bfs_main.cc:136-144: rng() % V from lcg(0) (lcg.h), rejecting
// vertices of degree 0.  Degrees are read from the device CSR on demand.
extern "C" int bd_pick_roots(bd_ctx *c, int64_t n, int64_t *roots) {
    if (!c || n < 0 || (n > 0 && !roots)) return BD_EINVAL;
    if (!c->d_off) return set_err(c, BD_ESTATE, "no graph");
    BD_HIP(c, hipSetDevice(c->device));
    std::vector<int64_t> off(c->nv + 1);
    BD_HIP(c, hipMemcpy(off.data(), c->d_off, sizeof(int64_t) * (c->nv + 1), hipMemcpyDeviceToHost));
    if (off[c->nv] == 0 && n > 0) return set_err(c, BD_ESTATE, "graph has no edges");
    uint64_t st = 1;  // lcg(0): the skip-ahead loop leaves ran = 1
    const uint64_t mul = 6364136223846793005ULL;
    for (int64_t i = 0; i < n; ++i) {
        int64_t s;
        do {
            st = mul * st + 1ULL;
            s = (int64_t)(st % (uint64_t)c->nv);
        } while (off[s + 1] - off[s] == 0);
        roots[i] = s;
    }
    return BD_OK;
}

extern "C" int bd_set_tuning(bd_ctx *c, const bd_tuning *t) {
    if (!c || !t) return BD_EINVAL;
    bd_tuning n = c->tune;
    if (t->block > 0) {
        if (t->block != 128 && t->block != 256 && t->block != 512)
            return set_err(c, BD_EINVAL, "block must be 128, 256 or 512");
        n.block = t->block;
    }
    if (t->bu_words > 0) n.bu_words = std::min(t->bu_words, 32);
    if (t->bu_cap > 0) n.bu_cap = t->bu_cap;
    if (t->grid_per_cu > 0) n.grid_per_cu = std::min(t->grid_per_cu, 64);
    if (t->wave_threshold > 0) n.wave_threshold = t->wave_threshold;
    if (t->task_threshold > 0) n.task_threshold = std::max(t->task_threshold, 16);
    if (t->task_edges > 0) n.task_edges = std::max(t->task_edges, 64);
    if (t->level_driver == 1 || t->level_driver == 2) n.level_driver = t->level_driver;
    if (t->time_steps >= 1 && t->time_steps <= 4) n.time_steps = t->time_steps;
    if (t->bu_head == 1 || t->bu_head == 2) n.bu_head = t->bu_head;
    if (t->step_kernels == 1 || t->step_kernels == 2) n.step_kernels = t->step_kernels;
    if (t->bu_grid_per_cu > 0) n.bu_grid_per_cu = std::min(t->bu_grid_per_cu, 64);
    if (t->learned_margin_milli > 0) n.learned_margin_milli = t->learned_margin_milli;
    if (t->ms_alpha > 0) n.ms_alpha = t->ms_alpha;
    if (t->learned_min_skew > 0) n.learned_min_skew = t->learned_min_skew;
    if (t->td_split_edges > 0) n.td_split_edges = t->td_split_edges;
    if (t->vertex_order > 4) return set_err(c, BD_EINVAL, "vertex_order must be 1 to 4");
    if (t->vertex_order > 0) n.vertex_order = t->vertex_order;
    if (t->bu_batch > 0) n.bu_batch = t->bu_batch >= 8 ? 8 : (t->bu_batch >= 4 ? 4 : 1);
    if (t->learned_near > 0) n.learned_near = std::min(t->learned_near, 1 << 16);
    bool realloc = (n.task_edges != c->tune.task_edges || n.task_threshold != c->tune.task_threshold ||
                    n.vertex_order != c->tune.vertex_order);
    c->tune = n;
    if (realloc && c->d_off) {
        BD_HIP(c, hipSetDevice(c->device));
        return alloc_search(c);  // rebuilds the search layout when its order changed
    }
    return BD_OK;
}

extern "C" int bd_get_tuning(bd_ctx *c, bd_tuning *t) {
    if (!c || !t) return BD_EINVAL;
    *t = c->tune;
    return BD_OK;
}

// ------------------------------------------------------------ measurement

namespace {

// A plain streaming copy (16-byte loads and stores, grid-stride): the
// reference point for the hardware traffic counters and for achievable HBM
// bandwidth.
__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src,
                                              uint4 *__restrict__ dst, int64_t n) {
    // four independent 16-byte loads in flight per thread
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride],
                    d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int bd_copy_bandwidth(bd_ctx *c, int64_t bytes, int reps, double *gbps) {
    if (!c || bytes < 16 || reps < 1) return BD_EINVAL;
    BD_HIP(c, hipSetDevice(c->device));
    const int64_t n = bytes / 16;
    void *src = nullptr, *dst = nullptr;
    int rc = dev_alloc(c, &src, sizeof(uint4) * n, "copy source");
    if (!rc) rc = dev_alloc(c, &dst, sizeof(uint4) * n, "copy target");
    struct guard {
        void *a, *b;
        ~guard() {
            dev_free(a);
            dev_free(b);
        }
    } g{src, dst};
    if (rc) return rc;
    BD_HIP(c, hipMemsetAsync(src, 1, sizeof(uint4) * n, c->stream));
    const int64_t grid = (int64_t)c->cu_count * 4;
    float best = 0;
    for (int r = 0; r < reps; ++r) {
        hipExtLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, c->ev[2], c->ev[3], 0,
                              (const uint4 *)src, (uint4 *)dst, n);
        BD_HIP(c, hipGetLastError());
        BD_HIP(c, hipEventSynchronize(c->ev[3]));
        float ms = 0;
        BD_HIP(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        if (r == 0 || ms < best) best = ms;
    }
    if (gbps) *gbps = best > 0 ? 2.0 * 16.0 * (double)n / (best * 1e-3) / 1e9 : 0.0;
    return BD_OK;
}

