// bd_mlp.hip -- training and inference kernels of the learned direction
// policy (no reference counterpart; it replaces the predicates at
// hybrid_bfs.cc:167 and :178-179).
//
//   k_mlp_infer_f32   one thread per sample, fp32 (the numerics reference)
//   k_mlp_infer_mfma  one wave per 16 samples, bf16 MFMA, fp32 accumulate
//   k_mlp_grad        one thread per sample; parameter gradients are summed
//                     in LDS and then once per block in HBM
//   k_adam            one thread per parameter
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "bd_internal.h"
#include "bd_mlp.h"

namespace bd {

namespace {

constexpr int kMaxParams = 32 * 8 + 32 + 32 * 32 + 32 + 2 * 32 + 2;

struct mlp_ptrs {
    const float *w1, *b1, *w2, *b2, *w3, *b3;
};

__device__ __forceinline__ mlp_ptrs split(const float *p, int h1, int h2) {
    mlp_ptrs m;
    m.w1 = p;
    m.b1 = m.w1 + h1 * kMlpIn;
    m.w2 = m.b1 + h1;
    m.b2 = m.w2 + h2 * h1;
    m.w3 = m.b2 + h2;
    m.b3 = m.w3 + kMlpOut * h2;
    return m;
}

__device__ void fwd_thread(const mlp_ptrs &m, int h1, int h2, const float *x, float *a1,
                           float *a2, float *y) {
    for (int j = 0; j < h1; ++j) {
        float s = m.b1[j];
        for (int i = 0; i < kMlpIn; ++i) s += m.w1[j * kMlpIn + i] * x[i];
        a1[j] = s > 0.f ? s : 0.f;
    }
    for (int j = 0; j < h2; ++j) {
        float s = m.b2[j];
        for (int i = 0; i < h1; ++i) s += m.w2[j * h1 + i] * a1[i];
        a2[j] = s > 0.f ? s : 0.f;
    }
    for (int o = 0; o < kMlpOut; ++o) {
        float s = m.b3[o];
        for (int i = 0; i < h2; ++i) s += m.w3[o * h2 + i] * a2[i];
        y[o] = s;
    }
}

__global__ void k_mlp_infer_f32(const float *__restrict__ p, int h1, int h2,
                                const float *__restrict__ x, int64_t n, float *__restrict__ y) {
    const mlp_ptrs m = split(p, h1, h2);
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        float a1[kMlpMaxHidden], a2[kMlpMaxHidden], xs[kMlpIn], out[kMlpOut];
        for (int i = 0; i < kMlpIn; ++i) xs[i] = x[s * kMlpIn + i];
        fwd_thread(m, h1, h2, xs, a1, a2, out);
        y[s * 2] = out[0];
        y[s * 2 + 1] = out[1];
    }
}

__global__ __launch_bounds__(64) void k_mlp_infer_mfma(const float *__restrict__ p, int h1, int h2,
                                                       const float *__restrict__ x, int64_t n,
                                                       float *__restrict__ y) {
    __shared__ float X[16][32];
    __shared__ float H[16][32];
    __shared__ float Y[16][32];
    const int l = threadIdx.x;
    for (int64_t base = (int64_t)blockIdx.x * 16; base < n; base += (int64_t)gridDim.x * 16) {
        for (int i = l; i < 16 * 32; i += 64) {
            const int r = i / 32, cidx = i % 32;
            const int64_t s = base + r;
            X[r][cidx] = (s < n && cidx < kMlpIn) ? x[s * kMlpIn + cidx] : 0.f;
        }
        wave_lds_sync();
        mlp_forward_rows_mfma(p, h1, h2, X, H, Y);
        if (l < 32) {
            const int r = l >> 1, o = l & 1;
            if (base + r < n) y[(base + r) * 2 + o] = Y[r][o];
        }
        wave_lds_sync();
    }
}

__global__ void k_mlp_grad(const float *__restrict__ p, int h1, int h2, const float *__restrict__ x,
                           const float *__restrict__ t, const float *__restrict__ mask, int64_t n,
                           float inv_count, float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float g[kMaxParams];
    __shared__ float s_loss;
    const int np = (int)mlp_nparams(h1, h2);
    for (int i = threadIdx.x; i < np; i += blockDim.x) g[i] = 0.f;
    if (threadIdx.x == 0) s_loss = 0.f;
    __syncthreads();
    const mlp_ptrs m = split(p, h1, h2);
    float *g_w1 = g, *g_b1 = g_w1 + h1 * kMlpIn, *g_w2 = g_b1 + h1;
    float *g_b2 = g_w2 + h2 * h1, *g_w3 = g_b2 + h2, *g_b3 = g_w3 + kMlpOut * h2;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        float xs[kMlpIn], a1[kMlpMaxHidden], a2[kMlpMaxHidden], y[kMlpOut], dy[kMlpOut];
        float d2[kMlpMaxHidden];
        for (int i = 0; i < kMlpIn; ++i) xs[i] = x[s * kMlpIn + i];
        fwd_thread(m, h1, h2, xs, a1, a2, y);
        float ls = 0.f;
        for (int o = 0; o < kMlpOut; ++o) {
            const float mk = mask ? mask[s * 2 + o] : 1.f;
            const float e = y[o] - t[s * 2 + o];
            ls += mk * e * e;
            dy[o] = 2.f * mk * e * inv_count;
        }
        atomicAdd(&s_loss, ls);
        for (int j = 0; j < h2; ++j) {
            float acc = 0.f;
            for (int o = 0; o < kMlpOut; ++o) acc += m.w3[o * h2 + j] * dy[o];
            d2[j] = a2[j] > 0.f ? acc : 0.f;
        }
        for (int o = 0; o < kMlpOut; ++o) {
            atomicAdd(&g_b3[o], dy[o]);
            for (int j = 0; j < h2; ++j) atomicAdd(&g_w3[o * h2 + j], dy[o] * a2[j]);
        }
        for (int k = 0; k < h2; ++k) {
            if (d2[k] == 0.f) continue;
            atomicAdd(&g_b2[k], d2[k]);
            for (int j = 0; j < h1; ++j) atomicAdd(&g_w2[k * h1 + j], d2[k] * a1[j]);
        }
        for (int j = 0; j < h1; ++j) {
            float acc = 0.f;
            for (int k = 0; k < h2; ++k) acc += m.w2[k * h1 + j] * d2[k];
            const float d1 = a1[j] > 0.f ? acc : 0.f;
            if (d1 == 0.f) continue;
            atomicAdd(&g_b1[j], d1);
            for (int i = 0; i < kMlpIn; ++i) atomicAdd(&g_w1[j * kMlpIn + i], d1 * xs[i]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += blockDim.x)
        if (g[i] != 0.f) atomicAdd(&grad[i], g[i]);
    if (threadIdx.x == 0) atomicAdd(loss, s_loss * inv_count);
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ mo,
                       float *__restrict__ ve, int np, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    mo[i] = b1 * mo[i] + (1.f - b1) * g[i];
    ve[i] = b2 * ve[i] + (1.f - b2) * g[i] * g[i];
    const float mh = mo[i] / c1, vh = ve[i] / c2;
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

struct dbuf {
    void *p = nullptr;
    ~dbuf() { dev_free(p); }
};

int check_mlp(bd_ctx *c) {
    if (!c->has_mlp) return set_err(c, BD_ESTATE, "no network loaded");
    return BD_OK;
}

int64_t grid_n(bd_ctx *c, int64_t n, int block) {
    return std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, c->cu_count * 4));
}

// Gradient of the masked mean-squared error into d_grad (zeroed here).
int device_grad(bd_ctx *c, const float *dx, const float *dt, const float *dm, int64_t n,
                double count, float *d_grad, float *d_loss) {
    const int np = (int)mlp_nparams(c->mlp_h1, c->mlp_h2);
    BD_HIP(c, hipMemsetAsync(d_grad, 0, sizeof(float) * np, c->stream));
    BD_HIP(c, hipMemsetAsync(d_loss, 0, sizeof(float), c->stream));
    const float inv = count > 0 ? (float)(1.0 / count) : 0.f;
    hipLaunchKernelGGL(k_mlp_grad, dim3(grid_n(c, n, 256)), dim3(256), 0, c->stream, c->d_mlp,
                       c->mlp_h1, c->mlp_h2, dx, dt, dm, n, inv, d_grad, d_loss);
    BD_HIP(c, hipGetLastError());
    return BD_OK;
}

double mask_count(const float *mask, int64_t n) {
    if (!mask) return 2.0 * (double)n;
    double s = 0;
    for (int64_t i = 0; i < 2 * n; ++i) s += mask[i];
    return s;
}

}  // namespace

}  // namespace bd

using namespace bd;

// The step kernels' operand table of the current network (one wave).
__global__ __launch_bounds__(64) void k_mlp_ops(const float *p, int h1, int h2, uint4 *ops) {
    mlp_ops_rec u;
    mlp_prefetch(p, h1, h2, u.r);
#pragma unroll
    for (int i = 0; i < kMlpOpsQuads; ++i) ops[threadIdx.x * kMlpOpsQuads + i] = u.q[i];
}

namespace bd {
// Rebuild the operand table after the parameters change.
int refresh_mlp_ops(bd_ctx *c) {
    if (!c->d_mlp_ops) {
        int rc = dev_alloc(c, (void **)&c->d_mlp_ops, sizeof(uint4) * 64 * kMlpOpsQuads,
                           "mlp operands");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_mlp_ops, dim3(1), dim3(64), 0, c->stream, c->d_mlp, c->mlp_h1, c->mlp_h2,
                       c->d_mlp_ops);
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipStreamSynchronize(c->stream));
    return BD_OK;
}
}  // namespace bd

extern "C" int bd_mlp_set(bd_ctx *c, const float *params, int h1, int h2) {
    if (!c || !params) return BD_EINVAL;
    if (h1 < 1 || h2 < 1 || h1 > kMlpMaxHidden || h2 > kMlpMaxHidden)
        return set_err(c, BD_EINVAL, "hidden sizes must be in [1, 32]");
    BD_HIP(c, hipSetDevice(c->device));
    const int64_t np = mlp_nparams(h1, h2);
    dev_free(c->d_mlp);
    c->d_mlp = nullptr;
    c->has_mlp = false;
    int rc = dev_alloc(c, (void **)&c->d_mlp, sizeof(float) * np, "mlp");
    if (rc) return rc;
    BD_HIP(c, hipMemcpy(c->d_mlp, params, sizeof(float) * np, hipMemcpyHostToDevice));
    c->mlp_h1 = h1;
    c->mlp_h2 = h2;
    c->h_mlp.assign(params, params + np);
    if ((rc = refresh_mlp_ops(c))) return rc;
    c->has_mlp = true;
    return BD_OK;
}

extern "C" int bd_mlp_get(bd_ctx *c, float *params, int64_t cap, int *h1, int *h2) {
    if (!c) return BD_EINVAL;
    int rc = check_mlp(c);
    if (rc) return rc;
    const int64_t np = mlp_nparams(c->mlp_h1, c->mlp_h2);
    if (h1) *h1 = c->mlp_h1;
    if (h2) *h2 = c->mlp_h2;
    if (params) {
        if (cap < np) return set_err(c, BD_EINVAL, "buffer too small");
        BD_HIP(c, hipSetDevice(c->device));
        BD_HIP(c, hipMemcpy(params, c->d_mlp, sizeof(float) * np, hipMemcpyDeviceToHost));
    }
    return BD_OK;
}

extern "C" int bd_mlp_features(const bd_level_rec *r, int64_t nv, int64_t m, int64_t prev_dir,
                               float *x8) {
    if (!r || !x8 || nv <= 0) return BD_EINVAL;
    mlp_features_dev(r->frontier, r->frontier_deg, r->max_frontier_deg, r->unvisited,
                     r->unvisited_deg, r->level, nv, m, prev_dir, x8);
    return BD_OK;
}

extern "C" int bd_mlp_infer(bd_ctx *c, const float *x, int64_t n, float *y, int use_mfma) {
    if (!c || n < 0 || (n > 0 && (!x || !y))) return BD_EINVAL;
    int rc = check_mlp(c);
    if (rc || n == 0) return rc;
    BD_HIP(c, hipSetDevice(c->device));
    dbuf dx, dy;
    if ((rc = dev_alloc(c, &dx.p, sizeof(float) * kMlpIn * n, "x")) ||
        (rc = dev_alloc(c, &dy.p, sizeof(float) * 2 * n, "y")))
        return rc;
    BD_HIP(c, hipMemcpy(dx.p, x, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice));
    if (use_mfma) {
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((n + 15) / 16, c->cu_count * 8));
        hipLaunchKernelGGL(k_mlp_infer_mfma, dim3(grid), dim3(64), 0, c->stream, c->d_mlp,
                           c->mlp_h1, c->mlp_h2, (const float *)dx.p, n, (float *)dy.p);
    } else {
        hipLaunchKernelGGL(k_mlp_infer_f32, dim3(grid_n(c, n, 256)), dim3(256), 0, c->stream,
                           c->d_mlp, c->mlp_h1, c->mlp_h2, (const float *)dx.p, n, (float *)dy.p);
    }
    BD_HIP(c, hipGetLastError());
    BD_HIP(c, hipStreamSynchronize(c->stream));
    BD_HIP(c, hipMemcpy(y, dy.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
    return BD_OK;
}

extern "C" int bd_mlp_loss_grad(bd_ctx *c, const float *x, const float *t, const float *mask,
                                int64_t n, float *grad, float *loss) {
    if (!c || n <= 0 || !x || !t) return BD_EINVAL;
    int rc = check_mlp(c);
    if (rc) return rc;
    BD_HIP(c, hipSetDevice(c->device));
    const int np = (int)mlp_nparams(c->mlp_h1, c->mlp_h2);
    dbuf dx, dt, dm, dg, dl;
    if ((rc = dev_alloc(c, &dx.p, sizeof(float) * kMlpIn * n, "x")) ||
        (rc = dev_alloc(c, &dt.p, sizeof(float) * 2 * n, "t")) ||
        (rc = dev_alloc(c, &dg.p, sizeof(float) * np, "grad")) ||
        (rc = dev_alloc(c, &dl.p, sizeof(float), "loss")))
        return rc;
    BD_HIP(c, hipMemcpy(dx.p, x, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice));
    BD_HIP(c, hipMemcpy(dt.p, t, sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    if (mask) {
        if ((rc = dev_alloc(c, &dm.p, sizeof(float) * 2 * n, "mask"))) return rc;
        BD_HIP(c, hipMemcpy(dm.p, mask, sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    }
    if ((rc = device_grad(c, (const float *)dx.p, (const float *)dt.p, (const float *)dm.p, n,
                          mask_count(mask, n), (float *)dg.p, (float *)dl.p)))
        return rc;
    BD_HIP(c, hipStreamSynchronize(c->stream));
    if (grad) BD_HIP(c, hipMemcpy(grad, dg.p, sizeof(float) * np, hipMemcpyDeviceToHost));
    if (loss) BD_HIP(c, hipMemcpy(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost));
    return BD_OK;
}

// Full-batch Adam on the device.  When no network is loaded one is created
// (h1 = h2 = 32) with seeded He-uniform weights.
extern "C" int bd_mlp_train(bd_ctx *c, const float *x, const float *t, const float *mask,
                            int64_t n, int epochs, float lr, uint32_t seed, float *loss) {
    if (!c || n <= 0 || !x || !t || epochs < 0) return BD_EINVAL;
    int rc;
    if (!c->has_mlp) {
        const int h1 = 32, h2 = 32;
        std::vector<float> p(mlp_nparams(h1, h2), 0.f);
        uint64_t st = 0x9E3779B97F4A7C15ULL ^ seed;
        auto unif = [&]() {
            st ^= st << 13;
            st ^= st >> 7;
            st ^= st << 17;
            return (float)((st >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
        };
        float *w1 = p.data(), *w2 = w1 + h1 * kMlpIn + h1, *w3 = w2 + h2 * h1 + h2;
        for (int i = 0; i < h1 * kMlpIn; ++i) w1[i] = unif() * sqrtf(6.f / kMlpIn);
        for (int i = 0; i < h2 * h1; ++i) w2[i] = unif() * sqrtf(6.f / h1);
        for (int i = 0; i < 2 * h2; ++i) w3[i] = unif() * sqrtf(6.f / h2);
        if ((rc = bd_mlp_set(c, p.data(), h1, h2))) return rc;
    }
    BD_HIP(c, hipSetDevice(c->device));
    const int np = (int)mlp_nparams(c->mlp_h1, c->mlp_h2);
    dbuf dx, dt, dm, dg, dl, mo, ve;
    if ((rc = dev_alloc(c, &dx.p, sizeof(float) * kMlpIn * n, "x")) ||
        (rc = dev_alloc(c, &dt.p, sizeof(float) * 2 * n, "t")) ||
        (rc = dev_alloc(c, &dg.p, sizeof(float) * np, "grad")) ||
        (rc = dev_alloc(c, &dl.p, sizeof(float), "loss")) ||
        (rc = dev_alloc(c, &mo.p, sizeof(float) * np, "adam m")) ||
        (rc = dev_alloc(c, &ve.p, sizeof(float) * np, "adam v")))
        return rc;
    BD_HIP(c, hipMemcpy(dx.p, x, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice));
    BD_HIP(c, hipMemcpy(dt.p, t, sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    if (mask) {
        if ((rc = dev_alloc(c, &dm.p, sizeof(float) * 2 * n, "mask"))) return rc;
        BD_HIP(c, hipMemcpy(dm.p, mask, sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    }
    BD_HIP(c, hipMemsetAsync(mo.p, 0, sizeof(float) * np, c->stream));
    BD_HIP(c, hipMemsetAsync(ve.p, 0, sizeof(float) * np, c->stream));
    const double count = mask_count(mask, n);
    for (int ep = 1; ep <= epochs; ++ep) {
        if ((rc = device_grad(c, (const float *)dx.p, (const float *)dt.p, (const float *)dm.p, n,
                              count, (float *)dg.p, (float *)dl.p)))
            return rc;
        const float c1 = 1.f - powf(0.9f, (float)ep), c2 = 1.f - powf(0.999f, (float)ep);
        hipLaunchKernelGGL(k_adam, dim3((np + 255) / 256), dim3(256), 0, c->stream, c->d_mlp,
                           (const float *)dg.p, (float *)mo.p, (float *)ve.p, np, lr, c1, c2);
        BD_HIP(c, hipGetLastError());
    }
    // final loss at the trained parameters
    if ((rc = device_grad(c, (const float *)dx.p, (const float *)dt.p, (const float *)dm.p, n,
                          count, (float *)dg.p, (float *)dl.p)))
        return rc;
    BD_HIP(c, hipStreamSynchronize(c->stream));
    if (loss) BD_HIP(c, hipMemcpy(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost));
    BD_HIP(c, hipMemcpy(c->h_mlp.data(), c->d_mlp, sizeof(float) * np, hipMemcpyDeviceToHost));
    return refresh_mlp_ops(c);
}
