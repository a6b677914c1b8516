// bd_mlp.h -- direction-cost network shared by the step kernels (one
// sample per step, on one wave) and the training/inference kernels.  8 inputs -> h1 -> h2 -> 2 outputs (predicted
// log2 microseconds of a top-down and of a bottom-up step), ReLU, h1,h2 <= 32.
// Parameter layout (row-major, out x in):
//   W1[h1*8] b1[h1] W2[h2*h1] b2[h2] W3[2*h2] b3[2]
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace bd {

constexpr int kMlpIn = 8;
constexpr int kMlpOut = 2;
constexpr int kMlpMaxHidden = 32;

__host__ __device__ inline int64_t mlp_nparams(int h1, int h2) {
    return (int64_t)h1 * kMlpIn + h1 + (int64_t)h2 * h1 + h2 + (int64_t)kMlpOut * h2 + kMlpOut;
}

__host__ __device__ inline float lg1p(float x) { return log2f(1.0f + (x > 0.f ? x : 0.f)); }

// Step features, all computable before the step from direction-independent
// quantities: |F|, sum deg(F), max deg(F), unvisited count and degree sum,
// level, average degree, and whether the previous step was bottom-up (a
// top-down step then reads F from the bitmap).  Single precision: the
// features feed a bf16 network and are evaluated once per step on the GPU.
__host__ __device__ inline void mlp_features_dev(int64_t frontier, int64_t frontier_deg,
                                                 int64_t max_frontier_deg, int64_t unvisited,
                                                 int64_t unvisited_deg, int64_t level, int64_t nv,
                                                 int64_t m, int64_t prev_dir, float *x) {
    const float lv = lg1p((float)nv);
    const float le = lg1p(2.0f * (float)m);
    x[0] = lg1p((float)frontier) / lv;
    x[1] = lg1p((float)frontier_deg) / le;
    x[2] = lg1p((float)unvisited) / lv;
    x[3] = lg1p((float)unvisited_deg) / le;
    x[4] = (float)(level < 64 ? level : 64) / 64.0f;
    x[5] = lg1p((float)max_frontier_deg) / lv;
    x[6] = lg1p(2.0f * (float)m / (float)(nv > 0 ? nv : 1)) / 8.0f;
    x[7] = prev_dir == 1 ? 1.0f : 0.0f;
}

// Feature `k` (0..7) of mlp_features_dev alone, so that lanes 0-7 of a wave
// compute the eight features side by side; the same arithmetic, in the same
// order, as mlp_features_dev.
__device__ inline float mlp_feature_k(int k, int64_t frontier, int64_t frontier_deg,
                                      int64_t max_frontier_deg, int64_t unvisited,
                                      int64_t unvisited_deg, int64_t level, int64_t nv, int64_t m,
                                      int64_t prev_dir) {
    const float lv = lg1p((float)nv);
    const float le = lg1p(2.0f * (float)m);
    switch (k) {
    case 0: return lg1p((float)frontier) / lv;
    case 1: return lg1p((float)frontier_deg) / le;
    case 2: return lg1p((float)unvisited) / lv;
    case 3: return lg1p((float)unvisited_deg) / le;
    case 4: return (float)(level < 64 ? level : 64) / 64.0f;
    case 5: return lg1p((float)max_frontier_deg) / lv;
    case 6: return lg1p(2.0f * (float)m / (float)(nv > 0 ? nv : 1)) / 8.0f;
    default: return prev_dir == 1 ? 1.0f : 0.0f;
    }
}

// Orders LDS accesses among the lanes of one wave.  The network code below
// runs on a single wave, so it never needs a workgroup barrier and may be
// called by one wave of a larger workgroup.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (short)(u >> 16);  // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// One dense layer on one wave with v_mfma_f32_16x16x16bf16_1k:
// Y[16 x nout] = act(X[16 x k] * W^T + b), X and Y in LDS (row stride 32).
// Operand layout of the 16x16x16 instruction (lane l, element j = 0..3):
//   A: row l%16, k 4*(l/16)+j     B: col l%16, k 4*(l/16)+j
//   D: col l%16, row 4*(l/16)+j
__device__ __forceinline__ void mlp_layer_mfma(const float *W, const float *b, int k, int nout,
                                               const float (*X)[32], float (*Y)[32], bool relu) {
    const int l = threadIdx.x & 63;
    const int r = l & 15, q = l >> 4;
    for (int nt = 0; nt * 16 < nout; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt * 16 < k; ++kt) {
            bf16x4 a, bw;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int kk = kt * 16 + 4 * q + j;
                const int n = nt * 16 + r;
                a[j] = f32_to_bf16(kk < k ? X[r][kk] : 0.f);
                bw[j] = f32_to_bf16((kk < k && n < nout) ? W[n * k + kk] : 0.f);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw, acc, 0, 0, 0);
        }
        const int n = nt * 16 + r;
        if (n < nout) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float v = acc[j] + b[n];
                Y[4 * q + j][n] = relu ? (v > 0.f ? v : 0.f) : v;
            }
        }
    }
}

// Forward pass of up to 16 samples (rows of X) on one wave with bf16 MFMA
// and fp32 accumulation.  X: [16][32] LDS, rows beyond n must be zero.
// Writes Y[16][32] (columns 0..1 are the outputs).
__device__ __forceinline__ void mlp_forward_rows_mfma(const float *p, int h1, int h2,
                                                      float (*X)[32], float (*H)[32],
                                                      float (*Y)[32]) {
    const float *w1 = p, *b1 = w1 + h1 * kMlpIn, *w2 = b1 + h1, *b2 = w2 + h2 * h1;
    const float *w3 = b2 + h2, *b3 = w3 + kMlpOut * h2;
    mlp_layer_mfma(w1, b1, kMlpIn, h1, X, H, true);
    wave_lds_sync();
    mlp_layer_mfma(w2, b2, h1, h2, H, X, true);
    wave_lds_sync();
    mlp_layer_mfma(w3, b3, h2, kMlpOut, X, Y, false);
    wave_lds_sync();
}

// Weights of a single-sample forward pass held in registers as bf16, in the
// operand layout of v_mfma_f32_16x16x16bf16_1k: lane l holds, for output
// tile nt and input tile kt, B[n = 16 nt + l%16][k = 16 kt + 4 (l/16) + j],
// j = 0..3, and the biases of its output columns.  Loading them is
// independent of the step's counters, so a step issues these loads with
// its other state loads and the forward pass itself touches no memory.
struct mlp_wave_regs {
    bf16x4 w1[2];     // layer 1 (k < 8: one input tile), output tiles 0, 1
    bf16x4 w2[2][2];  // layer 2 [nt][kt]
    bf16x4 w3[2];     // layer 3 (outputs 0, 1), [kt]
    float b1[2], b2[2], b3;
};

__device__ inline void mlp_prefetch(const float *p, int h1, int h2, mlp_wave_regs &w) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    const float *w1 = p, *b1 = w1 + h1 * kMlpIn, *w2 = b1 + h1, *b2 = w2 + h2 * h1;
    const float *w3 = b2 + h2, *b3 = w3 + kMlpOut * h2;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int n = 16 * nt + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int kk = 4 * q + j;
            w.w1[nt][j] = f32_to_bf16((n < h1 && kk < kMlpIn) ? w1[n * kMlpIn + kk] : 0.f);
        }
        w.b1[nt] = n < h1 ? b1[n] : 0.f;
#pragma unroll
        for (int kt = 0; kt < 2; ++kt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int kk = 16 * kt + 4 * q + j;
                w.w2[nt][kt][j] = f32_to_bf16((n < h2 && kk < h1) ? w2[n * h1 + kk] : 0.f);
            }
        w.b2[nt] = n < h2 ? b2[n] : 0.f;
    }
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int kk = 16 * kt + 4 * q + j;
            w.w3[kt][j] = f32_to_bf16((r < kMlpOut && kk < h2) ? w3[r * h2 + kk] : 0.f);
        }
    w.b3 = r < kMlpOut ? b3[r] : 0.f;
}

// Forward pass of one sample x (kMlpIn floats that every lane can read) on
// one wave, from prefetched weights.  Row 0 of each 16-row tile carries the
// sample.  The result tile holds row 0 in lanes 0..15 (column = lane), so a
// layer's activations reach the next layer's A operand (row 0, k = 4 (l/16)
// + j of tile kt) by a shuffle from lane 4 (l/16) + j.  Same rounding as
// mlp_forward_rows_mfma: bf16 operands, fp32 sums, bias and ReLU in fp32.
// Every lane returns y.
__device__ inline void mlp_eval(const mlp_wave_regs &w, const float *x, float *y) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int kk = 4 * q + j;
        a[j] = f32_to_bf16((r == 0 && kk < kMlpIn) ? x[kk] : 0.f);
    }
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    float h[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w.w1[nt], z, 0, 0, 0);
        const float v = d[0] + w.b1[nt];
        h[nt] = v > 0.f ? v : 0.f;
    }
    bf16x4 a2[2];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float v = __shfl(h[kt], 4 * q + j);
            a2[kt][j] = f32_to_bf16(r == 0 ? v : 0.f);
        }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[0], w.w2[nt][0], z, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[1], w.w2[nt][1], d, 0, 0, 0);
        const float v = d[0] + w.b2[nt];
        h[nt] = v > 0.f ? v : 0.f;
    }
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float v = __shfl(h[kt], 4 * q + j);
            a2[kt][j] = f32_to_bf16(r == 0 ? v : 0.f);
        }
    f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[0], w.w3[0], z, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[1], w.w3[1], d, 0, 0, 0);
    const float out = d[0] + w.b3;
    y[0] = __shfl(out, 0);
    y[1] = __shfl(out, 1);
}

// The prefetched weights as a table in device memory, 96 bytes per lane,
// built once per network (k_mlp_ops): a step loads its lane's record with
// six 16-byte loads issued beside its state loads.
constexpr int kMlpOpsQuads = 6;
union mlp_ops_rec {
    mlp_wave_regs r;
    uint4 q[kMlpOpsQuads];
};
static_assert(sizeof(mlp_wave_regs) <= sizeof(uint4) * kMlpOpsQuads, "operand record");

__device__ inline void mlp_load(const uint4 *ops, mlp_wave_regs &w) {
    const int l = threadIdx.x & 63;
    mlp_ops_rec u;
#pragma unroll
    for (int i = 0; i < kMlpOpsQuads; ++i) u.q[i] = ops[l * kMlpOpsQuads + i];
    w = u.r;
}

}  // namespace bd

