/*
 * oracle_mlp.c -- CPU forward/backward/Adam for the direction-cost MLP.
 * The reference has no learned component; this is the checker for the HIP
 * training and inference kernels.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NIN 8
#define NOUT 2

int64_t oracle_mlp_nparams(int h1, int h2) {
    return (int64_t)h1 * NIN + h1 + (int64_t)h2 * h1 + h2 + (int64_t)NOUT * h2 + NOUT;
}

typedef struct {
    const float *w1, *b1, *w2, *b2, *w3, *b3;
} mlp_view;

static mlp_view view_of(const float *p, int h1, int h2) {
    mlp_view v;
    v.w1 = p;
    v.b1 = v.w1 + h1 * NIN;
    v.w2 = v.b1 + h1;
    v.b2 = v.w2 + h2 * h1;
    v.w3 = v.b2 + h2;
    v.b3 = v.w3 + NOUT * h2;
    return v;
}

static void fwd_one(const mlp_view *v, int h1, int h2, const float *x, float *a1,
                    float *a2, float *y) {
    for (int j = 0; j < h1; ++j) {
        float s = v->b1[j];
        for (int i = 0; i < NIN; ++i) s += v->w1[j * NIN + i] * x[i];
        a1[j] = s > 0.f ? s : 0.f;
    }
    for (int j = 0; j < h2; ++j) {
        float s = v->b2[j];
        for (int i = 0; i < h1; ++i) s += v->w2[j * h1 + i] * a1[i];
        a2[j] = s > 0.f ? s : 0.f;
    }
    for (int o = 0; o < NOUT; ++o) {
        float s = v->b3[o];
        for (int i = 0; i < h2; ++i) s += v->w3[o * h2 + i] * a2[i];
        y[o] = s;
    }
}

void oracle_mlp_forward(const float *params, int h1, int h2, const float *x,
                        int64_t n, float *y) {
    mlp_view v = view_of(params, h1, h2);
    float *a1 = (float *)malloc(sizeof(float) * (size_t)h1);
    float *a2 = (float *)malloc(sizeof(float) * (size_t)h2);
    for (int64_t s = 0; s < n; ++s) fwd_one(&v, h1, h2, x + s * NIN, a1, a2, y + s * NOUT);
    free(a1);
    free(a2);
}

float oracle_mlp_loss_grad(const float *params, int h1, int h2, const float *x,
                           const float *target, const float *mask, int64_t n,
                           float *grad) {
    int64_t np = oracle_mlp_nparams(h1, h2);
    memset(grad, 0, sizeof(float) * (size_t)np);
    mlp_view v = view_of(params, h1, h2);
    float *g_w1 = grad, *g_b1 = g_w1 + h1 * NIN, *g_w2 = g_b1 + h1;
    float *g_b2 = g_w2 + h2 * h1, *g_w3 = g_b2 + h2, *g_b3 = g_w3 + NOUT * h2;
    float *a1 = (float *)malloc(sizeof(float) * (size_t)h1);
    float *a2 = (float *)malloc(sizeof(float) * (size_t)h2);
    float *d1 = (float *)malloc(sizeof(float) * (size_t)h1);
    float *d2 = (float *)malloc(sizeof(float) * (size_t)h2);
    double count = 0;
    for (int64_t s = 0; s < n * NOUT; ++s) count += mask ? mask[s] : 1.0f;
    double loss = 0;
    if (count > 0) {
        float inv = (float)(1.0 / count);
        for (int64_t s = 0; s < n; ++s) {
            const float *xs = x + s * NIN;
            float y[NOUT], dy[NOUT];
            fwd_one(&v, h1, h2, xs, a1, a2, y);
            for (int o = 0; o < NOUT; ++o) {
                float mk = mask ? mask[s * NOUT + o] : 1.0f;
                float e = y[o] - target[s * NOUT + o];
                loss += (double)(mk * e * e);
                dy[o] = 2.0f * mk * e * inv;
            }
            for (int j = 0; j < h2; ++j) {
                float acc = 0.f;
                for (int o = 0; o < NOUT; ++o) acc += v.w3[o * h2 + j] * dy[o];
                d2[j] = a2[j] > 0.f ? acc : 0.f;
            }
            for (int j = 0; j < h1; ++j) {
                float acc = 0.f;
                for (int k = 0; k < h2; ++k) acc += v.w2[k * h1 + j] * d2[k];
                d1[j] = a1[j] > 0.f ? acc : 0.f;
            }
            for (int o = 0; o < NOUT; ++o) {
                g_b3[o] += dy[o];
                for (int j = 0; j < h2; ++j) g_w3[o * h2 + j] += dy[o] * a2[j];
            }
            for (int k = 0; k < h2; ++k) {
                g_b2[k] += d2[k];
                for (int j = 0; j < h1; ++j) g_w2[k * h1 + j] += d2[k] * a1[j];
            }
            for (int j = 0; j < h1; ++j) {
                g_b1[j] += d1[j];
                for (int i = 0; i < NIN; ++i) g_w1[j * NIN + i] += d1[j] * xs[i];
            }
        }
        loss /= count;
    }
    free(a1);
    free(a2);
    free(d1);
    free(d2);
    return (float)loss;
}

void oracle_mlp_adam(float *params, const float *grad, float *m, float *v,
                     int64_t np, float lr, int64_t t) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.0f - powf(b1, (float)t);
    float c2 = 1.0f - powf(b2, (float)t);
    for (int64_t i = 0; i < np; ++i) {
        m[i] = b1 * m[i] + (1.0f - b1) * grad[i];
        v[i] = b2 * v[i] + (1.0f - b2) * grad[i] * grad[i];
        float mh = m[i] / c1;
        float vh = v[i] / c2;
        params[i] -= lr * mh / (sqrtf(vh) + eps);
    }
}
