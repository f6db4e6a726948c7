/*
 * oracle_cache.c -- CPU restatement of the radiance cache: hash-grid
 * encoding, MLP forward, relative-L2 loss and its gradient with respect to
 * the MLP weights and the grid entries.  TEST INFRASTRUCTURE ONLY (see
 * rt_oracle.h).
 *
 * The cache has no counterpart in the reference (it is one of the learned
 * parts added on top of the plain path tracer); this file restates the
 * design the GPU kernels implement:
 *   encoding (per level l, F = 2, T = 2^log2_table entries):
 *     p = clamp((pos - lo) * inv_extent, 0, 1)                    (FP32)
 *     x = p * res_l, c = min(floor(x), res_l - 1), f = x - c
 *     corner k = (k & 1, k >> 1 & 1, k >> 2): w_k = (wx * wy) * wz,
 *     index = dense_l ? vx + vy (res + 1) + vz (res + 1)^2
 *                     : (vx * 1 ^ vy * 2654435761 ^ vz * 805459861) & (T - 1)
 *     feature[2 l + j] = sum_k w_k table[l][index_k][j]  (k = 0..7 in order)
 *   input (64): features 0..31, outgoing dir 32..34, normal 35..37,
 *     roughness 38, albedo 39..41, 1 at 42, zeros
 *   h1 = relu(W0 x), h2 = relu(W1 h1), h3 = relu(W2 h2), h4 = relu(W3 h3),
 *   y = W4 h4 (3 outputs)
 *   loss = 1/(3n) sum (y - t)^2 / (y^2 + 0.01), y^2 held constant in the
 *   gradient
 * With `bf16` set the oracle rounds where the GPU stores bf16 (inputs,
 * weights, post-ReLU activations, dy and the pre-activation gradients) and
 * computes the encoding in FP32 in the GPU's operation order; products and
 * sums accumulate in double.  Without it everything is double (for finite
 * differences).
 */
#include "rt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { CI = 64, CH = 64, CL = 4, CO = 3 };
#define CACHE_EPS 0.01

int rto_cache_mlp_param_count(void) { return CL * CH * CI + CO * CH; }

static float unit_clampf(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

void rto_cache_corners(const rto_cache_grid* g, int l, const float p[3], uint32_t idx[8], float w[8]) {
    const float r = g->res[l];
    const int ri = (int)r;
    const float x = p[0] * r, y = p[1] * r, z = p[2] * r;
    int cx = (int)floorf(x), cy = (int)floorf(y), cz = (int)floorf(z);
    if (cx > ri - 1) cx = ri - 1;
    if (cy > ri - 1) cy = ri - 1;
    if (cz > ri - 1) cz = ri - 1;
    const float fx = x - (float)cx, fy = y - (float)cy, fz = z - (float)cz;
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    const uint32_t stride = (uint32_t)ri + 1u;
    const uint32_t tmask = (1u << g->log2_table) - 1u;
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        w[k] = (wx[dx] * wy[dy]) * wz[dz];
        const uint32_t vx = (uint32_t)(cx + dx), vy = (uint32_t)(cy + dy), vz = (uint32_t)(cz + dz);
        idx[k] = g->dense[l] ? vx + vy * stride + vz * stride * stride
                             : ((vx * 1u) ^ (vy * 2654435761u) ^ (vz * 805459861u)) & tmask;
    }
}

static void unit_pos(const rto_cache_grid* g, const float* in16, float p[3]) {
    for (int a = 0; a < 3; ++a) p[a] = unit_clampf((in16[a] - g->lo[a]) * g->inv_extent[a]);
}

static double rnd(double v, int bf16) { return bf16 ? (double)rto_bf16_round((float)v) : v; }

/* the 64 network inputs of one query */
static void features(const rto_cache_grid* g, const float* table, const float* in16, int bf16,
    double x[CI]) {
    float p[3];
    unit_pos(g, in16, p);
    for (int l = 0; l < 16; ++l) {
        double f[2] = {0.0, 0.0};
        if (l < g->levels) {
            uint32_t idx[8];
            float w[8];
            rto_cache_corners(g, l, p, idx, w);
            const float* t = table + ((size_t)l << g->log2_table) * 2u;
            if (bf16) {
                float f0 = 0.0f, f1 = 0.0f;
                for (int k = 0; k < 8; ++k) {
                    f0 = f0 + w[k] * t[(size_t)idx[k] * 2u];
                    f1 = f1 + w[k] * t[(size_t)idx[k] * 2u + 1u];
                }
                f[0] = f0;
                f[1] = f1;
            } else {
                for (int k = 0; k < 8; ++k) {
                    f[0] += (double)w[k] * t[(size_t)idx[k] * 2u];
                    f[1] += (double)w[k] * t[(size_t)idx[k] * 2u + 1u];
                }
            }
        }
        x[2 * l] = rnd(f[0], bf16);
        x[2 * l + 1] = rnd(f[1], bf16);
    }
    for (int k = 32; k < CI; ++k) x[k] = 0.0;
    x[32] = rnd(in16[4], bf16);
    x[33] = rnd(in16[5], bf16);
    x[34] = rnd(in16[6], bf16);
    x[35] = rnd(in16[8], bf16);
    x[36] = rnd(in16[9], bf16);
    x[37] = rnd(in16[10], bf16);
    x[38] = rnd(in16[7], bf16);
    x[39] = rnd(in16[12], bf16);
    x[40] = rnd(in16[13], bf16);
    x[41] = rnd(in16[14], bf16);
    x[42] = 1.0;
}

void rto_cache_features(const rto_cache_grid* g, const float* table, const float* in16, int bf16,
    float x_out[64]) {
    double x[CI];
    features(g, table, in16, bf16, x);
    for (int k = 0; k < CI; ++k) x_out[k] = (float)x[k];
}

typedef struct {
    double x[CI];
    double h[CL][CH];  /* h1..h4 */
    double y[CO];
} fwd_t;

static void forward(const double* W, const double x_in[CI], int bf16, fwd_t* f) {
    memcpy(f->x, x_in, sizeof(f->x));
    const double* in = f->x;
    for (int l = 0; l < CL; ++l) {
        const double* Wl = W + (size_t)l * CH * CI;
        for (int o = 0; o < CH; ++o) {
            double s = 0.0;
            for (int i = 0; i < CI; ++i) s += Wl[o * CI + i] * in[i];
            f->h[l][o] = rnd(s > 0.0 ? s : 0.0, bf16);
        }
        in = f->h[l];
    }
    const double* W4 = W + (size_t)CL * CH * CI;
    for (int c = 0; c < CO; ++c) {
        double s = 0.0;
        for (int i = 0; i < CH; ++i) s += W4[c * CH + i] * in[i];
        f->y[c] = s;
    }
}

static double* mlp_weights(const float* params, int bf16) {
    const int nm = rto_cache_mlp_param_count();
    double* W = (double*)malloc(sizeof(double) * (size_t)nm);
    if (!W) return NULL;
    for (int i = 0; i < nm; ++i) W[i] = rnd(params[i], bf16);
    return W;
}

int rto_cache_predict(const rto_cache_grid* g, const float* params, int n, const float* inputs,
    int bf16, double* out) {
    double* W = mlp_weights(params, bf16);
    if (!W) return -1;
    const float* table = params + rto_cache_mlp_param_count();
    for (int s = 0; s < n; ++s) {
        double x[CI];
        fwd_t f;
        features(g, table, inputs + (size_t)s * 16u, bf16, x);
        forward(W, x, bf16, &f);
        for (int c = 0; c < CO; ++c) out[(size_t)s * 3u + c] = f.y[c];
    }
    free(W);
    return 0;
}

double rto_cache_loss_grad(const rto_cache_grid* g, const float* params, int n, const float* inputs,
    const float* targets, int bf16, double* grad) {
    const int nm = rto_cache_mlp_param_count();
    const size_t ngrid = ((size_t)g->levels << g->log2_table) * 2u;
    double* W = mlp_weights(params, bf16);
    if (!W) return NAN;
    if (grad) memset(grad, 0, sizeof(double) * ((size_t)nm + ngrid));
    const float* table = params + nm;
    const double scale = n > 0 ? 1.0 / (3.0 * (double)n) : 0.0;
    double loss = 0.0;
    for (int s = 0; s < n; ++s) {
        double x[CI];
        fwd_t f;
        const float* in16 = inputs + (size_t)s * 16u;
        features(g, table, in16, bf16, x);
        forward(W, x, bf16, &f);
        double dy[CO];
        for (int c = 0; c < CO; ++c) {
            const double t = targets[(size_t)s * 4u + c];
            if (bf16) {
                /* FP32 operations in the GPU's order */
                const float yf = (float)f.y[c];
                const float d = yf - (float)t;
                const float den = yf * yf + (float)CACHE_EPS;
                loss += (double)((d * d) / den);
                dy[c] = rnd((double)(((2.0f * d) / den) * (float)scale), 1);
            } else {
                const double d = f.y[c] - t;
                const double den = f.y[c] * f.y[c] + CACHE_EPS;
                loss += d * d / den;
                dy[c] = 2.0 * d / den * scale;
            }
        }
        if (!grad) continue;
        /* output layer */
        const double* W4 = W + (size_t)CL * CH * CI;
        double* gW4 = grad + (size_t)CL * CH * CI;
        double dz[CH], dn[CH];
        for (int i = 0; i < CH; ++i) {
            double a = 0.0;
            for (int c = 0; c < CO; ++c) {
                a += W4[c * CH + i] * dy[c];
                gW4[c * CH + i] += dy[c] * f.h[CL - 1][i];
            }
            dz[i] = rnd(f.h[CL - 1][i] > 0.0 ? a : 0.0, bf16);
        }
        /* hidden layers l = 3, 2, 1: dz holds dz_{l+1} */
        for (int l = CL - 1; l >= 0; --l) {
            const double* Wl = W + (size_t)l * CH * CI;
            double* gWl = grad + (size_t)l * CH * CI;
            const double* xin = l == 0 ? f.x : f.h[l - 1];
            for (int o = 0; o < CH; ++o) {
                for (int i = 0; i < CI; ++i) gWl[o * CI + i] += dz[o] * xin[i];
            }
            if (l == 0) break;
            for (int i = 0; i < CI; ++i) {
                double a = 0.0;
                for (int o = 0; o < CH; ++o) a += Wl[o * CI + i] * dz[o];
                dn[i] = rnd(xin[i] > 0.0 ? a : 0.0, bf16);
            }
            memcpy(dz, dn, sizeof(dz));
        }
        /* dz = dz1: input gradient of the hash features, scattered to cells */
        float p[3];
        unit_pos(g, in16, p);
        double* gg = grad + nm;
        for (int l = 0; l < g->levels; ++l) {
            double dF[2];
            for (int j = 0; j < 2; ++j) {
                double a = 0.0;
                for (int o = 0; o < CH; ++o) a += W[(size_t)o * CI + 2 * l + j] * dz[o];
                dF[j] = a;
            }
            uint32_t idx[8];
            float w[8];
            rto_cache_corners(g, l, p, idx, w);
            double* gl = gg + ((size_t)l << g->log2_table) * 2u;
            for (int k = 0; k < 8; ++k) {
                gl[(size_t)idx[k] * 2u] += (double)w[k] * dF[0];
                gl[(size_t)idx[k] * 2u + 1u] += (double)w[k] * dF[1];
            }
        }
    }
    free(W);
    return loss * scale;
}
