/*
 * oracle_nn.c -- CPU restatement of the learned denoiser (forward, loss and
 * backward).  TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 *
 * Network (per pixel of each image; zero padding at image borders):
 *   features f (16 ch): log1p(beauty rgb), albedo rgb, decoded normal xyz,
 *                       log1p(depth) / 8, 6 zeros
 *   h0 = relu(conv3x3_d1(f;  16 -> 32))
 *   h1 = relu(conv3x3_d2(h0; 32 -> 32))
 *   h2 = relu(conv3x3_d4(h1; 32 -> 32))
 *   h3 = relu(conv3x3_d8(h2; 32 -> 32))
 *   out_log = log1p(beauty rgb) + W_head h3 + b_head         (1x1, 32 -> 3)
 *   output  = expm1(out_log) (clamped at 0)
 *   loss    = mean over pixels and channels of (out_log - log1p(target))^2
 * Parameter vector: per conv layer W[co][tap][ci] (tap = 3 ky + kx) then
 * b[co]; then W_head[c][k], b_head[c].
 *
 * With `bf16` set the oracle rounds exactly where the GPU stores bf16:
 * features, weights, post-ReLU activations and the backward pre-activation
 * gradients; everything else accumulates in FP32 (double here).
 */
#include "rt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { NN_CF = 16, NN_H = 32, NN_LAYERS = 4 };
static const int kDil[NN_LAYERS] = {1, 2, 4, 8};

static int layer_cin(int l) { return l == 0 ? NN_CF : NN_H; }
static size_t layer_w(int l) { return (size_t)NN_H * 9u * (size_t)layer_cin(l); }

int rto_nn_param_count(void) {
    size_t n = 0;
    for (int l = 0; l < NN_LAYERS; ++l) n += layer_w(l) + NN_H;
    n += 3u * NN_H + 3u;
    return (int)n;
}

/* round to nearest even bfloat16, returned as float */
static float bf16r(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

float rto_bf16_round(float x) { return bf16r(x); }

static size_t param_offset_w(int l) {
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += layer_w(i) + NN_H;
    return o;
}
static size_t param_offset_head(void) { return param_offset_w(NN_LAYERS); }

void rto_nn_features(int P, const float* beauty, const float* normal, const float* albedo,
    const float* depth, float* feat) {
    for (int p = 0; p < P; ++p) {
        float* f = feat + (size_t)p * NN_CF;
        for (int c = 0; c < 3; ++c) f[c] = bf16r(log1pf(fmaxf(beauty[4 * p + c], 0.0f)));
        for (int c = 0; c < 3; ++c) f[3 + c] = bf16r(fminf(fmaxf(albedo[4 * p + c], 0.0f), 1.0f));
        for (int c = 0; c < 3; ++c) f[6 + c] = bf16r(normal[4 * p + c] * 2.0f - 1.0f);
        f[9] = bf16r(log1pf(fmaxf(depth[p], 0.0f)) * 0.125f);
        for (int c = 10; c < NN_CF; ++c) f[c] = 0.0f;
    }
}

typedef struct { int img, x, y, w, h, base; } pix_t;

static pix_t locate(const rto_images* im, int p) {
    pix_t r;
    int i = 0;
    while (i + 1 < im->count && p >= im->offset[i + 1]) ++i;
    r.img = i;
    r.base = im->offset[i];
    r.w = im->width[i];
    r.h = (im->offset[i + 1] - im->offset[i]) / r.w;
    r.y = (p - r.base) / r.w;
    r.x = (p - r.base) - r.y * r.w;
    return r;
}

/* neighbour pixel for tap t at dilation d, or -1 outside the image */
static int tap_pixel(const pix_t* q, int t, int d) {
    const int x = q->x + (t % 3 - 1) * d, y = q->y + (t / 3 - 1) * d;
    if (x < 0 || x >= q->w || y < 0 || y >= q->h) return -1;
    return q->base + y * q->w + x;
}

/* loss (and gradient w.r.t. the parameters when grad != NULL) over all
 * pixels of the images; out4 (may be NULL) receives the denoised rgb */
double rto_nn_loss_grad(const rto_images* im, const float* params_in, const float* feat,
    const float* beauty, const float* target, float* out4, float* grad, int bf16) {
    const int P = im->offset[im->count];
    const int np = rto_nn_param_count();
    float* params = (float*)malloc(sizeof(float) * (size_t)np);
    for (int i = 0; i < np; ++i) params[i] = params_in[i];
    if (bf16) {
        for (int l = 0; l < NN_LAYERS; ++l) {
            float* W = params + param_offset_w(l);
            for (size_t i = 0; i < layer_w(l); ++i) W[i] = bf16r(W[i]);
        }
    }
    float* act[NN_LAYERS];
    for (int l = 0; l < NN_LAYERS; ++l) act[l] = (float*)calloc((size_t)P * NN_H, sizeof(float));
    /* forward */
    for (int l = 0; l < NN_LAYERS; ++l) {
        const int cin = layer_cin(l);
        const float* in = l == 0 ? feat : act[l - 1];
        const float* W = params + param_offset_w(l);
        const float* b = W + layer_w(l);
        for (int p = 0; p < P; ++p) {
            const pix_t q = locate(im, p);
            double acc[NN_H];
            for (int co = 0; co < NN_H; ++co) acc[co] = 0.0;
            for (int t = 0; t < 9; ++t) {
                const int s = tap_pixel(&q, t, kDil[l]);
                if (s < 0) continue;
                const float* x = in + (size_t)s * cin;
                for (int co = 0; co < NN_H; ++co) {
                    const float* w = W + ((size_t)co * 9 + t) * cin;
                    double a = 0.0;
                    for (int ci = 0; ci < cin; ++ci) a += (double)w[ci] * (double)x[ci];
                    acc[co] += a;
                }
            }
            for (int co = 0; co < NN_H; ++co) {
                float v = (float)(acc[co] + (double)b[co]);
                v = v > 0.0f ? v : 0.0f;
                act[l][(size_t)p * NN_H + co] = bf16 ? bf16r(v) : v;
            }
        }
    }
    const float* Wh = params + param_offset_head();
    const float* bh = Wh + 3 * NN_H;
    float* dout = (float*)calloc((size_t)P * 3, sizeof(float));
    double loss = 0.0;
    const double scale = 1.0 / (3.0 * (double)P);
    for (int p = 0; p < P; ++p) {
        const float* h = act[NN_LAYERS - 1] + (size_t)p * NN_H;
        for (int c = 0; c < 3; ++c) {
            double y = bh[c];
            for (int k = 0; k < NN_H; ++k) y += (double)Wh[c * NN_H + k] * (double)h[k];
            const float base = log1pf(fmaxf(beauty[4 * p + c], 0.0f));
            const float pred = (float)((double)base + y);
            if (out4) out4[4 * p + c] = fmaxf(expm1f(pred), 0.0f);
            if (target) {
                const float tl = log1pf(fmaxf(target[4 * p + c], 0.0f));
                const double e = (double)pred - (double)tl;
                loss += e * e * scale;
                dout[(size_t)p * 3 + c] = (float)(2.0 * e * scale);
            }
        }
        if (out4) out4[4 * p + 3] = 1.0f;
    }
    if (grad && target) {
        memset(grad, 0, sizeof(float) * (size_t)np);
        double* g = (double*)calloc((size_t)np, sizeof(double));
        /* head */
        float* dpre = (float*)calloc((size_t)P * NN_H, sizeof(float));
        const size_t oh = param_offset_head();
        for (int p = 0; p < P; ++p) {
            const float* h = act[NN_LAYERS - 1] + (size_t)p * NN_H;
            const float* d = dout + (size_t)p * 3;
            for (int c = 0; c < 3; ++c) {
                for (int k = 0; k < NN_H; ++k) g[oh + (size_t)c * NN_H + k] += (double)d[c] * (double)h[k];
                g[oh + 3 * NN_H + c] += d[c];
            }
            for (int k = 0; k < NN_H; ++k) {
                double s = 0.0;
                for (int c = 0; c < 3; ++c) s += (double)Wh[c * NN_H + k] * (double)d[c];
                const float v = h[k] > 0.0f ? (float)s : 0.0f;
                dpre[(size_t)p * NN_H + k] = bf16 ? bf16r(v) : v;
            }
        }
        /* conv layers, last to first */
        float* dprev = (float*)calloc((size_t)P * NN_H, sizeof(float));
        for (int l = NN_LAYERS - 1; l >= 0; --l) {
            const int cin = layer_cin(l);
            const float* in = l == 0 ? feat : act[l - 1];
            const float* W = params + param_offset_w(l);
            const size_t ow = param_offset_w(l);
            /* weights and bias */
            for (int p = 0; p < P; ++p) {
                const pix_t q = locate(im, p);
                const float* dp = dpre + (size_t)p * NN_H;
                for (int t = 0; t < 9; ++t) {
                    const int s = tap_pixel(&q, t, kDil[l]);
                    if (s < 0) continue;
                    const float* x = in + (size_t)s * cin;
                    for (int co = 0; co < NN_H; ++co) {
                        if (dp[co] == 0.0f) continue;
                        double* gw = g + ow + ((size_t)co * 9 + t) * cin;
                        for (int ci = 0; ci < cin; ++ci) gw[ci] += (double)dp[co] * (double)x[ci];
                    }
                }
                for (int co = 0; co < NN_H; ++co) g[ow + layer_w(l) + co] += dp[co];
            }
            if (l == 0) break;
            /* input gradient through the (transposed) convolution and ReLU */
            for (int p = 0; p < P; ++p) {
                const pix_t q = locate(im, p);
                double acc[NN_H];
                for (int ci = 0; ci < NN_H; ++ci) acc[ci] = 0.0;
                for (int t = 0; t < 9; ++t) {
                    /* pixel p feeds output pixel p - off(t) through tap t */
                    const int s = tap_pixel(&q, 8 - t, kDil[l]);
                    if (s < 0) continue;
                    const float* dp = dpre + (size_t)s * NN_H;
                    for (int co = 0; co < NN_H; ++co) {
                        if (dp[co] == 0.0f) continue;
                        const float* w = W + ((size_t)co * 9 + t) * cin;
                        for (int ci = 0; ci < cin; ++ci) acc[ci] += (double)w[ci] * (double)dp[co];
                    }
                }
                const float* h = act[l - 1] + (size_t)p * NN_H;
                for (int ci = 0; ci < NN_H; ++ci) {
                    const float v = h[ci] > 0.0f ? (float)acc[ci] : 0.0f;
                    dprev[(size_t)p * NN_H + ci] = bf16 ? bf16r(v) : v;
                }
            }
            float* t2 = dpre;
            dpre = dprev;
            dprev = t2;
        }
        for (int i = 0; i < np; ++i) grad[i] = (float)g[i];
        free(g);
        free(dpre);
        free(dprev);
    }
    free(dout);
    for (int l = 0; l < NN_LAYERS; ++l) free(act[l]);
    free(params);
    return loss;
}
