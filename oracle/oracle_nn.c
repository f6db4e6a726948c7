/*
 * oracle_nn.c -- CPU restatement of the learned denoiser (forward, loss and
 * backward).  TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 *
 * Network: a 4-level convolutional encoder-decoder, widths 32/48/64/96, 3x3
 * convolutions (zero padding at image borders), ReLU, skip connections:
 *   level 0: e0a = relu(conv(f; 16->32)), e0 = relu(conv(e0a; 32->32))
 *   level 1: e1a = relu(conv(pool(e0); 32->48)), e1 = relu(conv(e1a; 48->48))
 *   level 2: e2a = relu(conv(pool(e1); 48->64)), e2 = relu(conv(e2a; 64->64))
 *   level 3: b0 = relu(conv(pool(e2); 64->96)), b1 = relu(conv(b0)), b = relu(conv(b1))
 *   level 2: d2a = relu(conv([e2, up(b)]; 160->64)), d2 = relu(conv(d2a; 64->64))
 *   level 1: d1a = relu(conv([e1, up(d2)]; 112->48)), d1 = relu(conv(d1a; 48->48))
 *   level 0: d0a = relu(conv([e0, up(d1)]; 80->32)), d0 = relu(conv(d0a; 32->32))
 *   out_log = log1p(beauty rgb) + W_head d0 + b_head           (1x1, 32 -> 3)
 *   output  = expm1(out_log) (clamped at 0)
 *   loss    = mean over pixels and channels of w (out_log - log1p(t))^2,
 *             w = (1 + t)^2 / (t^2 + 1e-2), t = max(target, 0): the relative
 *             squared error (y - t)^2 / (t^2 + 1e-2) to first order in the
 *             log-space error
 * A level's image is ceil(w/2) x ceil(h/2) of the level above; pool is the
 * mean of the (1, 2 or 4) covered pixels; up(x) at a pixel is x at its
 * parent (x >> 1, y >> 1).  features f (16 ch): log1p(beauty rgb), albedo,
 * normal * 2 - 1, log1p(depth) / 8, 6 zeros.
 * Parameter vector: per convolution W[co][tap][ci] (tap = 3 ky + kx) then
 * b[co]; then W_head[c][k], b_head[c].
 *
 * Backward: the gradient of a convolution's input is computed raw and
 * rounded once; the producers take it as follows: a single input from the
 * previous convolution is masked by that convolution's ReLU; the upsampled
 * part of a decoder input is summed over the pixels a coarse pixel covers
 * (in (dy, dx) order) and masked; an encoder output e_l receives its skip
 * gradient plus the pooled path's gradient divided by the number of pixels
 * its parent covers, then the mask.
 * With `bf16` set the oracle rounds exactly where the GPU stores bf16:
 * features, weights, post-ReLU activations, pooled tensors, the raw input
 * gradients and the pre-activation gradients; products and sums accumulate
 * in double (FP32 where the GPU sums a handful of bf16 values).
 */
#include "rt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { NN_CF = 16, NN_OUT = 32, NN_CONVS = 15, NN_LEVELS = 4 };

typedef struct { int c0, c1, cout, level; } conv_t;
static const conv_t kConv[NN_CONVS] = {{16, 0, 32, 0}, {32, 0, 32, 0}, {32, 0, 48, 1},
    {48, 0, 48, 1}, {48, 0, 64, 2}, {64, 0, 64, 2}, {64, 0, 96, 3}, {96, 0, 96, 3}, {96, 0, 96, 3},
    {64, 96, 64, 2}, {64, 0, 64, 2}, {48, 64, 48, 1}, {48, 0, 48, 1}, {32, 48, 32, 0},
    {32, 0, 32, 0}};

static int cin_of(int l) { return kConv[l].c0 + kConv[l].c1; }
static size_t layer_params(int l) { return (size_t)kConv[l].cout * 9u * (size_t)cin_of(l) + (size_t)kConv[l].cout; }
static size_t layer_offset(int l) {
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += layer_params(i);
    return o;
}
static size_t head_offset(void) { return layer_offset(NN_CONVS); }

int rto_nn_param_count(void) { return (int)(head_offset() + 3u * NN_OUT + 3u); }

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

static float rnd(float v, int bf16) { return bf16 ? bf16r(v) : v; }

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

/* the batch at one level */
typedef struct { int count; int offset[17]; int width[16]; int height[16]; } lvl_t;

static void make_levels(const rto_images* im, lvl_t lv[NN_LEVELS]) {
    lv[0].count = im->count;
    for (int i = 0; i <= im->count; ++i) lv[0].offset[i] = im->offset[i];
    for (int i = 0; i < im->count; ++i) {
        lv[0].width[i] = im->width[i];
        lv[0].height[i] = (im->offset[i + 1] - im->offset[i]) / im->width[i];
    }
    for (int l = 1; l < NN_LEVELS; ++l) {
        lv[l].count = im->count;
        int o = 0;
        for (int i = 0; i < im->count; ++i) {
            lv[l].offset[i] = o;
            lv[l].width[i] = (lv[l - 1].width[i] + 1) / 2;
            lv[l].height[i] = (lv[l - 1].height[i] + 1) / 2;
            o += lv[l].width[i] * lv[l].height[i];
        }
        lv[l].offset[im->count] = o;
    }
}

static int npix(const lvl_t* L) { return L->offset[L->count]; }

typedef struct { int img, x, y; } pix_t;

static pix_t locate(const lvl_t* L, int p) {
    pix_t r;
    int i = 0;
    while (i + 1 < L->count && p >= L->offset[i + 1]) ++i;
    r.img = i;
    r.y = (p - L->offset[i]) / L->width[i];
    r.x = (p - L->offset[i]) - r.y * L->width[i];
    return r;
}

/* value of input channel ci of convolution l's input at (x, y) of image i,
 * or 0 outside the image */
static float conv_input(const lvl_t* L, const lvl_t* Lp, int i, int x, int y, int c0, int c1,
    const float* A, const float* B, int ci) {
    if (x < 0 || y < 0 || x >= L->width[i] || y >= L->height[i]) return 0.0f;
    if (ci < c0) return A[(size_t)(L->offset[i] + y * L->width[i] + x) * c0 + ci];
    return B[(size_t)(Lp->offset[i] + (y >> 1) * Lp->width[i] + (x >> 1)) * c1 + (ci - c0)];
}

static void conv_forward(const lvl_t* L, const lvl_t* Lp, const conv_t* c, const float* A,
    const float* B, const float* W, const float* b, int bf16, float* out) {
    const int cin = c->c0 + c->c1;
    const int P = npix(L);
    double* acc = (double*)malloc(sizeof(double) * (size_t)c->cout);
    for (int p = 0; p < P; ++p) {
        const pix_t q = locate(L, p);
        for (int co = 0; co < c->cout; ++co) acc[co] = 0.0;
        for (int t = 0; t < 9; ++t) {
            const int sx = q.x + t % 3 - 1, sy = q.y + t / 3 - 1;
            if (sx < 0 || sy < 0 || sx >= L->width[q.img] || sy >= L->height[q.img]) continue;
            for (int ci = 0; ci < cin; ++ci) {
                const float x = conv_input(L, Lp, q.img, sx, sy, c->c0, c->c1, A, B, ci);
                if (x == 0.0f) continue;
                for (int co = 0; co < c->cout; ++co)
                    acc[co] += (double)W[((size_t)co * 9 + t) * cin + ci] * (double)x;
            }
        }
        for (int co = 0; co < c->cout; ++co) {
            const float v = (float)(acc[co] + (double)b[co]);
            out[(size_t)p * c->cout + co] = rnd(v > 0.0f ? v : 0.0f, bf16);
        }
    }
    free(acc);
}

static void pool(const lvl_t* L, const lvl_t* Lp, int C, const float* in, float* out) {
    const int Pp = npix(Lp);
    for (int pp = 0; pp < Pp; ++pp) {
        const pix_t q = locate(Lp, pp);
        for (int c = 0; c < C; ++c) {
            float s = 0.0f;
            int n = 0;
            for (int dy = 0; dy < 2; ++dy) {
                for (int dx = 0; dx < 2; ++dx) {
                    const int x = 2 * q.x + dx, y = 2 * q.y + dy;
                    if (x < L->width[q.img] && y < L->height[q.img]) {
                        s += in[(size_t)(L->offset[q.img] + y * L->width[q.img] + x) * C + c];
                        ++n;
                    }
                }
            }
            out[(size_t)pp * C + c] = s / (float)n;
        }
    }
}

/* raw input gradient of convolution l: R[p][ci] = sum over taps t and
 * outputs co of W[co][t][ci] dz[p - off(t)][co] */
static void conv_input_grad(const lvl_t* L, const conv_t* c, const float* W, const float* dz,
    int bf16, float* R) {
    const int cin = c->c0 + c->c1;
    const int P = npix(L);
    double* acc = (double*)malloc(sizeof(double) * (size_t)cin);
    for (int p = 0; p < P; ++p) {
        const pix_t q = locate(L, p);
        for (int ci = 0; ci < cin; ++ci) acc[ci] = 0.0;
        for (int t = 0; t < 9; ++t) {
            const int sx = q.x - (t % 3 - 1), sy = q.y - (t / 3 - 1);
            if (sx < 0 || sy < 0 || sx >= L->width[q.img] || sy >= L->height[q.img]) continue;
            const float* d = dz + (size_t)(L->offset[q.img] + sy * L->width[q.img] + sx) * c->cout;
            for (int co = 0; co < c->cout; ++co) {
                if (d[co] == 0.0f) continue;
                const float* w = W + ((size_t)co * 9 + t) * cin;
                for (int ci = 0; ci < cin; ++ci) acc[ci] += (double)w[ci] * (double)d[co];
            }
        }
        for (int ci = 0; ci < cin; ++ci) R[(size_t)p * cin + ci] = rnd((float)acc[ci], bf16);
    }
    free(acc);
}

static void conv_weight_grad(const lvl_t* L, const lvl_t* Lp, const conv_t* c, const float* A,
    const float* B, const float* dz, double* g) {
    const int cin = c->c0 + c->c1;
    const int P = npix(L);
    for (int p = 0; p < P; ++p) {
        const pix_t q = locate(L, p);
        const float* d = dz + (size_t)p * c->cout;
        for (int t = 0; t < 9; ++t) {
            const int sx = q.x + t % 3 - 1, sy = q.y + t / 3 - 1;
            if (sx < 0 || sy < 0 || sx >= L->width[q.img] || sy >= L->height[q.img]) continue;
            for (int ci = 0; ci < cin; ++ci) {
                const float x = conv_input(L, Lp, q.img, sx, sy, c->c0, c->c1, A, B, ci);
                if (x == 0.0f) continue;
                for (int co = 0; co < c->cout; ++co) {
                    if (d[co] == 0.0f) continue;
                    g[((size_t)co * 9 + t) * cin + ci] += (double)d[co] * (double)x;
                }
            }
        }
        for (int co = 0; co < c->cout; ++co) g[(size_t)c->cout * 9 * cin + co] += d[co];
    }
}

/* dz = R where act > 0 (R and act with C channels at one level) */
static void mask(size_t n, const float* R, const float* act, float* dz) {
    for (size_t i = 0; i < n; ++i) dz[i] = act[i] > 0.0f ? R[i] : 0.0f;
}

/* the upsampled part R[:, off .. off + C) of a decoder input, summed over
 * the pixels each coarse pixel covers, masked by the coarse activation */
static void up_combine(const lvl_t* L, const lvl_t* Lp, int C, const float* R, int cr, int off,
    const float* act, int bf16, float* dz) {
    const int Pp = npix(Lp);
    for (int pp = 0; pp < Pp; ++pp) {
        const pix_t q = locate(Lp, pp);
        for (int c = 0; c < C; ++c) {
            float s = 0.0f;
            for (int dy = 0; dy < 2; ++dy) {
                for (int dx = 0; dx < 2; ++dx) {
                    const int x = 2 * q.x + dx, y = 2 * q.y + dy;
                    if (x < L->width[q.img] && y < L->height[q.img])
                        s += R[(size_t)(L->offset[q.img] + y * L->width[q.img] + x) * cr + off + c];
                }
            }
            dz[(size_t)pp * C + c] = act[(size_t)pp * C + c] > 0.0f ? rnd(s, bf16) : 0.0f;
        }
    }
}

/* encoder output gradient: skip part Rs[:, 0..C) plus the pooled path's
 * gradient Rp at the parent / covered pixels, masked */
static void skip_combine(const lvl_t* L, const lvl_t* Lp, int C, const float* Rs, int cs,
    const float* Rp, const float* act, int bf16, float* dz) {
    const int P = npix(L);
    for (int p = 0; p < P; ++p) {
        const pix_t q = locate(L, p);
        const int px = q.x >> 1, py = q.y >> 1;
        const int nx = L->width[q.img] - 2 * px < 2 ? L->width[q.img] - 2 * px : 2;
        const int ny = L->height[q.img] - 2 * py < 2 ? L->height[q.img] - 2 * py : 2;
        const int pp = Lp->offset[q.img] + py * Lp->width[q.img] + px;
        for (int c = 0; c < C; ++c) {
            const float v = Rs[(size_t)p * cs + c] + Rp[(size_t)pp * C + c] / (float)(nx * ny);
            dz[(size_t)p * C + c] = act[(size_t)p * C + c] > 0.0f ? rnd(v, bf16) : 0.0f;
        }
    }
}

/* loss (and gradient w.r.t. the parameters when grad != NULL) over all
 * pixels of the images; out4 (may be NULL) receives the denoised rgb */
double rto_nn_loss_grad(const rto_images* im, const float* params_in, const float* feat,
    const float* beauty, const float* target, float* out4, float* grad, int bf16) {
    lvl_t lv[NN_LEVELS];
    make_levels(im, lv);
    const int P = npix(&lv[0]);
    const int np = rto_nn_param_count();
    float* params = (float*)malloc(sizeof(float) * (size_t)np);
    for (int i = 0; i < np; ++i) params[i] = params_in[i];
    if (bf16) {
        for (int l = 0; l < NN_CONVS; ++l) {
            float* W = params + layer_offset(l);
            const size_t nw = (size_t)kConv[l].cout * 9u * (size_t)cin_of(l);
            for (size_t i = 0; i < nw; ++i) W[i] = bf16r(W[i]);
        }
    }
    float* act[NN_CONVS];
    float* pl[NN_LEVELS] = {NULL, NULL, NULL, NULL};
    for (int l = 0; l < NN_CONVS; ++l)
        act[l] = (float*)calloc((size_t)npix(&lv[kConv[l].level]) * kConv[l].cout, sizeof(float));
    for (int l = 1; l < NN_LEVELS; ++l)
        pl[l] = (float*)calloc((size_t)npix(&lv[l]) * kConv[2 * l - 1].cout, sizeof(float));
    /* forward */
    for (int l = 0; l < NN_CONVS; ++l) {
        const conv_t* c = &kConv[l];
        const float* A;
        const float* B = NULL;
        switch (l) {
            case 0: A = feat; break;
            case 2: A = pl[1]; break;
            case 4: A = pl[2]; break;
            case 6: A = pl[3]; break;
            case 9: A = act[5]; B = act[8]; break;
            case 11: A = act[3]; B = act[10]; break;
            case 13: A = act[1]; B = act[12]; break;
            default: A = act[l - 1]; break;
        }
        const int lvl = c->level;
        const float* W = params + layer_offset(l);
        conv_forward(&lv[lvl], &lv[lvl < NN_LEVELS - 1 ? lvl + 1 : lvl], c, A, B, W,
            W + (size_t)c->cout * 9 * cin_of(l), bf16, act[l]);
        if (l == 1 || l == 3 || l == 5) {
            pool(&lv[lvl], &lv[lvl + 1], c->cout, act[l], pl[lvl + 1]);
            if (bf16) {
                const size_t n = (size_t)npix(&lv[lvl + 1]) * c->cout;
                for (size_t i = 0; i < n; ++i) pl[lvl + 1][i] = bf16r(pl[lvl + 1][i]);
            }
        }
    }
    /* head */
    const float* Wh = params + head_offset();
    const float* bh = Wh + 3 * NN_OUT;
    float* dout = (float*)calloc((size_t)P * 3, sizeof(float));
    double loss = 0.0;
    const double scale = 1.0 / (3.0 * (double)P);
    const float* d0 = act[NN_CONVS - 1];
    for (int p = 0; p < P; ++p) {
        const float* h = d0 + (size_t)p * NN_OUT;
        for (int c = 0; c < 3; ++c) {
            double y = bh[c];
            for (int k = 0; k < NN_OUT; ++k) y += (double)Wh[c * NN_OUT + k] * (double)h[k];
            const float base = log1pf(fmaxf(beauty[4 * p + c], 0.0f));
            const float pred = (float)((double)base + y);
            if (out4) out4[4 * p + c] = fmaxf(expm1f(pred), 0.0f);
            if (target) {
                const float t = fmaxf(target[4 * p + c], 0.0f);
                const float tl = log1pf(t);
                const double w = ((1.0 + t) * (1.0 + t)) / ((double)t * t + 1e-2);
                const double e = (double)pred - (double)tl;
                loss += w * e * e * scale;
                dout[(size_t)p * 3 + c] = (float)(2.0 * w * e * scale);
            }
        }
        if (out4) out4[4 * p + 3] = 1.0f;
    }
    if (grad && target) {
        double* g = (double*)calloc((size_t)np, sizeof(double));
        float* dz[NN_CONVS];
        float* R[NN_CONVS];
        for (int l = 0; l < NN_CONVS; ++l) {
            const size_t n = (size_t)npix(&lv[kConv[l].level]);
            dz[l] = (float*)calloc(n * kConv[l].cout, sizeof(float));
            R[l] = (float*)calloc(n * cin_of(l), sizeof(float));
        }
        /* head */
        const size_t oh = head_offset();
        for (int p = 0; p < P; ++p) {
            const float* h = d0 + (size_t)p * NN_OUT;
            const float* d = dout + (size_t)p * 3;
            for (int c = 0; c < 3; ++c) {
                for (int k = 0; k < NN_OUT; ++k) g[oh + (size_t)c * NN_OUT + k] += (double)d[c] * (double)h[k];
                g[oh + 3 * NN_OUT + c] += d[c];
            }
            for (int k = 0; k < NN_OUT; ++k) {
                double s = 0.0;
                for (int c = 0; c < 3; ++c) s += (double)Wh[c * NN_OUT + k] * (double)d[c];
                const float v = h[k] > 0.0f ? (float)s : 0.0f;
                dz[NN_CONVS - 1][(size_t)p * NN_OUT + k] = rnd(v, bf16);
            }
        }
        /* convolutions, last to first */
        for (int l = NN_CONVS - 1; l >= 0; --l) {
            const conv_t* c = &kConv[l];
            const int lvl = c->level;
            const lvl_t* L = &lv[lvl];
            const lvl_t* Lp = &lv[lvl < NN_LEVELS - 1 ? lvl + 1 : lvl];
            const float* A;
            const float* B = NULL;
            switch (l) {
                case 0: A = feat; break;
                case 2: A = pl[1]; break;
                case 4: A = pl[2]; break;
                case 6: A = pl[3]; break;
                case 9: A = act[5]; B = act[8]; break;
                case 11: A = act[3]; B = act[10]; break;
                case 13: A = act[1]; B = act[12]; break;
                default: A = act[l - 1]; break;
            }
            conv_weight_grad(L, Lp, c, A, B, dz[l], g + layer_offset(l));
            if (l == 0) break;
            conv_input_grad(L, c, params + layer_offset(l), dz[l], bf16, R[l]);
            switch (l) {
                case 13: up_combine(&lv[0], &lv[1], 48, R[13], 80, 32, act[12], bf16, dz[12]); break;
                case 11: up_combine(&lv[1], &lv[2], 64, R[11], 112, 48, act[10], bf16, dz[10]); break;
                case 9: up_combine(&lv[2], &lv[3], 96, R[9], 160, 64, act[8], bf16, dz[8]); break;
                case 6: skip_combine(&lv[2], &lv[3], 64, R[9], 160, R[6], act[5], bf16, dz[5]); break;
                case 4: skip_combine(&lv[1], &lv[2], 48, R[11], 112, R[4], act[3], bf16, dz[3]); break;
                case 2: skip_combine(&lv[0], &lv[1], 32, R[13], 80, R[2], act[1], bf16, dz[1]); break;
                default:
                    mask((size_t)npix(L) * cin_of(l), R[l], act[l - 1], dz[l - 1]);
                    break;
            }
        }
        for (int i = 0; i < np; ++i) grad[i] = (float)g[i];
        free(g);
        for (int l = 0; l < NN_CONVS; ++l) {
            free(dz[l]);
            free(R[l]);
        }
    }
    free(dout);
    for (int l = 0; l < NN_CONVS; ++l) free(act[l]);
    for (int l = 1; l < NN_LEVELS; ++l) free(pl[l]);
    free(params);
    return loss;
}
