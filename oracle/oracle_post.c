/*
 * oracle_post.c -- CPU restatement of the per-frame post passes: temporal
 * reprojection/accumulation and the edge-aware denoise filter.
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 */
#include "rt_oracle.h"
#include "om.h"

/* resolve_reprojection_kernel, src/realtime/gpu/programs.cu:1955-2067, for
 * one camera image.  beauty / normal / depth are the current frame (updated
 * in place), h* the previous resolved frame of the same camera.  normal is
 * the 0.5 n + 0.5 encoding of store_output (programs.cu:1779-1783). */
void rto_resolve(const rto_camera* cam, const rto_camera* prev_cam, int history_length,
    int restir_enabled, float* beauty, float* normal, float* depth, const float* hbeauty,
    const float* hnormal, const float* hdepth, float* flow, float* trust) {
    const int w = cam->width, h = cam->height;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int p = y * w + x;
            flow[2 * p] = 0.0f;
            flow[2 * p + 1] = 0.0f;
            trust[p] = 0.0f;
            if (history_length <= 0 || prev_cam == NULL) continue;
            const float pxf = (float)x + 0.5f, pyf = (float)y + 0.5f;
            float o[3], d[3];
            rto_camera_ray(cam, (double)pxf, (double)pyf, o, d);
            const float cd = depth[p];
            const v3 wp = add3(ld3(o), scl3(ld3(d), cd));
            const v3 off = mk3(wp.x - (float)prev_cam->origin[0], wp.y - (float)prev_cam->origin[1],
                wp.z - (float)prev_cam->origin[2]);
            const v3 bx = mk3((float)prev_cam->basis_x[0], (float)prev_cam->basis_x[1], (float)prev_cam->basis_x[2]);
            const v3 by = mk3((float)prev_cam->basis_y[0], (float)prev_cam->basis_y[1], (float)prev_cam->basis_y[2]);
            const v3 bz = mk3((float)prev_cam->basis_z[0], (float)prev_cam->basis_z[1], (float)prev_cam->basis_z[2]);
            const float pcs[3] = {dot3(off, bx), dot3(off, by), dot3(off, bz)};
            float pp[2];
            rto_project(prev_cam, pcs, pp);
            const int finite = isfinite(pp[0]) && isfinite(pp[1]);
            if (finite) {
                flow[2 * p] = pxf - pp[0];
                flow[2 * p + 1] = pyf - pp[1];
            }
            const int qx = finite ? (int)floorf(pp[0]) : -1;
            const int qy = finite ? (int)floorf(pp[1]) : -1;
            if (!finite || qx < 0 || qx >= w || qy < 0 || qy >= h) continue;
            const int q = qy * w + qx;
            const v3 hn = mk3(hnormal[4 * q] * 2.0f - 1.0f, hnormal[4 * q + 1] * 2.0f - 1.0f,
                hnormal[4 * q + 2] * 2.0f - 1.0f);
            const v3 cn = mk3(normal[4 * p] * 2.0f - 1.0f, normal[4 * p + 1] * 2.0f - 1.0f,
                normal[4 * p + 2] * 2.0f - 1.0f);
            const float ndot = dot3(cn, hn);
            const float hd = hdepth[q];
            const float expected = len3(ld3(pcs));
            const float tol = fmaxf(0.02f, 0.05f * fmaxf(expected, hd));
            if (!(ndot > 0.95f && fabsf(expected - hd) < tol)) continue;
            trust[p] = 1.0f;
            const float blend = 1.0f / (float)(history_length + 1);
            if (!restir_enabled) {
                for (int c = 0; c < 3; ++c) {
                    const float cur = fminf(fmaxf(beauty[4 * p + c], 0.0f), 10.0f);
                    beauty[4 * p + c] = hbeauty[4 * q + c] + (cur - hbeauty[4 * q + c]) * blend;
                }
                beauty[4 * p + 3] = 1.0f;
            }
            for (int c = 0; c < 4; ++c)
                normal[4 * p + c] = hnormal[4 * q + c] + (normal[4 * p + c] - hnormal[4 * q + c]) * blend;
            depth[p] = hd + (cd - hd) * blend;
        }
    }
}

/* Edge-aware a-trous filter (5x5 B3-spline taps at spacing `step`), one pass
 * over one camera image.  Edge stops use only +, *, / so every platform
 * rounds them identically:
 *   normals  max(0, n_p.n_q)^16,
 *   depth    1 / (1 + (|z_p - z_q| / (0.1 step z_p))^2),
 *   albedo   1 / (1 + (|a_p - a_q|_1 / 0.1)^2),
 *   colour   1 / (1 + (|l_p - l_q| / (4 max(l_p, l_q) + 1e-3))^2).
 * Pixels without a surface (depth 0) pass through and are never taps. */
static const float kB3[3] = {0.375f, 0.25f, 0.0625f};

void rto_denoise_pass(int w, int h, int step, const float* in, const float* normal,
    const float* albedo, const float* depth, float* out) {
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int p = y * w + x;
            const float zp = depth[p];
            if (!(zp > 0.0f)) {
                for (int c = 0; c < 4; ++c) out[4 * p + c] = in[4 * p + c];
                continue;
            }
            const v3 cp = ld3(in + 4 * p);
            const v3 np = mk3(normal[4 * p] * 2.0f - 1.0f, normal[4 * p + 1] * 2.0f - 1.0f,
                normal[4 * p + 2] * 2.0f - 1.0f);
            const v3 ap = ld3(albedo + 4 * p);
            const float lp = 0.2126f * cp.x + 0.7152f * cp.y + 0.0722f * cp.z;
            const float zs = 0.1f * (float)step * zp;
            v3 sum = mk3(0.0f, 0.0f, 0.0f);
            float wsum = 0.0f;
            for (int dy = -2; dy <= 2; ++dy) {
                const int qy = y + dy * step;
                if (qy < 0 || qy >= h) continue;
                for (int dx = -2; dx <= 2; ++dx) {
                    const int qx = x + dx * step;
                    if (qx < 0 || qx >= w) continue;
                    const int q = qy * w + qx;
                    const float zq = depth[q];
                    if (!(zq > 0.0f)) continue;
                    const v3 cq = ld3(in + 4 * q);
                    const v3 nq = mk3(normal[4 * q] * 2.0f - 1.0f, normal[4 * q + 1] * 2.0f - 1.0f,
                        normal[4 * q + 2] * 2.0f - 1.0f);
                    const v3 aq = ld3(albedo + 4 * q);
                    const float lq = 0.2126f * cq.x + 0.7152f * cq.y + 0.0722f * cq.z;
                    float wn = fmaxf(dot3(np, nq), 0.0f);
                    wn = wn * wn; wn = wn * wn; wn = wn * wn; wn = wn * wn;
                    const float ez = fabsf(zp - zq) / zs;
                    const float wz = 1.0f / (1.0f + ez * ez);
                    const float ea = (fabsf(ap.x - aq.x) + fabsf(ap.y - aq.y) + fabsf(ap.z - aq.z)) / 0.1f;
                    const float wa = 1.0f / (1.0f + ea * ea);
                    const float el = fabsf(lp - lq) / (4.0f * fmaxf(lp, lq) + 1e-3f);
                    const float wl = 1.0f / (1.0f + el * el);
                    const float k = kB3[dx < 0 ? -dx : dx] * kB3[dy < 0 ? -dy : dy];
                    const float wt = k * wn * wz * wa * wl;
                    sum = add3(sum, scl3(cq, wt));
                    wsum += wt;
                }
            }
            if (wsum > 0.0f) {
                const float inv = 1.0f / wsum;
                out[4 * p] = sum.x * inv;
                out[4 * p + 1] = sum.y * inv;
                out[4 * p + 2] = sum.z * inv;
            } else {
                out[4 * p] = cp.x;
                out[4 * p + 1] = cp.y;
                out[4 * p + 2] = cp.z;
            }
            out[4 * p + 3] = 1.0f;
        }
    }
}

/* `iterations` passes at spacing 1, 2, 4, ...; `scratch` holds one image */
void rto_denoise(int w, int h, int iterations, const float* beauty, const float* normal,
    const float* albedo, const float* depth, float* out, float* scratch) {
    const size_t n = (size_t)w * (size_t)h * 4u;
    for (size_t i = 0; i < n; ++i) out[i] = beauty[i];
    for (int it = 0; it < iterations; ++it) {
        rto_denoise_pass(w, h, 1 << it, out, normal, albedo, depth, scratch);
        for (size_t i = 0; i < n; ++i) out[i] = scratch[i];
    }
}

/* linear_to_display_u8 of the viewer's tone-map kernel,
 * src/realtime/viewer/cuda_gl_presenter.cu:23-26: gamma 2 (square root) of
 * the radiance clamped below at 0 (fmaxf also sends a quiet NaN to 0) and
 * above at 0.999, scaled by 255 and rounded half away from zero (lroundf).
 * The host twin rt::linear_to_display_u8 (src/realtime/display_transfer.h:
 * 9-15) agrees on every non-NaN input. */
uint8_t rto_display_u8(float v) {
    const float d = fminf(sqrtf(fmaxf(v, 0.0f)), 0.999f);
    return (uint8_t)lroundf(d * 255.0f);
}

/* tone_map_to_rgba8, cuda_gl_presenter.cu:28-38: one float4 in, RGBA8 out,
 * alpha 255. */
void rto_tone_map_rgba8(const float* rgba, int64_t n, uint8_t* out) {
    for (int64_t i = 0; i < n; ++i) {
        out[4 * i + 0] = rto_display_u8(rgba[4 * i + 0]);
        out[4 * i + 1] = rto_display_u8(rgba[4 * i + 1]);
        out[4 * i + 2] = rto_display_u8(rgba[4 * i + 2]);
        out[4 * i + 3] = 255;
    }
}
