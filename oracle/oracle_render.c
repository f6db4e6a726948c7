/*
 * oracle_render.c -- CPU restatement of the realtime integrator.
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 *
 * Follows src/realtime/gpu/programs.cu:
 *   camera rays        :308-373 (pinhole32), :375-515 (equi62_lut1d), :2086-2099
 *   light selection    :559-619 (prim->light lookup replaces the O(L) scan)
 *   emission MIS       :621-647
 *   populate_material  :961-1009, direct response :1026-1055
 *   sample_direct_light:1057-1232, accumulate_direct_light :2249-2280
 *   russian roulette   :1764-1771, sample_bsdf :2451-2565
 *   radiance_kernel    :1828-1953, store_output :1773-1805
 * with the counter RNG of om.h: every bounce draws exactly two 4-tuples
 * (group 0: light select + light position; group 1: BSDF lobe/direction + RR);
 * rejection loops are replaced by closed-form samplers of the same
 * distribution (unit sphere from two uniforms, unit ball via cube root).
 */
#include "rt_oracle.h"
#include "om.h"
#include "oracle_internal.h"

#include <pthread.h>
#include <stdlib.h>

/* ------------------------------------------------------------- cameras */
typedef struct { double x, y; } d2;

static d2 distort_pinhole(const rto_camera* c, d2 xy) {
    const double x = xy.x, y = xy.y;
    const double x2 = x * x, y2 = y * y, xyt = x * y;
    const double r2 = x2 + y2, r4 = r2 * r2, r6 = r4 * r2;
    const double radial = 1.0 + c->k1 * r2 + c->k2 * r4 + c->k3 * r6;
    d2 r = {x * radial + c->p2 * (r2 + 2.0 * x2) + 2.0 * c->p1 * xyt,
        y * radial + c->p1 * (r2 + 2.0 * y2) + 2.0 * c->p2 * xyt};
    return r;
}

static int solve2(double a00, double a01, double a10, double a11, d2 rhs, d2* delta) {
    const double det = a00 * a11 - a01 * a10;
    if (fabs(det) < 1e-12) return 0;
    const double inv = 1.0 / det;
    delta->x = (a11 * rhs.x - a01 * rhs.y) * inv;
    delta->y = (-a10 * rhs.x + a00 * rhs.y) * inv;
    return 1;
}

/* programs.cu:335-373 */
static v3 unproject_pinhole(const rto_camera* c, double px, double py) {
    const d2 xyd = {(px - c->cx) / c->fx, (py - c->cy) / c->fy};
    d2 xy = xyd;
    for (int it = 0; it < 24; ++it) {
        const d2 dist = distort_pinhole(c, xy);
        const d2 err = {xyd.x - dist.x, xyd.y - dist.y};
        if (err.x * err.x + err.y * err.y < 1e-20) break;
        const double x = xy.x, y = xy.y;
        const double r2 = x * x + y * y, r4 = r2 * r2, r6 = r4 * r2;
        const double radial = 1.0 + c->k1 * r2 + c->k2 * r4 + c->k3 * r6;
        const double drdx = 2.0 * x * (c->k1 + 2.0 * c->k2 * r2 + 3.0 * c->k3 * r4);
        const double drdy = 2.0 * y * (c->k1 + 2.0 * c->k2 * r2 + 3.0 * c->k3 * r4);
        d2 delta;
        if (!solve2(radial + x * drdx + 6.0 * c->p2 * x + 2.0 * c->p1 * y,
                x * drdy + 2.0 * c->p1 * x + 2.0 * c->p2 * y,
                y * drdx + 2.0 * c->p1 * x + 2.0 * c->p2 * y,
                radial + y * drdy + 6.0 * c->p1 * y + 2.0 * c->p2 * x, err, &delta))
            break;
        xy.x += delta.x;
        xy.y += delta.y;
    }
    return norm3(mk3((float)xy.x, (float)xy.y, 1.0f));
}

static d2 equi_tangential(d2 xy, const double* tg) {
    const double xr = xy.x, yr = xy.y, xr2 = xr * xr, yr2 = yr * yr, xryr = xr * yr;
    d2 r = {xr + 2.0 * tg[0] * xryr + tg[1] * (xr2 + yr2 + 2.0 * xr2),
        yr + tg[0] * (xr2 + yr2 + 2.0 * yr2) + 2.0 * tg[1] * xryr};
    return r;
}

/* programs.cu:399-413 */
static d2 equi_remove_tangential(d2 xyd, const double* tg) {
    const d2 dist = equi_tangential(xyd, tg);
    d2 delta;
    const d2 rhs = {xyd.x - dist.x, xyd.y - dist.y};
    if (!solve2(1.0 + 2.0 * tg[0] * xyd.y + 6.0 * tg[1] * xyd.x,
            2.0 * tg[0] * xyd.x + 2.0 * tg[1] * xyd.y, 2.0 * tg[0] * xyd.x + 2.0 * tg[1] * xyd.y,
            1.0 + 6.0 * tg[0] * xyd.y + 2.0 * tg[1] * xyd.x, rhs, &delta))
        return xyd;
    d2 r = {xyd.x + delta.x, xyd.y + delta.y};
    return r;
}

/* programs.cu:415-433 */
static int equi_lut_theta(const rto_camera* c, double rd, double* theta) {
    if (rd <= 0.0 || c->lut_step <= 0.0) { *theta = 0.0; return 1; }
    const double pos = rd / c->lut_step;
    if (pos > 1023.0) return 0;
    if (pos >= 1023.0) { *theta = c->lut[1023]; return 1; }
    const int idx = (int)pos;
    const double alpha = pos - (double)idx;
    *theta = (1.0 - alpha) * c->lut[idx] + alpha * c->lut[idx + 1];
    return 1;
}

/* programs.cu:475-497 */
static v3 unproject_equi(const rto_camera* c, double px, double py) {
    const d2 xy = {(px - c->cx) / c->fx, (py - c->cy) / c->fy};
    if (xy.x * xy.x + xy.y * xy.y < 1e-12 * 1e-12) return mk3(0.0f, 0.0f, 1.0f);
    const d2 xr = equi_remove_tangential(xy, c->tangential);
    const double rd = sqrt(xr.x * xr.x + xr.y * xr.y);
    if (rd < 1e-12) return mk3(0.0f, 0.0f, 1.0f);
    double theta = 0.0;
    if (!equi_lut_theta(c, rd, &theta)) return norm3(mk3((float)xy.x, (float)xy.y, 1.0f));
    const double scale = tan(theta) / rd;
    return norm3(mk3((float)(xr.x * scale), (float)(xr.y * scale), 1.0f));
}

void rto_unproject(const rto_camera* cam, double px, double py, float dir_cam[3]) {
    const v3 d = cam->model == RTO_CAM_EQUI62_LUT1D ? unproject_equi(cam, px, py)
                                                    : unproject_pinhole(cam, px, py);
    dir_cam[0] = d.x; dir_cam[1] = d.y; dir_cam[2] = d.z;
}

/* programs.cu:521-532 */
static v3 cam_transform(const rto_camera* c, v3 dc) {
    const v3 bx = mk3((float)c->basis_x[0], (float)c->basis_x[1], (float)c->basis_x[2]);
    const v3 by = mk3((float)c->basis_y[0], (float)c->basis_y[1], (float)c->basis_y[2]);
    const v3 bz = mk3((float)c->basis_z[0], (float)c->basis_z[1], (float)c->basis_z[2]);
    return norm3(add3(add3(scl3(bx, dc.x), scl3(by, dc.y)), scl3(bz, dc.z)));
}

void rto_camera_ray(const rto_camera* cam, double px, double py, float origin[3], float dir[3]) {
    float dc[3];
    rto_unproject(cam, px, py, dc);
    const v3 d = cam_transform(cam, ld3(dc));
    origin[0] = (float)cam->origin[0]; origin[1] = (float)cam->origin[1]; origin[2] = (float)cam->origin[2];
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

/* direction_debug_kernel / encode_direction_channel (programs.cu:180-182,
 * 1807-1826): RGBA8 world directions of the pixel centres' rays */
void rto_direction_debug(const rto_camera* cam, uint8_t* rgba) {
    for (int y = 0; y < cam->height; ++y)
        for (int x = 0; x < cam->width; ++x) {
            float o[3], d[3];
            rto_camera_ray(cam, (double)((float)x + 0.5f), (double)((float)y + 0.5f), o, d);
            uint8_t* px = rgba + 4 * ((size_t)y * (size_t)cam->width + (size_t)x);
            for (int a = 0; a < 3; ++a) {
                float t = 0.5f * (d[a] + 1.0f);
                t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
                px[a] = (uint8_t)(255.0f * t);
            }
            px[3] = 255;
        }
}

/* programs.cu:375-473 (projection, used by reprojection) */
void rto_project(const rto_camera* c, const float dir_cam[3], float pixel[2]) {
    if (c->model == RTO_CAM_EQUI62_LUT1D) {
        const double dx = dir_cam[0], dy = dir_cam[1], dz = dir_cam[2];
        const double r = sqrt(dx * dx + dy * dy);
        if (r < 1e-12) { pixel[0] = (float)c->cx; pixel[1] = (float)c->cy; return; }
        const double theta = atan2(r, dz);
        double rd;
        if (theta <= 0.0 || c->lut_step <= 0.0) rd = 0.0;
        else if (theta >= c->lut[1023]) rd = 1023.0 * c->lut_step;
        else {
            rd = 1023.0 * c->lut_step;
            for (int i = 0; i < 1023; ++i) {
                if (c->lut[i] <= theta && theta <= c->lut[i + 1]) {
                    const double den = c->lut[i + 1] - c->lut[i];
                    const double a = den > 0.0 ? (theta - c->lut[i]) / den : 0.0;
                    rd = ((double)i + a) * c->lut_step;
                    break;
                }
            }
        }
        const d2 xyr = {dx / r * rd, dy / r * rd};
        const d2 xyd = equi_tangential(xyr, c->tangential);
        pixel[0] = (float)(xyd.x * c->fx + c->cx);
        pixel[1] = (float)(xyd.y * c->fy + c->cy);
        return;
    }
    if (fabsf(dir_cam[2]) < 1e-7f) { pixel[0] = -1.0f; pixel[1] = -1.0f; return; }
    const double inv_z = 1.0 / (double)dir_cam[2];
    const d2 xyu = {(double)dir_cam[0] * inv_z, (double)dir_cam[1] * inv_z};
    const d2 xyd = distort_pinhole(c, xyu);
    pixel[0] = (float)(xyd.x * c->fx + c->cx);
    pixel[1] = (float)(xyd.y * c->fy + c->cy);
}

/* host LUT builder, camera_models.cpp:64-100,228-248 */
static double equi_theta_dist(const double* radial, double theta) {
    const double t2 = theta * theta;
    double scale = 1.0, tp = t2;
    for (int i = 0; i < 6; ++i) { scale += radial[i] * tp; tp *= t2; }
    return theta * scale;
}
static double equi_theta_deriv(const double* radial, double theta) {
    const double t2 = theta * theta;
    double der = 1.0, tp = t2;
    for (int i = 0; i < 6; ++i) { der += (double)(2 * i + 3) * radial[i] * tp; tp *= t2; }
    return der;
}
static double equi_invert(const double* radial, double thetad) {
    double theta = cbrt(thetad > 0.0 ? thetad : 0.0);
    for (int it = 0; it < 30; ++it) {
        const double err = thetad - equi_theta_dist(radial, theta);
        if (fabs(err) < 1e-12) break;
        const double der = equi_theta_deriv(radial, theta);
        if (fabs(der) < 1e-12) break;
        const double delta = err / der;
        theta += delta;
        if (fabs(delta) < 1e-12 || fabs(delta) > 3.1416) break;
    }
    return theta;
}
void rto_make_equi_lut(rto_camera* c) {
    const double x = ((double)c->width + 10.0 - c->cx) / c->fx;
    const double y = ((double)c->height + 10.0 - c->cy) / c->fy;
    c->lut_step = sqrt(x * x + y * y) / 1024.0;
    for (int i = 0; i < 1024; ++i) c->lut[i] = equi_invert(c->radial, c->lut_step * (double)i);
}

static void pixel_position(int x, int y, int cam_index, int sample, uint32_t stream, int jitter,
    int width, double* px, double* py) {
    if (jitter) {
        float r[4];
        om_rng4((uint32_t)(y * width + x), (uint32_t)sample, (uint32_t)cam_index, 0u, OM_G_CAMERA,
            stream, r);
        *px = (double)((float)x + r[0]);
        *py = (double)((float)y + r[1]);
    } else {
        *px = (double)((float)x + 0.5f);
        *py = (double)((float)y + 0.5f);
    }
}

void rto_primary_rays(const rto_camera* cam, int cam_index, int sample, uint32_t stream, int jitter,
    rto_ray* rays) {
    for (int y = 0; y < cam->height; ++y) {
        for (int x = 0; x < cam->width; ++x) {
            double px, py;
            pixel_position(x, y, cam_index, sample, stream, jitter, cam->width, &px, &py);
            rto_ray* r = &rays[y * cam->width + x];
            rto_camera_ray(cam, px, py, r->o, r->d);
            r->tmin = 0.0f;
            r->tmax = OM_RAY_FAR;
        }
    }
}

/* ------------------------------------------------------------ lights */
static float lum_clamped(v3 c) {
    const v3 m = mk3(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
    return fmaxf(0.0f, 0.2126f * m.x + 0.7152f * m.y + 0.0722f * m.z);
}

/* average_texture (packed_scene_preparation.cpp:195-222) */
static v3 average_texture(const rto_scene* s, int index, int depth) {
    if (index < 0 || index >= s->texture_count || depth >= 8) return mk3(0, 0, 0);
    const rto_texture* t = &s->textures[index];
    if (t->type == 0) return mk3(fmaxf(t->color[0], 0.0f), fmaxf(t->color[1], 0.0f), fmaxf(t->color[2], 0.0f));
    if (t->type == 1) {
        const v3 a = average_texture(s, t->even_texture, depth + 1);
        const v3 b = average_texture(s, t->odd_texture, depth + 1);
        return mk3(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
    }
    if (t->type == 2 && t->image_width > 0 && t->image_height > 0) {
        const int64_t count = (int64_t)t->image_width * t->image_height;
        if (t->image_offset < 0 || t->image_offset + count > s->image_texel_count) return mk3(0, 0, 0);
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (int64_t i = 0; i < count; ++i) {
            const float* px = s->image_texels + 3 * ((int64_t)t->image_offset + i);
            sx += px[0];
            sy += px[1];
            sz += px[2];
        }
        const float n = (float)count;
        return mk3(fmaxf(sx / n, 0.0f), fmaxf(sy / n, 0.0f), fmaxf(sz / n, 0.0f));
    }
    if (t->type == 3) return mk3(0.5f, 0.5f, 0.5f);
    return mk3(0, 0, 0);
}

/* openpbr_source_to_linear (openpbr_core.h:322-333) */
static float srgb_channel_to_linear(float v) {
    return v <= 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f);
}
static v3 source_to_linear(v3 c, int space) {
    if (space != 2) return c;
    return mk3(srgb_channel_to_linear(c.x), srgb_channel_to_linear(c.y), srgb_channel_to_linear(c.z));
}

/* material_emission_luminance (packed_scene_preparation.cpp:224-248): a
 * textured emission counts with its texture's average */
static float material_emission_lum(const rto_scene* s, int m) {
    if (m < 0 || m >= s->material_count) return 0.0f;
    const rto_material* mt = &s->materials[m];
    if (mt->type == RTO_MAT_DIFFUSE_LIGHT) {
        if (s->material_emission_texture != NULL && s->material_emission_texture[m] >= 0)
            return lum_clamped(average_texture(s, s->material_emission_texture[m], 0));
        return lum_clamped(ld3(mt->emission));
    }
    if (mt->type != RTO_MAT_OPENPBR || mt->openpbr_index < 0 || mt->openpbr_index >= s->openpbr_count)
        return 0.0f;
    rto_openpbr p = s->openpbr[mt->openpbr_index];
    if (s->openpbr_textures != NULL && s->openpbr_textures[mt->openpbr_index].emission_color >= 0) {
        const rto_openpbr_textures* tb = &s->openpbr_textures[mt->openpbr_index];
        const v3 c = source_to_linear(average_texture(s, tb->emission_color, 0), tb->emission_color_space);
        p.emission_color[0] = c.x;
        p.emission_color[1] = c.y;
        p.emission_color[2] = c.z;
    }
    const v3 e = om_openpbr_emission(&p);
    return 0.2126f * fmaxf(0.0f, e.x) + 0.7152f * fmaxf(0.0f, e.y) + 0.0722f * fmaxf(0.0f, e.z);
}

static v3 background_mean(const rto_scene* s) {
    if (s->background_mode == RTO_BG_SKY_GRADIENT)
        return scl3(add3(ld3(s->background), ld3(s->background_top)), 0.5f);
    return ld3(s->background);
}

/* packed_scene_preparation.cpp:250-302; light types 0 sphere,1 quad,2 tri,3 env */
int rto_build_lights(const rto_scene* s, rto_light* out, int capacity) {
    int n = 0;
    const float kPi = OM_PI;
#define APPEND(TYPE, PRIM, W)                                                        \
    do {                                                                             \
        const float w_ = (W);                                                        \
        if (isfinite(w_) && w_ > 1e-12f) {                                           \
            if (n < capacity) { out[n].type = (TYPE); out[n].prim = (PRIM); out[n].selection_pdf = w_; out[n].cdf = 0.0f; } \
            ++n;                                                                     \
        }                                                                            \
    } while (0)
    for (int i = 0; i < s->sphere_count; ++i) {
        const float area = 4.0f * kPi * s->spheres[i].radius * s->spheres[i].radius;
        APPEND(0, i, area * material_emission_lum(s, s->spheres[i].material));
    }
    for (int i = 0; i < s->quad_count; ++i) {
        const float area = len3(cross3(ld3(s->quads[i].edge_u), ld3(s->quads[i].edge_v)));
        APPEND(1, i, area * material_emission_lum(s, s->quads[i].material));
    }
    for (int i = 0; i < s->triangle_count; ++i) {
        const rto_triangle* t = &s->triangles[i];
        const float area = 0.5f * len3(cross3(sub3(ld3(t->p1), ld3(t->p0)), sub3(ld3(t->p2), ld3(t->p0))));
        APPEND(2, i, area * material_emission_lum(s, t->material));
    }
    APPEND(3, -1, 4.0f * kPi * lum_clamped(background_mean(s)));
#undef APPEND
    if (n > capacity) return n;
    float total = 0.0f;
    for (int i = 0; i < n; ++i) total += out[i].selection_pdf;
    float cdf = 0.0f;
    for (int i = 0; i < n; ++i) {
        out[i].selection_pdf /= total;
        cdf += out[i].selection_pdf;
        out[i].cdf = cdf;
    }
    if (n > 0) out[n - 1].cdf = 1.0f;
    return n;
}

/* -------------------------------------------------------------- context */
typedef struct {
    const rto_scene* s;
    const rto_bvh* bvh;
    rto_light lights[4096];
    int light_count;
    int32_t* light_of_prim; /* global prim id -> light index or -1 */
    int env_light;
    float* a_sel;           /* analytic light selection pdf */
    float* a_cdf;
    int a_count;
} ctx_t;

/* rto_analytic_lights_compile: analytic_light_compiler.cpp:178-240 (world
 * area, delta, selection weight per type) and :15-39 (normalize, close the
 * cdf at the last light with positive probability), in FP64 */
int rto_analytic_lights_compile(rto_analytic_light* lights, int count) {
    const double pi = 3.14159265358979323846;
    double* w = (double*)malloc(sizeof(double) * (size_t)(count > 0 ? count : 1));
    double total = 0.0;
    for (int i = 0; i < count; ++i) {
        rto_analytic_light* L = &lights[i];
        if (L->type < RTO_ALIGHT_SPHERE || L->type > RTO_ALIGHT_DOME) { free(w); return -1; }
        /* columns of the local-to-world map */
        const double bx[3] = {L->basis_x[0], L->basis_x[1], L->basis_x[2]};
        const double by[3] = {L->basis_y[0], L->basis_y[1], L->basis_y[2]};
        const double bz[3] = {L->basis_z[0], L->basis_z[1], L->basis_z[2]};
        const double* col[3] = {bx, by, bz};
        const double cx = bx[1] * by[2] - bx[2] * by[1];
        const double cy = bx[2] * by[0] - bx[0] * by[2];
        const double cz = bx[0] * by[1] - bx[1] * by[0];
        const double cross_norm = sqrt(cx * cx + cy * cy + cz * cz);
        double scale = 1.0, area = 0.0;
        if (L->type == RTO_ALIGHT_SPHERE || L->type == RTO_ALIGHT_CYLINDER) {
            /* require_similarity_scale (:114-128): the Gram matrix is s^2 I */
            double g[3][3], s2 = 0.0, dev = 0.0;
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q)
                    g[r][q] = col[r][0] * col[q][0] + col[r][1] * col[q][1] + col[r][2] * col[q][2];
            s2 = (g[0][0] + g[1][1] + g[2][2]) / 3.0;
            if (!isfinite(s2) || s2 <= 0.0) { free(w); return -1; }
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) {
                    const double e = fabs(g[r][q] - (r == q ? s2 : 0.0));
                    if (e > dev) dev = e;
                }
            if (dev > 1e-6 * (s2 > 1.0 ? s2 : 1.0)) { free(w); return -1; }
            scale = sqrt(s2);
        }
        switch (L->type) {
            case RTO_ALIGHT_SPHERE: area = 4.0 * pi * pow((double)L->radius * scale, 2.0); break;
            case RTO_ALIGHT_DISK: area = pi * (double)L->radius * (double)L->radius * cross_norm; break;
            case RTO_ALIGHT_RECT: area = (double)L->width * (double)L->height * cross_norm; break;
            case RTO_ALIGHT_CYLINDER:
                area = 2.0 * pi * ((double)L->radius * scale) * ((double)L->length * scale);
                break;
            default: break;
        }
        if (L->type <= RTO_ALIGHT_CYLINDER) L->world_area = (float)area;
        if (L->type == RTO_ALIGHT_DISTANT) L->delta = L->cos_theta_max >= 1.0f ? 1 : 0;
        double lum = 0.2126 * (double)L->radiance[0] + 0.7152 * (double)L->radiance[1]
                     + 0.0722 * (double)L->radiance[2];
        if (lum < 0.0) lum = 0.0;
        double x;
        if (L->type == RTO_ALIGHT_DOME) x = 4.0 * pi * lum;
        else if (L->type == RTO_ALIGHT_DISTANT)
            x = (L->delta ? 1.0 : 2.0 * pi * (1.0 - (double)L->cos_theta_max)) * lum;
        else x = area * lum;
        x = isfinite(x) ? (x > 0.0 ? x : 0.0) : 0.0;
        w[i] = x;
        total += x;
    }
    double acc = 0.0;
    int last = -1;
    for (int i = 0; i < count; ++i) {
        const double p = total > 0.0 ? w[i] / total : 0.0;
        acc += p;
        lights[i].selection_pdf = (float)p;
        lights[i].cdf = (float)acc;
        if (p > 0.0) last = i;
    }
    if (total > 0.0)
        for (int i = last; i < count; ++i) lights[i].cdf = 1.0f;
    free(w);
    return 0;
}

static int prim_global(const rto_scene* s, int type, int idx) {
    if (type == 0) return idx;
    if (type == 1) return s->sphere_count + idx;
    return s->sphere_count + s->quad_count + idx;
}

static int ctx_init(ctx_t* c, const rto_scene* s, const rto_bvh* bvh) {
    c->s = s;
    c->bvh = bvh;
    c->light_count = rto_build_lights(s, c->lights, 4096);
    if (c->light_count > 4096) return -1;
    const int np = s->sphere_count + s->quad_count + s->triangle_count;
    c->light_of_prim = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
    for (int i = 0; i < np; ++i) c->light_of_prim[i] = -1;
    c->env_light = -1;
    for (int i = 0; i < c->light_count; ++i) {
        if (c->lights[i].type == 3) c->env_light = i;
        else c->light_of_prim[prim_global(s, c->lights[i].type, c->lights[i].prim)] = i;
    }
    /* the lights' compiled selection pdf / cdf, read as given */
    c->a_count = s->analytic_light_count;
    c->a_sel = (float*)malloc(sizeof(float) * (size_t)(c->a_count > 0 ? c->a_count : 1));
    c->a_cdf = (float*)malloc(sizeof(float) * (size_t)(c->a_count > 0 ? c->a_count : 1));
    for (int i = 0; i < c->a_count; ++i) {
        c->a_sel[i] = s->analytic_lights[i].selection_pdf;
        c->a_cdf[i] = s->analytic_lights[i].cdf;
    }
    return 0;
}
static void ctx_free(ctx_t* c) {
    free(c->light_of_prim);
    free(c->a_sel);
    free(c->a_cdf);
}

static v3 background_at(const rto_scene* s, v3 d) {
    if (s->background_mode == RTO_BG_SKY_GRADIENT) {
        const v3 dn = norm3(d);
        const float t = 0.5f * (dn.y + 1.0f);
        return add3(scl3(ld3(s->background), 1.0f - t), scl3(ld3(s->background_top), t));
    }
    return ld3(s->background);
}

/* --------------------------------------------------------------- hit info */
typedef struct {
    float t;
    v3 pos, gn, sn, base, emission;
    float fuzz, ior;
    float tex_u, tex_v;
    int mat_type, openpbr, prim, prim_type, prim_index, front;
} hit_t;

/* a texture at a point, with the texture coordinates u, v */
static v3 texture_at(const rto_scene* s, int index, float u, float v, v3 p) {
    const float pf[3] = {p.x, p.y, p.z};
    float c[3];
    rto_eval_texture(s, index, u, v, pf, c);
    return ld3(c);
}

/* openpbr_apply_color_input / openpbr_apply_scalar_input
 * (openpbr_core.h:335-353) from a texture at the hit */
static void apply_color_binding(const rto_scene* s, float dst[3], int tex, int space, const hit_t* h) {
    if (tex < 0) return;
    const v3 c = source_to_linear(texture_at(s, tex, h->tex_u, h->tex_v, h->pos), space);
    dst[0] = c.x;
    dst[1] = c.y;
    dst[2] = c.z;
}
static void apply_scalar_binding(const rto_scene* s, float* dst, int tex, const hit_t* h) {
    if (tex < 0) return;
    *dst = texture_at(s, tex, h->tex_u, h->tex_v, h->pos).x;
}

static void set_face(v3 d, v3 outward, hit_t* h) {
    h->front = dot3(d, outward) < 0.0f;
    h->gn = outward;
    h->sn = h->front ? outward : neg3(outward);
}

/* a hit record's primitive id: >= 0 a surface, -1 none, OM_MEDIUM_HIT - i
 * (<= -2) a scatter event inside medium i */
#define OM_MEDIUM_HIT (-2)

static void make_hit(const rto_scene* s, v3 o, v3 d, float t, int prim, hit_t* h) {
    h->t = t;
    h->prim = prim;
    int mat;
    if (prim <= OM_MEDIUM_HIT) {
        /* the medium candidate of intersect_scene (programs.cu:1736-1748) */
        const rto_medium* md = &s->media[OM_MEDIUM_HIT - prim];
        h->prim = -1;
        h->prim_type = -1; h->prim_index = -1;
        h->pos = add3(o, scl3(d, t));
        h->front = 1;
        h->gn = scl3(norm3(d), -1.0f);
        h->sn = h->gn;
        mat = md->material_index;
    } else if (prim < s->sphere_count) {
        const rto_sphere* sp = &s->spheres[prim];
        h->prim_type = 0; h->prim_index = prim;
        h->pos = add3(o, scl3(d, t));
        set_face(d, div3(sub3(h->pos, ld3(sp->center)), sp->radius), h);
        mat = sp->material;
    } else if (prim < s->sphere_count + s->quad_count) {
        const int qi = prim - s->sphere_count;
        const rto_quad* q = &s->quads[qi];
        h->prim_type = 1; h->prim_index = qi;
        const v3 n = cross3(ld3(q->edge_u), ld3(q->edge_v));
        const v3 normal = div3(n, sqrtf(lensq3(n)));
        h->pos = add3(o, scl3(d, t));
        set_face(d, normal, h);
        mat = q->material;
    } else {
        const int ti = prim - s->sphere_count - s->quad_count;
        const rto_triangle* tr = &s->triangles[ti];
        h->prim_type = 2; h->prim_index = ti;
        const v3 p0 = ld3(tr->p0);
        const v3 n = norm3(cross3(sub3(ld3(tr->p1), p0), sub3(ld3(tr->p2), p0)));
        h->pos = add3(o, scl3(d, t));
        set_face(d, n, h);
        mat = tr->material;
        if (s->triangle_normals != NULL && s->triangle_normals[9 * (int64_t)ti] == s->triangle_normals[9 * (int64_t)ti]) {
            /* interpolated vertex normals, programs.cu:1535-1545 */
            const float* vn = s->triangle_normals + 9 * (int64_t)ti;
            const v3 e1 = sub3(ld3(tr->p1), p0), e2 = sub3(ld3(tr->p2), p0);
            const v3 pvec = cross3(d, e2);
            const float inv_det = 1.0f / dot3(e1, pvec);
            const v3 tvec = sub3(o, p0);
            const float bu = dot3(tvec, pvec) * inv_det;
            const float bv = dot3(d, cross3(tvec, e1)) * inv_det;
            const float bw = 1.0f - bu - bv;
            v3 sh = add3(scl3(ld3(vn), bw), add3(scl3(ld3(vn + 3), bu), scl3(ld3(vn + 6), bv)));
            if (!(fabsf(sh.x) < 1e-7f && fabsf(sh.y) < 1e-7f && fabsf(sh.z) < 1e-7f)) {
                sh = norm3(sh);
                if (dot3(sh, n) < 0.0f) sh = scl3(sh, -1.0f);
                h->sn = h->front ? sh : scl3(sh, -1.0f);
            }
        }
    }
    /* the intersectors' texture coordinates (programs.cu:1447,1485,1524-1530);
     * 0 for a medium event */
    h->tex_u = 0.0f;
    h->tex_v = 0.0f;
    if (h->prim_type >= 0) {
        const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
        const float pf[3] = {h->pos.x, h->pos.y, h->pos.z};
        rto_surface_uv(s, of, df, h->prim_type, h->prim_index, pf, &h->tex_u, &h->tex_v);
    }
    /* populate_material, programs.cu:961-1009 */
    if (mat < 0 || mat >= s->material_count) {
        h->base = mk3(0, 0, 0); h->emission = mk3(0, 0, 0); h->mat_type = 0; h->ior = 1.0f;
        h->fuzz = 0.0f; h->openpbr = -1;
        return;
    }
    const rto_material* m = &s->materials[mat];
    h->openpbr = m->openpbr_index;
    h->mat_type = m->type;
    h->fuzz = m->fuzz;
    h->ior = m->ior;
    /* the albedo and emission textures (programs.cu:978-981); -1 keeps the constant */
    h->base = ld3(m->albedo);
    if (s->material_albedo_texture != NULL && s->material_albedo_texture[mat] >= 0)
        h->base = texture_at(s, s->material_albedo_texture[mat], h->tex_u, h->tex_v, h->pos);
    h->emission = ld3(m->emission);
    if (s->material_emission_texture != NULL && s->material_emission_texture[mat] >= 0)
        h->emission = texture_at(s, s->material_emission_texture[mat], h->tex_u, h->tex_v, h->pos);
    if (h->mat_type == RTO_MAT_DIELECTRIC) { h->base = mk3(0.92f, 0.95f, 1.0f); h->emission = mk3(0, 0, 0); }
    if (h->mat_type == RTO_MAT_DIFFUSE_LIGHT && !h->front) h->emission = mk3(0, 0, 0);
    if (h->mat_type != RTO_MAT_DIFFUSE_LIGHT) h->emission = mk3(0, 0, 0);
    if (h->mat_type == RTO_MAT_OPENPBR) {
        if (m->openpbr_index < 0 || m->openpbr_index >= s->openpbr_count) {
            h->mat_type = -1; h->base = mk3(0, 0, 0); h->emission = mk3(0, 0, 0);
            return;
        }
        /* the base and emission colour bindings (programs.cu:1001-1008) */
        rto_openpbr p = s->openpbr[m->openpbr_index];
        if (s->openpbr_textures != NULL) {
            const rto_openpbr_textures* tb = &s->openpbr_textures[m->openpbr_index];
            apply_color_binding(s, p.base_color, tb->base_color, tb->base_color_space, h);
            apply_color_binding(s, p.emission_color, tb->emission_color, tb->emission_color_space, h);
        }
        h->base = ld3(p.base_color);
        h->emission = om_openpbr_emission(&p);
    }
}

/* evaluate_openpbr_scattering_material (programs.cu:944-959): the record
 * with the hit's base colour and the bound specular / transmission colours,
 * metalness and roughness at the hit */
static rto_openpbr scattering_material(const rto_scene* s, const hit_t* h) {
    rto_openpbr p = s->openpbr[h->openpbr];
    if (s->openpbr_textures != NULL) {
        const rto_openpbr_textures* tb = &s->openpbr_textures[h->openpbr];
        if (tb->base_color >= 0) {
            p.base_color[0] = h->base.x;
            p.base_color[1] = h->base.y;
            p.base_color[2] = h->base.z;
        }
        apply_color_binding(s, p.specular_color, tb->specular_color, tb->specular_color_space, h);
        apply_color_binding(s, p.transmission_color, tb->transmission_color, tb->transmission_color_space, h);
        apply_scalar_binding(s, &p.base_metalness, tb->base_metalness, h);
        apply_scalar_binding(s, &p.specular_roughness, tb->specular_roughness, h);
    }
    return p;
}

/* ------------------------------------------------------------- path state */
typedef struct {
    v3 o, d, throughput, radiance;
    v3 prev_pos, prev_n;
    float prev_pdf;
    int prev_mat, prev_prim, prev_valid, prev_delta, alive;
    int inside;   /* OpenPBR index of the subsurface medium the path is in, -1 none */
    rto_openpbr inside_mat;   /* that material as it was where the walk began
                                 (PathState::subsurface_medium, programs.cu:2545-2548) */
} path_t;

static float power_heuristic(float a, float b) {
    if (a <= 0.0f) return 0.0f;
    const float a2 = a * a, b2 = b * b;
    return a2 / (a2 + b2);
}
static float cone_pdf(float cos_max) {
    const float sa = 6.28318530717958648f * (1.0f - cos_max);
    return sa > 1e-12f ? 1.0f / sa : 0.0f;
}
static float area_pdf(float area, float dsq, float acos_) {
    if (area <= 1e-12f || dsq <= 0.0f || acos_ <= 1e-8f) return 0.0f;
    return dsq / (area * acos_);
}
#define UNIFORM_SPHERE_PDF 0.07957747154594767f

/* programs.cu:571-582 */
static float eff_selection_pdf(const ctx_t* c, int light, int excluded_prim) {
    const int ex = excluded_prim < 0 ? -1 : c->light_of_prim[excluded_prim];
    if (ex == light) return 0.0f;
    const float elig = ex < 0 ? 1.0f : 1.0f - c->lights[ex].selection_pdf;
    return elig > 1e-8f ? c->lights[light].selection_pdf / elig : 0.0f;
}

/* conditional_light_pdf_for_hit (programs.cu:584-619) times the effective
 * selection pdf: the solid-angle density with which next-event estimation
 * from (prev_pos, prev_n, prev_mat, prev_prim) picks the point the ray in
 * direction `dir` hit (h; NULL with env for a miss) */
static float light_pdf_for_hit(const ctx_t* c, v3 prev_pos, v3 prev_n, int prev_mat, int prev_prim,
    v3 dir, const hit_t* h, int env, int* light_out) {
    const int light = env ? c->env_light : c->light_of_prim[h->prim];
    *light_out = light;
    if (light < 0) return 0.0f;
    float cond;
    if (env) {
        cond = prev_mat == RTO_MAT_ISOTROPIC ? UNIFORM_SPHERE_PDF
                                             : fmaxf(dot3(prev_n, dir), 0.0f) / OM_PI;
    } else {
        const rto_scene* s = c->s;
        if (!h->front) cond = 0.0f;
        else if (h->prim_type == 0) {
            const rto_sphere* sp = &s->spheres[h->prim_index];
            const float dsq = lensq3(sub3(ld3(sp->center), prev_pos));
            const float rsq = sp->radius * sp->radius;
            cond = dsq <= rsq ? 0.0f : cone_pdf(sqrtf(fmaxf(0.0f, 1.0f - rsq / dsq)));
        } else {
            const float dsq = lensq3(sub3(h->pos, prev_pos));
            const float ac = fmaxf(dot3(h->gn, scl3(dir, -1.0f)), 0.0f);
            float area;
            if (h->prim_type == 1) {
                const rto_quad* q = &s->quads[h->prim_index];
                area = len3(cross3(ld3(q->edge_u), ld3(q->edge_v)));
            } else {
                const rto_triangle* t = &s->triangles[h->prim_index];
                area = 0.5f * len3(cross3(sub3(ld3(t->p1), ld3(t->p0)), sub3(ld3(t->p2), ld3(t->p0))));
            }
            cond = area_pdf(area, dsq, ac);
        }
    }
    const float sel = eff_selection_pdf(c, light, prev_prim);
    return sel * cond;
}

/* emission_mis_weight (programs.cu:621-647) */
static float emission_mis(const ctx_t* c, const path_t* st, const hit_t* h, int env) {
    if (!st->prev_valid || st->prev_delta) return 1.0f;
    if (!env && h->prim < 0) return 1.0f;   /* a medium event: no primitive, no light */
    int light;
    const float lpdf = light_pdf_for_hit(c, st->prev_pos, st->prev_n, st->prev_mat, st->prev_prim,
        norm3(st->d), h, env, &light);
    if (light < 0) return 1.0f;
    return power_heuristic(st->prev_pdf, lpdf);
}

typedef struct { v3 dir, emission; float dist, pdf; int infinite, valid, delta; } light_sample_t;

/* programs.cu:539-550 with the own sin/cos */
static v3 cone_dir(v3 axis, float cos_max, float u0, float u1) {
    const float ct = 1.0f - u0 * (1.0f - cos_max);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    om_sincos_turn(u1, &sp, &cp);
    const v3 helper = fabsf(axis.x) > 0.9f ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f);
    const v3 tangent = norm3(cross3(helper, axis));
    const v3 bitangent = cross3(axis, tangent);
    return norm3(add3(scl3(axis, ct), add3(scl3(tangent, st * cp), scl3(bitangent, st * sp))));
}

/* sample_packed_light / sample_packed_analytic_light (light_sampling.h:87-116):
 * the first entry whose CDF exceeds u clamped to [0, 1), else the last;
 * cdf values are `stride` floats apart */
static int pick_cdf(const float* cdf, int stride, int n, float u) {
    if (cdf == NULL || n <= 0) return -1;
    const float s = u < 0.0f ? 0.0f : (u < 1.0f ? u : 0.99999994f);
    for (int i = 0; i < n; ++i)
        if (s < cdf[(size_t)i * (size_t)stride]) return i;
    return n - 1;
}

/* sample_uniform_triangle (light_sampling.h:118-125) */
static void triangle_barycentrics(float u0, float u1, float* w0, float* w1, float* w2) {
    const float su = sqrtf(u0 < 0.0f ? 0.0f : (u0 < 1.0f ? u0 : 1.0f));
    *w0 = 1.0f - su;
    *w1 = su * (1.0f - u1);
    *w2 = su * u1;
}

/* programs.cu:1057-1232 */
static light_sample_t sample_light(const ctx_t* c, const hit_t* h, const float r[4]) {
    light_sample_t ls;
    memset(&ls, 0, sizeof(ls));
    const rto_scene* s = c->s;
    const int ex = h->prim >= 0 ? c->light_of_prim[h->prim] : -1;
    const float elig = ex < 0 ? 1.0f : 1.0f - c->lights[ex].selection_pdf;
    if (elig <= 1e-8f) return ls;
    int li = -1;
    if (ex < 0) {
        li = pick_cdf(&c->lights[0].cdf, (int)(sizeof(rto_light) / sizeof(float)), c->light_count,
            r[0]);
    } else {
        const float target = r[0] * elig;
        float acc = 0.0f;
        for (int i = 0; i < c->light_count; ++i) {
            if (i == ex) continue;
            acc += c->lights[i].selection_pdf;
            if (target < acc) { li = i; break; }
        }
        if (li < 0)
            for (int i = c->light_count - 1; i >= 0; --i)
                if (i != ex) { li = i; break; }
    }
    if (li < 0) return ls;
    const rto_light* L = &c->lights[li];
    const float sel = L->selection_pdf / elig;
    if (sel <= 0.0f) return ls;
    const v3 sp = h->pos;
    if (L->type == 3) {
        if (h->mat_type == RTO_MAT_ISOTROPIC) {
            ls.dir = om_sphere_dir(r[1], r[2]);
            ls.pdf = sel * UNIFORM_SPHERE_PDF;
        } else {
            ls.dir = add3(h->sn, om_sphere_dir(r[1], r[2]));
            if (near_zero3(ls.dir)) ls.dir = h->sn;
            ls.dir = norm3(ls.dir);
            const float cs = fmaxf(dot3(h->sn, ls.dir), 0.0f);
            ls.pdf = sel * cs / OM_PI;
        }
        ls.emission = background_at(s, ls.dir);
        ls.dist = OM_RAY_FAR;
        ls.infinite = 1;
        ls.valid = ls.pdf > 0.0f;
        return ls;
    }
    float cond = 0.0f;
    int mat = -1;
    /* the sampled point and its texture coordinates (programs.cu:1154,1166-1167,1193-1199) */
    v3 lpos = mk3(0, 0, 0);
    float tu = 0.0f, tv = 0.0f;
    if (L->type == 0) {
        const rto_sphere* sph = &s->spheres[L->prim];
        const v3 center = ld3(sph->center);
        const v3 tc = sub3(center, sp);
        const float dsq = lensq3(tc);
        const float rsq = sph->radius * sph->radius;
        if (dsq <= rsq * (1.0f + 1e-6f)) return ls;
        const float cmax = sqrtf(fmaxf(0.0f, 1.0f - rsq / dsq));
        ls.dir = cone_dir(norm3(tc), cmax, r[1], r[2]);
        const v3 oc = sub3(sp, center);
        const float hb = dot3(oc, ls.dir);
        const float disc = hb * hb - (lensq3(oc) - rsq);
        if (disc <= 0.0f) return ls;
        ls.dist = -hb - sqrtf(disc);
        if (ls.dist <= OM_RAY_EPS) return ls;
        mat = sph->material;
        cond = cone_pdf(cmax);
        lpos = add3(sp, scl3(ls.dir, ls.dist));
        const v3 n = norm3(sub3(lpos, center));
        /* sphere_uv (programs.cu:188-193) */
        tu = (atan2f(-n.z, n.x) + OM_PI) / (2.0f * OM_PI);
        tv = acosf(fminf(fmaxf(-n.y, -1.0f), 1.0f)) / OM_PI;
    } else if (L->type == 1) {
        const rto_quad* q = &s->quads[L->prim];
        const v3 eu = ld3(q->edge_u), ev = ld3(q->edge_v);
        const v3 pos = add3(ld3(q->origin), add3(scl3(eu, r[1]), scl3(ev, r[2])));
        const v3 an = cross3(eu, ev);
        const float area = len3(an);
        if (area <= 1e-12f) return ls;
        const v3 normal = div3(an, area);
        mat = q->material;
        lpos = pos;
        tu = r[1];
        tv = r[2];
        const v3 tl = sub3(pos, sp);
        const float dsq = lensq3(tl);
        if (dsq <= 1e-12f) { memset(&ls, 0, sizeof(ls)); return ls; }
        ls.dist = sqrtf(dsq);
        ls.dir = div3(tl, ls.dist);
        cond = area_pdf(area, dsq, dot3(normal, scl3(ls.dir, -1.0f)));
    } else {
        const rto_triangle* t = &s->triangles[L->prim];
        const v3 p0 = ld3(t->p0), p1 = ld3(t->p1), p2 = ld3(t->p2);
        float w0, w1, w2;
        triangle_barycentrics(r[1], r[2], &w0, &w1, &w2);
        const v3 pos = add3(scl3(p0, w0), add3(scl3(p1, w1), scl3(p2, w2)));
        const v3 an = cross3(sub3(p1, p0), sub3(p2, p0));
        const float da = len3(an);
        if (da <= 1e-12f) return ls;
        const v3 normal = div3(an, da);
        mat = t->material;
        lpos = pos;
        tu = w1;
        tv = w2;
        if (s->triangle_uvs != NULL && s->triangle_uvs[6 * (int64_t)L->prim] == s->triangle_uvs[6 * (int64_t)L->prim]) {
            const float* uv = s->triangle_uvs + 6 * (int64_t)L->prim;
            tu = uv[0] * w0 + uv[2] * w1 + uv[4] * w2;
            tv = uv[1] * w0 + uv[3] * w1 + uv[5] * w2;
        }
        const v3 tl = sub3(pos, sp);
        const float dsq = lensq3(tl);
        if (dsq <= 1e-12f) { memset(&ls, 0, sizeof(ls)); return ls; }
        ls.dist = sqrtf(dsq);
        ls.dir = div3(tl, ls.dist);
        cond = area_pdf(0.5f * da, dsq, dot3(normal, scl3(ls.dir, -1.0f)));
    }
    if (cond <= 0.0f || mat < 0 || mat >= s->material_count) { memset(&ls, 0, sizeof(ls)); return ls; }
    const rto_material* m = &s->materials[mat];
    /* evaluate_material_emission (programs.cu:1013-1024) at the sampled point */
    if (m->type == RTO_MAT_OPENPBR) {
        if (m->openpbr_index >= 0 && m->openpbr_index < s->openpbr_count) {
            rto_openpbr p = s->openpbr[m->openpbr_index];
            if (s->openpbr_textures != NULL && s->openpbr_textures[m->openpbr_index].emission_color >= 0) {
                const rto_openpbr_textures* tb = &s->openpbr_textures[m->openpbr_index];
                const v3 c = source_to_linear(texture_at(s, tb->emission_color, tu, tv, lpos),
                    tb->emission_color_space);
                p.emission_color[0] = c.x;
                p.emission_color[1] = c.y;
                p.emission_color[2] = c.z;
            }
            ls.emission = om_openpbr_emission(&p);
        } else {
            ls.emission = mk3(0, 0, 0);
        }
    } else if (s->material_emission_texture != NULL && s->material_emission_texture[mat] >= 0) {
        ls.emission = texture_at(s, s->material_emission_texture[mat], tu, tv, lpos);
    } else {
        ls.emission = ld3(m->emission);
    }
    ls.pdf = sel * cond;
    ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
    return ls;
}

/* programs.cu:1026-1055 */
static v3 direct_response(const ctx_t* c, const hit_t* h, const path_t* st, v3 ldir, float* pdf) {
    *pdf = 0.0f;
    if (h->mat_type == RTO_MAT_LAMBERTIAN) {
        const float cs = fmaxf(dot3(h->sn, ldir), 0.0f);
        *pdf = cs / OM_PI;
        return scl3(h->base, *pdf);
    }
    if (h->mat_type == RTO_MAT_ISOTROPIC) {
        *pdf = UNIFORM_SPHERE_PDF;
        return scl3(h->base, *pdf);
    }
    if (h->mat_type != RTO_MAT_OPENPBR) return mk3(0, 0, 0);
    const rto_openpbr p = scattering_material(c->s, h);
    v3 value;
    om_openpbr_eval(&p, h->gn, neg3(norm3(st->d)), ldir, st->throughput, &value, pdf);
    const float cs = fabsf(dot3(h->sn, ldir));
    return scl3(value, cs);
}

/* programs.cu:2249-2280 */
static void direct_light(const ctx_t* c, const hit_t* h, const float r[4], int bsdf_avail, path_t* st) {
    if (h->mat_type == RTO_MAT_DIFFUSE_LIGHT || c->light_count <= 0) return;
    const light_sample_t L = sample_light(c, h, r);
    if (!L.valid) return;
    float bpdf = 0.0f;
    const v3 resp = direct_response(c, h, st, L.dir, &bpdf);
    if (maxc3(resp) <= 0.0f) return;
    const v3 so = add3(h->pos, scl3(L.dir, OM_RAY_EPS * 2.0f));
    const float max_t = L.infinite ? OM_RAY_FAR : L.dist - OM_RAY_EPS * 3.0f;
    if (max_t <= OM_RAY_EPS || om_occluded(c->s, c->bvh, so, L.dir, OM_RAY_EPS, max_t)) return;
    const float mis = bsdf_avail ? power_heuristic(L.pdf, bpdf) : 1.0f;
    const v3 direct = scl3(mul3(L.emission, resp), mis / L.pdf);
    st->radiance = add3(st->radiance, mul3(st->throughput, direct));
}

/* ------------------------------------------------------ analytic lights */
static int alight_finite(int type) {
    return type == RTO_ALIGHT_SPHERE || type == RTO_ALIGHT_DISK || type == RTO_ALIGHT_RECT
           || type == RTO_ALIGHT_CYLINDER;
}

/* analytic_light_normal (programs.cu:655-659) */
static v3 alight_normal(const rto_analytic_light* L) {
    return norm3(scl3(cross3(ld3(L->basis_x), ld3(L->basis_y)), -1.0f));
}

/* analytic_plane_coordinates (programs.cu:661-677) */
static int alight_plane_xy(const rto_analytic_light* L, v3 p, float* lx, float* ly) {
    const v3 bx = ld3(L->basis_x), by = ld3(L->basis_y);
    const v3 local = sub3(p, ld3(L->position));
    const float xx = dot3(bx, bx), xy = dot3(bx, by), yy = dot3(by, by);
    const float det = xx * yy - xy * xy;
    if (det <= 1e-16f) return 0;
    const float px = dot3(local, bx), py = dot3(local, by);
    *lx = (px * yy - py * xy) / det;
    *ly = (py * xx - px * xy) / det;
    return 1;
}

/* try_intersect_analytic_light (programs.cu:679-793): sphere (radius scaled
 * by |basis_x|), disk / rect on the plane through the position with normal
 * -(basis_x x basis_y), and the cylinder's side; front faces only, the
 * nearer root in (tmin, tmax), else the farther */
static int alight_hit(const rto_analytic_light* L, v3 o, v3 d, float tmin, float tmax, float* t,
    v3* pos, v3* n) {
    const v3 center = ld3(L->position);
    if (L->type == RTO_ALIGHT_SPHERE) {
        const float radius = L->radius * len3(ld3(L->basis_x));
        const v3 oc = sub3(o, center);
        const float a = lensq3(d);
        const float hb = dot3(oc, d);
        const float c = lensq3(oc) - radius * radius;
        const float disc = hb * hb - a * c;
        if (a <= 1e-16f || disc < 0.0f) return 0;
        const float rt = sqrtf(disc);
        float root = (-hb - rt) / a;
        if (root <= tmin || root >= tmax) root = (-hb + rt) / a;
        if (root <= tmin || root >= tmax) return 0;
        const v3 p = add3(o, scl3(d, root));
        const v3 nn = norm3(sub3(p, center));
        if (dot3(d, nn) >= 0.0f) return 0;
        *t = root;
        *pos = p;
        *n = nn;
        return 1;
    }
    if (L->type == RTO_ALIGHT_DISK || L->type == RTO_ALIGHT_RECT) {
        const v3 nn = alight_normal(L);
        const float den = dot3(d, nn);
        if (den >= -1e-8f) return 0;
        const float tt = dot3(sub3(center, o), nn) / den;
        if (tt <= tmin || tt >= tmax) return 0;
        const v3 p = add3(o, scl3(d, tt));
        float lx = 0.0f, ly = 0.0f;
        if (!alight_plane_xy(L, p, &lx, &ly)) return 0;
        const int inside = L->type == RTO_ALIGHT_DISK
                               ? lx * lx + ly * ly <= L->radius * L->radius
                               : fabsf(lx) <= 0.5f * L->width && fabsf(ly) <= 0.5f * L->height;
        if (!inside) return 0;
        *t = tt;
        *pos = p;
        *n = nn;
        return 1;
    }
    if (L->type != RTO_ALIGHT_CYLINDER) return 0;
    const v3 bx = ld3(L->basis_x), by = ld3(L->basis_y), bz = ld3(L->basis_z);
    const float scale = len3(bx);
    if (scale <= 1e-8f) return 0;
    const v3 ax = div3(bx, scale), ay = div3(by, scale), az = norm3(bz);
    const v3 lo = sub3(o, center);
    const float ox = dot3(lo, ax), oy = dot3(lo, ay), oz = dot3(lo, az);
    const float dx = dot3(d, ax), dy = dot3(d, ay), dz = dot3(d, az);
    const float radius = L->radius * scale;
    const float a = dx * dx + dy * dy;
    const float hb = ox * dx + oy * dy;
    const float c = ox * ox + oy * oy - radius * radius;
    const float disc = hb * hb - a * c;
    if (a <= 1e-16f || disc < 0.0f) return 0;
    const float rt = sqrtf(disc);
    float root = (-hb - rt) / a;
    float z = oz + root * dz;
    const float half_len = 0.5f * L->length * len3(bz);
    if (root <= tmin || root >= tmax || fabsf(z) > half_len) {
        root = (-hb + rt) / a;
        z = oz + root * dz;
    }
    if (root <= tmin || root >= tmax || fabsf(z) > half_len) return 0;
    const float x = ox + root * dx, y = oy + root * dy;
    const v3 nn = norm3(add3(scl3(ax, x), scl3(ay, y)));
    if (dot3(d, nn) >= 0.0f) return 0;
    *t = root;
    *pos = add3(o, scl3(d, root));
    *n = nn;
    return 1;
}

/* intersect_analytic_lights (programs.cu:795-813): finite lights that are
 * not points */
static int intersect_analytic(const ctx_t* c, v3 o, v3 d, float tmin, float tmax, int* index) {
    float closest = tmax;
    int found = -1;
    for (int i = 0; i < c->a_count; ++i) {
        const rto_analytic_light* L = &c->s->analytic_lights[i];
        if (!alight_finite(L->type) || L->treat_as_point != 0) continue;
        float t;
        v3 p, n;
        if (!alight_hit(L, o, d, tmin, closest, &t, &p, &n)) continue;
        closest = t;
        found = i;
    }
    *index = found;
    return found >= 0;
}

/* analytic_light_pdf_for_hit (programs.cu:815-834) */
static float alight_pdf_for_hit(const rto_analytic_light* L, v3 origin, v3 dir, v3 hit_pos, v3 hit_n) {
    if (L->delta != 0 || L->treat_as_point != 0 || L->treat_as_line != 0) return 0.0f;
    if (L->type == RTO_ALIGHT_SPHERE) {
        const float radius = L->radius * len3(ld3(L->basis_x));
        const float dsq = lensq3(sub3(ld3(L->position), origin));
        const float rsq = radius * radius;
        if (dsq <= rsq) return 0.0f;
        return cone_pdf(sqrtf(fmaxf(0.0f, 1.0f - rsq / dsq)));
    }
    const float dsq = lensq3(sub3(hit_pos, origin));
    const float ac = fmaxf(dot3(hit_n, scl3(dir, -1.0f)), 0.0f);
    return area_pdf(L->world_area, dsq, ac);
}

/* analytic_emission_mis_weight (programs.cu:836-849), with the light's hit
 * along the path's ray (o, d) */
static float analytic_mis(const ctx_t* c, const path_t* st, int li) {
    if (!st->prev_valid || st->prev_delta || li < 0) return 1.0f;
    const rto_analytic_light* L = &c->s->analytic_lights[li];
    float t = 0.0f;
    v3 p = mk3(0, 0, 0), n = mk3(0, 0, 0);
    if (L->type != RTO_ALIGHT_SPHERE) (void)alight_hit(L, st->o, st->d, OM_RAY_EPS, OM_RAY_FAR, &t, &p, &n);
    const float lp = L->selection_pdf * alight_pdf_for_hit(L, st->prev_pos, norm3(st->d), p, n);
    return power_heuristic(st->prev_pdf, lp);
}

/* analytic_infinite_direction / _pdf / _radiance (programs.cu:851-890) */
static v3 analytic_infinite(const ctx_t* c, const path_t* st) {
    v3 sum = mk3(0, 0, 0);
    const v3 dir = norm3(st->d);
    for (int i = 0; i < c->a_count; ++i) {
        const rto_analytic_light* L = &c->s->analytic_lights[i];
        if (L->type == RTO_ALIGHT_DISTANT) {
            const v3 axis = norm3(ld3(L->basis_z));
            const float threshold = L->delta != 0 ? 0.999999f : L->cos_theta_max;
            if (!(dot3(axis, dir) >= threshold)) continue;
        } else if (L->type != RTO_ALIGHT_DOME) {
            continue;
        }
        float w = 1.0f;
        if (st->prev_valid && !st->prev_delta) {
            const float cond = L->delta != 0 ? 0.0f
                               : (L->type == RTO_ALIGHT_DOME ? UNIFORM_SPHERE_PDF : cone_pdf(L->cos_theta_max));
            if (cond <= 0.0f) continue;
            w = power_heuristic(st->prev_pdf, L->selection_pdf * cond);
        }
        sum = add3(sum, scl3(ld3(L->radiance), w));
    }
    return sum;
}

static int has_infinite_light(const ctx_t* c) {
    for (int i = 0; i < c->a_count; ++i)
        if (!alight_finite(c->s->analytic_lights[i].type)) return 1;
    return 0;
}

/* sample_packed_analytic_light (light_sampling.h:104-116) */
static int pick_analytic(const ctx_t* c, float u) {
    if (c->a_count <= 0) return -1;
    return pick_cdf(c->a_cdf, 1, c->a_count, u);
}

/* sample_analytic_direct_light_candidate (programs.cu:1234-1346) */
static light_sample_t analytic_candidate(const ctx_t* c, v3 sp, int li, float u0, float u1) {
    light_sample_t ls;
    memset(&ls, 0, sizeof(ls));
    if (li < 0 || li >= c->a_count) return ls;
    const rto_analytic_light* L = &c->s->analytic_lights[li];
    if (L->selection_pdf <= 0.0f) return ls;
    const v3 emission = ld3(L->radiance);
    if (L->type == RTO_ALIGHT_DOME) {
        ls.dir = om_sphere_dir(u0, u1);
        ls.dist = OM_RAY_FAR;
        ls.emission = emission;
        ls.pdf = L->selection_pdf * UNIFORM_SPHERE_PDF;
        ls.infinite = 1;
        ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
        return ls;
    }
    if (L->type == RTO_ALIGHT_DISTANT) {
        const v3 axis = norm3(ld3(L->basis_z));
        ls.dir = L->delta != 0 ? axis : cone_dir(axis, L->cos_theta_max, u0, u1);
        ls.dist = OM_RAY_FAR;
        ls.emission = emission;
        ls.pdf = L->selection_pdf * (L->delta != 0 ? 1.0f : cone_pdf(L->cos_theta_max));
        ls.infinite = 1;
        ls.delta = L->delta != 0;
        ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
        return ls;
    }
    const v3 center = ld3(L->position);
    if (L->type == RTO_ALIGHT_SPHERE) {
        const v3 tc = sub3(center, sp);
        const float dsq = lensq3(tc);
        const float radius = L->radius * len3(ld3(L->basis_x));
        const float rsq = radius * radius;
        if (dsq <= rsq * (1.0f + 1e-6f)) return ls;
        if (L->treat_as_point != 0) {
            ls.dist = sqrtf(dsq);
            ls.dir = div3(tc, ls.dist);
            const float scale = L->world_area > 1e-12f ? L->world_area : 1.0f;
            ls.emission = scl3(emission, scale / dsq);
            ls.pdf = L->selection_pdf;
            ls.delta = 1;
            ls.valid = maxc3(ls.emission) > 0.0f;
            return ls;
        }
        const float cmax = sqrtf(fmaxf(0.0f, 1.0f - rsq / dsq));
        const v3 dir = cone_dir(norm3(tc), cmax, u0, u1);
        float t;
        v3 p, n;
        if (!alight_hit(L, sp, dir, OM_RAY_EPS, OM_RAY_FAR, &t, &p, &n)) return ls;
        ls.dir = dir;
        ls.emission = emission;
        ls.dist = t;
        ls.pdf = L->selection_pdf * cone_pdf(cmax);
        ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
        return ls;
    }
    v3 pos, nrm;
    if (L->type == RTO_ALIGHT_DISK) {
        const float radial = L->radius * sqrtf(u0);
        float sphi, cphi;
        om_sincos_turn(u1, &sphi, &cphi);
        pos = add3(center, add3(scl3(ld3(L->basis_x), radial * cphi), scl3(ld3(L->basis_y), radial * sphi)));
        nrm = alight_normal(L);
    } else if (L->type == RTO_ALIGHT_RECT) {
        pos = add3(center, add3(scl3(ld3(L->basis_x), (u0 - 0.5f) * L->width),
                                scl3(ld3(L->basis_y), (u1 - 0.5f) * L->height)));
        nrm = alight_normal(L);
    } else {
        float sphi, cphi;
        om_sincos_turn(u0, &sphi, &cphi);
        const float rx = L->radius * cphi, ry = L->radius * sphi;
        pos = add3(center, add3(add3(scl3(ld3(L->basis_x), rx), scl3(ld3(L->basis_y), ry)),
                                scl3(ld3(L->basis_z), (u1 - 0.5f) * L->length)));
        nrm = norm3(add3(scl3(ld3(L->basis_x), rx), scl3(ld3(L->basis_y), ry)));
    }
    const v3 tl = sub3(pos, sp);
    const float dsq = lensq3(tl);
    if (dsq <= 1e-12f) return ls;
    ls.dist = sqrtf(dsq);
    ls.dir = div3(tl, ls.dist);
    const float ac = dot3(nrm, scl3(ls.dir, -1.0f));
    ls.emission = emission;
    ls.pdf = L->selection_pdf * area_pdf(L->world_area, dsq, ac);
    ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
    return ls;
}

/* ------------------------------------------------------------ media */
/* hit_medium_boundary (programs.cu:1356-1419) */
static int medium_boundary(const rto_medium* m, v3 o, v3 d, float tmin, float tmax, float* entry,
    float* exit_) {
    const v3 r0 = ld3(m->rotation_row0), r1 = ld3(m->rotation_row1), r2 = ld3(m->rotation_row2);
    const v3 sh = sub3(o, ld3(m->translation));
    const v3 lo = mk3(dot3(r0, sh), dot3(r1, sh), dot3(r2, sh));
    const v3 ldir = mk3(dot3(r0, d), dot3(r1, d), dot3(r2, d));
    if (m->boundary_type == RTO_MEDIUM_SPHERE) {
        const v3 oc = sub3(lo, ld3(m->local_center_or_min));
        const float a = lensq3(ldir);
        const float hb = dot3(oc, ldir);
        const float c = lensq3(oc) - m->radius * m->radius;
        const float disc = hb * hb - a * c;
        if (disc <= 0.0f) return 0;
        const float sq = sqrtf(disc);
        float t0 = (-hb - sq) / a, t1 = (-hb + sq) / a;
        if (t0 > t1) { const float tt = t0; t0 = t1; t1 = tt; }
        *entry = fmaxf(t0, tmin);
        *exit_ = fminf(t1, tmax);
        return *entry < *exit_;
    }
    if (m->boundary_type != RTO_MEDIUM_BOX) return 0;
    float t0 = tmin, t1 = tmax;
    const float org[3] = {lo.x, lo.y, lo.z}, dir[3] = {ldir.x, ldir.y, ldir.z};
    for (int a = 0; a < 3; ++a) {
        if (fabsf(dir[a]) < 1e-8f) {
            if (org[a] < m->local_center_or_min[a] || org[a] > m->local_max[a]) return 0;
            continue;
        }
        float a0 = (m->local_center_or_min[a] - org[a]) / dir[a];
        float a1 = (m->local_max[a] - org[a]) / dir[a];
        if (a0 > a1) { const float tt = a0; a0 = a1; a1 = tt; }
        t0 = fmaxf(t0, a0);
        t1 = fminf(t1, a1);
        if (t0 >= t1) return 0;
    }
    *entry = t0;
    *exit_ = t1;
    return 1;
}

/* the media part of intersect_scene (programs.cu:1711-1753): medium i draws
 * its uniform from the path's group 48 + i / 4, component i % 4; returns the
 * scattering medium (*closest moved to the event) or -1 */
static int sample_media(const rto_scene* s, v3 o, v3 d, float tmin, float* closest, uint32_t pixel,
    uint32_t sample, uint32_t cam, uint32_t bounce, uint32_t stream) {
    const float ray_len = len3(d);
    if (!(ray_len > 1e-8f)) return -1;
    int event = -1;
    float u4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = 0; i < s->medium_count; ++i) {
        if ((i & 3) == 0) om_rng4(pixel, sample, cam, bounce, OM_G_MEDIUM0 + (uint32_t)(i >> 2), stream, u4);
        const rto_medium* m = &s->media[i];
        if (m->density <= 0.0f) continue;
        float entry = 0.0f, exit_ = 0.0f;
        if (!medium_boundary(m, o, d, tmin, *closest, &entry, &exit_)) continue;
        entry = fmaxf(entry, tmin);
        exit_ = fminf(exit_, *closest);
        if (entry >= exit_) continue;
        const float inside = (exit_ - entry) * ray_len;
        const float dist = -om_log_fixed(fmaxf(u4[i & 3], 1e-6f)) / m->density;
        if (dist >= inside) continue;
        *closest = entry + dist / ray_len;
        event = i;
    }
    return event;
}

float rto_log_fixed(float x) { return om_log_fixed(x); }

/* accumulate_analytic_direct_light (programs.cu:2282-2313) */
static void analytic_direct_light(const ctx_t* c, const hit_t* h, const float r[4], int bsdf_avail,
    path_t* st) {
    if (h->mat_type == RTO_MAT_DIFFUSE_LIGHT || c->a_count <= 0) return;
    const int li = pick_analytic(c, r[0]);
    const light_sample_t L = analytic_candidate(c, h->pos, li, r[1], r[2]);
    if (!L.valid) return;
    float bpdf = 0.0f;
    const v3 resp = direct_response(c, h, st, L.dir, &bpdf);
    if (maxc3(resp) <= 0.0f) return;
    const v3 so = add3(h->pos, scl3(L.dir, OM_RAY_EPS * 2.0f));
    const float max_t = L.infinite ? OM_RAY_FAR : L.dist - OM_RAY_EPS * 3.0f;
    if (max_t <= OM_RAY_EPS || om_occluded(c->s, c->bvh, so, L.dir, OM_RAY_EPS, max_t)) return;
    const float mis = L.delta || !bsdf_avail ? 1.0f : power_heuristic(L.pdf, bpdf);
    const v3 direct = scl3(mul3(L.emission, resp), mis / L.pdf);
    st->radiance = add3(st->radiance, mul3(st->throughput, direct));
}

/* --------------------------------------------------------------- ReSTIR DI */
/* reservoir math: src/common/restir_di.h:54-199 */
static int rs_finite_pos(float v) { return v > 0.0f && isfinite(v); }
static float rs_clamp_unit(float v) { return v < 0.0f ? 0.0f : (v < 1.0f ? v : 0.99999994f); }

static void rs_init(rto_reservoir* r) {
    memset(r, 0, sizeof(*r));
    r->light_index = -1;
    r->material_type = -1;
    r->primitive_type = -1;
    r->primitive_index = -1;
}

int rto_restir_update(rto_reservoir* r, int light, float u0, float u1, float target, float weight,
    int mult, float u) {
    if (light < 0 || mult <= 0) return 0;
    r->candidate_count += mult;
    if (!rs_finite_pos(target) || !rs_finite_pos(weight)) return 0;
    const float next = r->weight_sum + weight;
    if (!rs_finite_pos(next)) return 0;
    const int sel = rs_clamp_unit(u) * next < weight;
    if (sel) {
        r->light_index = light;
        r->sample_u0 = u0;
        r->sample_u1 = u1;
        r->selected_target = target;
        r->valid = 1;
    }
    r->weight_sum = next;
    return sel;
}

void rto_restir_finalize(rto_reservoir* r) {
    r->estimator_weight = 0.0f;
    if (r->valid == 0 || r->candidate_count <= 0 || !rs_finite_pos(r->selected_target)
        || !rs_finite_pos(r->weight_sum)) {
        r->valid = 0;
        return;
    }
    r->estimator_weight = r->weight_sum / ((float)r->candidate_count * r->selected_target);
    if (!rs_finite_pos(r->estimator_weight)) {
        r->valid = 0;
        r->estimator_weight = 0.0f;
    }
}

void rto_restir_finalize_basic(rto_reservoir* r, float sst, float sum, int count) {
    r->estimator_weight = 0.0f;
    r->bias_correction_source_count = count > 0 ? count : 0;
    if (r->valid == 0 || count <= 0 || !rs_finite_pos(r->selected_target)
        || !rs_finite_pos(r->weight_sum) || !rs_finite_pos(sst) || !rs_finite_pos(sum)) {
        r->valid = 0;
        return;
    }
    r->estimator_weight = r->weight_sum * sst / (r->selected_target * sum);
    if (!rs_finite_pos(r->estimator_weight)) {
        r->valid = 0;
        r->estimator_weight = 0.0f;
    }
}

int rto_restir_merge(rto_reservoir* cur, const rto_reservoir* prev, float target, int max_count,
    float u, int temporal, int* selected) {
    *selected = 0;
    if (prev->valid == 0 || prev->candidate_count <= 0 || !rs_finite_pos(prev->estimator_weight))
        return 0;
    int mult = prev->candidate_count;
    if (max_count > 0 && mult > max_count) mult = max_count;
    const float cw = rs_finite_pos(target) ? target * prev->estimator_weight * (float)mult : 0.0f;
    *selected = rto_restir_update(cur, prev->light_index, prev->sample_u0, prev->sample_u1, target,
        cw, mult, u);
    if (temporal) {
        cur->temporal_candidate_count += mult;
        cur->age = prev->age + 1;
    } else {
        cur->spatial_candidate_count += mult;
    }
    return mult;
}

/* restir_temporal_surface_valid (restir_di.h:137-153) */
static int rs_surface_valid(const rto_reservoir* cur, const rto_reservoir* prev, float max_dist,
    float min_dot, int max_age) {
    if (prev->valid == 0 || prev->light_index < 0 || prev->age < 0 || prev->age >= max_age
        || cur->material_type != prev->material_type || cur->primitive_type != prev->primitive_type
        || cur->primitive_index != prev->primitive_index)
        return 0;
    const float lim = max_dist > 0.0f ? max_dist : 0.0f;
    const float x = cur->position[0] - prev->position[0];
    const float y = cur->position[1] - prev->position[1];
    const float z = cur->position[2] - prev->position[2];
    const float d2 = x * x + y * y + z * z;
    const float nd = cur->normal[0] * prev->normal[0] + cur->normal[1] * prev->normal[1]
                     + cur->normal[2] * prev->normal[2];
    return d2 <= lim * lim && nd >= min_dot;
}

static void rs_surface(const hit_t* h, rto_reservoir* r) {
    r->position[0] = h->pos.x; r->position[1] = h->pos.y; r->position[2] = h->pos.z;
    r->normal[0] = h->sn.x; r->normal[1] = h->sn.y; r->normal[2] = h->sn.z;
    r->material_type = h->mat_type;
    r->primitive_type = h->prim_type;
    r->primitive_index = h->prim_index;
}

/* restir_unshadowed_numerator (programs.cu:2123-2145) */
static v3 rs_numerator(const ctx_t* c, const hit_t* h, const path_t* st, const light_sample_t* L,
    int bsdf_avail, float* target) {
    *target = 0.0f;
    if (!L->valid || L->pdf <= 0.0f) return mk3(0, 0, 0);
    float bpdf = 0.0f;
    const v3 resp = direct_response(c, h, st, L->dir, &bpdf);
    if (maxc3(resp) <= 0.0f) return mk3(0, 0, 0);
    const float mis = L->delta || !bsdf_avail ? 1.0f : power_heuristic(L->pdf, bpdf);
    const v3 num = scl3(mul3(L->emission, resp), mis);
    *target = fmaxf(0.0f, 0.2126f * num.x + 0.7152f * num.y + 0.0722f * num.z);
    return num;
}

/* restir_previous_pixel (programs.cu:2147-2171) */
static int rs_previous_pixel(const rto_restir_io* io, const rto_camera* cam, v3 pos, int* index) {
    *index = -1;
    if (!io || io->history_length <= 0 || !io->prev_cam || !io->history) return 0;
    const rto_camera* pc = io->prev_cam;
    const v3 origin = mk3((float)pc->origin[0], (float)pc->origin[1], (float)pc->origin[2]);
    const v3 off = sub3(pos, origin);
    const v3 bx = mk3((float)pc->basis_x[0], (float)pc->basis_x[1], (float)pc->basis_x[2]);
    const v3 by = mk3((float)pc->basis_y[0], (float)pc->basis_y[1], (float)pc->basis_y[2]);
    const v3 bz = mk3((float)pc->basis_z[0], (float)pc->basis_z[1], (float)pc->basis_z[2]);
    const float cp[3] = {dot3(off, bx), dot3(off, by), dot3(off, bz)};
    float px[2];
    rto_project(pc, cp, px);
    if (!isfinite(px[0]) || !isfinite(px[1])) return 0;
    const int x = (int)floorf(px[0]);
    const int y = (int)floorf(px[1]);
    if (x < 0 || x >= cam->width || y < 0 || y >= cam->height) return 0;
    *index = y * cam->width + x;
    return 1;
}

/* restir_target_at_previous_surface (programs.cu:2210-2247) */
static float rs_target_at_source(const ctx_t* c, const rto_restir_io* io, const rto_reservoir* src,
    const rto_reservoir* sel, int bsdf_avail) {
    if (!io || !io->prev_cam) return 0.0f;
    const rto_camera* pc = io->prev_cam;
    const v3 sp = ld3(src->position);
    const v3 origin = mk3((float)pc->origin[0], (float)pc->origin[1], (float)pc->origin[2]);
    const v3 off = sub3(sp, origin);
    const float dist = len3(off);
    if (!isfinite(dist) || dist <= OM_RAY_EPS) return 0.0f;
    path_t ss;
    memset(&ss, 0, sizeof(ss));
    ss.throughput = mk3(1.0f, 1.0f, 1.0f);   /* a fresh PathState */
    ss.inside = -1;
    ss.o = origin;
    ss.d = scl3(off, 1.0f / dist);
    float t, u, v;
    int prim;
    if (!om_closest(c->s, c->bvh, ss.o, ss.d, OM_RAY_EPS, dist + 0.05f, &t, &prim, &u, &v)) return 0.0f;
    hit_t h;
    make_hit(c->s, ss.o, ss.d, t, prim, &h);
    if (h.mat_type != src->material_type || h.prim_type != src->primitive_type
        || h.prim_index != src->primitive_index)
        return 0.0f;
    const float tol = fmaxf(0.01f, 0.02f * dist);
    if (lensq3(sub3(h.pos, sp)) > tol * tol || dot3(h.sn, ld3(src->normal)) < 0.95f) return 0.0f;
    const light_sample_t L = analytic_candidate(c, h.pos, sel->light_index, sel->sample_u0, sel->sample_u1);
    float target = 0.0f;
    rs_numerator(c, &h, &ss, &L, bsdf_avail, &target);
    return target;
}

typedef struct {
    const rto_restir_io* io;
    const rto_camera* cam;
    const rto_profile* pr;
    uint32_t pixel, sample, cam_index, stream;
    int x, y;
} rs_ctx;

/* accumulate_restir_analytic_direct_light (programs.cu:2315-2449); spatial
 * neighbours come from the fixed 8-ring of the reference, or, when
 * restir_spatial_radius > 1, from uniform offsets in a disk of that radius */
static void restir_direct(const ctx_t* c, const hit_t* h, const rs_ctx* rc, int bsdf_avail, path_t* st) {
    if (h->mat_type == RTO_MAT_DIFFUSE_LIGHT || c->a_count <= 0) return;
    const rto_profile* pr = rc->pr;
    rto_reservoir r;
    rs_init(&r);
    const int req = pr->restir_initial_candidates > 0 ? pr->restir_initial_candidates : 1;
    const int ncand = req < 32 ? req : 32;
    for (int k = 0; k < ncand; ++k) {
        float q[4];
        om_rng4(rc->pixel, rc->sample, rc->cam_index, 0u, (uint32_t)(OM_G_RESTIR0 + k), rc->stream, q);
        const int li = pick_analytic(c, q[0]);
        const light_sample_t L = analytic_candidate(c, h->pos, li, q[1], q[2]);
        float target = 0.0f;
        rs_numerator(c, h, st, &L, bsdf_avail, &target);
        const float w = L.pdf > 0.0f ? target / L.pdf : 0.0f;
        rto_restir_update(&r, li, q[1], q[2], target, w, 1, q[3]);
    }
    const int canonical = r.candidate_count;
    rto_reservoir sources[9];
    int mults[9];
    int nsrc = 0, selected_source = -1;
    float sel_u[12], off_u[16];
    for (int g = 0; g < 3; ++g) om_rng4(rc->pixel, rc->sample, rc->cam_index, 0u, (uint32_t)(OM_G_RESTIR0 + 32 + g), rc->stream, sel_u + 4 * g);
    for (int g = 0; g < 4; ++g) om_rng4(rc->pixel, rc->sample, rc->cam_index, 0u, (uint32_t)(OM_G_RESTIR0 + 35 + g), rc->stream, off_u + 4 * g);
    rto_reservoir cur_surface;
    rs_init(&cur_surface);
    rs_surface(h, &cur_surface);
    if (pr->restir_temporal_reuse != 0) {
        int pidx;
        if (rs_previous_pixel(rc->io, rc->cam, h->pos, &pidx)) {
            const rto_reservoir* prev = &rc->io->history[pidx];
            if (rs_surface_valid(&cur_surface, prev, fmaxf(0.005f, 0.01f * h->t), 0.95f, pr->restir_max_history_age)) {
                const light_sample_t L = analytic_candidate(c, h->pos, prev->light_index, prev->sample_u0, prev->sample_u1);
                float target = 0.0f;
                rs_numerator(c, h, st, &L, bsdf_avail, &target);
                int sel = 0;
                const int mult = rto_restir_merge(&r, prev, target, pr->restir_max_temporal_candidates, sel_u[0], 1, &sel);
                if (mult > 0) {
                    sources[nsrc] = *prev;
                    mults[nsrc] = mult;
                    if (sel) selected_source = nsrc;
                    ++nsrc;
                }
            }
        }
    }
    const int nsp = pr->restir_spatial_neighbors < 8 ? pr->restir_spatial_neighbors : 8;
    if (nsp > 0) {
        int ridx;
        if (rs_previous_pixel(rc->io, rc->cam, h->pos, &ridx)) {
            static const int ring[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
            const int rot = (rc->x * 3 + rc->y * 5 + (int)rc->stream) & 7;
            const int bx = ridx % rc->cam->width, by = ridx / rc->cam->width;
            for (int j = 0; j < nsp; ++j) {
                int dx, dy;
                if (pr->restir_spatial_radius > 1) {
                    const float rad = (float)pr->restir_spatial_radius * sqrtf(off_u[2 * j]);
                    float sp_, cp_;
                    om_sincos_turn(off_u[2 * j + 1], &sp_, &cp_);
                    dx = (int)floorf(rad * cp_ + 0.5f);
                    dy = (int)floorf(rad * sp_ + 0.5f);
                } else {
                    dx = ring[(rot + j) & 7][0];
                    dy = ring[(rot + j) & 7][1];
                }
                const int nx = bx + dx, ny = by + dy;
                if (nx < 0 || nx >= rc->cam->width || ny < 0 || ny >= rc->cam->height) continue;
                const rto_reservoir* prev = &rc->io->history[ny * rc->cam->width + nx];
                if (!rs_surface_valid(&cur_surface, prev, fmaxf(0.01f, 0.02f * h->t), 0.95f, pr->restir_max_history_age)) continue;
                const light_sample_t L = analytic_candidate(c, h->pos, prev->light_index, prev->sample_u0, prev->sample_u1);
                float target = 0.0f;
                rs_numerator(c, h, st, &L, bsdf_avail, &target);
                int sel = 0;
                const int mult = rto_restir_merge(&r, prev, target, pr->restir_max_spatial_candidates, sel_u[1 + j], 0, &sel);
                if (mult > 0 && nsrc < 9) {
                    sources[nsrc] = *prev;
                    mults[nsrc] = mult;
                    if (sel) selected_source = nsrc;
                    ++nsrc;
                }
            }
        }
    }
    rs_surface(h, &r);
    if (pr->restir_bias_correction == 1 && nsrc > 0 && r.valid != 0) {
        float sst = r.selected_target;
        float sum = r.selected_target * (float)canonical;
        for (int i = 0; i < nsrc; ++i) {
            const float t = rs_target_at_source(c, rc->io, &sources[i], &r, bsdf_avail);
            sum += t * (float)mults[i];
            if (selected_source == i) sst = t;
        }
        rto_restir_finalize_basic(&r, sst, sum, nsrc);
    } else {
        rto_restir_finalize(&r);
    }
    if (rc->io && rc->io->out && (int)rc->sample == (pr->samples_per_pixel > 0 ? pr->samples_per_pixel : 1) - 1)
        rc->io->out[rc->pixel] = r;
    if (r.valid == 0) return;
    const light_sample_t L = analytic_candidate(c, h->pos, r.light_index, r.sample_u0, r.sample_u1);
    float target = 0.0f;
    const v3 num = rs_numerator(c, h, st, &L, bsdf_avail, &target);
    if (!L.valid || target <= 0.0f) return;
    const v3 so = add3(h->pos, scl3(L.dir, OM_RAY_EPS * 2.0f));
    const float max_t = L.infinite ? OM_RAY_FAR : L.dist - OM_RAY_EPS * 3.0f;
    if (max_t <= OM_RAY_EPS || om_occluded(c->s, c->bvh, so, L.dir, OM_RAY_EPS, max_t)) return;
    const v3 direct = scl3(num, r.estimator_weight);
    st->radiance = add3(st->radiance, mul3(st->throughput, direct));
}

static void record_scatter(path_t* st, const hit_t* h, float pdf, int delta) {
    st->prev_pos = h->pos;
    st->prev_n = h->sn;
    st->prev_pdf = pdf;
    st->prev_mat = h->mat_type;
    st->prev_prim = h->prim;
    st->prev_valid = 1;
    st->prev_delta = delta;
}

static float schlick(float cosine, float ri) {
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 *= r0;
    const float x = 1.0f - cosine;
    const float x2 = x * x;
    return r0 + (1.0f - r0) * (x2 * x2 * x);
}

/* programs.cu:2463-2565 */
static void sample_bsdf(const ctx_t* c, const hit_t* h, const float r[4], path_t* st) {
    if (h->mat_type == RTO_MAT_DIFFUSE_LIGHT) { st->alive = 0; return; }
    if (h->mat_type == RTO_MAT_LAMBERTIAN) {
        v3 sc = add3(h->sn, om_sphere_dir(r[0], r[1]));
        if (near_zero3(sc)) sc = h->sn;
        st->o = add3(h->pos, scl3(h->sn, OM_RAY_EPS));
        st->d = norm3(sc);
        st->throughput = mul3(st->throughput, h->base);
        const float cs = fmaxf(dot3(h->sn, st->d), 0.0f);
        record_scatter(st, h, cs / OM_PI, 0);
        return;
    }
    if (h->mat_type == RTO_MAT_METAL) {
        const v3 refl = reflect3(norm3(st->d), h->sn);
        const v3 ball = scl3(om_sphere_dir(r[0], r[1]), om_cbrt01(r[2]));
        const v3 fz = scl3(ball, fminf(fmaxf(h->fuzz, 0.0f), 1.0f));
        const v3 sc = norm3(add3(refl, fz));
        if (dot3(sc, h->sn) <= 0.0f) { st->alive = 0; return; }
        st->o = add3(h->pos, scl3(h->sn, OM_RAY_EPS));
        st->d = sc;
        st->throughput = mul3(st->throughput, h->base);
        record_scatter(st, h, 0.0f, 1);
        return;
    }
    if (h->mat_type == RTO_MAT_DIELECTRIC) {
        const float ratio = h->front ? (1.0f / fmaxf(h->ior, 1e-4f)) : fmaxf(h->ior, 1e-4f);
        const v3 unit = norm3(st->d);
        const float ct = fminf(dot3(neg3(unit), h->sn), 1.0f);
        const float stt = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        const int cannot = ratio * stt > 1.0f;
        v3 dir;
        if (cannot || schlick(ct, ratio) > r[0]) dir = reflect3(unit, h->sn);
        else dir = refract3(unit, h->sn, ratio);
        st->o = add3(h->pos, scl3(dir, OM_RAY_EPS));
        st->d = norm3(dir);
        st->throughput = mul3(st->throughput, h->base);
        record_scatter(st, h, 0.0f, 1);
        return;
    }
    if (h->mat_type == RTO_MAT_OPENPBR) {
        const rto_openpbr p = scattering_material(c->s, h);
        rto_openpbr_sample smp;
        om_openpbr_sample(&p, h->gn, neg3(norm3(st->d)), r[0], r[1], r[2], st->throughput, &smp);
        if (!smp.valid) { st->alive = 0; return; }
        const v3 dir = ld3(smp.wi);
        st->o = add3(h->pos, scl3(dir, OM_RAY_EPS));
        st->d = norm3(dir);
        st->throughput = mul3(st->throughput, ld3(smp.weight));
        record_scatter(st, h, smp.pdf, smp.delta != 0);
        /* subsurface_entry / subsurface_exit (programs.cu:2545-2551) */
        if (smp.event == 8) {
            st->inside = h->openpbr;
            st->inside_mat = p;
        } else if (smp.event == 9) {
            st->inside = -1;
        }
        return;
    }
    if (h->mat_type == RTO_MAT_ISOTROPIC) {
        const v3 dir = om_sphere_dir(r[0], r[1]);
        st->o = add3(h->pos, scl3(dir, OM_RAY_EPS));
        st->d = dir;
        st->throughput = mul3(st->throughput, h->base);
        record_scatter(st, h, UNIFORM_SPHERE_PDF, 0);
        return;
    }
    st->alive = 0;
}

typedef struct { v3 radiance, normal, albedo; float depth; int aux; } path_out;

/* programs.cu:1828-1953 (one sample of one pixel) */
static void trace_one(const ctx_t* c, const rto_camera* cam, int cam_index, const rto_profile* pr,
    uint32_t stream, const rto_restir_io* io, int x, int y, int sample, path_out* out,
    rto_render_stats* stats) {
    const rto_scene* s = c->s;
    const uint32_t pixel = (uint32_t)(y * cam->width + x);
    double px, py;
    pixel_position(x, y, cam_index, sample, stream, pr->pixel_jitter, cam->width, &px, &py);
    path_t st;
    memset(&st, 0, sizeof(st));
    float o[3], d[3];
    rto_camera_ray(cam, px, py, o, d);
    st.o = ld3(o);
    st.d = ld3(d);
    st.throughput = mk3(1, 1, 1);
    st.radiance = mk3(0, 0, 0);
    st.prev_delta = 1;
    st.prev_prim = -1;
    st.prev_mat = -1;
    st.alive = 1;
    st.inside = -1;
    out->aux = 0;
    out->normal = mk3(0, 0, 1);
    out->albedo = mk3(0, 0, 0);
    out->depth = 0.0f;
    const int max_b = pr->max_bounces > 0 ? pr->max_bounces : 1;
    for (int b = 0; b < max_b && st.alive; ++b) {
        float rl[4], rb[4];
        om_rng4(pixel, (uint32_t)sample, (uint32_t)cam_index, (uint32_t)b, OM_G_LIGHT, stream, rl);
        om_rng4(pixel, (uint32_t)sample, (uint32_t)cam_index, (uint32_t)b, OM_G_BSDF, stream, rb);
        float t, u, v;
        int prim;
        if (stats) stats->rays_closest++;
        int found = om_closest_counted(s, c->bvh, st.o, st.d, OM_RAY_EPS, OM_RAY_FAR, &t, &prim,
            &u, &v, stats ? &stats->node_visits : NULL, stats ? &stats->prim_tests : NULL);
        if (s->medium_count > 0) {
            float closest = found ? t : OM_RAY_FAR;
            const int mi = sample_media(s, st.o, st.d, OM_RAY_EPS, &closest, pixel, (uint32_t)sample,
                (uint32_t)cam_index, (uint32_t)b, stream);
            if (mi >= 0) {
                t = closest;
                prim = OM_MEDIUM_HIT - mi;
                found = 1;
            }
        }
        int ali = -1;
        if (c->a_count > 0 && intersect_analytic(c, st.o, st.d, OM_RAY_EPS, found ? t : OM_RAY_FAR, &ali)) {
            const float w = analytic_mis(c, &st, ali);
            st.radiance = add3(st.radiance,
                mul3(st.throughput, scl3(ld3(s->analytic_lights[ali].radiance), w)));
            break;
        }
        if (!found) {
            const float w = emission_mis(c, &st, NULL, 1);
            st.radiance = add3(st.radiance, mul3(st.throughput, scl3(background_at(s, st.d), w)));
            if (c->a_count > 0 && has_infinite_light(c))
                st.radiance = add3(st.radiance, mul3(st.throughput, analytic_infinite(c, &st)));
            break;
        }
        hit_t h;
        make_hit(s, st.o, st.d, t, prim, &h);
        if (st.inside >= 0) {
            /* the OpenPBR subsurface random walk (programs.cu:1878-1901): only
             * the medium's own back face continues the path */
            if (h.mat_type != RTO_MAT_OPENPBR || h.openpbr != st.inside || h.front) break;
            float rw[4], seg[5];
            om_rng4(pixel, (uint32_t)sample, (uint32_t)cam_index, (uint32_t)b, OM_G_SUBSURFACE, stream, rw);
            const rto_openpbr* pm = &st.inside_mat;
            const float rlen = sqrtf(lensq3(st.d));
            om_openpbr_segment(pm, t * rlen, rw[0], st.throughput, seg);
            st.throughput = mul3(st.throughput, ld3(seg));
            if (seg[4] != 0.0f) {
                const v3 dir = om_openpbr_hg(pm, norm3(st.d), rw[1], rw[2]);
                st.o = add3(st.o, scl3(st.d, seg[3] / fmaxf(rlen, 1e-8f)));
                st.d = dir;
                st.o = add3(st.o, scl3(st.d, OM_RAY_EPS));
                st.prev_valid = 0;
                continue;
            }
        }
        if (b == 0) {
            out->normal = h.sn;
            out->albedo = h.base;
            out->depth = h.t;
            out->aux = 1;
        }
        const float ew = emission_mis(c, &st, &h, 0);
        st.radiance = add3(st.radiance, mul3(st.throughput, scl3(h.emission, ew)));
        const int bsdf_avail = b + 1 < max_b;
        if (stats && h.mat_type != RTO_MAT_DIFFUSE_LIGHT && c->light_count > 0) stats->rays_shadow++;
        direct_light(c, &h, rl, bsdf_avail, &st);
        const int min_lights = pr->restir_min_analytic_lights > 0 ? pr->restir_min_analytic_lights : 1;
        if (pr->enable_restir_di && b == 0 && c->a_count >= min_lights) {
            rs_ctx rc = {io, cam, pr, pixel, (uint32_t)sample, (uint32_t)cam_index, stream, x, y};
            restir_direct(c, &h, &rc, bsdf_avail, &st);
        } else {
            float ra[4];
            om_rng4(pixel, (uint32_t)sample, (uint32_t)cam_index, (uint32_t)b, OM_G_ALIGHT, stream, ra);
            analytic_direct_light(c, &h, ra, bsdf_avail, &st);
        }
        if (!bsdf_avail) break;
        sample_bsdf(c, &h, rb, &st);
        if (st.alive && b >= pr->rr_start_bounce) {
            const float p = fminf(fmaxf(maxc3(st.throughput), 0.05f), 0.95f);
            if (rb[3] > p) { st.alive = 0; break; }
            st.throughput = div3(st.throughput, p);
        }
    }
    out->radiance = st.radiance;
}

void rto_trace_path(const rto_scene* s, const rto_bvh* bvh, const rto_camera* cam, int cam_index,
    const rto_profile* pr, uint32_t stream, int x, int y, int sample, float radiance[3]) {
    ctx_t c;
    if (ctx_init(&c, s, bvh) != 0) return;
    path_out po;
    trace_one(&c, cam, cam_index, pr, stream, NULL, x, y, sample, &po, NULL);
    radiance[0] = po.radiance.x; radiance[1] = po.radiance.y; radiance[2] = po.radiance.z;
    ctx_free(&c);
}

/* --------------------------------------------------------------- render */
typedef struct {
    const ctx_t* c; const rto_camera* cam; int cam_index; const rto_profile* pr; uint32_t stream;
    const rto_restir_io* io;
    float *beauty, *normal, *albedo, *depth;
    int row_begin, row_end;
    volatile int next_row;
    pthread_mutex_t lock;
    rto_render_stats stats;
} render_job;

static void* render_worker(void* arg) {
    render_job* j = (render_job*)arg;
    const rto_camera* cam = j->cam;
    const int spp = j->pr->samples_per_pixel > 0 ? j->pr->samples_per_pixel : 1;
    rto_render_stats local = {0, 0, 0, 0, 0};
    while (1) {
        pthread_mutex_lock(&j->lock);
        const int y = j->next_row++;
        pthread_mutex_unlock(&j->lock);
        if (y >= j->row_end) break;
        for (int x = 0; x < cam->width; ++x) {
            v3 beauty = mk3(0, 0, 0), nsum = mk3(0, 0, 0), asum = mk3(0, 0, 0);
            float dsum = 0.0f;
            int aux = 0;
            for (int s = 0; s < spp; ++s) {
                path_out po;
                trace_one(j->c, cam, j->cam_index, j->pr, j->stream, j->io, x, y, s, &po, &local);
                local.paths++;
                beauty = add3(beauty, po.radiance);
                if (po.aux) {
                    nsum = add3(nsum, po.normal);
                    asum = add3(asum, po.albedo);
                    dsum += po.depth;
                    ++aux;
                }
            }
            /* programs.cu:1944-1952 + store_output :1773-1805 */
            const float inv_spp = 1.0f / (float)spp;
            const v3 bavg = scl3(beauty, inv_spp);
            const float inv_aux = aux > 0 ? 1.0f / (float)aux : 0.0f;
            const v3 navg = aux > 0 ? norm3(scl3(nsum, inv_aux)) : mk3(0, 0, 1);
            const v3 aavg = aux > 0 ? scl3(asum, inv_aux) : mk3(0, 0, 0);
            const float davg = aux > 0 ? dsum * inv_aux : 0.0f;
            const size_t p = (size_t)y * (size_t)cam->width + (size_t)x;
            if (j->beauty) { float* q = j->beauty + 4 * p; q[0] = bavg.x; q[1] = bavg.y; q[2] = bavg.z; q[3] = 1.0f; }
            if (j->normal) {
                const v3 e = clamp3(add3(scl3(navg, 0.5f), mk3(0.5f, 0.5f, 0.5f)), 0.0f, 1.0f);
                float* q = j->normal + 4 * p; q[0] = e.x; q[1] = e.y; q[2] = e.z; q[3] = 1.0f;
            }
            if (j->albedo) {
                const v3 e = clamp3(aavg, 0.0f, 1.0f);
                float* q = j->albedo + 4 * p; q[0] = e.x; q[1] = e.y; q[2] = e.z; q[3] = 1.0f;
            }
            if (j->depth) j->depth[p] = davg;
        }
    }
    pthread_mutex_lock(&j->lock);
    j->stats.paths += local.paths;
    j->stats.rays_closest += local.rays_closest;
    j->stats.rays_shadow += local.rays_shadow;
    j->stats.node_visits += local.node_visits;
    j->stats.prim_tests += local.prim_tests;
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

int rto_render_camera(const rto_scene* s, const rto_bvh* bvh, const rto_camera* cam, int cam_index,
    const rto_profile* pr, uint32_t stream, float* beauty, float* normal, float* albedo,
    float* depth, int threads, int row_begin, int row_end, rto_render_stats* stats) {
    return rto_render_camera_ex(s, bvh, cam, cam_index, pr, stream, NULL, beauty, normal, albedo,
        depth, threads, row_begin, row_end, stats);
}

int rto_render_camera_ex(const rto_scene* s, const rto_bvh* bvh, const rto_camera* cam,
    int cam_index, const rto_profile* pr, uint32_t stream, const rto_restir_io* io, float* beauty,
    float* normal, float* albedo, float* depth, int threads, int row_begin, int row_end,
    rto_render_stats* stats) {
    static ctx_t c;  /* large; one render at a time per process */
    static pthread_mutex_t once = PTHREAD_MUTEX_INITIALIZER;
    pthread_mutex_lock(&once);
    if (ctx_init(&c, s, bvh) != 0) { pthread_mutex_unlock(&once); return -1; }
    if (row_end <= 0 || row_end > cam->height) row_end = cam->height;
    if (row_begin < 0) row_begin = 0;
    render_job j;
    memset(&j, 0, sizeof(j));
    j.c = &c; j.cam = cam; j.cam_index = cam_index; j.pr = pr; j.stream = stream; j.io = io;
    j.beauty = beauty; j.normal = normal; j.albedo = albedo; j.depth = depth;
    j.row_begin = row_begin; j.row_end = row_end; j.next_row = row_begin;
    pthread_mutex_init(&j.lock, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, render_worker, &j);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&j.lock);
    if (stats) *stats = j.stats;
    ctx_free(&c);
    pthread_mutex_unlock(&once);
    return 0;
}

/* ------------------------------------------------------ test entry points */
/* The integrator's light-sampling densities, CDF selection and scattering,
 * exposed so that tests can pin them against the reference's known answers,
 * the reference headers compiled in place, and closed forms. */
float rto_power_heuristic(float a, float b) { return power_heuristic(a, b); }
float rto_cone_pdf(float cos_max) { return cone_pdf(cos_max); }
float rto_area_pdf(float area, float dsq, float abs_cos) { return area_pdf(area, dsq, abs_cos); }
float rto_uniform_sphere_pdf(void) { return UNIFORM_SPHERE_PDF; }
int rto_pick_cdf(const float* cdf, int n, float u) { return pick_cdf(cdf, 1, n, u); }
void rto_triangle_barycentrics(float u0, float u1, float w[3]) {
    triangle_barycentrics(u0, u1, &w[0], &w[1], &w[2]);
}

static void test_hit(const rto_scene* s, int prim, const float pos[3], const float gn[3],
    const float d[3], hit_t* h) {
    memset(h, 0, sizeof(*h));
    const int ns = s->sphere_count, nq = s->quad_count;
    h->prim = prim;
    h->prim_type = prim < ns ? 0 : (prim < ns + nq ? 1 : 2);
    h->prim_index = prim < ns ? prim : (prim < ns + nq ? prim - ns : prim - ns - nq);
    h->pos = ld3(pos);
    set_face(ld3(d), ld3(gn), h);
    const int mat = h->prim_type == 0 ? s->spheres[h->prim_index].material
                  : h->prim_type == 1 ? s->quads[h->prim_index].material
                                      : s->triangles[h->prim_index].material;
    const rto_material* m = &s->materials[mat];
    h->mat_type = m->type;
    h->base = ld3(m->albedo);
    h->fuzz = m->fuzz;
    h->ior = m->ior;
    h->openpbr = m->openpbr_index;
    if (h->mat_type == RTO_MAT_DIELECTRIC) h->base = mk3(0.92f, 0.95f, 1.0f);
    if (h->mat_type == RTO_MAT_OPENPBR) h->base = ld3(s->openpbr[m->openpbr_index].base_color);
}

/* NEE from a point of primitive `prim` (global id) with outward geometric
 * normal gn, seen along d: out = dir[3], emission[3], dist, pdf; returns
 * the light index (or -1 when the sample is invalid) */
int rto_sample_light_at(const rto_scene* s, const rto_bvh* bvh, int prim, const float pos[3],
    const float gn[3], const float d[3], const float r[4], float out[8]) {
    ctx_t* c = (ctx_t*)malloc(sizeof(ctx_t));
    if (c == NULL || ctx_init(c, s, bvh) != 0) { free(c); return -2; }
    hit_t h;
    test_hit(s, prim, pos, gn, d, &h);
    const light_sample_t L = sample_light(c, &h, r);
    out[0] = L.dir.x; out[1] = L.dir.y; out[2] = L.dir.z;
    out[3] = L.emission.x; out[4] = L.emission.y; out[5] = L.emission.z;
    out[6] = L.dist; out[7] = L.pdf;
    int li = -1;
    if (L.valid) {
        /* the light the sample landed on: trace it (-3 when something else,
         * another light included, is hit first) */
        float t, u, v;
        int hp;
        const int found = om_closest(s, bvh, h.pos, L.dir, OM_RAY_EPS, OM_RAY_FAR, &t, &hp, &u, &v);
        if (L.infinite) li = found ? -3 : c->env_light;
        else if (!found || t < L.dist * 0.999f) li = -3;
        else li = c->light_of_prim[hp];
    }
    ctx_free(c);
    free(c);
    return L.valid ? (li >= 0 ? li : -3) : -1;
}

/* the hit record of the ray o + t d on global primitive prim (populate_material,
 * textures included): out = base colour[3], emission[3], texture u, v */
void rto_hit_material(const rto_scene* s, const float o[3], const float d[3], float t, int prim,
    float out[8]) {
    hit_t h;
    make_hit(s, ld3(o), ld3(d), t, prim, &h);
    out[0] = h.base.x; out[1] = h.base.y; out[2] = h.base.z;
    out[3] = h.emission.x; out[4] = h.emission.y; out[5] = h.emission.z;
    out[6] = h.tex_u; out[7] = h.tex_v;
}

/* the NEE density of the surface point (or environment) hit by the ray
 * (pos, dir) leaving primitive `prim` (-1: none) of material type prev_mat;
 * returns -1 when the ray hits no light */
float rto_light_pdf_for_hit(const rto_scene* s, const rto_bvh* bvh, int prim, int prev_mat,
    const float pos[3], const float sn[3], const float dir[3]) {
    ctx_t* c = (ctx_t*)malloc(sizeof(ctx_t));
    if (c == NULL || ctx_init(c, s, bvh) != 0) { free(c); return -2.0f; }
    float t, u, v;
    int hp, light = -1;
    float pdf;
    const v3 o = ld3(pos), d = norm3(ld3(dir));
    if (om_closest(s, bvh, o, d, OM_RAY_EPS, OM_RAY_FAR, &t, &hp, &u, &v)) {
        hit_t h;
        make_hit(s, o, d, t, hp, &h);
        pdf = light_pdf_for_hit(c, o, ld3(sn), prev_mat, prim, d, &h, 0, &light);
    } else {
        pdf = light_pdf_for_hit(c, o, ld3(sn), prev_mat, prim, d, NULL, 1, &light);
    }
    ctx_free(c);
    free(c);
    return light < 0 ? -1.0f : pdf;
}

/* one BSDF scattering event at a point of primitive `prim` with outward
 * geometric normal gn, reached along d: out = new direction[3], throughput
 * factor[3], pdf, delta; returns 1 when the path continues */
int rto_scatter(const rto_scene* s, int prim, const float pos[3], const float gn[3],
    const float d[3], const float r[4], float out[8]) {
    ctx_t c;
    memset(&c, 0, sizeof(c));
    c.s = s;
    hit_t h;
    test_hit(s, prim, pos, gn, d, &h);
    path_t st;
    memset(&st, 0, sizeof(st));
    st.d = ld3(d);
    st.throughput = mk3(1.0f, 1.0f, 1.0f);
    st.alive = 1;
    sample_bsdf(&c, &h, r, &st);
    out[0] = st.d.x; out[1] = st.d.y; out[2] = st.d.z;
    out[3] = st.throughput.x; out[4] = st.throughput.y; out[5] = st.throughput.z;
    out[6] = st.prev_pdf;
    out[7] = (float)st.prev_delta;
    return st.alive;
}
