/*
 * oracle_scenes.c -- seeded synthetic scenes (independent second copy used
 * to pin the product's scene generators byte for byte).
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 *
 * Draws come from PCG32 (O'Neill 2014, pcg32_srandom_r(seed, 54), output
 * XSH-RR); a uniform float is (x >> 8) * 2^-24.  Draw order is part of the
 * scene definition.
 *   random_spheres : ground + 22x22 small spheres + 3 large spheres
 *   cornell_mesh   : 555-unit Cornell room (5 walls + light quad), icosphere
 *                    subdivided 6x (81,920 triangles), one OpenPBR material
 *   instanced_mesh : 1000 rigid instances of a 1024-triangle mesh + 256
 *                    analytic sphere lights
 * Camera construction (look-at, vertical fov) is host math in FP64.
 */
#include "rt_oracle.h"
#include "om.h"

#include <stdlib.h>

typedef struct { uint64_t state, inc; } pcg32_t;

uint32_t rto_pcg32_next(uint64_t* state, uint64_t inc) {
    const uint64_t old = *state;
    *state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}
static void pcg_seed(pcg32_t* r, uint64_t seed) {
    r->state = 0u;
    r->inc = (54u << 1u) | 1u;
    rto_pcg32_next(&r->state, r->inc);
    r->state += seed;
    rto_pcg32_next(&r->state, r->inc);
}
static float pcg_u(pcg32_t* r) { return (float)(rto_pcg32_next(&r->state, r->inc) >> 8) * (1.0f / 16777216.0f); }

typedef struct {
    rto_scene scene;
    int32_t camera_count;
    int32_t pad;
    rto_camera cameras[4];
    rto_profile profile;
} rto_generated;

static void set3(float* d, float x, float y, float z) { d[0] = x; d[1] = y; d[2] = z; }

static void openpbr_defaults(rto_openpbr* m) {
    memset(m, 0, sizeof(*m));
    m->base_weight = 1.0f; set3(m->base_color, 0.8f, 0.8f, 0.8f);
    m->specular_weight = 1.0f; set3(m->specular_color, 1.0f, 1.0f, 1.0f);
    m->specular_roughness = 0.3f; m->specular_ior = 1.5f;
    set3(m->transmission_color, 1.0f, 1.0f, 1.0f);
    m->transmission_dispersion_abbe_number = 20.0f;
    set3(m->subsurface_color, 0.8f, 0.8f, 0.8f); m->subsurface_radius = 1.0f;
    set3(m->subsurface_radius_scale, 1.0f, 0.5f, 0.25f);
    set3(m->fuzz_color, 1.0f, 1.0f, 1.0f); m->fuzz_roughness = 0.5f;
    set3(m->coat_color, 1.0f, 1.0f, 1.0f); m->coat_ior = 1.6f; m->coat_darkening = 1.0f;
    m->thin_film_thickness = 0.5f; m->thin_film_ior = 1.4f;
    set3(m->emission_color, 1.0f, 1.0f, 1.0f);
    m->geometry_opacity = 1.0f;
}

/* pinhole32 look-at camera; OpenCV axes (x right, y down, z forward) */
static void lookat_pinhole(rto_camera* c, int w, int h, double vfov_deg, const double from[3],
    const double at[3], const double up[3]) {
    memset(c, 0, sizeof(*c));
    c->enabled = 1; c->width = w; c->height = h; c->model = RTO_CAM_PINHOLE32;
    double z[3] = {at[0] - from[0], at[1] - from[1], at[2] - from[2]};
    double zl = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (int i = 0; i < 3; ++i) z[i] /= zl;
    double x[3] = {z[1] * up[2] - z[2] * up[1], z[2] * up[0] - z[0] * up[2], z[0] * up[1] - z[1] * up[0]};
    double xl = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (int i = 0; i < 3; ++i) x[i] /= xl;
    double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
    for (int i = 0; i < 3; ++i) {
        c->origin[i] = from[i]; c->basis_x[i] = x[i]; c->basis_y[i] = y[i]; c->basis_z[i] = z[i];
    }
    const double vfov = vfov_deg * 3.14159265358979323846 / 180.0;
    const double f = 0.5 * (double)h / tan(0.5 * vfov);
    c->fx = f; c->fy = f; c->cx = 0.5 * (double)w; c->cy = 0.5 * (double)h;
}

static void profile_defaults(rto_profile* p, int spp, int depth) {
    memset(p, 0, sizeof(*p));
    p->samples_per_pixel = spp; p->max_bounces = depth; p->enable_denoise = 0;
    p->rr_start_bounce = depth; /* no roulette unless asked */
    p->accumulation_reset_rotation_deg = 2.0; p->accumulation_reset_translation = 0.05;
    p->restir_initial_candidates = 4; p->restir_temporal_reuse = 1; p->restir_max_history_age = 20;
    p->restir_max_temporal_candidates = 64; p->restir_spatial_neighbors = 1;
    p->restir_max_spatial_candidates = 4; p->restir_bias_correction = 1;
    p->restir_min_analytic_lights = 16; p->pixel_jitter = 1; p->restir_spatial_radius = 30;
}

static int gen_random_spheres(rto_generated* g, uint64_t seed) {
    const int n = 1 + 22 * 22 + 3;
    rto_sphere* sp = (rto_sphere*)calloc((size_t)n, sizeof(rto_sphere));
    rto_material* mt = (rto_material*)calloc((size_t)n, sizeof(rto_material));
    pcg32_t r;
    pcg_seed(&r, seed);
    int k = 0;
    set3(sp[k].center, 0.0f, -1000.0f, 0.0f); sp[k].radius = 1000.0f; sp[k].material = k;
    mt[k].type = RTO_MAT_LAMBERTIAN; set3(mt[k].albedo, 0.5f, 0.5f, 0.5f); mt[k].ior = 1.0f; mt[k].openpbr_index = -1;
    ++k;
    for (int a = -11; a < 11; ++a) {
        for (int b = -11; b < 11; ++b) {
            const float choose = pcg_u(&r);
            const float u = pcg_u(&r);
            const float v = pcg_u(&r);
            set3(sp[k].center, (float)a + 0.9f * u, 0.2f, (float)b + 0.9f * v);
            sp[k].radius = 0.2f;
            sp[k].material = k;
            mt[k].openpbr_index = -1;
            mt[k].ior = 1.0f;
            if (choose < 0.8f) {
                mt[k].type = RTO_MAT_LAMBERTIAN;
                for (int ch = 0; ch < 3; ++ch) { const float x0 = pcg_u(&r); const float x1 = pcg_u(&r); mt[k].albedo[ch] = x0 * x1; }
            } else if (choose < 0.95f) {
                mt[k].type = RTO_MAT_METAL;
                for (int ch = 0; ch < 3; ++ch) mt[k].albedo[ch] = 0.5f + 0.5f * pcg_u(&r);
                mt[k].fuzz = 0.5f * pcg_u(&r);
            } else {
                mt[k].type = RTO_MAT_DIELECTRIC;
                mt[k].ior = 1.5f;
            }
            ++k;
        }
    }
    set3(sp[k].center, 0.0f, 1.0f, 0.0f); sp[k].radius = 1.0f; sp[k].material = k;
    mt[k].type = RTO_MAT_DIELECTRIC; mt[k].ior = 1.5f; mt[k].openpbr_index = -1; ++k;
    set3(sp[k].center, -4.0f, 1.0f, 0.0f); sp[k].radius = 1.0f; sp[k].material = k;
    mt[k].type = RTO_MAT_LAMBERTIAN; set3(mt[k].albedo, 0.4f, 0.2f, 0.1f); mt[k].ior = 1.0f; mt[k].openpbr_index = -1; ++k;
    set3(sp[k].center, 4.0f, 1.0f, 0.0f); sp[k].radius = 1.0f; sp[k].material = k;
    mt[k].type = RTO_MAT_METAL; set3(mt[k].albedo, 0.7f, 0.6f, 0.5f); mt[k].fuzz = 0.0f; mt[k].ior = 1.0f; mt[k].openpbr_index = -1; ++k;
    g->scene.sphere_count = n;
    g->scene.material_count = n;
    g->scene.spheres = sp;
    g->scene.materials = mt;
    g->scene.background_mode = RTO_BG_SKY_GRADIENT;
    set3(g->scene.background, 1.0f, 1.0f, 1.0f);
    set3(g->scene.background_top, 0.5f, 0.7f, 1.0f);
    g->camera_count = 1;
    const double from[3] = {13, 2, 3}, at[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    lookat_pinhole(&g->cameras[0], 400, 225, 20.0, from, at, up);
    profile_defaults(&g->profile, 32, 10);
    return 0;
}

/* ------------------------------------------------------------ icosphere */
typedef struct { uint64_t* keys; int32_t* vals; uint32_t cap; } edge_map;

static int32_t edge_mid(edge_map* m, double (*verts)[3], int* nverts, int32_t a, int32_t b) {
    const uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    const uint64_t key = ((uint64_t)lo << 32) | hi;
    uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ULL) >> 40) & (m->cap - 1u);
    while (m->vals[h] >= 0) {
        if (m->keys[h] == key) return m->vals[h];
        h = (h + 1u) & (m->cap - 1u);
    }
    double p[3];
    for (int i = 0; i < 3; ++i) p[i] = 0.5 * (verts[a][i] + verts[b][i]);
    const double l = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const int32_t id = (*nverts)++;
    for (int i = 0; i < 3; ++i) verts[id][i] = p[i] / l;
    m->keys[h] = key;
    m->vals[h] = id;
    return id;
}

/* unit icosphere with `levels` 4:1 subdivisions; outward winding */
int rto_icosphere(int levels, double (**verts_out)[3], int32_t (**faces_out)[3], int* nv_out, int* nf_out) {
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
        {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int32_t bf[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9},
        {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    int nf = 20;
    for (int l = 0; l < levels; ++l) nf *= 4;
    const int maxv = nf / 2 + 2;
    double (*verts)[3] = (double (*)[3])malloc(sizeof(double) * 3 * (size_t)maxv);
    int32_t (*faces)[3] = (int32_t (*)[3])malloc(sizeof(int32_t) * 3 * (size_t)nf);
    int nv = 12;
    for (int i = 0; i < 12; ++i) {
        const double l = sqrt(base[i][0] * base[i][0] + base[i][1] * base[i][1] + base[i][2] * base[i][2]);
        for (int k = 0; k < 3; ++k) verts[i][k] = base[i][k] / l;
    }
    int cur = 20;
    for (int i = 0; i < 20; ++i) {
        int32_t a = bf[i][0], b = bf[i][1], c = bf[i][2];
        double e1[3], e2[3], cr[3], ce[3];
        for (int k = 0; k < 3; ++k) { e1[k] = verts[b][k] - verts[a][k]; e2[k] = verts[c][k] - verts[a][k]; ce[k] = verts[a][k] + verts[b][k] + verts[c][k]; }
        cr[0] = e1[1] * e2[2] - e1[2] * e2[1]; cr[1] = e1[2] * e2[0] - e1[0] * e2[2]; cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
        if (cr[0] * ce[0] + cr[1] * ce[1] + cr[2] * ce[2] < 0.0) { const int32_t tmp = b; b = c; c = tmp; }
        faces[i][0] = a; faces[i][1] = b; faces[i][2] = c;
    }
    int32_t (*next)[3] = (int32_t (*)[3])malloc(sizeof(int32_t) * 3 * (size_t)nf);
    for (int l = 0; l < levels; ++l) {
        edge_map m;
        m.cap = 1u;
        while (m.cap < (uint32_t)(cur * 4)) m.cap <<= 1;
        m.keys = (uint64_t*)calloc(m.cap, sizeof(uint64_t));
        m.vals = (int32_t*)malloc(sizeof(int32_t) * m.cap);
        for (uint32_t i = 0; i < m.cap; ++i) m.vals[i] = -1;
        int nn = 0;
        for (int f = 0; f < cur; ++f) {
            const int32_t a = faces[f][0], b = faces[f][1], c = faces[f][2];
            const int32_t ab = edge_mid(&m, verts, &nv, a, b);
            const int32_t bc = edge_mid(&m, verts, &nv, b, c);
            const int32_t ca = edge_mid(&m, verts, &nv, c, a);
            next[nn][0] = a; next[nn][1] = ab; next[nn][2] = ca; ++nn;
            next[nn][0] = b; next[nn][1] = bc; next[nn][2] = ab; ++nn;
            next[nn][0] = c; next[nn][1] = ca; next[nn][2] = bc; ++nn;
            next[nn][0] = ab; next[nn][1] = bc; next[nn][2] = ca; ++nn;
        }
        memcpy(faces, next, sizeof(int32_t) * 3 * (size_t)nn);
        cur = nn;
        free(m.keys);
        free(m.vals);
    }
    free(next);
    *verts_out = verts; *faces_out = faces; *nv_out = nv; *nf_out = cur;
    return 0;
}

static void quad_set(rto_quad* q, float ox, float oy, float oz, float ux, float uy, float uz,
    float vx, float vy, float vz, int mat) {
    set3(q->origin, ox, oy, oz); set3(q->edge_u, ux, uy, uz); set3(q->edge_v, vx, vy, vz); q->material = mat;
}

static int gen_cornell_mesh(rto_generated* g, uint64_t seed) {
    pcg32_t r;
    pcg_seed(&r, seed);
    rto_material* mt = (rto_material*)calloc(5, sizeof(rto_material));
    for (int i = 0; i < 5; ++i) { mt[i].openpbr_index = -1; mt[i].ior = 1.0f; }
    set3(mt[0].albedo, 0.65f, 0.05f, 0.05f);
    set3(mt[1].albedo, 0.73f, 0.73f, 0.73f);
    set3(mt[2].albedo, 0.12f, 0.45f, 0.15f);
    mt[3].type = RTO_MAT_DIFFUSE_LIGHT; set3(mt[3].emission, 15.0f, 15.0f, 15.0f);
    set3(mt[3].albedo, 0.0f, 0.0f, 0.0f);
    mt[4].type = RTO_MAT_OPENPBR; mt[4].openpbr_index = 0;
    rto_openpbr* op = (rto_openpbr*)calloc(1, sizeof(rto_openpbr));
    openpbr_defaults(op);
    for (int ch = 0; ch < 3; ++ch) op->base_color[ch] = pcg_u(&r);
    op->specular_roughness = 0.05f + 0.85f * pcg_u(&r);
    op->base_metalness = pcg_u(&r) < 0.5f ? 0.0f : 1.0f;
    op->specular_ior = 1.5f;
    op->coat_weight = 0.5f * pcg_u(&r);
    for (int ch = 0; ch < 3; ++ch) mt[4].albedo[ch] = op->base_color[ch];

    rto_quad* q = (rto_quad*)calloc(6, sizeof(rto_quad));
    quad_set(&q[0], 555, 0, 0, 0, 555, 0, 0, 0, 555, 2);
    quad_set(&q[1], 0, 0, 0, 0, 555, 0, 0, 0, 555, 0);
    quad_set(&q[2], 343, 554, 332, -130, 0, 0, 0, 0, -105, 3);
    quad_set(&q[3], 0, 0, 0, 555, 0, 0, 0, 0, 555, 1);
    quad_set(&q[4], 555, 555, 555, -555, 0, 0, 0, 0, -555, 1);
    quad_set(&q[5], 0, 0, 555, 555, 0, 0, 0, 555, 0, 1);

    double (*verts)[3];
    int32_t (*faces)[3];
    int nv, nf;
    rto_icosphere(6, &verts, &faces, &nv, &nf);
    const double center[3] = {277.5, 277.5, 277.5};
    const double radius = 150.0;
    float (*pos)[3] = (float (*)[3])malloc(sizeof(float) * 3 * (size_t)nv);
    for (int i = 0; i < nv; ++i) {
        const float u = pcg_u(&r);
        const double rr = radius * (1.0 + 0.05 * (2.0 * (double)u - 1.0));
        for (int k = 0; k < 3; ++k) pos[i][k] = (float)(center[k] + verts[i][k] * rr);
    }
    rto_triangle* tr = (rto_triangle*)calloc((size_t)nf, sizeof(rto_triangle));
    for (int f = 0; f < nf; ++f) {
        memcpy(tr[f].p0, pos[faces[f][0]], 12);
        memcpy(tr[f].p1, pos[faces[f][1]], 12);
        memcpy(tr[f].p2, pos[faces[f][2]], 12);
        tr[f].material = 4;
    }
    free(pos); free(verts); free(faces);
    g->scene.quad_count = 6; g->scene.quads = q;
    g->scene.triangle_count = nf; g->scene.triangles = tr;
    g->scene.material_count = 5; g->scene.materials = mt;
    g->scene.openpbr_count = 1; g->scene.openpbr = op;
    g->scene.background_mode = RTO_BG_CONSTANT;
    set3(g->scene.background, 0.0f, 0.0f, 0.0f);
    g->camera_count = 4;
    const double up[3] = {0, 1, 0};
    for (int c = 0; c < 4; ++c) {
        const double th = (double)c * 0.5 * 3.14159265358979323846;
        const double from[3] = {center[0] + 800.0 * sin(th), center[1], center[2] - 800.0 * cos(th)};
        lookat_pinhole(&g->cameras[c], 640, 480, 40.0, from, center, up);
    }
    profile_defaults(&g->profile, 16, 8);
    return 0;
}

/* cornell_random (id 4): Cornell room with 1-8 displaced icospheres (subdivision 4),
 * one random OpenPBR material each, a random ceiling light (see the
 * library's generate_cornell_random for the draw order) and four cameras on
 * an arc in front of the open side */
static int gen_cornell_random(rto_generated* g, uint64_t seed) {
    pcg32_t r;
    pcg_seed(&r, seed);
    const float lw = 80.0f + 80.0f * pcg_u(&r);
    const float ld = 70.0f + 70.0f * pcg_u(&r);
    const float lx = 20.0f + (515.0f - lw) * pcg_u(&r);
    const float lz = 20.0f + (515.0f - ld) * pcg_u(&r);
    const float radiance = 8.0f + 12.0f * pcg_u(&r);
    int spheres = 1 + (int)(8.0f * pcg_u(&r));
    if (spheres > 8) spheres = 8;
    rto_material* mt = (rto_material*)calloc((size_t)(4 + spheres), sizeof(rto_material));
    for (int i = 0; i < 4 + spheres; ++i) { mt[i].openpbr_index = -1; mt[i].ior = 1.0f; }
    set3(mt[0].albedo, 0.65f, 0.05f, 0.05f);
    set3(mt[1].albedo, 0.73f, 0.73f, 0.73f);
    set3(mt[2].albedo, 0.12f, 0.45f, 0.15f);
    mt[3].type = RTO_MAT_DIFFUSE_LIGHT; set3(mt[3].emission, radiance, radiance, radiance);
    rto_quad* q = (rto_quad*)calloc(6, sizeof(rto_quad));
    quad_set(&q[0], 555, 0, 0, 0, 555, 0, 0, 0, 555, 2);
    quad_set(&q[1], 0, 0, 0, 0, 555, 0, 0, 0, 555, 0);
    quad_set(&q[2], lx + lw, 554, lz + ld, -lw, 0, 0, 0, 0, -ld, 3);
    quad_set(&q[3], 0, 0, 0, 555, 0, 0, 0, 0, 555, 1);
    quad_set(&q[4], 555, 555, 555, -555, 0, 0, 0, 0, -555, 1);
    quad_set(&q[5], 0, 0, 555, 555, 0, 0, 0, 555, 0, 1);
    double (*verts)[3];
    int32_t (*faces)[3];
    int nv, nf;
    rto_icosphere(4, &verts, &faces, &nv, &nf);
    rto_openpbr* op = (rto_openpbr*)calloc((size_t)spheres, sizeof(rto_openpbr));
    rto_triangle* tr = (rto_triangle*)calloc((size_t)nf * (size_t)spheres, sizeof(rto_triangle));
    float (*pos)[3] = (float (*)[3])malloc(sizeof(float) * 3 * (size_t)nv);
    for (int m = 0; m < spheres; ++m) {
        const double radius = 40.0 + 70.0 * (double)pcg_u(&r);
        double center[3];
        for (int k = 0; k < 3; ++k) center[k] = radius + (555.0 - 2.0 * radius) * (double)pcg_u(&r);
        openpbr_defaults(&op[m]);
        for (int ch = 0; ch < 3; ++ch) op[m].base_color[ch] = pcg_u(&r);
        op[m].specular_roughness = 0.05f + 0.85f * pcg_u(&r);
        op[m].base_metalness = pcg_u(&r) < 0.5f ? 0.0f : 1.0f;
        op[m].specular_ior = 1.5f;
        op[m].coat_weight = 0.5f * pcg_u(&r);
        mt[4 + m].type = RTO_MAT_OPENPBR;
        mt[4 + m].openpbr_index = m;
        for (int ch = 0; ch < 3; ++ch) mt[4 + m].albedo[ch] = op[m].base_color[ch];
        for (int i = 0; i < nv; ++i) {
            const float u = pcg_u(&r);
            const double rr = radius * (1.0 + 0.05 * (2.0 * (double)u - 1.0));
            for (int k = 0; k < 3; ++k) pos[i][k] = (float)(center[k] + verts[i][k] * rr);
        }
        for (int f = 0; f < nf; ++f) {
            rto_triangle* t = &tr[(size_t)m * nf + f];
            memcpy(t->p0, pos[faces[f][0]], 12);
            memcpy(t->p1, pos[faces[f][1]], 12);
            memcpy(t->p2, pos[faces[f][2]], 12);
            t->material = 4 + m;
        }
    }
    free(pos); free(verts); free(faces);
    g->scene.quad_count = 6; g->scene.quads = q;
    g->scene.triangle_count = nf * spheres; g->scene.triangles = tr;
    g->scene.material_count = 4 + spheres; g->scene.materials = mt;
    g->scene.openpbr_count = spheres; g->scene.openpbr = op;
    g->scene.background_mode = RTO_BG_CONSTANT;
    set3(g->scene.background, 0.0f, 0.0f, 0.0f);
    g->camera_count = 4;
    const double center[3] = {277.5, 277.5, 277.5}, up[3] = {0, 1, 0};
    const double azimuth[4] = {-24.0, -8.0, 8.0, 24.0}, height[4] = {100.0, -60.0, 60.0, -100.0};
    for (int c = 0; c < 4; ++c) {
        const double th = azimuth[c] * (3.14159265358979323846 / 180.0);
        const double from[3] = {center[0] + 800.0 * sin(th), center[1] + height[c],
                                center[2] - 800.0 * cos(th)};
        lookat_pinhole(&g->cameras[c], 640, 480, 40.0, from, center, up);
    }
    profile_defaults(&g->profile, 16, 8);
    return 0;
}

/* equi62_lut1d camera with the reference's default 120 degree horizontal
 * field of view (camera_models.cpp:196-226) and a 1024-entry theta LUT */
static void lookat_equi(rto_camera* c, int w, int h, const double radial[6], const double from[3],
    const double at[3], const double up[3]) {
    lookat_pinhole(c, w, h, 60.0, from, at, up);
    c->model = RTO_CAM_EQUI62_LUT1D;
    const double hfov = 120.0 * 3.14159265358979323846 / 180.0;
    const double f = 0.5 * (double)w / (0.5 * hfov);
    c->fx = f; c->fy = f; c->cx = 0.5 * (double)w; c->cy = 0.5 * (double)h;
    for (int i = 0; i < 6; ++i) c->radial[i] = radial[i];
    c->tangential[0] = 0.0; c->tangential[1] = 0.0;
    rto_make_equi_lut(c);
}

static double gauss(pcg32_t* r, double sigma) {
    const double u1 = 1.0 - (double)pcg_u(r);
    const double u2 = (double)pcg_u(r);
    return sigma * sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

/* four cameras on a radius-70 orbit around the cube centre, 90 degrees apart;
 * frame f of `frames` advances the orbit by 2*pi*f/frames */
static void instanced_rig_sized(uint64_t seed, int frame, int frames, int wp, int hp, int we,
    int he, rto_camera* cams) {
    pcg32_t r;
    pcg_seed(&r, seed ^ 0x5bd1e995ULL);
    /* k_i ~ N(0, 0.01), redrawn until theta_d(theta) is monotone on
     * [0, 1.6] rad (derivative >= 0.2) so the LUT inverse is well defined */
    double radial[2][6];
    for (int k = 0; k < 2; ++k) {
        for (int tries = 0; tries < 10000; ++tries) {
            for (int i = 0; i < 6; ++i) radial[k][i] = gauss(&r, 0.01);
            int ok = 1;
            for (int s = 0; s <= 64 && ok; ++s) {
                const double th = 1.6 * (double)s / 64.0, t2 = th * th;
                double der = 1.0, tp = t2;
                for (int i = 0; i < 6; ++i) { der += (double)(2 * i + 3) * radial[k][i] * tp; tp *= t2; }
                ok = der >= 0.2;
            }
            if (ok) break;
        }
    }
    const double center[3] = {50.0, 50.0, 50.0}, up[3] = {0.0, 1.0, 0.0};
    for (int k = 0; k < 4; ++k) {
        const double th = 2.0 * 3.14159265358979323846 * (double)frame / (double)(frames > 0 ? frames : 1)
                          + (double)k * 0.5 * 3.14159265358979323846;
        const double from[3] = {center[0] + 70.0 * cos(th), center[1] + 15.0, center[2] + 70.0 * sin(th)};
        if (k < 2) lookat_pinhole(&cams[k], wp, hp, 60.0, from, center, up);
        else lookat_equi(&cams[k], we, he, radial[k - 2], from, center, up);
    }
}

/* instanced_mesh_restir (id 2): 2 pinhole 640x480 + 2 equi 1024x1024;
 * radiance_cache_realtime (id 3): all four 1280x720 */
static void instanced_rig(uint64_t seed, int frame, int frames, rto_camera* cams) {
    instanced_rig_sized(seed, frame, frames, 640, 480, 1024, 1024, cams);
}

int rto_generate_rig(int config, uint64_t seed, int frame, int frames, rto_camera* cams, int* count) {
    if (config == 2) instanced_rig(seed, frame, frames, cams);
    else if (config == 3) instanced_rig_sized(seed, frame, frames, 1280, 720, 1280, 720, cams);
    else return -1;
    *count = 4;
    return 0;
}

static int gen_instanced_mesh(rto_generated* g, uint64_t seed) {
    pcg32_t r;
    pcg_seed(&r, seed);
    /* torus 32 x 16 segments = 1024 triangles, tube radius jittered per vertex */
    enum { NU = 32, NV = 16, NINST = 1000, NMAT = 16, NLIGHT = 256 };
    double verts[NU * NV][3];
    for (int i = 0; i < NU; ++i) {
        for (int j = 0; j < NV; ++j) {
            const double u = 2.0 * 3.14159265358979323846 * (double)i / NU;
            const double v = 2.0 * 3.14159265358979323846 * (double)j / NV;
            const double tube = 0.35 * (1.0 + 0.1 * (2.0 * (double)pcg_u(&r) - 1.0));
            verts[i * NV + j][0] = (1.0 + tube * cos(v)) * cos(u);
            verts[i * NV + j][1] = tube * sin(v);
            verts[i * NV + j][2] = (1.0 + tube * cos(v)) * sin(u);
        }
    }
    int faces[NU * NV * 2][3];
    int nf = 0;
    for (int i = 0; i < NU; ++i) {
        for (int j = 0; j < NV; ++j) {
            const int a = i * NV + j, b = ((i + 1) % NU) * NV + j;
            const int c = ((i + 1) % NU) * NV + (j + 1) % NV, d = i * NV + (j + 1) % NV;
            faces[nf][0] = a; faces[nf][1] = b; faces[nf][2] = c; ++nf;
            faces[nf][0] = a; faces[nf][1] = c; faces[nf][2] = d; ++nf;
        }
    }
    rto_material* mt = (rto_material*)calloc(NMAT, sizeof(rto_material));
    for (int m = 0; m < NMAT; ++m) {
        mt[m].type = RTO_MAT_LAMBERTIAN;
        mt[m].ior = 1.0f;
        mt[m].openpbr_index = -1;
        for (int ch = 0; ch < 3; ++ch) mt[m].albedo[ch] = 0.2f + 0.7f * pcg_u(&r);
    }
    rto_triangle* tr = (rto_triangle*)calloc((size_t)NINST * (size_t)nf, sizeof(rto_triangle));
    for (int k = 0; k < NINST; ++k) {
        /* uniform rotation from a random unit quaternion (Shoemake 1992) */
        const double u1 = (double)pcg_u(&r), u2 = (double)pcg_u(&r), u3 = (double)pcg_u(&r);
        const double s1 = sqrt(1.0 - u1), s2 = sqrt(u1);
        const double qx = s1 * sin(2.0 * 3.14159265358979323846 * u2);
        const double qy = s1 * cos(2.0 * 3.14159265358979323846 * u2);
        const double qz = s2 * sin(2.0 * 3.14159265358979323846 * u3);
        const double qw = s2 * cos(2.0 * 3.14159265358979323846 * u3);
        const double R[3][3] = {
            {1.0 - 2.0 * (qy * qy + qz * qz), 2.0 * (qx * qy - qz * qw), 2.0 * (qx * qz + qy * qw)},
            {2.0 * (qx * qy + qz * qw), 1.0 - 2.0 * (qx * qx + qz * qz), 2.0 * (qy * qz - qx * qw)},
            {2.0 * (qx * qz - qy * qw), 2.0 * (qy * qz + qx * qw), 1.0 - 2.0 * (qx * qx + qy * qy)}};
        const double t[3] = {100.0 * (double)pcg_u(&r), 100.0 * (double)pcg_u(&r), 100.0 * (double)pcg_u(&r)};
        const int mat = (int)(pcg_u(&r) * (float)NMAT) % NMAT;
        const double scale = 2.0;
        for (int f = 0; f < nf; ++f) {
            rto_triangle* o = &tr[(size_t)k * (size_t)nf + (size_t)f];
            float* dst[3] = {o->p0, o->p1, o->p2};
            for (int vi = 0; vi < 3; ++vi) {
                const double* p = verts[faces[f][vi]];
                for (int a = 0; a < 3; ++a)
                    dst[vi][a] = (float)(scale * (R[a][0] * p[0] + R[a][1] * p[1] + R[a][2] * p[2]) + t[a]);
            }
            o->material = mat;
        }
    }
    rto_analytic_light* al = (rto_analytic_light*)calloc(NLIGHT, sizeof(rto_analytic_light));
    for (int i = 0; i < NLIGHT; ++i) {
        al[i].type = RTO_ALIGHT_SPHERE;
        for (int a = 0; a < 3; ++a) al[i].position[a] = 100.0f * pcg_u(&r);
        al[i].radius = 0.1f + 0.4f * pcg_u(&r);
        const float rad = 1.0f + 19.0f * pcg_u(&r);
        set3(al[i].radiance, rad, rad, rad);
        set3(al[i].basis_x, 1.0f, 0.0f, 0.0f);
        set3(al[i].basis_y, 0.0f, 1.0f, 0.0f);
        set3(al[i].basis_z, 0.0f, 0.0f, 1.0f);
        al[i].width = al[i].height = al[i].length = 1.0f;
        al[i].cos_theta_max = 1.0f;
    }
    (void)rto_analytic_lights_compile(al, NLIGHT);
    g->scene.triangle_count = NINST * nf;
    g->scene.triangles = tr;
    g->scene.material_count = NMAT;
    g->scene.materials = mt;
    g->scene.analytic_light_count = NLIGHT;
    g->scene.analytic_lights = al;
    g->scene.background_mode = RTO_BG_CONSTANT;
    set3(g->scene.background, 0.0f, 0.0f, 0.0f);
    g->camera_count = 4;
    instanced_rig(seed, 0, 60, g->cameras);
    profile_defaults(&g->profile, 1, 2);
    g->profile.enable_restir_di = 1;
    g->profile.restir_initial_candidates = 32;
    g->profile.restir_spatial_neighbors = 5;
    g->profile.restir_spatial_radius = 30;
    return 0;
}

/* four-camera surround rig of a realtime view preset (world z up): the body
 * camera at yaw offsets 0, 90, -90, 180 degrees
 * (src/realtime/viewer/four_camera_rig.cpp:55-95, body_pose.cpp:36-85) */
static void surround_rig(const double pos[3], double yaw_deg, double pitch_deg, double fx, double fy,
    double cx, double cy, int w, int h, rto_camera* cams) {
    const double offsets[4] = {0.0, 90.0, -90.0, 180.0};
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < 4; ++i) {
        rto_camera* c = &cams[i];
        memset(c, 0, sizeof(*c));
        c->enabled = 1; c->width = w; c->height = h; c->model = RTO_CAM_PINHOLE32;
        const double yaw = (yaw_deg + offsets[i]) * pi / 180.0;
        const double pitch = -pitch_deg * pi / 180.0;
        /* Rx(yaw) Ry(-pitch) (0, 0, -1), then the body basis columns
         * (0,0,1), (-1,0,0), (0,-1,0) */
        const double b[3] = {-sin(pitch), sin(yaw) * cos(pitch), -cos(yaw) * cos(pitch)};
        double f[3] = {-b[1], -b[2], b[0]};
        const double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        for (int k = 0; k < 3; ++k) f[k] /= fl;
        double r[3] = {f[1], -f[0], 0.0};   /* forward x up */
        const double rl = sqrt(r[0] * r[0] + r[1] * r[1]);
        if (rl * rl < 1e-12) {
            r[0] = cos(yaw); r[1] = sin(yaw); r[2] = 0.0;
        } else {
            for (int k = 0; k < 3; ++k) r[k] /= rl;
        }
        double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
        const double ul = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (int k = 0; k < 3; ++k) u[k] /= ul;
        for (int k = 0; k < 3; ++k) {
            c->origin[k] = pos[k]; c->basis_x[k] = r[k]; c->basis_y[k] = -u[k]; c->basis_z[k] = f[k];
        }
        c->fx = fx; c->fy = fy; c->cx = cx; c->cy = cy;
    }
}

/* legacy renderer axes to world: (x, y, z) -> (x, -z, y) */
static void legacy_quad(rto_quad* q, double ox, double oy, double oz, double ux, double uy, double uz,
    double vx, double vy, double vz, int mat) {
    quad_set(q, (float)ox, (float)-oz, (float)oy, (float)ux, (float)-uz, (float)uy, (float)vx,
        (float)-vz, (float)vy, mat);
}

static void legacy_sphere(rto_sphere* s, double x, double y, double z, double radius, int mat) {
    set3(s->center, (float)x, (float)-z, (float)y);
    s->radius = (float)radius;
    s->material = mat;
}

static rto_material* builtin_materials(int n) {
    rto_material* mt = (rto_material*)calloc((size_t)n, sizeof(rto_material));
    for (int i = 0; i < n; ++i) { mt[i].openpbr_index = -1; mt[i].ior = 1.0f; }
    return mt;
}

static void finish_builtin(rto_generated* g, rto_sphere* sp, int ns, rto_quad* q, int nq,
    rto_material* mt, int nm) {
    g->scene.sphere_count = ns; g->scene.spheres = sp;
    g->scene.quad_count = nq; g->scene.quads = q;
    g->scene.material_count = nm; g->scene.materials = mt;
    g->scene.background_mode = RTO_BG_CONSTANT;
    set3(g->scene.background, 0.0f, 0.0f, 0.0f);
    g->camera_count = 4;
    profile_defaults(&g->profile, 16, 8);
}

/* smoke (id 5): the reference's builtin "smoke" scene (shared_scene_builders.cpp:475-487)
 * and view (body at (0, -2.5, 0.25), yaw 0, fx = fy = 360) */
static int gen_smoke(rto_generated* g) {
    rto_material* mt = builtin_materials(2);
    set3(mt[0].albedo, 0.75f, 0.25f, 0.2f);
    mt[1].type = RTO_MAT_DIFFUSE_LIGHT; set3(mt[1].emission, 10.0f, 10.0f, 10.0f);
    rto_sphere* sp = (rto_sphere*)calloc(1, sizeof(rto_sphere));
    legacy_sphere(&sp[0], 0.0, 0.0, -1.0, 0.5, 0);
    rto_quad* q = (rto_quad*)calloc(1, sizeof(rto_quad));
    legacy_quad(&q[0], -0.75, 1.25, -1.5, 1.5, 0.0, 0.0, 0.0, 0.0, 1.5, 1);
    finish_builtin(g, sp, 1, q, 1, mt, 2);
    const double pos[3] = {0.0, -2.5, 0.25};
    surround_rig(pos, 0.0, 0.0, 360.0, 360.0, 320.0, 240.0, 640, 480, g->cameras);
    return 0;
}

/* final_room (id 6): the reference's builtin "final_room" (shared_scene_builders.cpp:489-550)
 * and view (default spawn pose (0, -0.8, 0.35), fx = 480, fy = 360) */
static int gen_final_room(rto_generated* g) {
    rto_material* mt = builtin_materials(5);
    set3(mt[0].albedo, 0.73f, 0.73f, 0.73f);
    set3(mt[1].albedo, 0.30f, 0.70f, 0.35f);
    set3(mt[2].albedo, 0.72f, 0.25f, 0.22f);
    set3(mt[3].albedo, 0.25f, 0.35f, 0.75f);
    mt[4].type = RTO_MAT_DIFFUSE_LIGHT; set3(mt[4].emission, 10.0f, 10.0f, 10.0f);
    rto_quad* q = (rto_quad*)calloc(9, sizeof(rto_quad));
    legacy_quad(&q[0], -4.0, -1.0, -4.0, 0.0, 0.0, 8.0, 8.0, 0.0, 0.0, 0);
    legacy_quad(&q[1], -4.0, 3.5, -4.0, 8.0, 0.0, 0.0, 0.0, 0.0, 8.0, 0);
    legacy_quad(&q[2], -4.0, -1.0, -4.0, 0.0, 4.5, 0.0, 0.0, 0.0, 8.0, 1);
    legacy_quad(&q[3], 4.0, -1.0, -4.0, 0.0, 0.0, 8.0, 0.0, 4.5, 0.0, 2);
    legacy_quad(&q[4], -4.0, -1.0, -4.0, 8.0, 0.0, 0.0, 0.0, 4.5, 0.0, 0);
    legacy_quad(&q[5], -4.0, -1.0, 4.0, 0.0, 4.5, 0.0, 8.0, 0.0, 0.0, 3);
    legacy_quad(&q[6], -1.0, 3.15, -1.0, 2.0, 0.0, 0.0, 0.0, 0.0, 2.0, 4);
    legacy_quad(&q[7], -3.2, -0.25, -3.0, 0.0, 0.0, 1.8, 1.8, 0.0, 0.0, 0);
    legacy_quad(&q[8], 1.2, 0.15, 1.0, 0.0, 0.0, 1.6, 1.6, 0.0, 0.0, 0);
    rto_sphere* sp = (rto_sphere*)calloc(8, sizeof(rto_sphere));
    legacy_sphere(&sp[0], 0.0, 0.25, -1.2, 0.55, 0);
    legacy_sphere(&sp[1], -1.6, 0.35, 1.7, 0.55, 0);
    legacy_sphere(&sp[2], -3.1, 1.0, 0.8, 0.55, 0);
    legacy_sphere(&sp[3], 3.0, 1.35, -0.9, 0.65, 0);
    legacy_sphere(&sp[4], 1.1, 1.1, -3.0, 0.60, 0);
    legacy_sphere(&sp[5], -0.8, 2.55, 2.2, 0.45, 0);
    legacy_sphere(&sp[6], 0.9, 0.55, -0.1, 0.45, 0);
    legacy_sphere(&sp[7], -0.35, 0.4, -1.15, 0.35, 0);
    finish_builtin(g, sp, 8, q, 9, mt, 5);
    const double pos[3] = {0.0, -0.8, 0.35};
    surround_rig(pos, 0.0, 0.0, 480.0, 360.0, 320.0, 240.0, 640, 480, g->cameras);
    return 0;
}

int rto_generate_scene(int config, uint64_t seed, rto_generated* g) {
    memset(g, 0, sizeof(*g));
    switch (config) {
        case 0: return gen_random_spheres(g, seed);
        case 1: return gen_cornell_mesh(g, seed);
        case 2: return gen_instanced_mesh(g, seed);
        case 4: return gen_cornell_random(g, seed);
        case 5: return gen_smoke(g);
        case 6: return gen_final_room(g);
        case 3: {
            /* the instanced_mesh_restir scene, four 1280x720 cameras, paths of depth 8 with
             * Russian roulette from bounce 3 (the radiance-cache workload) */
            const int rc = gen_instanced_mesh(g, seed);
            if (rc != 0) return rc;
            instanced_rig_sized(seed, 0, 60, 1280, 720, 1280, 720, g->cameras);
            g->profile.max_bounces = 8;
            g->profile.rr_start_bounce = 3;
            return 0;
        }
        default: return -1;
    }
}

void rto_free_generated(rto_generated* g) {
    free((void*)g->scene.spheres); free((void*)g->scene.quads); free((void*)g->scene.triangles);
    free((void*)g->scene.materials); free((void*)g->scene.openpbr); free((void*)g->scene.analytic_lights);
    memset(g, 0, sizeof(*g));
}
