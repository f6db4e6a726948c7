/*
 * oracle_texture.c -- CPU restatement of texture evaluation at a surface
 * hit: the PackedTexture table (constant, checker, image, turbulence noise)
 * and the hit's texture coordinates.  TEST INFRASTRUCTURE ONLY (see
 * rt_oracle.h).  Same FP32 operation order as the HIP path.
 */
#include "rt_oracle.h"
#include "om.h"

#define TEX_RESOLVE_DEPTH 8   /* kTextureResolveDepth, programs.cu:94 */
static const float kTexPi = 3.14159265358979323846f;

/* sin(x) as a fixed FP32 operation sequence: three-part Cody-Waite
 * reduction by pi/2 and the Cephes sinf / cosf minimax polynomials
 * (S. L. Moshier) on [-pi/4, pi/4].  Replaces sinf in the noise texture
 * (programs.cu:194-197,296) so devices and restatements agree bit for bit. */
float rto_noise_sin(float x) {
    if (!isfinite(x)) return NAN;
    const float k = rintf(x * 0.636619772367581343f);
    const float r = ((x - k * 1.5703125f) - k * 4.837512969970703125e-4f) - k * 7.54978995489188216e-8f;
    const float z = r * r;
    const float sp = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
    const float cp = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
                   - 0.5f * z + 1.0f;
    int q = (int)fmodf(k, 4.0f);
    if (q < 0) q += 4;
    return q == 0 ? sp : (q == 1 ? cp : (q == 2 ? -sp : -cp));
}

/* hash_noise / turbulence_noise, programs.cu:194-208 */
static float hash_noise(v3 p) {
    const float n = rto_noise_sin(dot3(p, mk3(12.9898f, 78.233f, 37.719f))) * 43758.5453f;
    return (n - floorf(n)) * 2.0f - 1.0f;
}

float rto_turbulence(const float p_in[3]) {
    v3 p = ld3(p_in);
    float accum = 0.0f, weight = 1.0f;
    for (int i = 0; i < 7; ++i) {
        accum += weight * hash_noise(p);
        p = scl3(p, 2.0f);
        weight *= 0.5f;
    }
    return fabsf(accum);
}

/* address_texture_coordinate, programs.cu:210-223 */
static float address_coord(float c, int mode, int* valid) {
    if (mode == 0 && (c < 0.0f || c > 1.0f)) {
        *valid = 0;
        return 0.0f;
    }
    if (mode == 2) return c - floorf(c);
    if (mode == 3) {
        const float period = c - 2.0f * floorf(0.5f * c);
        return period <= 1.0f ? period : 2.0f - period;
    }
    return fminf(fmaxf(c, 0.0f), 1.0f);
}

/* image_texel, programs.cu:225-234 */
static v3 image_texel(const rto_scene* s, const rto_texture* t, int x, int y) {
    x = x < 0 ? 0 : (x >= t->image_width ? t->image_width - 1 : x);
    y = y < 0 ? 0 : (y >= t->image_height ? t->image_height - 1 : y);
    const int64_t i = (int64_t)t->image_offset + (int64_t)y * t->image_width + x;
    if (i < 0 || i >= s->image_texel_count) return mk3(0.0f, 1.0f, 1.0f);
    return ld3(s->image_texels + 3 * i);
}

/* sample_image_texture, programs.cu:236-266 */
static v3 sample_image(const rto_scene* s, const rto_texture* t, float u, float v) {
    if (s->image_texels == NULL || t->image_width <= 0 || t->image_height <= 0) return mk3(0.0f, 1.0f, 1.0f);
    int valid = 1;
    const float au = address_coord(u, t->u_address_mode, &valid);
    const float av = 1.0f - address_coord(v, t->v_address_mode, &valid);
    if (!valid) return mk3(0.0f, 0.0f, 0.0f);
    if (t->filter_type == 0) {
        int x = (int)(au * (float)t->image_width);
        int y = (int)(av * (float)t->image_height);
        x = x >= t->image_width ? t->image_width - 1 : x;
        y = y >= t->image_height ? t->image_height - 1 : y;
        return image_texel(s, t, x, y);
    }
    const float x = au * (float)(t->image_width - 1);
    const float y = av * (float)(t->image_height - 1);
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const float tx = x - (float)x0, ty = y - (float)y0;
    const v3 top = add3(scl3(image_texel(s, t, x0, y0), 1.0f - tx), scl3(image_texel(s, t, x0 + 1, y0), tx));
    const v3 bottom = add3(scl3(image_texel(s, t, x0, y0 + 1), 1.0f - tx),
                           scl3(image_texel(s, t, x0 + 1, y0 + 1), tx));
    return add3(scl3(top, 1.0f - ty), scl3(bottom, ty));
}

/* floor(v) as an int clamped to +-2^30 */
static int cell_index(float v) {
    float f = floorf(v);
    if (!(f >= -1073741824.0f)) f = -1073741824.0f;
    if (f > 1073741824.0f) f = 1073741824.0f;
    return (int)f;
}

/* evaluate_texture, programs.cu:268-302 ((x + y + z) % 2 == 0 is the
 * parity of x ^ y ^ z, which cannot overflow) */
void rto_eval_texture(const rto_scene* s, int index, float u, float v, const float p_in[3], float out[3]) {
    const v3 p = ld3(p_in);
    v3 r = mk3(1.0f, 0.0f, 1.0f);
    for (int depth = 0; depth < TEX_RESOLVE_DEPTH; ++depth) {
        if (index < 0 || index >= s->texture_count || s->textures == NULL) break;
        const rto_texture* t = &s->textures[index];
        if (t->type == 0) { r = ld3(t->color); break; }
        if (t->type == 1) {
            if (fabsf(t->scale) < 1e-8f) {
                index = t->even_texture;
                continue;
            }
            const int x = cell_index(p.x / t->scale), y = cell_index(p.y / t->scale),
                      z = cell_index(p.z / t->scale);
            index = (((x ^ y ^ z) & 1) == 0) ? t->even_texture : t->odd_texture;
            continue;
        }
        if (t->type == 2) { r = sample_image(s, t, u, v); break; }
        if (t->type == 3) {
            const float value = 0.5f * (1.0f + rto_noise_sin(t->scale * p.z + 10.0f * rto_turbulence(p_in)));
            r = mk3(value, value, value);
            break;
        }
        break;
    }
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* texture coordinates of a hit at pos = o + d t (programs.cu:1447,1485,
 * 1528); prim_type 0 sphere, 1 quad, 2 triangle */
void rto_surface_uv(const rto_scene* s, const float o_in[3], const float d_in[3], int prim_type,
    int prim_index, const float pos_in[3], float* u, float* v) {
    const v3 o = ld3(o_in), d = ld3(d_in), pos = ld3(pos_in);
    if (prim_type == 0) {
        const rto_sphere* sp = &s->spheres[prim_index];
        const v3 n = div3(sub3(pos, ld3(sp->center)), sp->radius);
        const float theta = acosf(fminf(fmaxf(-n.y, -1.0f), 1.0f));
        const float phi = atan2f(-n.z, n.x) + kTexPi;
        *u = phi / (2.0f * kTexPi);
        *v = theta / kTexPi;
    } else if (prim_type == 1) {
        const rto_quad* q = &s->quads[prim_index];
        const v3 eu = ld3(q->edge_u), ev = ld3(q->edge_v);
        const v3 n = cross3(eu, ev);
        const v3 w = div3(n, lensq3(n));
        const v3 rel = sub3(pos, ld3(q->origin));
        *u = dot3(w, cross3(rel, ev));
        *v = dot3(w, cross3(eu, rel));
    } else {
        const rto_triangle* tr = &s->triangles[prim_index];
        const v3 p0 = ld3(tr->p0);
        const v3 e1 = sub3(ld3(tr->p1), p0), e2 = sub3(ld3(tr->p2), p0);
        const v3 pvec = cross3(d, e2);
        const float inv_det = 1.0f / dot3(e1, pvec);
        const v3 tvec = sub3(o, p0);
        const float bu = dot3(tvec, pvec) * inv_det;
        const float bv = dot3(d, cross3(tvec, e1)) * inv_det;
        *u = bu;
        *v = bv;
        if (s->triangle_uvs != NULL) {
            /* uv0 w + uv1 u + uv2 v, programs.cu:1524-1530 */
            const float* t = s->triangle_uvs + 6 * (int64_t)prim_index;
            if (t[0] == t[0]) {
                const float w = 1.0f - bu - bv;
                *u = t[0] * w + t[2] * bu + t[4] * bv;
                *v = t[1] * w + t[3] * bu + t[5] * bv;
            }
        }
    }
}
