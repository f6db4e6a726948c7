// rt_texture.h -- texture evaluation at a surface hit: the scene's table of
// PackedTexture records (constant, checker, image, turbulence noise) and the
// hit's texture coordinates.
#pragma once

#include "rt_internal.h"
#include "rt_math.h"

namespace rtk {

// kTextureResolveDepth (src/realtime/gpu/programs.cu:94)
constexpr int kTextureResolveDepth = 8;
constexpr float kTexPi = 3.14159265358979323846f;

// sin(x) as a fixed sequence of FP32 operations, so that the noise texture
// is the same on every device and in every restatement: reduction by pi/2
// in three parts (Cody and Waite; 1.5703125 has 8 significant bits, so
// k * 1.5703125 is exact for |k| < 2^16) and the single-precision minimax
// polynomials of the Cephes library's sinf / cosf (S. L. Moshier) on
// [-pi/4, pi/4].  A non-finite argument gives NaN.
__device__ __forceinline__ float noise_sin(float x) {
    if (!isfinite(x)) return __builtin_nanf("");
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

// hash_noise / turbulence_noise (programs.cu:194-208)
__device__ __forceinline__ float hash_noise(f3 p) {
    const float n = noise_sin(dot3(p, mk3(12.9898f, 78.233f, 37.719f))) * 43758.5453f;
    return (n - floorf(n)) * 2.0f - 1.0f;
}

__device__ __forceinline__ float turbulence(f3 p) {
    float accum = 0.0f, weight = 1.0f;
    for (int i = 0; i < 7; ++i) {
        accum += weight * hash_noise(p);
        p = p * 2.0f;
        weight *= 0.5f;
    }
    return fabsf(accum);
}

// address_texture_coordinate (programs.cu:210-223): 0 border, 1 clamp,
// 2 repeat, 3 mirror
__device__ __forceinline__ float address_coord(float c, int mode, bool& valid) {
    if (mode == 0 && (c < 0.0f || c > 1.0f)) {
        valid = false;
        return 0.0f;
    }
    if (mode == 2) return c - floorf(c);
    if (mode == 3) {
        const float period = c - 2.0f * floorf(0.5f * c);
        return period <= 1.0f ? period : 2.0f - period;
    }
    return fminf(fmaxf(c, 0.0f), 1.0f);
}

// image_texel (programs.cu:225-234): clamped to the image, cyan outside the
// texel array
__device__ __forceinline__ f3 image_texel(const SceneView& s, const rt_texture& t, int x, int y) {
    x = x < 0 ? 0 : (x >= t.image_width ? t.image_width - 1 : x);
    y = y < 0 ? 0 : (y >= t.image_height ? t.image_height - 1 : y);
    const int64_t i = (int64_t)t.image_offset + (int64_t)y * t.image_width + x;
    if (i < 0 || i >= s.texel_count) return mk3(0.0f, 1.0f, 1.0f);
    const float* p = s.texels + 3 * i;
    return mk3(p[0], p[1], p[2]);
}

// sample_image_texture (programs.cu:236-266): v runs up the image
__device__ inline f3 sample_image(const SceneView& s, const rt_texture& t, float u, float v) {
    if (s.texels == nullptr || t.image_width <= 0 || t.image_height <= 0) return mk3(0.0f, 1.0f, 1.0f);
    bool valid = true;
    const float au = address_coord(u, t.u_address_mode, valid);
    const float av = 1.0f - address_coord(v, t.v_address_mode, valid);
    if (!valid) return mk3(0.0f, 0.0f, 0.0f);
    if (t.filter_type == 0) {
        int x = (int)(au * (float)t.image_width);
        int y = (int)(av * (float)t.image_height);
        x = x >= t.image_width ? t.image_width - 1 : x;
        y = y >= t.image_height ? t.image_height - 1 : y;
        return image_texel(s, t, x, y);
    }
    const float x = au * (float)(t.image_width - 1);
    const float y = av * (float)(t.image_height - 1);
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const float tx = x - (float)x0, ty = y - (float)y0;
    const f3 top = image_texel(s, t, x0, y0) * (1.0f - tx) + image_texel(s, t, x0 + 1, y0) * tx;
    const f3 bottom = image_texel(s, t, x0, y0 + 1) * (1.0f - tx) + image_texel(s, t, x0 + 1, y0 + 1) * tx;
    return top * (1.0f - ty) + bottom * ty;
}

// floor(v) as an int, clamped to +-2^30 so the checker's cell index stays
// defined for any position
__device__ __forceinline__ int cell_index(float v) {
    float f = floorf(v);
    if (!(f >= -1073741824.0f)) f = -1073741824.0f;
    if (f > 1073741824.0f) f = 1073741824.0f;
    return (int)f;
}

// evaluate_texture (programs.cu:268-302): checkers resolve through their
// even / odd children at most kTextureResolveDepth deep; a bad index or a
// deeper chain gives magenta
__device__ inline f3 eval_texture(const SceneView& s, int index, float u, float v, f3 p) {
    for (int depth = 0; depth < kTextureResolveDepth; ++depth) {
        if (index < 0 || index >= s.texture_count || s.textures == nullptr) return mk3(1.0f, 0.0f, 1.0f);
        const rt_texture t = s.textures[index];
        if (t.type == RT_TEX_CONSTANT) return ld3(t.color);
        if (t.type == RT_TEX_CHECKER) {
            if (fabsf(t.scale) < 1e-8f) {
                index = t.even_texture;
                continue;
            }
            const int x = cell_index(p.x / t.scale), y = cell_index(p.y / t.scale),
                      z = cell_index(p.z / t.scale);
            index = (((x ^ y ^ z) & 1) == 0) ? t.even_texture : t.odd_texture;
            continue;
        }
        if (t.type == RT_TEX_IMAGE) return sample_image(s, t, u, v);
        if (t.type == RT_TEX_NOISE) {
            const float value = 0.5f * (1.0f + noise_sin(t.scale * p.z + 10.0f * turbulence(p)));
            return mk3(value, value, value);
        }
        return mk3(1.0f, 0.0f, 1.0f);
    }
    return mk3(1.0f, 0.0f, 1.0f);
}

// sphere_uv (programs.cu:188-193) of a unit outward normal
__device__ __forceinline__ void sphere_uv(f3 n, float* u, float* v) {
    const float theta = acosf(fminf(fmaxf(-n.y, -1.0f), 1.0f));
    const float phi = atan2f(-n.z, n.x) + kTexPi;
    *u = phi / (2.0f * kTexPi);
    *v = theta / kTexPi;
}

// texture coordinates of a hit at pos = o + d * t (programs.cu:1447,1485,
// 1524-1530): sphere_uv of the outward normal; the quad's edge coordinates;
// the triangle's barycentrics (recomputed with the intersectors'
// arithmetic), interpolating its per-vertex coordinates when it has them
__device__ inline void surface_uv(const SceneView& s, f3 o, f3 d, int prim_type, int prim_index,
    f3 pos, float* u, float* v) {
    if (prim_type == 0) {
        const rt_sphere sp = s.spheres[prim_index];
        sphere_uv(div3(pos - ld3(sp.center), sp.radius), u, v);
    } else if (prim_type == 1) {
        const rt_quad q = s.quads[prim_index];
        const f3 eu = ld3(q.edge_u), ev = ld3(q.edge_v);
        const f3 n = cross3(eu, ev);
        const f3 w = div3(n, lensq3(n));
        const f3 rel = pos - ld3(q.origin);
        *u = dot3(w, cross3(rel, ev));
        *v = dot3(w, cross3(eu, rel));
    } else {
        const rt_triangle tr = s.triangles[prim_index];
        const f3 p0 = ld3(tr.p0);
        const f3 e1 = ld3(tr.p1) - p0, e2 = ld3(tr.p2) - p0;
        const f3 pvec = cross3(d, e2);
        const float inv_det = 1.0f / dot3(e1, pvec);
        const f3 tvec = o - p0;
        const float bu = dot3(tvec, pvec) * inv_det;
        const float bv = dot3(d, cross3(tvec, e1)) * inv_det;
        *u = bu;
        *v = bv;
        if (s.tri_uv != nullptr) {
            const float* t = s.tri_uv + 6 * (int64_t)prim_index;
            if (t[0] == t[0]) {   // not NaN: has_texcoords
                const float w = 1.0f - bu - bv;
                *u = t[0] * w + t[2] * bu + t[4] * bv;
                *v = t[1] * w + t[3] * bu + t[5] * bv;
            }
        }
    }
}

}  // namespace rtk
