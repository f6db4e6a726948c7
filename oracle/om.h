/*
 * om.h -- oracle-internal FP32 vector math, RNG and elementary functions.
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 *
 * Vector helpers restate src/realtime/gpu/device_math.h:19-88 operation for
 * operation (div3 multiplies by the reciprocal; normalize3 returns 0 when
 * the length is <= 1e-8), so that, compiled with -ffp-contract=off, every
 * result is the IEEE-754 result of the same operation sequence.
 */
#ifndef RT_ORACLE_OM_H
#define RT_ORACLE_OM_H

#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float x, y, z; } v3;

static inline v3 mk3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld3(const float* p) { v3 r = {p[0], p[1], p[2]}; return r; }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl3(v3 v, float s) { return mk3(v.x * s, v.y * s, v.z * s); }
static inline v3 div3(v3 v, float s) { const float inv = 1.0f / s; return mk3(v.x * inv, v.y * inv, v.z * inv); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float lensq3(v3 v) { return dot3(v, v); }
static inline float len3(v3 v) { return sqrtf(lensq3(v)); }
static inline v3 norm3(v3 v) {
    const float l = len3(v);
    if (l <= 1e-8f) return mk3(0.0f, 0.0f, 0.0f);
    return div3(v, l);
}
static inline v3 neg3(v3 v) { return mk3(-v.x, -v.y, -v.z); }
/* fmaxf/fminf on non-NaN operands (device_math.h:80-88) */
static inline float maxf_(float a, float b) { return fmaxf(a, b); }
static inline float minf_(float a, float b) { return fminf(a, b); }
static inline float maxc3(v3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
static inline v3 clamp3(v3 v, float lo, float hi) {
    return mk3(fminf(fmaxf(v.x, lo), hi), fminf(fmaxf(v.y, lo), hi), fminf(fmaxf(v.z, lo), hi));
}
static inline v3 reflect3(v3 i, v3 n) { return sub3(i, scl3(n, 2.0f * dot3(i, n))); }
static inline v3 refract3(v3 i, v3 n, float eta) {
    const float cos_theta = fminf(dot3(neg3(i), n), 1.0f);
    const v3 perp = scl3(add3(i, scl3(n, cos_theta)), eta);
    const float k = 1.0f - lensq3(perp);
    const v3 par = scl3(n, -sqrtf(fmaxf(k, 0.0f)));
    return add3(perp, par);
}
static inline int near_zero3(v3 v) {
    return fabsf(v.x) < 1e-7f && fabsf(v.y) < 1e-7f && fabsf(v.z) < 1e-7f;
}
/* selects used where the sign of zero must be deterministic (BVH bounds) */
static inline float sel_min(float a, float b) { return a < b ? a : b; }
static inline float sel_max(float a, float b) { return a > b ? a : b; }

/* ---- counter RNG: pcg4d, Jarzynski & Olano, "Hash Functions for GPU
 * Rendering", JCGT 9(3), 2020.  Key = (pixel, sample, camera<<16 |
 * bounce<<6 | group, stream); one hash yields four 24-bit uniforms. ---- */
static inline void om_pcg4d(uint32_t v[4]) {
    v[0] = v[0] * 1664525u + 1013904223u;
    v[1] = v[1] * 1664525u + 1013904223u;
    v[2] = v[2] * 1664525u + 1013904223u;
    v[3] = v[3] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
    v[0] ^= v[0] >> 16; v[1] ^= v[1] >> 16; v[2] ^= v[2] >> 16; v[3] ^= v[3] >> 16;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
}
static inline float om_u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
static inline void om_rng4(uint32_t pixel, uint32_t sample, uint32_t camera, uint32_t bounce,
    uint32_t group, uint32_t stream, float out[4]) {
    uint32_t v[4] = {pixel, sample, (camera << 16) | (bounce << 6) | group, stream};
    om_pcg4d(v);
    out[0] = om_u01(v[0]); out[1] = om_u01(v[1]); out[2] = om_u01(v[2]); out[3] = om_u01(v[3]);
}
/* RNG groups within one bounce */
enum { OM_G_LIGHT = 0, OM_G_BSDF = 1, OM_G_ALIGHT = 2, OM_G_SUBSURFACE = 4, OM_G_RESTIR0 = 8, OM_G_MEDIUM0 = 48,
       OM_G_CAMERA = 63 };

/* ---- sin/cos of a whole turn, own Taylor polynomials.
 * Returns sin(2*pi*u), cos(2*pi*u).  q = 4u is exact, k = nearest quadrant,
 * f = q - k in [-0.5, 0.5] is exact, r = f*pi/2 in [-pi/4, pi/4]; Taylor
 * series to r^11 (sin) and r^12 (cos), truncation error < 3e-11. ---- */
static inline void om_sincos_turn(float u, float* s, float* c) {
    const float q = u * 4.0f;
    const float k = floorf(q + 0.5f);
    const float f = q - k;
    const float r = f * 1.57079637f;
    const float r2 = r * r;
    const float sp = r * (1.0f + r2 * (-0.166666672f + r2 * (0.00833333377f
                         + r2 * (-0.000198412701f + r2 * (2.75573188e-06f
                         + r2 * (-2.50521079e-08f))))));
    const float cp = 1.0f + r2 * (-0.5f + r2 * (0.0416666679f + r2 * (-0.00138888892f
                         + r2 * (2.48015876e-05f + r2 * (-2.75573188e-07f + r2 * 2.08767570e-09f)))));
    const int quadrant = ((int)k) & 3;
    switch (quadrant) {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}
/* cube root of x in [0, 1]: Kahan's exponent-division seed + 4 Newton steps */
static inline float om_cbrt01(float x) {
    if (!(x > 0.0f)) return 0.0f;
    uint32_t i;
    memcpy(&i, &x, 4);
    i = i / 3u + 709921077u;
    float y;
    memcpy(&y, &i, 4);
    for (int it = 0; it < 4; ++it) y = (2.0f * y + x / (y * y)) * 0.333333343f;
    return y;
}
/* uniform direction on the unit sphere from two uniforms */
/* the fixed-sequence natural logarithm of the GPU (Cephes logf method:
 * x = m 2^e, m in [sqrt(1/2), sqrt(2)), degree-9 polynomial in m - 1) */
static inline float om_log_fixed(float x) {
    uint32_t b;
    memcpy(&b, &x, sizeof(b));
    int e = (int)((b >> 23) & 0xffu) - 126;
    const uint32_t mb = (b & 0x807fffffu) | 0x3f000000u;
    float m;
    memcpy(&m, &mb, sizeof(m));
    if (m < 0.70710678118654752f) {
        e -= 1;
        m = m + m - 1.0f;
    } else {
        m = m - 1.0f;
    }
    const float z = m * m;
    float y = 7.0376836292e-2f;
    y = y * m - 1.1514610310e-1f;
    y = y * m + 1.1676998740e-1f;
    y = y * m - 1.2420140846e-1f;
    y = y * m + 1.4249322787e-1f;
    y = y * m - 1.6668057665e-1f;
    y = y * m + 2.0000714765e-1f;
    y = y * m - 2.4999993993e-1f;
    y = y * m + 3.3333331174e-1f;
    y = y * m * z;
    const float fe = (float)e;
    y = y + -2.12194440e-4f * fe;
    y = y + -0.5f * z;
    float r = m + y;
    r = r + 0.693359375f * fe;
    return r;
}

static inline v3 om_sphere_dir(float u0, float u1) {
    const float z = 1.0f - 2.0f * u0;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float s, c;
    om_sincos_turn(u1, &s, &c);
    return mk3(r * c, r * s, z);
}

#define OM_PI 3.14159265358979323846f
#define OM_RAY_EPS 1e-3f
#define OM_RAY_FAR 1e30f
#define OM_LEAF_BIT 0x80000000u
#define OM_EMPTY 0xFFFFFFFFu

#endif
