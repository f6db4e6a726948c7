// rt_math.h -- device FP32 vector math, counter RNG and elementary functions
// for the gfx950 path tracer.
//
// The vector helpers keep the operation order of the reference's
// src/realtime/gpu/device_math.h:19-88 (div3 multiplies by the reciprocal,
// normalize3 returns 0 for length <= 1e-8).  All HIP code is compiled with
// -ffp-contract=off and correctly rounded division/sqrt, so a given sequence
// of operations yields the IEEE-754 result on the GPU and on any other
// conforming implementation of the same sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtk {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator*(f3 v, float s) { return mk3(v.x * s, v.y * s, v.z * s); }
__device__ __forceinline__ f3 operator-(f3 v) { return mk3(-v.x, -v.y, -v.z); }
__device__ __forceinline__ f3 div3(f3 v, float s) {
    const float inv = 1.0f / s;
    return mk3(v.x * inv, v.y * inv, v.z * inv);
}
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float lensq3(f3 v) { return dot3(v, v); }
__device__ __forceinline__ float len3(f3 v) { return sqrtf(lensq3(v)); }
__device__ __forceinline__ f3 normalize3(f3 v) {
    const float l = len3(v);
    if (l <= 1e-8f) return mk3(0.0f, 0.0f, 0.0f);
    return div3(v, l);
}
__device__ __forceinline__ float maxc3(f3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
__device__ __forceinline__ f3 clamp3(f3 v, float lo, float hi) {
    return mk3(fminf(fmaxf(v.x, lo), hi), fminf(fmaxf(v.y, lo), hi), fminf(fmaxf(v.z, lo), hi));
}
__device__ __forceinline__ f3 reflect3(f3 i, f3 n) { return i - n * (2.0f * dot3(i, n)); }
__device__ __forceinline__ f3 refract3(f3 i, f3 n, float eta) {
    const float cos_theta = fminf(dot3(-i, n), 1.0f);
    const f3 perp = (i + n * cos_theta) * eta;
    const float k = 1.0f - lensq3(perp);
    const f3 par = n * (-sqrtf(fmaxf(k, 0.0f)));
    return perp + par;
}
__device__ __forceinline__ bool near_zero3(f3 v) {
    return fabsf(v.x) < 1e-7f && fabsf(v.y) < 1e-7f && fabsf(v.z) < 1e-7f;
}
__device__ __forceinline__ f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
__device__ __forceinline__ f3 xyz(float4 v) { return mk3(v.x, v.y, v.z); }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// select-based min/max: the sign of zero follows the operand order, which
// keeps BVH bounds bit-identical to a sequential build
__device__ __forceinline__ float sel_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float sel_max(float a, float b) { return a > b ? a : b; }

// ---- counter RNG: pcg4d (Jarzynski & Olano, JCGT 9(3), 2020).
// key = (pixel, sample, camera<<16 | bounce<<6 | group, stream); one hash
// gives four 24-bit uniforms in [0, 1).
__device__ __forceinline__ uint4 pcg4d(uint4 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
    v.x ^= v.x >> 16; v.y ^= v.y >> 16; v.z ^= v.z >> 16; v.w ^= v.w >> 16;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
    return v;
}
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
__device__ __forceinline__ float4 rng4(uint32_t pixel, uint32_t sample, uint32_t camera,
    uint32_t bounce, uint32_t group, uint32_t stream) {
    const uint4 h = pcg4d(make_uint4(pixel, sample, (camera << 16) | (bounce << 6) | group, stream));
    return make_float4(u01(h.x), u01(h.y), u01(h.z), u01(h.w));
}
enum : uint32_t { kGroupLight = 0, kGroupBsdf = 1, kGroupAnalytic = 2, kGroupCache = 3, kGroupSubsurface = 4, kGroupRestir0 = 8,
    kGroupMedium0 = 48, kGroupCamera = 63 };

// sin/cos of a whole turn u (angle 2*pi*u) by quadrant reduction in turns
// (exact) and Taylor polynomials on [-pi/4, pi/4] (truncation < 3e-11)
__host__ __device__ __forceinline__ void sincos_turn(float u, float* s, float* c) {
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
    if (quadrant == 0) { *s = sp; *c = cp; }
    else if (quadrant == 1) { *s = cp; *c = -sp; }
    else if (quadrant == 2) { *s = -sp; *c = -cp; }
    else { *s = -cp; *c = sp; }
}
// cube root on [0, 1]: exponent-division seed, four Newton steps
__device__ __forceinline__ float cbrt01(float x) {
    if (!(x > 0.0f)) return 0.0f;
    float y = __uint_as_float(__float_as_uint(x) / 3u + 709921077u);
    for (int it = 0; it < 4; ++it) y = (2.0f * y + x / (y * y)) * 0.333333343f;
    return y;
}
__device__ __forceinline__ f3 sphere_dir(float u0, float u1) {
    const float z = 1.0f - 2.0f * u0;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float s, c;
    sincos_turn(u1, &s, &c);
    return mk3(r * c, r * s, z);
}

constexpr float kPi = 3.14159265358979323846f;

// Natural logarithm by a fixed operation sequence (the Cephes logf method:
// x = m 2^e with m in [sqrt(1/2), sqrt(2)), a degree-9 polynomial in m - 1
// and the split ln 2 = 0.693359375 - 2.12194440e-4), so the GPU and the CPU
// restatement round alike; for finite normal x > 0 (free-flight sampling
// uses x in [1e-6, 1]).
__device__ __forceinline__ float log_fixed(float x) {
    const uint32_t b = __float_as_uint(x);
    int e = (int)((b >> 23) & 0xffu) - 126;
    float m = __uint_as_float((b & 0x807fffffu) | 0x3f000000u);   // [0.5, 1)
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
constexpr float kRayEps = 1e-3f;
constexpr float kRayFar = 1e30f;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr float kUniformSpherePdf = 0.07957747154594767f;

}  // namespace rtk
