// rt_traverse.h -- stack-in-LDS closest-hit and any-hit traversal of the
// binary LBVH, with sphere / quad / triangle intersection.
//
// Replaces intersect_accelerated_surfaces + try_hit_* of the reference
// (src/realtime/gpu/programs.cu:1421-1547, 1549-1652), which walks a
// per-thread 64-entry private stack, tests one box per node fetch, runs
// material evaluation for every candidate hit and has no any-hit early exit.
// Here one 64-byte node fetch yields both child boxes; leaves hold one
// primitive record (48 B) inline in Morton order; the deferred-node stack
// (node index + entry distance) lives in LDS, lane-strided so that the 64
// lanes of a wave hit 64 distinct banks; popped entries farther than the
// current best hit are skipped.  Materials are evaluated once, after the
// nearest hit, by the shade stage.
//
// Closest-hit result: the lexicographic minimum of (t, primitive id) over all
// primitives with tmin < t < tmax, independent of traversal order.
#pragma once

#include "rt_internal.h"
#include "rt_math.h"

namespace rtk {

struct SlabPre {
    float ix, iy, iz;
    float ox, oy, oz;  // origin * inverse direction
};

__device__ __forceinline__ float safe_inv(float d) {
    const float dv = fabsf(d) > 1e-12f ? d : copysignf(1e-12f, d);
    return 1.0f / dv;
}

__device__ __forceinline__ SlabPre slab_setup(f3 o, f3 d) {
    SlabPre p;
    p.ix = safe_inv(d.x);
    p.iy = safe_inv(d.y);
    p.iz = safe_inv(d.z);
    p.ox = o.x * p.ix;
    p.oy = o.y * p.iy;
    p.oz = o.z * p.iz;
    return p;
}

// box entry/exit by fma slabs; the exit distance is widened by 8e-7
// relative so a box never culls a primitive hit lying on its surface
__device__ __forceinline__ bool slab_hit(const SlabPre& p, float lx, float ly, float lz, float hx,
    float hy, float hz, float tmin, float tbest, float* tnear) {
    const float t0x = fmaf(lx, p.ix, -p.ox), t1x = fmaf(hx, p.ix, -p.ox);
    const float t0y = fmaf(ly, p.iy, -p.oy), t1y = fmaf(hy, p.iy, -p.oy);
    const float t0z = fmaf(lz, p.iz, -p.oz), t1z = fmaf(hz, p.iz, -p.oz);
    const float nx = fminf(t0x, t1x), ny = fminf(t0y, t1y), nz = fminf(t0z, t1z);
    const float fx = fmaxf(t0x, t1x), fy = fmaxf(t0y, t1y), fz = fmaxf(t0z, t1z);
    const float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
    const float tf = fminf(fminf(fx, fy), fminf(fz, tbest)) * 1.0000008f;
    *tnear = tn;
    return tn <= tf;
}

// programs.cu:1421-1449
__device__ __forceinline__ float hit_sphere(f3 c, float radius, f3 o, f3 d, float tmin, float tmax) {
    const f3 oc = o - c;
    const float a = lensq3(d);
    const float half_b = dot3(oc, d);
    const float cc = lensq3(oc) - radius * radius;
    const float disc = half_b * half_b - a * cc;
    if (disc <= 0.0f) return -1.0f;
    const float sq = sqrtf(disc);
    float root = (-half_b - sq) / a;
    if (root <= tmin || root >= tmax) {
        root = (-half_b + sq) / a;
        if (root <= tmin || root >= tmax) return -1.0f;
    }
    return root;
}

// programs.cu:1451-1488
__device__ __forceinline__ float hit_quad(f3 qo, f3 eu, f3 ev, f3 o, f3 d, float tmin, float tmax,
    float* pu, float* pv) {
    const f3 n = cross3(eu, ev);
    const float nls = lensq3(n);
    if (nls <= 1e-12f) return -1.0f;
    const f3 normal = div3(n, sqrtf(nls));
    const float denom = dot3(normal, d);
    if (fabsf(denom) < 1e-6f) return -1.0f;
    const float t = dot3(qo - o, normal) / denom;
    if (t <= tmin || t >= tmax) return -1.0f;
    const f3 p = o + d * t;
    const f3 rel = p - qo;
    const f3 w = div3(n, nls);
    const float u = dot3(w, cross3(rel, ev));
    const float v = dot3(w, cross3(eu, rel));
    if (u < 0.0f || u > 1.0f || v < 0.0f || v > 1.0f) return -1.0f;
    *pu = u;
    *pv = v;
    return t;
}

// programs.cu:1490-1520 (Moller-Trumbore; edges precomputed at build time)
__device__ __forceinline__ float hit_triangle(f3 p0, f3 e1, f3 e2, f3 o, f3 d, float tmin,
    float tmax, float* pu, float* pv) {
    const f3 pvec = cross3(d, e2);
    const float det = dot3(e1, pvec);
    if (fabsf(det) <= 1e-8f) return -1.0f;
    const float inv_det = 1.0f / det;
    const f3 tvec = o - p0;
    const float u = dot3(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const f3 qvec = cross3(tvec, e1);
    const float v = dot3(d, qvec) * inv_det;
    if (v < 0.0f || (u + v) > 1.0f) return -1.0f;
    const float t = dot3(e2, qvec) * inv_det;
    if (t <= tmin || t >= tmax) return -1.0f;
    *pu = u;
    *pv = v;
    return t;
}

__device__ __forceinline__ float hit_leaf(const float4* leaves, uint32_t leaf, f3 o, f3 d,
    float tmin, float tmax, uint32_t* id, float* u, float* v) {
    const float4 a = leaves[3u * leaf];
    const uint32_t tag = __float_as_uint(a.w);
    *id = tag & kIdMask;
    const uint32_t type = tag >> 30;
    *u = 0.0f;
    *v = 0.0f;
    if (type == kPrimTriangle) {
        const float4 b = leaves[3u * leaf + 1u];
        const float4 c = leaves[3u * leaf + 2u];
        return hit_triangle(xyz(a), xyz(b), xyz(c), o, d, tmin, tmax, u, v);
    }
    if (type == kPrimSphere) {
        const float4 b = leaves[3u * leaf + 1u];
        return hit_sphere(xyz(a), b.x, o, d, tmin, tmax);
    }
    const float4 b = leaves[3u * leaf + 1u];
    const float4 c = leaves[3u * leaf + 2u];
    return hit_quad(xyz(a), xyz(b), xyz(c), o, d, tmin, tmax, u, v);
}

struct TraceResult {
    float t;
    uint32_t prim;
    float u, v;
    bool found;
};

// node visits and primitive tests of one traversal (algorithmic traffic)
struct TraverseCounts {
    uint32_t visits;
    uint32_t tests;
    uint32_t sphere_tests;
};

// ANY = true: stop at the first hit in (tmin, tmax).
// COUNT = true: report node visits / primitive tests in *cnt.
// stack_node / stack_t: this lane's slot 0 in LDS; entry k at [k * stride].
template <bool ANY, bool COUNT = false>
__device__ __forceinline__ TraceResult traverse(const SceneView& s, f3 o, f3 d, float tmin,
    float tmax, uint32_t* stack_node, float* stack_t, int stride, TraverseCounts* cnt = nullptr) {
    TraceResult r{tmax, 0u, 0.0f, 0.0f, false};
    uint32_t nv = 0, nt = 0, ns = 0;
    if (s.prim_count > 0) {
        const SlabPre p = slab_setup(o, d);
        int sp = 0;
        uint32_t node = 0;
        const float4* nodes = s.nodes;
        while (true) {
            const float4 n0 = nodes[4u * node];
            const float4 n1 = nodes[4u * node + 1u];
            const float4 n2 = nodes[4u * node + 2u];
            const uint4 n3 = reinterpret_cast<const uint4*>(nodes)[4u * node + 3u];
            ++nv;
            float tn0 = 0.0f, tn1 = 0.0f;
            const bool h0 = n3.x != kEmpty
                            && slab_hit(p, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, tmin, r.t, &tn0);
            const bool h1 = n3.y != kEmpty
                            && slab_hit(p, n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, tmin, r.t, &tn1);
            const bool l0 = h0 && (n3.x & kLeafBit);
            const bool l1 = h1 && (n3.y & kLeafBit);
            if (l0) {
                uint32_t id;
                float u, v;
                ++nt;
                const float t = hit_leaf(s.leaves, n3.x & ~kLeafBit, o, d, tmin, tmax, &id, &u, &v);
                if (COUNT && (int)id < s.sphere_count) ++ns;
                if (t >= 0.0f && (t < r.t || (t == r.t && r.found && id < r.prim))) {
                    r.t = t; r.prim = id; r.u = u; r.v = v; r.found = true;
                }
                if (ANY && r.found) break;
            }
            if (l1) {
                uint32_t id;
                float u, v;
                ++nt;
                const float t = hit_leaf(s.leaves, n3.y & ~kLeafBit, o, d, tmin, tmax, &id, &u, &v);
                if (COUNT && (int)id < s.sphere_count) ++ns;
                if (t >= 0.0f && (t < r.t || (t == r.t && r.found && id < r.prim))) {
                    r.t = t; r.prim = id; r.u = u; r.v = v; r.found = true;
                }
                if (ANY && r.found) break;
            }
            const bool i0 = h0 && !(n3.x & kLeafBit);
            const bool i1 = h1 && !(n3.y & kLeafBit);
            if (i0 && i1) {
                const bool near1 = tn1 < tn0;
                stack_node[sp * stride] = near1 ? n3.x : n3.y;
                stack_t[sp * stride] = near1 ? tn0 : tn1;
                ++sp;
                node = near1 ? n3.y : n3.x;
                continue;
            }
            if (i0) { node = n3.x; continue; }
            if (i1) { node = n3.y; continue; }
            bool got = false;
            while (sp > 0) {
                --sp;
                if (stack_t[sp * stride] <= r.t) {
                    node = stack_node[sp * stride];
                    got = true;
                    break;
                }
            }
            if (!got) break;
        }
    }
    if (COUNT) {
        cnt->visits = nv;
        cnt->tests = nt;
        cnt->sphere_tests = ns;
    }
    return r;
}

}  // namespace rtk
