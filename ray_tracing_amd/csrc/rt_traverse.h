// rt_traverse.h -- closest-hit and any-hit traversal of the 4-wide quantized
// BVH, with sphere / quad / triangle intersection.
//
// Replaces intersect_accelerated_surfaces + try_hit_* of the reference
// (src/realtime/gpu/programs.cu:1421-1547, 1549-1652), which walks a
// per-thread 64-entry private stack, tests one box per node fetch, runs
// material evaluation for every candidate hit and has no any-hit early exit.
// Here one 64-byte node fetch yields four quantized child boxes; leaves are
// ranges of up to two 48-byte primitive records in Morton order; the
// deferred-node stack keeps its top entries in LDS (lane-strided, so the 64
// lanes of a wave touch distinct banks) and spills deeper ones to HBM;
// popped entries farther than the current best hit are skipped; lanes that
// finish refill from the ray queue.  Materials are evaluated once, after
// the nearest hit, by the shade stage.
//
// Closest-hit result: the lexicographic minimum of (t, primitive id) over all
// primitives with tmin < t < tmax, independent of traversal order.
#pragma once

#include "rt_internal.h"
#include "rt_math.h"

namespace rtk {

__device__ __forceinline__ float safe_inv(float d) {
    const float dv = fabsf(d) > 1e-12f ? d : copysignf(1e-12f, d);
    return 1.0f / dv;
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

// one primitive record (48 bytes, loaded whole so the type branch does not
// wait on a second round trip)
struct PrimRec {
    float4 a, b, c;
};

__device__ __forceinline__ PrimRec load_prim(const float4* leaves, uint32_t leaf) {
    return PrimRec{leaves[3u * leaf], leaves[3u * leaf + 1u], leaves[3u * leaf + 2u]};
}

__device__ __forceinline__ float hit_prim(const PrimRec& p, f3 o, f3 d, float tmin, float tmax,
    uint32_t* id, float* u, float* v) {
    const uint32_t tag = __float_as_uint(p.a.w);
    *id = tag & kIdMask;
    const uint32_t type = tag >> 30;
    *u = 0.0f;
    *v = 0.0f;
    if (type == kPrimTriangle) return hit_triangle(xyz(p.a), xyz(p.b), xyz(p.c), o, d, tmin, tmax, u, v);
    if (type == kPrimSphere) return hit_sphere(xyz(p.a), p.b.x, o, d, tmin, tmax);
    return hit_quad(xyz(p.a), xyz(p.b), xyz(p.c), o, d, tmin, tmax, u, v);
}

__device__ __forceinline__ float hit_leaf(const float4* leaves, uint32_t leaf, f3 o, f3 d,
    float tmin, float tmax, uint32_t* id, float* u, float* v) {
    return hit_prim(load_prim(leaves, leaf), o, d, tmin, tmax, id, u, v);
}

struct TraceResult {
    float t;
    uint32_t prim;
    float u, v;
    bool found;
};

// box tests and primitive tests of one traversal (algorithmic traffic)
struct TraverseCounts {
    uint32_t visits;  // child boxes tested (one canonical BVH node each)
    uint32_t tests;
    uint32_t sphere_tests;
};

// ---------------------------------------------------------------------------
// 4-wide quantized BVH traversal.
//
// Order (shared with the CPU oracle so node / primitive counts agree): at a
// node all valid child boxes are tested against the current best distance;
// leaf ranges that pass are intersected in slot order; internal children
// that pass are ranked by entry distance (ties: lower slot first), the
// nearest is visited next and the others are pushed farthest first; popped
// entries farther than the best hit are skipped.
//
// Per-lane deferred-node stack: entries 0..kStackLds-1 in LDS (lane-strided
// uint2 = node, entry distance), deeper entries in a per-thread HBM spill
// area.  A 4-wide tree of L levels never holds more than 3 (L - 1) entries.

struct TraceStack {
    uint2* lds;             // this lane's entry 0; entry k at lds[k * BLOCK]
    uint2* spill;           // this thread's entry kStackLds; then spill_stride apart
    uint32_t spill_stride;
};

struct WideRay {
    f3 o, d;
    float ix, iy, iz;
    float tmin, tmax;
    bool nx, ny, nz;   // the inverse direction is negative: the box's upper plane is entered first
    bool any;          // any-hit query: the first hit ends the traversal
};

__device__ __forceinline__ WideRay make_wide_ray(f3 o, f3 d, float tmin, float tmax,
    bool any = false) {
    WideRay r;
    r.o = o;
    r.d = d;
    r.ix = safe_inv(d.x);
    r.iy = safe_inv(d.y);
    r.iz = safe_inv(d.z);
    r.tmin = tmin;
    r.tmax = tmax;
    r.nx = r.ix < 0.0f;
    r.ny = r.iy < 0.0f;
    r.nz = r.iz < 0.0f;
    r.any = any;
    return r;
}

struct WideState {
    uint32_t node;
    int sp;
    TraceResult best;
};

__device__ __forceinline__ void wide_begin(WideState& w, float tmax) {
    w.node = 0u;
    w.sp = 0;
    w.best = TraceResult{tmax, 0u, 0.0f, 0.0f, false};
}

template <int BLOCK>
__device__ __forceinline__ void stack_push(const TraceStack& st, int& sp, int limit, uint32_t node,
    float tn) {
    if (sp >= limit) return;  // cannot happen for a well-formed tree (3 (L-1) < limit)
    const uint2 e = make_uint2(node, __float_as_uint(tn));
    if (sp < kStackLds) st.lds[sp * BLOCK] = e;
    else st.spill[(uint32_t)(sp - kStackLds) * st.spill_stride] = e;
    ++sp;
}

template <int BLOCK>
__device__ __forceinline__ uint2 stack_pop(const TraceStack& st, int& sp) {
    --sp;
    return sp < kStackLds ? st.lds[sp * BLOCK] : st.spill[(uint32_t)(sp - kStackLds) * st.spill_stride];
}

__device__ __forceinline__ float qbyte(uint32_t word, int k) {
    return (float)((word >> (8 * k)) & 255u);
}

// branch-free: the comparisons combine with bitwise operators and the
// record updates are selects, so diverging lanes add no exec-mask branches
__device__ __forceinline__ void consider_hit(TraceResult& b, uint32_t id, float t, float u, float v) {
    const bool take = (t >= 0.0f) & ((t < b.t) | ((t == b.t) & b.found & (id < b.prim)));
    b.t = take ? t : b.t;
    b.prim = take ? id : b.prim;
    b.u = take ? u : b.u;
    b.v = take ? v : b.v;
    b.found = b.found | take;
}

// One node visit (then pops until the next node).  Returns true when the
// traversal is finished.
template <int BLOCK, bool ANY, bool COUNT>
__device__ __forceinline__ bool wide_step(const SceneView& s, const WideRay& r, WideState& w,
    const TraceStack& st, TraverseCounts& cnt) {
    const float4* np = s.nodes4 + 4u * w.node;
    const float4 v0 = np[0];
    const uint4 v1 = reinterpret_cast<const uint4*>(np)[1];
    const uint4 v2 = reinterpret_cast<const uint4*>(np)[2];
    const uint4 v3 = reinterpret_cast<const uint4*>(np)[3];
    const uint32_t meta = __float_as_uint(v0.w);
    if (COUNT) cnt.visits += meta >> 24;
    const float sx = r.ix * __uint_as_float((meta & 255u) << 23);
    const float sy = r.iy * __uint_as_float(((meta >> 8) & 255u) << 23);
    const float sz = r.iz * __uint_as_float(((meta >> 16) & 255u) << 23);
    const float ox = (v0.x - r.o.x) * r.ix;
    const float oy = (v0.y - r.o.y) * r.iy;
    const float oz = (v0.z - r.o.z) * r.iz;
    const float slack = 4e-7f * ((fabsf(ox) + fabsf(oy)) + fabsf(oz));
    const uint32_t ch[4] = {v1.x, v1.y, v1.z, v1.w};
    // per axis, the quantized plane the ray enters first: the lower bound
    // when the inverse direction is positive, else the upper one.  fma is
    // monotone in q for a fixed scale, so this selects exactly the
    // min / max of the two plane distances.
    const uint32_t lx = r.nx ? v2.w : v2.x, hx = r.nx ? v2.x : v2.w;
    const uint32_t ly = r.ny ? v3.x : v2.y, hy = r.ny ? v2.y : v3.x;
    const uint32_t lz = r.nz ? v3.y : v2.z, hz = r.nz ? v2.z : v3.y;
    float tn[4];
    bool hit[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float nx = fmaf(qbyte(lx, k), sx, ox), fx = fmaf(qbyte(hx, k), sx, ox);
        const float ny = fmaf(qbyte(ly, k), sy, oy), fy = fmaf(qbyte(hy, k), sy, oy);
        const float nz = fmaf(qbyte(lz, k), sz, oz), fz = fmaf(qbyte(hz, k), sz, oz);
        const float n = fmaxf(fmaxf(nx, ny), fmaxf(nz, r.tmin));
        const float f = fminf(fminf(fx, fy), fminf(fz, w.best.t)) * 1.0000024f + slack;
        tn[k] = n;
        hit[k] = (ch[k] != kEmpty) & (n <= f);
    }
    // leaf ranges in slot order, as one per-lane loop over the primitives of
    // every hit leaf (a wave runs the longest lane's list once instead of
    // one loop per slot)
    uint32_t leaves = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) leaves |= (hit[k] & ((ch[k] & kLeafBit) != 0u)) ? (1u << k) : 0u;
    uint32_t prim = 0u, left = 0u;
    while (leaves != 0u || left != 0u) {
        if (left == 0u) {
            const int k = __builtin_ctz(leaves);
            leaves &= leaves - 1u;
            const uint32_t c = k == 0 ? ch[0] : k == 1 ? ch[1] : k == 2 ? ch[2] : ch[3];
            prim = c & 0x0FFFFFFFu;
            left = ((c >> 28) & 7u) + 1u;
        }
        uint32_t id;
        float u, v;
        const float t = hit_leaf(s.leaves, prim, r.o, r.d, r.tmin, r.tmax, &id, &u, &v);
        if (COUNT) {
            ++cnt.tests;
            if ((int)id < s.sphere_count) ++cnt.sphere_tests;
        }
        consider_hit(w.best, id, t, u, v);
        if ((ANY || r.any) && w.best.found) return true;
        ++prim;
        --left;
    }
    // internal children: stable rank by entry distance (ties: lower slot
    // first), computed without branches: for j < k child j goes first when
    // its distance is not larger, for j > k when it is smaller
    int in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) in[k] = (int)(hit[k] & ((ch[k] & kLeafBit) == 0u));
    int rank[4];
    int m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int rk = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j == k) continue;
            rk += in[j] & (int)(j < k ? tn[j] <= tn[k] : tn[j] < tn[k]);
        }
        rank[k] = in[k] ? rk : 4;
        m += in[k];
    }
    if (m > 0) {
#pragma unroll
        for (int rr = 3; rr >= 1; --rr) {
            if (rr < m) {
                uint32_t ref = 0u;
                float key = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    ref = rank[k] == rr ? ch[k] : ref;
                    key = rank[k] == rr ? tn[k] : key;
                }
                stack_push<BLOCK>(st, w.sp, s.stack4_entries, ref, key);
            }
        }
        uint32_t next = 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k) next = rank[k] == 0 ? ch[k] : next;
        w.node = next;
        return false;
    }
    while (w.sp > 0) {
        const uint2 e = stack_pop<BLOCK>(st, w.sp);
        if (__uint_as_float(e.y) <= w.best.t) {
            w.node = e.x;
            return false;
        }
    }
    return true;
}

// one complete traversal by one lane
template <int BLOCK, bool ANY, bool COUNT = false>
__device__ __forceinline__ TraceResult trace_wide(const SceneView& s, f3 o, f3 d, float tmin,
    float tmax, const TraceStack& st, TraverseCounts* cnt = nullptr) {
    TraverseCounts c{0u, 0u, 0u};
    WideState w;
    wide_begin(w, tmax);
    if (s.prim_count > 0 && s.node4_count > 0) {
        const WideRay r = make_wide_ray(o, d, tmin, tmax);
        while (!wide_step<BLOCK, ANY, COUNT>(s, r, w, st, c)) {
        }
    }
    if (COUNT && cnt) *cnt = c;
    return w.best;
}

__device__ __forceinline__ uint64_t lanes_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Persistent traversal over a queue of `count` rays.  Each wave claims the
// queue in batches (one atomic per batch) and hands them to its idle lanes
// once at least `refill_min` lanes are idle: refill_min = 64 runs whole
// coherent batches (classic persistent threads), smaller values let a lane
// take a new ray as soon as its own is done.  A batch is 64 entries, or 128
// when the queue holds at least 32 batches of 64 per wave of the grid: the
// claims are same-address atomics, which serialise at one memory channel and
// limit the large launches, while a larger batch would leave a longer tail
// of unclaimed work at the end of the small ones.
// PARTS > 1 splits the queue into PARTS contiguous partitions with a cursor
// each (fetch[p * kCursorStride], one cache line apart): a wave claims from
// partition blockIdx % PARTS, then moves on to the next ones as they drain.
// Workgroups are dispatched to the XCDs round robin, so with PARTS = 8 each
// XCD starts on its own eighth of the queue: neighbouring rays (pixels) meet
// the same BVH nodes in one L2, and the claims spread over eight addresses.
//   load(q)               : ray of queue entry q
//   done(q, result, cnt)  : store the result of queue entry q
constexpr uint32_t kCursorStride = 64;  // words between partition cursors
template <int BLOCK, bool ANY, bool COUNT, int PARTS = 1, class Load, class Done>
__device__ __forceinline__ void trace_queue(const SceneView& s, uint32_t count, uint32_t* fetch,
    const TraceStack& st, int refill_min, Load load, Done done) {
    const int lane = (int)(threadIdx.x & 63u);
    bool active = false;
    bool exhausted = false;
    uint32_t batch_base = 0u, batch_left = 0u;  // wave-uniform
    uint32_t q = 0u;
    WideRay r{};
    WideState w{};
    TraverseCounts c{0u, 0u, 0u};
    const bool has_tree = s.prim_count > 0 && s.node4_count > 0;
    if (count == 0u) return;  // no wave touches the fetch counter of an empty queue
    const uint32_t waves = gridDim.x * (uint32_t)(BLOCK / 64);
    const uint32_t batch = count / 64u / waves >= 32u ? 128u : 64u;
    uint32_t part = PARTS > 1 ? blockIdx.x % (uint32_t)PARTS : 0u;  // wave-uniform
    uint32_t parts_left = (uint32_t)PARTS;
    while (true) {
        const uint64_t idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (nidle >= refill_min || nidle == 64) {
            while (batch_left == 0u && !exhausted) {
                // claim a batch of the current partition; a partition whose
                // cursor has passed its end sends the wave on to the next one
                const uint32_t lo = (uint32_t)((uint64_t)count * part / (uint32_t)PARTS);
                const uint32_t n = (uint32_t)((uint64_t)count * (part + 1u) / (uint32_t)PARTS) - lo;
                uint32_t base = 0u;
                if (lane == 0) base = atomicAdd(fetch + part * kCursorStride, batch);
                base = (uint32_t)__shfl((int)base, 0);
                batch_base = lo + base;
                batch_left = base < n ? (n - base < batch ? n - base : batch) : 0u;
                if (base + batch >= n) {
                    part = part + 1u == (uint32_t)PARTS ? 0u : part + 1u;
                    if (--parts_left == 0u) exhausted = true;
                }
            }
            const uint32_t take = (uint32_t)nidle < batch_left ? (uint32_t)nidle : batch_left;
            if (take > 0u) {
                const uint32_t rank = (uint32_t)__popcll(idle & lanes_below(lane));
                if (!active && rank < take) {
                    q = batch_base + rank;
                    r = load(q);
                    wide_begin(w, r.tmax);
                    c = TraverseCounts{0u, 0u, 0u};
                    active = true;
                    if (!has_tree) {
                        done(q, w.best, c);
                        active = false;
                    }
                }
                batch_base += take;
                batch_left -= take;
            }
        }
        if (__ballot(active) == 0ull) {
            if (exhausted && batch_left == 0u) break;
            continue;
        }
        if (active && wide_step<BLOCK, ANY, COUNT>(s, r, w, st, c)) {
            done(q, w.best, c);
            active = false;
        }
    }
}

}  // namespace rtk
