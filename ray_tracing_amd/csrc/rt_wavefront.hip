// rt_wavefront.hip -- GPU-resident wavefront path tracer for gfx950.
//
// Replaces the reference's per-camera megakernel radiance_kernel
// (src/realtime/gpu/programs.cu:1828-1953, one thread = one pixel running
// spp x bounces with traversal, shading and shadow rays inlined) with stage
// kernels over SoA path state in HBM:
//   raygen     : one path slot per (camera, pixel, sample) of a chunk, every
//                camera of the rig in the same launch
//   extend     : closest hit of every active path (stack-in-LDS traversal)
//   shade      : emission + MIS, next-event estimation (enqueues one shadow
//                ray with its precomputed contribution), BSDF sampling,
//                Russian roulette, compaction into the next active queue
//   connect    : any-hit with early exit for the shadow queue; adds the
//                contribution of unoccluded rays
//   accumulate : per pixel, sums the chunk's samples in sample order
//   finalize   : averages and encodes beauty / normal / albedo / depth
// extend, shade and connect are persistent: a fixed grid of a few blocks per
// CU strides over the queue whose length it reads from device memory, so the
// per-frame launch sequence is fixed and can be captured in a hipGraph.
// Floating-point operations per path happen in the same order as the
// single-path integrator, so a path's radiance does not depend on the queue
// order in which the stages visit it.
#include <atomic>

#include "rt_camera.h"
#include "rt_kernels.h"
#include "rt_restir.h"
#include "rt_shading.h"
#include "rt_traverse.h"

namespace rtk {

namespace {

enum : uint32_t { kCntQueue0 = 0, kCntQueue1 = 1, kCntShadow = 2, kCntShadow2 = 3,
    kCntStatExtend = 4, kCntStatShadow = 6, kCntVisits = 8, kCntTests = 10, kCntSphereTests = 12,
    kCntRestir = 14,
    kCntCacheQuery = 20 /* cache queries of the chunk */, kCntCacheRecord = 21 /* training
    records of the frame */, kCntCacheDone = 22 /* records with targets */,
    kCntCacheQueries = 23 /* cache queries of the frame */,
    kCntAnyVisits = 24, kCntAnyTests = 26, kCntAnySphereTests = 28 /* u64 traversal stats of
    the any-hit rays */, kCntPbr = 30 /* OpenPBR hits handed from the legacy-material shade
    pass to the OpenPBR pass */, kCounterWords = 31,
    kCntCursor = 64 /* kTraceParts ray-fetch cursors of the traversal launch, kCursorStride
    words apart */ };
constexpr int kTraceParts = 8;
static_assert(kCntCursor + kTraceParts * kCursorStride <= kCounterAlloc, "cursor layout");
static_assert(kCntCacheRecord == kStatCacheRecords && kCntCacheQueries == kStatCacheQueries
    && kCounterWords == kStatWords, "counter layout");

// prev_flags: bit 0 valid, bit 1 delta, bits 8-11 material type + 1 of the
// previous vertex, bits 12-30 the OpenPBR index + 1 of the subsurface medium
// the path is inside (0: none)
enum : int32_t { kPrevValid = 1, kPrevDelta = 2, kPrevMedium = 12 };

// feature bits of the shade kernel's variants
// (kShadePbrFull: every OpenPBR lobe, not only the opaque core;
// kShadePbrTex: OpenPBR inputs bound to textures)
enum : uint32_t { kShadePbr = 1, kShadeAnalytic = 2, kShadeCache = 4, kShadeTexture = 8,
    kShadeAll = 15, kShadePbrFull = 16, kShadePbrTex = 32 };

struct SlotId {
    uint32_t pixel;   // pixel index inside its camera
    uint32_t sample;
    uint32_t camera;        // slot in the rig (pixel layout, history)
    uint32_t global_pixel;
    uint32_t key;           // the slot's camera key and sample stream (RNG identity)
    uint32_t stream;
};

// the frame's per-slot RNG identity, read once per wave (uniform loads)
struct SlotRng {
    uint4 key;
    uint4 stream;
};
__device__ __forceinline__ SlotRng slot_rng(const FrameParams& fp) {
    SlotRng r;
    r.key = make_uint4(fp.dyn->camera_key[0], fp.dyn->camera_key[1], fp.dyn->camera_key[2],
        fp.dyn->camera_key[3]);
    r.stream = make_uint4(fp.dyn->stream[0], fp.dyn->stream[1], fp.dyn->stream[2], fp.dyn->stream[3]);
    return r;
}
__device__ __forceinline__ uint32_t pick4(uint4 v, uint32_t c) {
    return c == 0u ? v.x : (c == 1u ? v.y : (c == 2u ? v.z : v.w));
}

__device__ __forceinline__ SlotId slot_id(const CameraView& cams, const FrameParams& fp,
    const SlotRng& rng, uint32_t slot) {
    const uint32_t P = (uint32_t)fp.total_pixels;
    const uint32_t layer = slot / P;
    const uint32_t p = slot - layer * P;
    uint32_t c = 0;
    while (c + 1 < (uint32_t)cams.count && p >= (uint32_t)cams.pixel_offset[c + 1]) ++c;
    SlotId id;
    id.pixel = p - (uint32_t)cams.pixel_offset[c];
    id.sample = (uint32_t)fp.first_sample + layer;
    id.camera = c;
    id.global_pixel = p;
    id.key = pick4(rng.key, c);
    id.stream = pick4(rng.stream, c);
    return id;
}

// wave-aggregated append: one atomic per wave
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0ull) return 0u;
    const int lane = (int)(threadIdx.x & 63u);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = (uint32_t)__shfl((int)base, leader);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    return base + (uint32_t)__popcll(mask & lt);
}

// Queue appends of a 256-thread block for N queues at once: one atomic per
// queue per block, all N in flight together, instead of one per queue per
// wave in sequence (the shade stage's appends are same-address atomics that
// serialise at one memory channel). Lanes keep their order inside a wave and
// waves their order inside the block. Every thread of the block must call it
// (the grid-stride loops give every wave the same trip count); `lds` holds
// four rows of N words.
template <int N>
__device__ __forceinline__ void block_append(const bool (&pred)[N], uint32_t* counters,
    const uint32_t (&counter)[N], uint32_t (&pos)[N], uint32_t (*lds)[N]) {
    const int lane = (int)(threadIdx.x & 63u);
    const int w = (int)(threadIdx.x >> 6);
    uint64_t m[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        m[i] = __ballot(pred[i]);
        if (lane == 0) lds[w][i] = (uint32_t)__popcll(m[i]);
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)N) {
        const int i = (int)threadIdx.x;
        uint32_t ci = 0u;
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (j == i) ci = counter[j];
        const uint32_t c0 = lds[0][i], c1 = lds[1][i], c2 = lds[2][i], c3 = lds[3][i];
        const uint32_t total = c0 + c1 + c2 + c3;
        const uint32_t base = total != 0u ? atomicAdd(counters + ci, total) : 0u;
        lds[0][i] = base;
        lds[1][i] = base + c0;
        lds[2][i] = base + c0 + c1;
        lds[3][i] = base + c0 + c1 + c2;
    }
    __syncthreads();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int i = 0; i < N; ++i) pos[i] = lds[w][i] + (uint32_t)__popcll(m[i] & lt);
}

__global__ void k_reset_counters(uint32_t* counters) {
    if (threadIdx.x < kCounterWords) counters[threadIdx.x] = 0u;
    if (threadIdx.x < kTraceParts) counters[kCntCursor + threadIdx.x * kCursorStride] = 0u;
}

__global__ void k_clear_accum(int pixels, float4* a, float4* b, float4* c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    a[i] = z;
    b[i] = z;
    c[i] = z;
}

// zero a buffer of n 32-bit words (a fill kernel rather than a memset node,
// so the frame's stage sequence can be captured in a graph)
__global__ void k_zero_words(uint32_t* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ __launch_bounds__(256) void k_raygen(CameraView cams, FrameParams fp, WavefrontBuffers b,
    int paths) {
    const SlotRng slot_rng_v = slot_rng(fp);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kTraceParts) b.counters[kCntCursor + i * kCursorStride] = 0u;
    if (i == 0) {
        b.counters[kCntQueue0] = (uint32_t)paths;
        b.counters[kCntQueue1] = 0u;
        b.counters[kCntShadow] = 0u;
        b.counters[kCntShadow2] = 0u;
        b.counters[kCntRestir] = 0u;
        b.counters[kCntCacheQuery] = 0u;
    }
    if (i >= paths) return;
    const SlotId id = slot_id(cams, fp, slot_rng_v, (uint32_t)i);
    const rt_camera& cam = cams.cams[id.camera];
    const int w = cams.width[id.camera];
    const int y = (int)id.pixel / w;
    const int x = (int)id.pixel - y * w;
    double px, py;
    if (fp.jitter) {
        const float4 r = rng4(id.pixel, id.sample, id.key, 0u, kGroupCamera, id.stream);
        px = (double)((float)x + r.x);
        py = (double)((float)y + r.y);
    } else {
        px = (double)((float)x + 0.5f);
        py = (double)((float)y + 0.5f);
    }
    f3 o, d;
    camera_ray(cam, px, py, &o, &d);
    b.ray_o[i] = f4(o, 0.0f);
    b.ray_d[i] = f4(d, 0.0f);
    b.throughput[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    b.radiance[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    b.prev_flags[i] = kPrevDelta;
    b.prev_n[i] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    b.aux_normal[i] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    b.aux_albedo[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    b.queue[0][i] = (uint32_t)i;
}

template <int BLOCK>
__device__ __forceinline__ TraceStack lane_stack(uint2* lds, uint2* spill) {
    TraceStack st;
    st.lds = lds + threadIdx.x;
    st.spill_stride = gridDim.x * BLOCK;
    st.spill = spill + (size_t)blockIdx.x * BLOCK + threadIdx.x;
    return st;
}

// One traversal launch per bounce: the closest-hit extension rays of this
// bounce (path queue `cur`, when `ext` is set) and the any-hit shadow rays
// the previous bounce queued (emissive next-event queue, then the analytic /
// ReSTIR queue), claimed from one fetch cursor in that order.  A bounce then
// costs one launch, whose length is set by its slowest ray, instead of
// three.  Hits go to the path slots; a shadow ray's visibility goes to the w
// of its contribution, which k_apply_shadows adds in queue order.
template <int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace_fused(SceneView s, WavefrontBuffers b, int cur,
    int ext, int refill_min) {
    __shared__ uint2 lds[kStackLds * BLOCK];
    const uint32_t ne = ext ? b.counters[cur] : 0u;
    const uint32_t n1 = b.counters[kCntShadow];
    const uint32_t n2 = b.counters[kCntShadow2];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kCntStatExtend),
            (unsigned long long)ne);
        atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kCntStatShadow),
            (unsigned long long)(n1 + n2));
    }
    const TraceStack st = lane_stack<BLOCK>(lds, b.spill);
    const uint32_t* queue = b.queue[cur];
    TraverseCounts tot{0u, 0u, 0u}, tot_any{0u, 0u, 0u};
    trace_queue<BLOCK, false, COUNT, kTraceParts>(s, ne + n1 + n2, b.counters + kCntCursor, st,
        refill_min,
        [&](uint32_t q) {
            if (q < ne) {
                const uint32_t slot = queue[q];
                return make_wide_ray(xyz(b.ray_o[slot]), xyz(b.ray_d[slot]), kRayEps, kRayFar);
            }
            q -= ne;
            const bool second = q >= n1;
            const uint32_t i = second ? q - n1 : q;
            const float4 o = (second ? b.shadow2_o : b.shadow_o)[i];
            const float4 d = (second ? b.shadow2_d : b.shadow_d)[i];
            return make_wide_ray(xyz(o), xyz(d), kRayEps, o.w, true);
        },
        [&](uint32_t q, const TraceResult& r, const TraverseCounts& c) {
            if (q < ne) {
                const uint32_t slot = queue[q];
                b.hit[slot] = make_float4(r.t, __int_as_float(r.found ? (int)r.prim : -1),
                    __int_as_float(-1), r.u);
                if (COUNT) {
                    tot.visits += c.visits;
                    tot.tests += c.tests;
                    tot.sphere_tests += c.sphere_tests;
                }
                return;
            }
            q -= ne;
            const bool second = q >= n1;
            const uint32_t i = second ? q - n1 : q;
            (second ? b.shadow2_c : b.shadow_c)[i].w = r.found ? 0.0f : 1.0f;
            if (COUNT) {
                tot_any.visits += c.visits;
                tot_any.tests += c.tests;
                tot_any.sphere_tests += c.sphere_tests;
            }
        });
    if (COUNT) {
        atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kCntVisits),
            (unsigned long long)tot.visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kCntTests),
            (unsigned long long)tot.tests);
        atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kCntSphereTests),
            (unsigned long long)tot.sphere_tests);
        atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kCntAnyVisits),
            (unsigned long long)tot_any.visits);
        atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kCntAnyTests),
            (unsigned long long)tot_any.tests);
        atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kCntAnySphereTests),
            (unsigned long long)tot_any.sphere_tests);
    }
}

// the unoccluded shadow rays of one queue add their contribution to their
// path's radiance (one entry per path and queue, so no two lanes share a
// slot); the emissive queue is applied before the analytic / ReSTIR one
__global__ __launch_bounds__(256) void k_apply_shadows(WavefrontBuffers b, int which) {
    const uint32_t count = b.counters[which == 0 ? kCntShadow : kCntShadow2];
    const float4* sd = which == 0 ? b.shadow_d : b.shadow2_d;
    const float4* sc = which == 0 ? b.shadow_c : b.shadow2_c;
    for (uint32_t q = blockIdx.x * 256u + threadIdx.x; q < count; q += gridDim.x * 256u) {
        const float4 c = sc[q];
        if (c.w != 0.0f) {
            const uint32_t slot = __float_as_uint(sd[q].w);
            const float4 rad = b.radiance[slot];
            b.radiance[slot] = f4(xyz(rad) + xyz(c), 0.0f);
        }
    }
}

// after a traversal launch and its shadow contributions: the shadow queues
// are empty for the coming shade stage, and the fetch cursor rewinds
__global__ void k_after_trace(uint32_t* counters) {
    if (threadIdx.x == 0) {
        counters[kCntShadow] = 0u;
        counters[kCntShadow2] = 0u;
    }
    if (threadIdx.x < kTraceParts) counters[kCntCursor + threadIdx.x * kCursorStride] = 0u;
}

// after a bounce's shade (and ReSTIR) stage: the consumed path queue and the
// ReSTIR deferral queue are empty
__global__ void k_after_shade(uint32_t* counters, int cur) {
    if (threadIdx.x == 0) {
        counters[cur] = 0u;
        counters[kCntRestir] = 0u;
        counters[kCntPbr] = 0u;
    }
}

// the radiance cache's input at a surface hit: position (w = path slot),
// outgoing direction and roughness, shading normal, albedo
template <bool PTEX>
__device__ __forceinline__ CacheInput cache_input(const SceneView& s, const HitInfo& h, f3 rd,
    uint32_t slot) {
    const f3 wo = -normalize3(rd);
    float rough = 1.0f;
    if (h.mat_type == RT_MAT_METAL) rough = h.fuzz;
    else if (h.mat_type == RT_MAT_DIELECTRIC) rough = 0.0f;
    else if (h.mat_type == RT_MAT_OPENPBR) {
        rough = s.openpbr[h.openpbr].specular_roughness;
        if (PTEX && s.pbr_tex != nullptr && s.pbr_tex[h.openpbr].specular_roughness >= 0)
            rough = eval_texture(s, s.pbr_tex[h.openpbr].specular_roughness, h.tu, h.tv, h.pos).x;
    }
    CacheInput c;
    c.a = make_float4(h.pos.x, h.pos.y, h.pos.z, __uint_as_float(slot));
    c.b = make_float4(wo.x, wo.y, wo.z, rough);
    c.c = make_float4(h.sn.x, h.sn.y, h.sn.z, 0.0f);
    c.d = make_float4(fminf(fmaxf(h.base.x, 0.0f), 1.0f), fminf(fmaxf(h.base.y, 0.0f), 1.0f),
        fminf(fmaxf(h.base.z, 0.0f), 1.0f), 0.0f);
    return c;
}

// One shade pass over a queue of paths, compiled for a feature set F
// (kShade* bits). Leaving features out shrinks the register budget, so more
// waves are in flight: the kernel is bound by the latency of its gathers.
// Without kShadePbr (LITE) the pass hands paths whose hit is OpenPBR to the
// pbr queue, untouched, and the full pass shades them next. Each path's
// result depends only on its own slot, so the split changes queue order,
// never values.
template <uint32_t F>
__device__ __forceinline__ void shade_pass(const SceneView& s, const CameraView& cams,
    const FrameParams& fp, const WavefrontBuffers& b, int bounce, int cur, const uint32_t* queue,
    const uint32_t* count_ptr) {
    constexpr bool LITE = (F & kShadePbr) == 0;
    constexpr int PBR = LITE ? kPbrOff : ((F & kShadePbrFull) != 0 ? kPbrFull : kPbrCore);
    constexpr bool ANALYTIC = (F & kShadeAnalytic) != 0;
    constexpr bool CACHE = (F & kShadeCache) != 0;
    constexpr bool TEX = (F & kShadeTexture) != 0;
    constexpr bool PTEX = (F & kShadePbrTex) != 0;
    const SlotRng slot_rng_v = slot_rng(fp);
    const uint32_t count = *count_ptr;
    uint32_t* next_queue = b.queue[cur ^ 1];
    __shared__ uint32_t append_lds[4][5];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t nq = (count + stride - 1u) / stride * stride;  // uniform trip count per wave
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += stride) {
        const bool active = q < count;
        bool shadow = false, shadow2 = false, survive = false, defer = false, pbr = false;
        uint32_t slot = 0u;
        float4 sh_o = make_float4(0, 0, 0, 0), sh_d = sh_o, sh_c = sh_o;
        float4 s2_o = sh_o, s2_d = sh_o, s2_c = sh_o;
        if (active) {
            slot = queue[q];
            const SlotId id = slot_id(cams, fp, slot_rng_v, slot);
            const f3 ro = xyz(b.ray_o[slot]);
            const f3 rd = xyz(b.ray_d[slot]);
            const float4 hv = b.hit[slot];
            f3 thr = xyz(b.throughput[slot]);
            f3 rad = xyz(b.radiance[slot]);
            const int32_t flags = b.prev_flags[slot];
            const float4 ppos = b.prev_pos[slot];
            const float4 pn = b.prev_n[slot];
            PathPrev pv;
            pv.valid = (flags & kPrevValid) != 0;
            pv.delta = (flags & kPrevDelta) != 0;
            pv.mat = ((flags >> 8) & 15) - 1;
            const int inside = (flags >> kPrevMedium) - 1;   // subsurface medium, -1 none
            pv.pos = xyz(ppos);
            pv.pdf = ppos.w;
            pv.n = xyz(pn);
            pv.prim = __float_as_int(pn.w);
            const int prim = __float_as_int(hv.y);
            const int ali = __float_as_int(hv.z);
            if (ANALYTIC && ali >= 0) {
                // path ends on an analytic light (programs.cu:1857-1866)
                const float w = analytic_mis(s, pv, ali, ro, rd);
                rad = rad + thr * (ld3(s.analytic[ali].radiance) * w);
            } else if (prim == -1) {
                // a miss: background, then the dome / distant lights (programs.cu:1867-1875)
                const float w = emission_mis(s, pv, rd, nullptr, true);
                rad = rad + thr * (background_at(s, rd) * w);
                if (ANALYTIC && s.a_infinite_count > 0) rad = rad + thr * analytic_infinite(s, pv, rd);
            } else {
                const float4 rl = rng4(id.pixel, id.sample, id.key, (uint32_t)bounce,
                    kGroupLight, id.stream);
                const float4 rb = rng4(id.pixel, id.sample, id.key, (uint32_t)bounce,
                    kGroupBsdf, id.stream);
                HitInfo h;
                make_hit<TEX>(s, ro, rd, hv.x, prim, h);
                if (LITE && h.mat_type == RT_MAT_OPENPBR) {
                    pbr = true;
                    goto handed_over;
                }
                // inside an OpenPBR subsurface medium (programs.cu:1878-1901): a
                // hit on anything but the medium's own back face ends the path;
                // otherwise a free-flight segment either scatters inside (a new
                // ray along the phase function, with no vertex for MIS) or
                // reaches the boundary, which is shaded as usual
                bool walk_ended = false;
                if (inside >= 0) {
                    walk_ended = true;
                    if (PBR == kPbrFull && h.mat_type == RT_MAT_OPENPBR && h.openpbr == inside
                        && !h.front) {
                        const float4 rw = rng4(id.pixel, id.sample, id.key, (uint32_t)bounce,
                            kGroupSubsurface, id.stream);
                        const rt_openpbr_material& pm = s.openpbr[inside];
                        const float rl = len3(rd);
                        const OpenPbrSegment sg = openpbr_segment(pm, hv.x * rl, rw.x, thr);
                        thr = thr * sg.weight;
                        walk_ended = sg.scattered;
                        if (sg.scattered && bounce + 1 < fp.max_bounces) {
                            const f3 dir = openpbr_hg(pm, normalize3(rd), rw.y, rw.z);
                            const f3 o = (ro + rd * (sg.distance / fmaxf(rl, 1e-8f))) + dir * kRayEps;
                            survive = true;
                            b.ray_o[slot] = f4(o, 0.0f);
                            b.ray_d[slot] = f4(dir, 0.0f);
                            b.throughput[slot] = f4(thr, 0.0f);
                            b.prev_flags[slot] = (inside + 1) << kPrevMedium;
                        }
                    }
                }
                if (walk_ended) {
                } else {
                if (bounce == 0) {
                    b.aux_normal[slot] = f4(h.sn, 1.0f);
                    b.aux_albedo[slot] = f4(h.base, h.t);
                }
                const float ew = emission_mis(s, pv, rd, &h, false);
                rad = rad + thr * (h.emission * ew);
                // radiance cache at the query vertex: training paths leave a
                // record and go on with the plain integrator; once the cache
                // is usable the other paths end here with its prediction
                bool cache_end = false;
                if (CACHE && bounce == fp.cache_bounce && h.mat_type != RT_MAT_DIFFUSE_LIGHT) {
                    const float4 rc = rng4(id.pixel, id.sample, id.key, 0u, kGroupCache, id.stream);
                    const bool train = rc.x < fp.cache_state->train_fraction;
                    const bool query = !train && fp.cache_state->usable != 0;
                    CacheInput ci{};
                    if (train || query) {
                        ci = cache_input<PTEX>(s, h, rd, slot);
                        if (train) {
                            const uint32_t k = atomicAdd(&b.counters[kCntCacheRecord], 1u);
                            if ((int)k < fp.cache_record_cap) {
                                CacheRecord* R = fp.cache_records + k;
                                R->in = ci;
                                R->thr = f4(thr, __uint_as_float(slot));
                                R->rad = f4(rad, 0.0f);
                            }
                        }
                    }
                    const uint32_t qpos = wave_append(query, &b.counters[kCntCacheQuery]);
                    if (query) {
                        fp.cache_queries[qpos] = ci;
                        cache_end = true;
                    }
                }
                const bool bsdf_avail = bounce + 1 < fp.max_bounces && !cache_end;
                // next-event estimation toward emissive geometry (programs.cu:2249-2280)
                if (!cache_end && h.mat_type != RT_MAT_DIFFUSE_LIGHT && s.light_count > 0) {
                    const LightSample ls = sample_light<TEX>(s, h, rl);
                    if (ls.valid) {
                        float bpdf = 0.0f;
                        const f3 resp = direct_response<PBR, PTEX>(s, h, rd, ls.dir, thr, &bpdf);
                        if (maxc3(resp) > 0.0f) {
                            const f3 so = h.pos + ls.dir * (kRayEps * 2.0f);
                            const float max_t = ls.infinite ? kRayFar : ls.dist - kRayEps * 3.0f;
                            if (max_t > kRayEps) {
                                const float mis = bsdf_avail ? power_heuristic(ls.pdf, bpdf) : 1.0f;
                                const f3 direct = (ls.emission * resp) * (mis / ls.pdf);
                                shadow = true;
                                sh_o = f4(so, max_t);
                                sh_d = f4(ls.dir, __uint_as_float(slot));
                                sh_c = f4(thr * direct, 0.0f);
                            }
                        }
                    }
                }
                // analytic lights: ReSTIR stage at the first vertex, otherwise one
                // light sample (programs.cu:2282-2313)
                if (!ANALYTIC) {
                } else if (fp.restir != 0 && bounce == 0) {
                    if (h.mat_type != RT_MAT_DIFFUSE_LIGHT && s.analytic_count > 0) {
                        defer = true;
                        b.restir_o[slot] = f4(ro, 0.0f);
                        b.restir_d[slot] = f4(rd, hv.x);
                        b.restir_thr[slot] = f4(thr, __int_as_float(prim));
                    }
                } else if (!cache_end && h.mat_type != RT_MAT_DIFFUSE_LIGHT && s.analytic_count > 0) {
                    const float4 ra = rng4(id.pixel, id.sample, id.key, (uint32_t)bounce,
                        kGroupAnalytic, id.stream);
                    const int li = pick_analytic(s, ra.x);
                    const LightSample ls = analytic_candidate(s, h.pos, li, ra.y, ra.z);
                    if (ls.valid) {
                        float bpdf = 0.0f;
                        const f3 resp = direct_response<PBR, PTEX>(s, h, rd, ls.dir, thr, &bpdf);
                        if (maxc3(resp) > 0.0f) {
                            const float max_t = ls.infinite ? kRayFar : ls.dist - kRayEps * 3.0f;
                            if (max_t > kRayEps) {
                                const float mis = ls.delta || !bsdf_avail ? 1.0f : power_heuristic(ls.pdf, bpdf);
                                const f3 direct = (ls.emission * resp) * (mis / ls.pdf);
                                shadow2 = true;
                                s2_o = f4(h.pos + ls.dir * (kRayEps * 2.0f), max_t);
                                s2_d = f4(ls.dir, __uint_as_float(slot));
                                s2_c = f4(thr * direct, 0.0f);
                            }
                        }
                    }
                }
                if (bsdf_avail) {
                    const BsdfResult bs = sample_bsdf<PBR, PTEX>(s, h, rd, rb, thr);
                    if (bs.alive) {
                        thr = thr * bs.weight;
                        bool alive = true;
                        if (bounce >= fp.rr_start_bounce) {
                            const float p = fminf(fmaxf(maxc3(thr), 0.05f), 0.95f);
                            if (rb.w > p) alive = false;
                            else thr = div3(thr, p);
                        }
                        if (alive) {
                            survive = true;
                            b.ray_o[slot] = f4(bs.o, 0.0f);
                            b.ray_d[slot] = f4(bs.d, 0.0f);
                            b.throughput[slot] = f4(thr, 0.0f);
                            b.prev_pos[slot] = f4(h.pos, bs.pdf);
                            b.prev_n[slot] = f4(h.sn, __int_as_float(h.prim));
                            // subsurface_entry / subsurface_exit (programs.cu:2545-2551)
                            const int next_inside = bs.event == kPbrSubsurfaceEntry ? h.openpbr
                                                    : (bs.event == kPbrSubsurfaceExit ? -1 : inside);
                            b.prev_flags[slot] = kPrevValid | (bs.delta ? kPrevDelta : 0)
                                                 | ((h.mat_type + 1) << 8)
                                                 | ((next_inside + 1) << kPrevMedium);
                        }
                    }
                }
                }
            }
            b.radiance[slot] = f4(rad, 0.0f);
        }
    handed_over:
        const bool pred[5] = {survive, shadow, shadow2, defer, pbr};
        const uint32_t counter[5] = {(uint32_t)(cur ^ 1), kCntShadow, kCntShadow2, kCntRestir,
            kCntPbr};
        uint32_t pos[5];
        block_append<5>(pred, b.counters, counter, pos, append_lds);
        if (survive) next_queue[pos[0]] = slot;
        if (shadow) {
            b.shadow_o[pos[1]] = sh_o;
            b.shadow_d[pos[1]] = sh_d;
            b.shadow_c[pos[1]] = sh_c;
        }
        if (shadow2) {
            b.shadow2_o[pos[2]] = s2_o;
            b.shadow2_d[pos[2]] = s2_d;
            b.shadow2_c[pos[2]] = s2_c;
        }
        if (defer) b.restir_queue[pos[3]] = slot;
        if (pbr) b.pbr_queue[pos[4]] = slot;
    }
}

// three waves per SIMD for the core variants (their natural budget is a few
// registers above it); the variant with every OpenPBR lobe takes what it needs
template <uint32_t F>
__global__ __launch_bounds__(256, (F & kShadePbrFull) != 0 ? 1 : 3) void k_shade(SceneView s,
    CameraView cams, FrameParams fp,
    WavefrontBuffers b, int bounce, int cur, const uint32_t* queue, const uint32_t* count_ptr) {
    shade_pass<F>(s, cams, fp, b, bounce, cur, queue, count_ptr);
}

// the plain legacy-material pass fits 96 registers: five waves per SIMD
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_shade_plain(
    SceneView s, CameraView cams, FrameParams fp, WavefrontBuffers b, int bounce, int cur,
    const uint32_t* queue, const uint32_t* count_ptr) {
    shade_pass<0>(s, cams, fp, b, bounce, cur, queue, count_ptr);
}

// ReSTIR DI at the first vertex for the deferred paths; its shadow ray joins
// the second shadow queue
// analytic lights are not in the BVH: nearest front-facing one before the
// surface hit (programs.cu:1861-1862), for every path of the queue.  A
// separate convergent pass: inside the persistent traversal loop the light
// loop would run once per finishing lane.
// After the closest-hit traversal of the extension rays: scatter events in
// homogeneous media before the surface hit (a medium event replaces the
// hit), then the nearest analytic light before whatever the ray reached
// (intersect_scene + intersect_analytic_lights, programs.cu:1654-1757,1856-1866)
__global__ __launch_bounds__(256) void k_analytic_hits(SceneView s, CameraView cams, FrameParams fp,
    WavefrontBuffers b, int bounce, int cur) {
    const SlotRng slot_rng_v = slot_rng(fp);
    const uint32_t count = b.counters[cur];
    const uint32_t* queue = b.queue[cur];
    for (uint32_t q = blockIdx.x * 256u + threadIdx.x; q < count; q += gridDim.x * 256u) {
        const uint32_t slot = queue[q];
        float4 h = b.hit[slot];
        const f3 o = xyz(b.ray_o[slot]), d = xyz(b.ray_d[slot]);
        bool found = __float_as_int(h.y) >= 0;
        float closest = found ? h.x : kRayFar;
        if (s.medium_count > 0) {
            const SlotId id = slot_id(cams, fp, slot_rng_v, slot);
            const int mi = sample_media(s, o, d, kRayEps, &closest, id.pixel, id.sample, id.key,
                (uint32_t)bounce, id.stream);
            if (mi >= 0) {
                h.x = closest;
                h.y = __int_as_float(kMediumHit - mi);
                found = true;
            }
        }
        const int ali = s.analytic_count > 0 ? intersect_analytic(s, o, d, kRayEps, found ? closest : kRayFar) : -1;
        h.z = __int_as_float(ali);
        b.hit[slot] = h;
    }
}

template <int BLOCK, int PBR, bool PTEX>
__global__ __launch_bounds__(BLOCK) void k_restir(SceneView s, CameraView cams, FrameParams fp,
    RestirView rv, WavefrontBuffers b) {
    const SlotRng slot_rng_v = slot_rng(fp);
    __shared__ uint2 lds[kStackLds * BLOCK];
    const uint32_t count = b.counters[kCntRestir];
    const TraceStack st = lane_stack<BLOCK>(lds, b.spill);
    const uint32_t stride = gridDim.x * BLOCK;
    const uint32_t nq = (count + stride - 1u) / stride * stride;
    for (uint32_t q = blockIdx.x * BLOCK + threadIdx.x; q < nq; q += stride) {
        const bool active = q < count;
        bool shadow = false;
        float4 s_o = make_float4(0, 0, 0, 0), s_d = s_o, s_c = s_o;
        if (active) {
            const uint32_t slot = b.restir_queue[q];
            const SlotId id = slot_id(cams, fp, slot_rng_v, slot);
            const float4 o4 = b.restir_o[slot];
            const float4 d4 = b.restir_d[slot];
            const float4 t4 = b.restir_thr[slot];
            HitInfo h;
            make_hit(s, xyz(o4), xyz(d4), d4.w, __float_as_int(t4.w), h);
            const int w = cams.width[id.camera];
            const int hgt = (cams.pixel_offset[id.camera + 1] - cams.pixel_offset[id.camera]) / w;
            const int y = (int)id.pixel / w;
            const int x = (int)id.pixel - y * w;
            Reservoir* out = (int)id.sample == rv.spp - 1 && rv.current != nullptr
                                 ? rv.current + id.global_pixel
                                 : nullptr;
            const bool bsdf_avail = 1 < fp.max_bounces;
            const RestirOut ro = restir_shade<BLOCK, PBR, PTEX>(s, rv, h, xyz(d4), xyz(t4), (int)id.camera, id.key,
                cams.pixel_offset[id.camera], x, y, w, hgt,
                id.pixel, id.sample, id.stream, bsdf_avail, out, st);
            if (ro.shadow) {
                shadow = true;
                s_o = f4(ro.origin, ro.max_t);
                s_d = f4(ro.dir, __uint_as_float(slot));
                s_c = f4(xyz(t4) * (ro.numerator * ro.W), 0.0f);
            }
        }
        const uint32_t spos = wave_append(shadow, &b.counters[kCntShadow2]);
        if (shadow) {
            b.shadow2_o[spos] = s_o;
            b.shadow2_d[spos] = s_d;
            b.shadow2_c[spos] = s_c;
        }
    }
}

// targets of the chunk's training records: (final path radiance - radiance
// at the query vertex) / throughput there, per channel
__global__ __launch_bounds__(256) void k_cache_collect(WavefrontBuffers b, CacheRecord* rec, int cap) {
    const uint32_t done = b.counters[kCntCacheDone];
    const uint32_t end = min(b.counters[kCntCacheRecord], (uint32_t)cap);
    for (uint32_t i = done + blockIdx.x * blockDim.x + threadIdx.x; i < end; i += gridDim.x * blockDim.x) {
        CacheRecord* R = rec + i;
        const float4 thr = R->thr, rad = R->rad;
        const float4 fin = b.radiance[__float_as_uint(thr.w)];
        R->target = make_float4(thr.x > 0.0f ? (fin.x - rad.x) / thr.x : 0.0f,
            thr.y > 0.0f ? (fin.y - rad.y) / thr.y : 0.0f, thr.z > 0.0f ? (fin.z - rad.z) / thr.z : 0.0f,
            0.0f);
    }
}

__global__ void k_cache_chunk_done(uint32_t* counters, int cap) {
    if (threadIdx.x != 0) return;
    counters[kCntCacheDone] = min(counters[kCntCacheRecord], (uint32_t)cap);
    counters[kCntCacheQueries] += counters[kCntCacheQuery];
}

__global__ __launch_bounds__(256) void k_accumulate(FrameParams fp, WavefrontBuffers b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= fp.total_pixels) return;
    float4 ab = b.acc_beauty[p], an = b.acc_normal[p], aa = b.acc_albedo[p];
    f3 beauty = xyz(ab), nsum = xyz(an), asum = xyz(aa);
    float aux = an.w, dsum = aa.w;
    for (int layer = 0; layer < fp.layer_count; ++layer) {
        const size_t slot = (size_t)layer * (size_t)fp.total_pixels + (size_t)p;
        beauty = beauty + xyz(b.radiance[slot]);
        const float4 n = b.aux_normal[slot];
        if (n.w > 0.0f) {
            const float4 a = b.aux_albedo[slot];
            nsum = nsum + xyz(n);
            asum = asum + xyz(a);
            dsum += a.w;
            aux += 1.0f;
        }
    }
    b.acc_beauty[p] = f4(beauty, 0.0f);
    b.acc_normal[p] = f4(nsum, aux);
    b.acc_albedo[p] = f4(asum, dsum);
}

// programs.cu:1944-1952 + store_output :1773-1805
__global__ __launch_bounds__(256) void k_finalize(int pixels, int spp, const float4* acc_beauty,
    const float4* acc_normal, const float4* acc_albedo, float4* beauty, float4* normal,
    float4* albedo, float* depth) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    const float inv_spp = 1.0f / (float)spp;
    const float4 ab = acc_beauty[p], an = acc_normal[p], aa = acc_albedo[p];
    const int aux = (int)an.w;
    const float inv_aux = aux > 0 ? 1.0f / (float)aux : 0.0f;
    const f3 bavg = xyz(ab) * inv_spp;
    const f3 navg = aux > 0 ? normalize3(xyz(an) * inv_aux) : mk3(0.0f, 0.0f, 1.0f);
    const f3 aavg = aux > 0 ? xyz(aa) * inv_aux : mk3(0.0f, 0.0f, 0.0f);
    const float davg = aux > 0 ? aa.w * inv_aux : 0.0f;
    beauty[p] = f4(bavg, 1.0f);
    normal[p] = f4(clamp3(navg * 0.5f + mk3(0.5f, 0.5f, 0.5f), 0.0f, 1.0f), 1.0f);
    albedo[p] = f4(clamp3(aavg, 0.0f, 1.0f), 1.0f);
    depth[p] = davg;
}

}  // namespace

hipError_t wf_reset_counters(uint32_t* counters, hipStream_t st) {
    k_reset_counters<<<1, 64, 0, st>>>(counters);
    return hipGetLastError();
}

hipError_t wf_clear_accum(int pixels, WavefrontBuffers& b, hipStream_t st) {
    k_clear_accum<<<(pixels + 255) / 256, 256, 0, st>>>(pixels, b.acc_beauty, b.acc_normal,
        b.acc_albedo);
    return hipGetLastError();
}

hipError_t wf_zero(void* p, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 4u;
    if (n == 0) return hipSuccess;
    k_zero_words<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(static_cast<uint32_t*>(p), n);
    return hipGetLastError();
}

hipError_t wf_raygen(const CameraView& cams, const FrameParams& fp, WavefrontBuffers& b, int paths,
    hipStream_t st) {
    k_raygen<<<(paths + 255) / 256, 256, 0, st>>>(cams, fp, b, paths);
    return hipGetLastError();
}

template <bool COUNT>
static void launch_fused(const SceneView& sv, WavefrontBuffers& b, int cur, int ext, StageLaunch l,
    hipStream_t st) {
    if (l.block == 64) k_trace_fused<64, COUNT><<<l.grid, 64, 0, st>>>(sv, b, cur, ext, l.refill_min);
    else if (l.block == 256) k_trace_fused<256, COUNT><<<l.grid, 256, 0, st>>>(sv, b, cur, ext, l.refill_min);
    else k_trace_fused<128, COUNT><<<l.grid, 128, 0, st>>>(sv, b, cur, ext, l.refill_min);
}

hipError_t wf_trace(const SceneView& s, const CameraView& cams, const FrameParams& fp,
    WavefrontBuffers& b, int bounce, bool extension, bool count, StageLaunch l, int apply_grid,
    hipStream_t st) {
    const int cur = bounce & 1;
    const int ext = extension ? 1 : 0;
    if (count) launch_fused<true>(s, b, cur, ext, l, st);
    else launch_fused<false>(s, b, cur, ext, l, st);
    if (extension && (s.analytic_count > 0 || s.medium_count > 0))
        k_analytic_hits<<<l.grid, 256, 0, st>>>(s, cams, fp, b, bounce, cur);
    const unsigned g = (unsigned)(apply_grid > 0 ? apply_grid : 1);
    k_apply_shadows<<<g, 256, 0, st>>>(b, 0);
    k_apply_shadows<<<g, 256, 0, st>>>(b, 1);
    k_after_trace<<<1, 64, 0, st>>>(b.counters);
    return hipGetLastError();
}

hipError_t wf_end_bounce(WavefrontBuffers& b, int bounce, hipStream_t st) {
    k_after_shade<<<1, 64, 0, st>>>(b.counters, bounce & 1);
    return hipGetLastError();
}

hipError_t wf_shade(const SceneView& s, const CameraView& cams, const FrameParams& fp,
    WavefrontBuffers& b, int bounce, StageLaunch l, StageLaunch lite, int mode, hipStream_t st) {
    const int cur = bounce & 1;
    // the full pass: every OpenPBR lobe only when a material uses more than
    // the opaque core
    auto full = [&](const uint32_t* queue, const uint32_t* count) {
        if (s.pbr_tex != nullptr) {
            if (s.pbr_full != 0)
                k_shade<kShadeAll | kShadePbrFull | kShadePbrTex><<<l.grid, 256, 0, st>>>(s, cams, fp, b,
                    bounce, cur, queue, count);
            else
                k_shade<kShadeAll | kShadePbrTex><<<l.grid, 256, 0, st>>>(s, cams, fp, b, bounce, cur,
                    queue, count);
        } else if (s.pbr_full != 0) {
            k_shade<kShadeAll | kShadePbrFull><<<l.grid, 256, 0, st>>>(s, cams, fp, b, bounce, cur,
                queue, count);
        } else {
            k_shade<kShadeAll><<<l.grid, 256, 0, st>>>(s, cams, fp, b, bounce, cur, queue, count);
        }
    };
    if (mode == 0) {
        full(b.queue[cur], b.counters + cur);
        return hipGetLastError();
    }
    // the legacy-material pass: its plain variant when the scene and frame
    // use no analytic light, radiance cache, albedo or emission texture
    const bool plain = s.analytic_count == 0 && s.a_infinite_count == 0 && fp.cache_bounce < 0
                       && s.mat_albedo_tex == nullptr && s.mat_emission_tex == nullptr
                       && s.pbr_tex == nullptr;
    if (plain)
        k_shade_plain<<<lite.grid, 256, 0, st>>>(s, cams, fp, b, bounce, cur, b.queue[cur],
            b.counters + cur);
    else
        k_shade<kShadeAll & ~kShadePbr><<<l.grid, 256, 0, st>>>(s, cams, fp, b, bounce, cur,
            b.queue[cur], b.counters + cur);
    if (mode == 2) full(b.pbr_queue, b.counters + kCntPbr);
    return hipGetLastError();
}

static int resident_blocks(const void* kernel, int block) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, 0) != hipSuccess || n < 1)
        n = 1;
    return n;
}

int wf_restir_blocks_per_cu() {
    static std::atomic<int> cached{0};
    int n = cached.load(std::memory_order_relaxed);
    if (n == 0) {
        n = resident_blocks(reinterpret_cast<const void*>(k_restir<128, kPbrCore, false>), 128);
        cached.store(n, std::memory_order_relaxed);
    }
    return n;
}

// k_shade and k_restir are grid-stride loops with an equal share per wave, so
// their grids are exactly what the device keeps resident (a larger grid
// leaves a second, partial round of waves behind the first)

int wf_shade_blocks_per_cu(bool lite) {
    static std::atomic<int> cached[2]{{0}, {0}};
    int n = cached[lite].load(std::memory_order_relaxed);
    if (n == 0) {
        n = resident_blocks(lite ? reinterpret_cast<const void*>(k_shade_plain)
                                 : reinterpret_cast<const void*>(k_shade<kShadeAll>), 256);
        cached[lite].store(n, std::memory_order_relaxed);
    }
    return n;
}

// the ReSTIR stage runs 128-thread blocks on `l.grid` blocks, which the caller
// keeps within the traversal spill area (trace grid x trace block threads)
hipError_t wf_restir(const SceneView& s, const CameraView& cams, const FrameParams& fp,
    const RestirView& rv, WavefrontBuffers& b, StageLaunch l, hipStream_t st) {
    const unsigned g = (unsigned)(l.grid > 0 ? l.grid : 1);
    if (s.pbr_tex != nullptr) {
        if (s.pbr_full != 0) k_restir<128, kPbrFull, true><<<g, 128, 0, st>>>(s, cams, fp, rv, b);
        else k_restir<128, kPbrCore, true><<<g, 128, 0, st>>>(s, cams, fp, rv, b);
    } else if (s.pbr_full != 0) {
        k_restir<128, kPbrFull, false><<<g, 128, 0, st>>>(s, cams, fp, rv, b);
    } else {
        k_restir<128, kPbrCore, false><<<g, 128, 0, st>>>(s, cams, fp, rv, b);
    }
    return hipGetLastError();
}

hipError_t wf_cache_chunk(const CacheGridView& g, const CachePacked& pk, const FrameParams& fp,
    WavefrontBuffers& b, int path_cap, int grid, hipStream_t st) {
    hipError_t e = cache_infer_paths(g, pk, fp.cache_queries, b.counters + kCntCacheQuery, path_cap,
        b.throughput, b.radiance, grid, st);
    if (e != hipSuccess) return e;
    k_cache_collect<<<256, 256, 0, st>>>(b, fp.cache_records, fp.cache_record_cap);
    k_cache_chunk_done<<<1, 64, 0, st>>>(b.counters, fp.cache_record_cap);
    return hipGetLastError();
}

const uint32_t* wf_cache_record_count(const WavefrontBuffers& b) { return b.counters + kCntCacheRecord; }

hipError_t wf_accumulate(const FrameParams& fp, WavefrontBuffers& b, hipStream_t st) {
    k_accumulate<<<(fp.total_pixels + 255) / 256, 256, 0, st>>>(fp, b);
    return hipGetLastError();
}

hipError_t wf_finalize(int pixels, int spp, const WavefrontBuffers& b, float4* beauty,
    float4* normal, float4* albedo, float* depth, hipStream_t st) {
    k_finalize<<<(pixels + 255) / 256, 256, 0, st>>>(pixels, spp, b.acc_beauty, b.acc_normal,
        b.acc_albedo, beauty, normal, albedo, depth);
    return hipGetLastError();
}

}  // namespace rtk
