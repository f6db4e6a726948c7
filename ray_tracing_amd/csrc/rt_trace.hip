// rt_trace.hip -- closest-hit / any-hit kernels over caller-provided rays
// (test hooks replacing intersect_scene / is_occluded,
// src/realtime/gpu/programs.cu:1654-1762).  One thread per ray, deferred
// node stack in dynamic LDS sized from the BVH's measured depth.
#include "rt_kernels.h"
#include "rt_traverse.h"

namespace rtk {

namespace {

constexpr int kTraceBlock = 128;

template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace(SceneView s, const rt_ray* rays, int64_t n,
    rt_hit* hits, int32_t* occluded) {
    extern __shared__ uint32_t lds[];
    const int64_t i = (int64_t)blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t* stack_node = lds + threadIdx.x;
    float* stack_t = reinterpret_cast<float*>(lds + kTraceBlock * s.stack_depth) + threadIdx.x;
    const rt_ray r = rays[i];
    const TraceResult res = traverse<ANY>(s, ld3(r.o), ld3(r.d), r.tmin, r.tmax, stack_node,
        stack_t, kTraceBlock);
    if (ANY) {
        occluded[i] = res.found ? 1 : 0;
    } else {
        rt_hit h;
        h.t = res.found ? res.t : -1.0f;
        h.prim = res.found ? (int32_t)res.prim : -1;
        h.u = res.u;
        h.v = res.v;
        hits[i] = h;
    }
}

}  // namespace

hipError_t trace_closest_launch(const SceneView& s0, int stack_depth, const rt_ray* rays,
    int64_t n, rt_hit* hits, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    SceneView s = s0;
    s.stack_depth = stack_depth;
    const size_t lds = (size_t)stack_depth * kTraceBlock * 8u;
    const unsigned grid = (unsigned)((n + kTraceBlock - 1) / kTraceBlock);
    k_trace<false><<<grid, kTraceBlock, lds, st>>>(s, rays, n, hits, nullptr);
    return hipGetLastError();
}

hipError_t trace_any_launch(const SceneView& s0, int stack_depth, const rt_ray* rays, int64_t n,
    int32_t* occluded, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    SceneView s = s0;
    s.stack_depth = stack_depth;
    const size_t lds = (size_t)stack_depth * kTraceBlock * 8u;
    const unsigned grid = (unsigned)((n + kTraceBlock - 1) / kTraceBlock);
    k_trace<true><<<grid, kTraceBlock, lds, st>>>(s, rays, n, nullptr, occluded);
    return hipGetLastError();
}

}  // namespace rtk
