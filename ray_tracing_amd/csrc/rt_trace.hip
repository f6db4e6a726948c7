// rt_trace.hip -- closest-hit / any-hit kernels over caller-provided rays
// (test hooks replacing intersect_scene / is_occluded,
// src/realtime/gpu/programs.cu:1654-1762).  The same persistent 4-wide
// traversal as the renderer's extend / connect stages.
#include "rt_kernels.h"
#include "rt_traverse.h"

namespace rtk {

namespace {

template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace(SceneView s, const rt_ray* rays, int64_t n,
    rt_hit* hits, int32_t* occluded, uint32_t* fetch, uint2* spill) {
    __shared__ uint2 lds[kStackLds * kTraceBlock];
    TraceStack st;
    st.lds = lds + threadIdx.x;
    st.spill_stride = gridDim.x * kTraceBlock;
    st.spill = spill + (size_t)blockIdx.x * kTraceBlock + threadIdx.x;
    trace_queue<kTraceBlock, ANY, false>(s, (uint32_t)n, fetch, st, 16,
        [&](uint32_t q) {
            const rt_ray r = rays[q];
            return make_wide_ray(ld3(r.o), ld3(r.d), r.tmin, r.tmax);
        },
        [&](uint32_t q, const TraceResult& res, const TraverseCounts&) {
            if (ANY) {
                occluded[q] = res.found ? 1 : 0;
            } else {
                rt_hit h;
                h.t = res.found ? res.t : -1.0f;
                h.prim = res.found ? (int32_t)res.prim : -1;
                h.u = res.u;
                h.v = res.v;
                hits[q] = h;
            }
        });
}

}  // namespace

hipError_t trace_closest_launch(const SceneView& s, int grid, const rt_ray* rays, int64_t n,
    rt_hit* hits, uint32_t* fetch, uint2* spill, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_trace<false><<<grid, kTraceBlock, 0, st>>>(s, rays, n, hits, nullptr, fetch, spill);
    return hipGetLastError();
}

hipError_t trace_any_launch(const SceneView& s, int grid, const rt_ray* rays, int64_t n,
    int32_t* occluded, uint32_t* fetch, uint2* spill, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_trace<true><<<grid, kTraceBlock, 0, st>>>(s, rays, n, nullptr, occluded, fetch, spill);
    return hipGetLastError();
}

}  // namespace rtk
