// rt_trace.hip -- closest-hit / any-hit kernels over caller-provided rays
// (test hooks replacing intersect_scene / is_occluded,
// src/realtime/gpu/programs.cu:1654-1762).  The same persistent 4-wide
// traversal as the renderer's extend / connect stages.
#include <mutex>

#include "rt_kernels.h"
#include "rt_traverse.h"

namespace rtk {

// Scratch of the allocation-free closest-hit entry point: the ray-fetch
// cursor in the first 256 bytes, the traversal stack spill after it.  Each
// device holds its own copy for the life of the process; host code reaches
// it only through the address hipGetSymbolAddress returns for that device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t trace_scratch_init(int ordinal) {
    if (ordinal < 0 || ordinal >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (g_scratch_ptr[ordinal] != nullptr) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));   // current device
    if (e != hipSuccess) return e;
    g_scratch_ptr[ordinal] = ptr;
    return hipSuccess;
}

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

int trace_closest_noalloc(const SceneView& s, int ordinal, int cu_count, const rt_ray* rays,
    int64_t n, rt_hit* hits, hipStream_t st, hipError_t* err) {
    *err = hipSuccess;
    if (n <= 0) return 0;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_lock);
        if (ordinal >= 0 && ordinal < kMaxDevices) base = g_scratch_ptr[ordinal];
    }
    if (base == nullptr) return 1;
    // the spill a block needs; the grid is cut to what the scratch holds
    const size_t entries = (size_t)(s.stack4_entries > kStackLds ? s.stack4_entries - kStackLds : 1);
    const size_t per_block = entries * (size_t)kTraceBlock * sizeof(uint2);
    const size_t fit = (kNoAllocScratchBytes - 256u) / per_block;
    if (fit == 0) return 2;
    const int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(blocks,
        (int64_t)cu_count * 8), (int64_t)fit));
    uint32_t* fetch = static_cast<uint32_t*>(base);
    uint2* spill = reinterpret_cast<uint2*>(static_cast<unsigned char*>(base) + 256);
    if ((*err = wf_zero(fetch, 4, st)) != hipSuccess) return 3;   // fill kernel, not a memset
    if ((*err = trace_closest_launch(s, grid, rays, n, hits, fetch, spill, st)) != hipSuccess) return 3;
    return 0;
}

hipError_t trace_any_launch(const SceneView& s, int grid, const rt_ray* rays, int64_t n,
    int32_t* occluded, uint32_t* fetch, uint2* spill, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_trace<true><<<grid, kTraceBlock, 0, st>>>(s, rays, n, nullptr, occluded, fetch, spill);
    return hipGetLastError();
}

}  // namespace rtk
