// rt_kernels.h -- host-side launch entry points of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_cache.h"
#include "rt_internal.h"

namespace rtk {

struct BvhBuildScratch {
    float4* pbox;          // 2 x float4 per primitive
    uint32_t* keys[2];
    uint32_t* vals[2];
    uint32_t* counts;      // 16 x sort blocks
    uint32_t* leaf_parent;
    uint32_t* flags;
    uint32_t* cb;          // centroid bounds (6) + max depth (1)
};

hipError_t bvh_build(const SceneView& s, int n, BvhBuildScratch& w, float4* nodes, float4* leaves,
    uint32_t* sorted_out, uint32_t* morton_out, hipStream_t st);
// node bounds and leaf records of new primitive positions over the topology
// and Morton order of the last build (leaf_parent from that build)
hipError_t bvh_refit(const SceneView& s, int n, BvhBuildScratch& w, float4* nodes, float4* leaves,
    const uint32_t* sorted, const uint32_t* leaf_parent, hipStream_t st);

struct WideScratch {
    uint32_t* is_wide;   // per binary node: roots a 4-wide node
    uint32_t* index;     // exclusive scan of is_wide: 4-wide node index
    uint32_t* front[2];  // top-down frontiers
    uint32_t* counts;    // two frontier sizes, then the level count (3 words)
    uint32_t* tiles;     // scan_tiles(nbin) tile sums of the node numbering scan
};
constexpr int kScanTile = 4096;
inline int scan_tiles(int n) { return (n + kScanTile - 1) / kScanTile; }
// collapse the binary LBVH (nbin nodes, root 0) into the 4-wide quantized
// BVH; reports its node count and number of levels (synchronises `st`)
hipError_t bvh_collapse_wide(const float4* nodes, int nbin, WideScratch& w, float4* nodes4,
    int* wide_count, int* wide_levels, hipStream_t st);

// test-hook traversal over device arrays of rt_ray / rt_hit
// (persistent grid of `grid` blocks of kTraceBlock threads; `fetch` is one
// zeroed counter; `spill` holds (stack entries - kStackLds) x grid x kTraceBlock)
constexpr int kTraceBlock = 128;
hipError_t trace_closest_launch(const SceneView& s, int grid, const rt_ray* rays, int64_t n,
    rt_hit* hits, uint32_t* fetch, uint2* spill, hipStream_t st);
hipError_t trace_any_launch(const SceneView& s, int grid, const rt_ray* rays, int64_t n,
    int32_t* occluded, uint32_t* fetch, uint2* spill, hipStream_t st);

// allocation-free closest hit: a file-scope device array (fetch cursor +
// stack spill) instead of allocated scratch; trace_scratch_init resolves
// its address once per device (with that device current)
constexpr size_t kNoAllocScratchBytes = (16u << 20) + 256u;
constexpr int kMaxDevices = 64;
hipError_t trace_scratch_init(int ordinal);
// 0: launched; 1: scratch not resolved for this device; 2: one block's
// stack spill exceeds the scratch; 3: a launch failed (*err)
int trace_closest_noalloc(const SceneView& s, int ordinal, int cu_count, const rt_ray* rays,
    int64_t n, rt_hit* hits, hipStream_t st, hipError_t* err);

// ---- wavefront pipeline
struct WavefrontBuffers {
    // per path slot (capacity = queue capacity)
    float4* ray_o;      // origin.xyz, unused
    float4* ray_d;      // direction.xyz, unused
    float4* hit;        // t, prim bits, u, v
    float4* throughput; // xyz, unused
    float4* radiance;   // xyz, unused
    float4* prev_pos;   // xyz, bsdf pdf
    float4* prev_n;     // xyz, bits(prev prim)
    int32_t* prev_flags;// bit0 valid, bit1 delta, bits 8.. material type + 1
    float4* aux_normal; // bounce-0 shading normal, w = 1 if captured
    float4* aux_albedo; // bounce-0 base colour, w = depth
    uint32_t* queue[2]; // active path slots (ping-pong)
    uint32_t* pbr_queue;// OpenPBR hits handed from the legacy-material shade pass
    float4* shadow_o;   // origin.xyz, max_t
    float4* shadow_d;   // direction.xyz, bits(path slot)
    float4* shadow_c;   // contribution.xyz
    // second shadow queue (analytic lights / ReSTIR), traced after the first
    float4* shadow2_o;
    float4* shadow2_d;
    float4* shadow2_c;
    // bounce-0 paths deferred to the ReSTIR stage
    uint32_t* restir_queue;
    float4* restir_o;   // incoming ray origin
    float4* restir_d;   // incoming ray direction, w = hit t
    float4* restir_thr; // throughput before BSDF sampling, w = bits(prim)
    uint32_t* counters; // [0]/[1] queue counts, [2]/[3] shadow counts, [4..15] u64 ray /
                        // traversal stats, [16] ray-fetch cursor of the traversal launch,
                        // [20..23] radiance cache, [24..29] u64 any-hit traversal stats
    uint2* spill;       // traversal stack spill: (stack entries - kStackLds) x trace threads
    // per pixel (all cameras)
    float4* acc_beauty;
    float4* acc_normal; // xyz sum, w = aux count
    float4* acc_albedo; // xyz sum, w = depth sum
};

struct FrameParams {
    int32_t spp;
    int32_t max_bounces;
    int32_t rr_start_bounce;
    int32_t jitter;
    const FrameDyn* dyn;    // per-frame values in device memory
    int32_t total_pixels;
    int32_t layer_count;    // samples per pixel in this chunk
    int32_t first_sample;   // sample index of the chunk's first layer
    int32_t restir;         // bounce-0 analytic lighting goes through the ReSTIR stage
    // radiance cache (cache_bounce = -1 without one)
    int32_t cache_bounce;            // path vertex of queries and training records
    const CacheState* cache_state;   // usable flag, training share (device memory)
    CacheInput* cache_queries;       // capacity = path capacity
    CacheRecord* cache_records;
    int32_t cache_record_cap;
};

struct StageLaunch {
    int grid;
    int block;
    int refill_min = 64;  // idle lanes of a wave before it takes new rays (trace stages)
};

hipError_t wf_reset_counters(uint32_t* counters, hipStream_t st);
// the camera-model debug image: RGBA8 encoded world directions of the pixel
// centres' rays (width x height words; the camera in device memory)
hipError_t wf_direction_debug(const rt_camera* cam, uint32_t* rgba, int width, int height,
    hipStream_t st);
hipError_t wf_raygen(const CameraView& cams, const FrameParams& fp, WavefrontBuffers& b,
    int paths, hipStream_t st);
// one traversal launch: the extension rays of `bounce` (when `extension`)
// and the shadow rays the previous shade stage queued, then the analytic
// light hits of the extension rays and the unoccluded shadow contributions
// added to their paths in queue order (emissive, then analytic / ReSTIR)
hipError_t wf_trace(const SceneView& s, const CameraView& cams, const FrameParams& fp,
    WavefrontBuffers& b, int bounce, bool extension, bool count, StageLaunch l, int apply_grid,
    hipStream_t st);
// after the shade / ReSTIR stages of `bounce`: its path queue has been consumed
hipError_t wf_end_bounce(WavefrontBuffers& b, int bounce, hipStream_t st);
// resident blocks per CU of the shade (256 threads; `lite`: its plain
// legacy-material variant) and ReSTIR (128 threads) kernels: their launch
// grids per CU
int wf_shade_blocks_per_cu(bool lite = false);
int wf_restir_blocks_per_cu();
// mode 0: one full pass; 1: the legacy-material pass alone (the scene has
// no OpenPBR material); 2: the legacy pass, then the full pass over the
// OpenPBR hits it handed over
hipError_t wf_shade(const SceneView& s, const CameraView& cams, const FrameParams& fp,
    WavefrontBuffers& b, int bounce, StageLaunch l, StageLaunch lite, int mode, hipStream_t st);
hipError_t wf_restir(const SceneView& s, const CameraView& cams,
    const FrameParams& fp, const RestirView& rv, WavefrontBuffers& b, StageLaunch l,
    hipStream_t st);
hipError_t wf_accumulate(const FrameParams& fp, WavefrontBuffers& b, hipStream_t st);
// radiance cache, after a chunk's bounces: cache predictions for the paths
// that ended at the query vertex, targets of the chunk's training records
hipError_t wf_cache_chunk(const CacheGridView& g, const CachePacked& pk, const FrameParams& fp,
    WavefrontBuffers& b, int path_cap, int grid, hipStream_t st);
// device counters of the cache: records written this frame, queries answered
const uint32_t* wf_cache_record_count(const WavefrontBuffers& b);
constexpr int kStatWords = 31;  // counter words downloaded after a frame
constexpr int kCounterAlloc = 1024;  // counter words allocated (queue counts, stats, cursors)
constexpr int kStatCacheRecords = 21, kStatCacheQueries = 23;
hipError_t wf_clear_accum(int pixels, WavefrontBuffers& b, hipStream_t st);
// zero `bytes` (a multiple of 4) of device memory with a fill kernel
hipError_t wf_zero(void* p, size_t bytes, hipStream_t st);
// temporal reprojection of every camera with history (per camera: frames of
// history so far, previous camera valid; previous cameras at cams.cams[4 + c])
struct ResolveParams {
    const FrameDyn* dyn;    // history lengths / previous-camera flags
    int32_t restir_enabled;
};
hipError_t wf_resolve(const CameraView& cams, const ResolveParams& rp, float4* beauty,
    float4* normal, float* depth, const float4* hbeauty, const float4* hnormal,
    const float* hdepth, float2* flow, float* trust, int total, hipStream_t st);

// the edge-aware a-trous denoise filter, spacing 1, 2, 4, ... per pass;
// with `gate` set only the cameras c with gate[c] != 0 are filtered (the
// others keep what `out` holds)
constexpr int kDenoiseIterations = 4;
hipError_t wf_denoise_filter(const CameraView& cams, int iterations, const float4* beauty,
    const float4* normal, const float4* albedo, const float* depth, float4* out,
    float4* scratch, int total, hipStream_t st, const int32_t* gate = nullptr);

// Guard of the learned denoiser, in device memory: per camera, the
// statistics of the network's output against its 1-spp input over the
// surface pixels, the decision to fall back to the a-trous filter, and the
// budgets the decision applies.
struct NnGuard {
    float acc[4][6];         // non-finite outputs, sum log1p L(in), sum log1p L(out),
                             // HF(in), HF(out), pixels; cleared by each decision
    int32_t fallback[4];     // 1: this frame's camera output comes from the filter
    float metric[4][3];      // non-finite outputs, mean shift, residual noise (last frame)
    uint32_t frames_checked;
    uint32_t frames_fallback;
    float max_mean_shift;    // |mean log1p L(out) - mean log1p L(in)| / mean log1p L(in)
    float max_residual_noise;// HF(out) / HF(in); HF = sum of |L(p) - L(right neighbour)|
};
// statistics and decision for `cams.count` cameras (no host synchronisation)
hipError_t wf_nn_guard(const CameraView& cams, const float4* in, const float4* out,
    const float* depth, NnGuard* guard, hipStream_t st);

// the viewer's tone map: float4 linear radiance -> RGBA8 (alpha 255)
hipError_t wf_tone_map_rgba8(const float4* src, uint32_t* dst, int64_t n, hipStream_t st);

hipError_t wf_finalize(int pixels, int spp, const WavefrontBuffers& b, float4* beauty,
    float4* normal, float4* albedo, float* depth, hipStream_t st);

}  // namespace rtk
