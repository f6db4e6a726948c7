// rt_cache.h -- online radiance cache: multiresolution hash-grid encoding and
// a fully fused MLP, trained every frame on path suffixes of the plain
// integrator and queried at a fixed path vertex to end paths early.
//
// Encoding (per level l of L, F = 2 features, table of T = 2^log2_table
// entries per level):
//   p01 = clamp((pos - lo) * inv_extent, 0, 1)           per axis, FP32
//   x = p01 * res_l; c = min(floor(x), res_l - 1); f = x - c
//   corner k (bit 0 = x, 1 = y, 2 = z): v = c + bits(k),
//     index = dense ? v.x + v.y (res_l + 1) + v.z (res_l + 1)^2        if (res_l + 1)^3 <= T
//           : (v.x * 1 ^ v.y * 2654435761 ^ v.z * 805459861) & (T - 1) otherwise
//     w_k = wx * wy * wz with w = f or 1 - f, multiplied in that order
//   feature[2 l + j] = sum over k = 0..7 (in order) of w_k * table[l][index][j]
//   res_l = floor(base * b^l), b = exp((ln max - ln base) / (L - 1))   (host, double)
//
// Network input (64, bf16): 32 hash features (levels beyond L are zero),
// outgoing direction (3), shading normal (3), roughness, albedo (3), a
// constant 1 (the first layer's bias), zeros.  Four hidden layers of 64 with
// ReLU and no bias, linear RGB output:
//   z1 = W0 x, h1 = relu(z1), ..., h4 = relu(W3 h3), y = W4 h4   (W4: 3 x 64)
// Loss over n samples: (1 / 3n) sum (y - t)^2 / (y_sg^2 + 0.01), with y_sg
// held constant in the gradient (relative L2).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtk {

constexpr int kCacheIn = 64;
constexpr int kCacheHidden = 64;
constexpr int kCacheHiddenLayers = 4;
constexpr int kCacheOut = 3;
constexpr int kCacheMaxLevels = 16;
constexpr int kCacheFeatures = 2;
constexpr int kCacheSamplesPerBlock = 256;   // 4 waves x 64 samples
constexpr int kCacheGradSamples = 128;       // samples per weight-gradient block
constexpr float kCacheLossEps = 0.01f;

// input channel layout
constexpr int kCacheChDir = 32, kCacheChNormal = 35, kCacheChRough = 38, kCacheChAlbedo = 39,
              kCacheChOne = 42;

// FP32 parameter vector: W0..W3 [64 out][64 in], W4 [3][64], then the grid
// [L][T][2]
__host__ __device__ constexpr int cache_layer_offset(int l) { return l * kCacheHidden * kCacheIn; }
__host__ __device__ constexpr int cache_mlp_params() {
    return kCacheHiddenLayers * kCacheHidden * kCacheIn + kCacheOut * kCacheHidden;
}

struct CacheGridView {
    const float* table;   // [L][T][2]
    float* grad;          // same layout (training only)
    uint8_t* touched;     // [L][T]: 1 where a sample of this step wrote a gradient (training only)
    int32_t levels;
    uint32_t tmask;       // T - 1
    int32_t log2_table;
    float res[kCacheMaxLevels];
    int32_t dense[kCacheMaxLevels];
    float lo[3], inv_extent[3];
};

// one query / training input, 64 B:
//   a = pos.xyz, bits(path slot); b = outgoing direction.xyz, roughness;
//   c = shading normal.xyz, 0;    d = albedo.rgb, 0
struct CacheInput {
    float4 a, b, c, d;
};

// device weights in the kernels' layouts (bf16)
struct CachePacked {
    uint16_t* fwd[kCacheHiddenLayers + 1];   // W [out][64]; W4 padded to 16 rows
    uint16_t* bwd[kCacheHiddenLayers + 1];   // W^T [64 in][out]; W4^T padded to 32 columns
};

// training buffers (capacity samples)
struct CacheTrainScratch {
    uint16_t* act;        // bf16 [5][cap][64]: x, h1..h4
    uint16_t* dz;         // bf16 [5][cap][64]: dz1..dz4, dy (first 3 columns)
    float4* pred;         // [cap]
    float* loss_partial;  // [cap / 256]
    float* grad_partial;  // [cap / kCacheGradSamples][64][64]
};

// per-frame state in device memory (read by the shade kernel)
struct CacheState {
    float loss_ema;       // EMA of the relative loss on fresh samples
    float last_loss;
    int32_t frames;       // frames trained
    int32_t usable;       // the renderer terminates paths with the cache
    int32_t steps;        // Adam steps taken
    float train_fraction; // share of pixel samples traced as training paths
    int32_t pad[2];
};

// One training record per training path, written at the query vertex:
// in = the input features, thr = throughput before scattering there,
// rad = path radiance after that vertex's emission; after the path ends
// `target` = (final radiance - rad) / thr.
struct CacheRecord {
    CacheInput in;
    float4 thr;           // xyz, w = bits(path slot)
    float4 rad;           // xyz
    float4 target;        // xyz
};

hipError_t cache_pack(const float* params, const CachePacked& pk, hipStream_t st);
// predictions for n inputs (host count) into out (float4, w = 0)
hipError_t cache_infer(const CacheGridView& g, const CachePacked& pk, const CacheInput* in, int n,
    float4* out, int grid, hipStream_t st);
// renderer variant: count on the device; adds throughput * max(y, 0) to the
// path radiance of each query's slot
hipError_t cache_infer_paths(const CacheGridView& g, const CachePacked& pk, const CacheInput* in,
    const uint32_t* count, int cap, const float4* throughput, float4* radiance, int grid,
    hipStream_t st);
// One training pass over batch `batch` of `batches` equal parts of the first
// min(*count, cap) samples: forward with the current weights, loss into
// *loss, MLP gradient into grad[0 .. cache_mlp_params()), grid gradient
// added into g.grad (zero on entry).  Inputs are strided by `in_stride`
// float4s, targets (xyz) by `t_stride`.
hipError_t cache_train_pass(const CacheGridView& g, const CachePacked& pk,
    const CacheTrainScratch& w, const float4* in, int in_stride, const float4* target,
    int t_stride, const uint32_t* count, int cap, int batch, int batches, float* grad, float* loss,
    int grid, hipStream_t st);
// Adam with the step number state->steps + 1 (device memory), then
// state->steps += 1; nothing when `count` (optional) is zero.  The n_mlp
// MLP parameters are updated densely; the grid entries only where
// g.touched is set (lazy, sparse Adam: the moments of untouched entries are
// left as they are), and those entries' gradients and flags are cleared, so
// the grid gradient is zero again for the next step.
hipError_t cache_adam(int64_t n_mlp, float* params, const float* grad, float* m, float* v, float lr,
    const CacheGridView& g, CacheState* state, const uint32_t* count, hipStream_t st);
hipError_t cache_fill(float* p, size_t n, float value, hipStream_t st);
// after a frame's training with *count > 0 records: EMA of the first
// batch's loss, frames += 1, usable = frames >= warmup and
// loss_ema <= max_rel_loss (a NaN loss never passes); with `adapt` the
// training share moves toward 1.1 x `target` records a frame (at most 2x
// per frame, within [1e-5, 1])
hipError_t cache_update_state(CacheState* state, const float* first_loss, const uint32_t* count,
    int warmup, float max_rel_loss, int adapt, int target, hipStream_t st);

}  // namespace rtk
