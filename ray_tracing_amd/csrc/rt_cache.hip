// rt_cache.hip -- radiance cache kernels: hash-grid encoding, fully fused
// BF16 MFMA MLP (gfx950 v_mfma_f32_16x16x32_bf16, FP32 accumulation) for
// inference and training, weight gradients, Adam.
//
// The MLP runs transposed: a layer computes Z^T = W X^T with the weights as
// the MFMA A operand (rows = output channels) and 16 samples as the columns
// of the B operand, so each lane's accumulator holds 4 consecutive output
// channels of one sample and is written to the LDS activation row of that
// sample as one 8-byte store.  A block owns 256 samples (4 waves x 64);
// every wave keeps its 64 x 64 bf16 activation tile in LDS (row stride 72
// elements) and overwrites it layer by layer.  Training stores each layer's
// activations (bf16) for the backward pass, which runs the same layout with
// the transposed weights; weight gradients are GEMMs over samples written as
// per-block partial sums reduced in a fixed order; the grid gradient is
// scattered with FP32 atomics and the cells it reaches are flagged, so that
// Adam visits only those (sparse, lazy Adam) and leaves the gradient zero.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rt_bf16.h"
#include "rt_cache.h"

namespace rtk {
namespace {

constexpr int kRow = 72;  // LDS activation row stride (bf16 elements)

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

struct Range {
    int begin, m;
};

__device__ __forceinline__ Range batch_range(const uint32_t* count, int n, int cap, int batch,
    int batches) {
    int all = count ? (int)min(*count, (uint32_t)cap) : n;
    if (all < 0) all = 0;
    const int b0 = (int)((int64_t)all * batch / batches);
    const int b1 = (int)((int64_t)all * (batch + 1) / batches);
    return Range{b0, b1 - b0};
}

__device__ __forceinline__ float unit_clamp(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// the 8 corners of level l around p (in [0,1]^3): table indices and weights
__device__ __forceinline__ void grid_corners(const CacheGridView& g, int l, float px, float py,
    float pz, uint32_t idx[8], float w[8]) {
    const float r = g.res[l];
    const int ri = (int)r;
    const float x = px * r, y = py * r, z = pz * r;
    const int cx = min((int)floorf(x), ri - 1);
    const int cy = min((int)floorf(y), ri - 1);
    const int cz = min((int)floorf(z), ri - 1);
    const float fx = x - (float)cx, fy = y - (float)cy, fz = z - (float)cz;
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    const uint32_t stride = (uint32_t)ri + 1u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        w[k] = (wx[dx] * wy[dy]) * wz[dz];
        const uint32_t vx = (uint32_t)(cx + dx), vy = (uint32_t)(cy + dy), vz = (uint32_t)(cz + dz);
        idx[k] = g.dense[l] ? vx + vy * stride + vz * stride * stride
                            : ((vx * 1u) ^ (vy * 2654435761u) ^ (vz * 805459861u)) & g.tmask;
    }
}

__device__ __forceinline__ void unit_position(const CacheGridView& g, float4 a, float p[3]) {
    p[0] = unit_clamp((a.x - g.lo[0]) * g.inv_extent[0]);
    p[1] = unit_clamp((a.y - g.lo[1]) * g.inv_extent[1]);
    p[2] = unit_clamp((a.z - g.lo[2]) * g.inv_extent[2]);
}

// the 64 network inputs of one query (bf16)
__device__ __forceinline__ void cache_features(const CacheGridView& g, const float4* q, uint16_t x[64]) {
    const float4 a = q[0], b = q[1], c = q[2], d = q[3];
    float p[3];
    unit_position(g, a, p);
#pragma unroll
    for (int l = 0; l < kCacheMaxLevels; ++l) {
        float f0 = 0.0f, f1 = 0.0f;
        if (l < g.levels) {
            uint32_t idx[8];
            float w[8];
            grid_corners(g, l, p[0], p[1], p[2], idx, w);
            const float2* t = reinterpret_cast<const float2*>(g.table) + ((size_t)l << g.log2_table);
            float2 v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = t[idx[k]];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                f0 = f0 + w[k] * v[k].x;
                f1 = f1 + w[k] * v[k].y;
            }
        }
        x[2 * l] = to_bf16(f0);
        x[2 * l + 1] = to_bf16(f1);
    }
    x[kCacheChDir] = to_bf16(b.x);
    x[kCacheChDir + 1] = to_bf16(b.y);
    x[kCacheChDir + 2] = to_bf16(b.z);
    x[kCacheChNormal] = to_bf16(c.x);
    x[kCacheChNormal + 1] = to_bf16(c.y);
    x[kCacheChNormal + 2] = to_bf16(c.z);
    x[kCacheChRough] = to_bf16(b.w);
    x[kCacheChAlbedo] = to_bf16(d.x);
    x[kCacheChAlbedo + 1] = to_bf16(d.y);
    x[kCacheChAlbedo + 2] = to_bf16(d.z);
    x[kCacheChOne] = 0x3F80u;  // 1.0
#pragma unroll
    for (int k = kCacheChOne + 1; k < kCacheIn; ++k) x[k] = 0;
}

__device__ __forceinline__ void store_row(uint16_t* dst, const uint16_t x[64]) {
    uint4* d = reinterpret_cast<uint4*>(dst);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(x);
#pragma unroll
    for (int v = 0; v < 8; ++v) d[v] = make_uint4(w[4 * v], w[4 * v + 1], w[4 * v + 2], w[4 * v + 3]);
}

__device__ __forceinline__ void copy_row(uint16_t* dst, const uint16_t* src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int v = 0; v < 8; ++v) d[v] = s[v];
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d) {
    return make_uint2((uint32_t)to_bf16(a) | ((uint32_t)to_bf16(b) << 16),
        (uint32_t)to_bf16(c) | ((uint32_t)to_bf16(d) << 16));
}

struct FwdArgs {
    CacheGridView g;
    CachePacked pk;
    const float4* in;
    int in_stride;          // float4s between inputs
    const uint32_t* count;  // device count (or null: n)
    int n, cap, batch, batches;
    float4* out;            // MODE 0 / 2: prediction per local index
    const float4* thr;      // MODE 1: path throughput
    float4* radiance;       // MODE 1: path radiance
    uint16_t* act;          // MODE 2: [5][cap][64]
};

// MODE 0: out[i] = y; MODE 1: radiance[slot] += thr[slot] * max(y, 0);
// MODE 2: training forward (stores x, h1..h4 and y)
template <int MODE>
__global__ __launch_bounds__(256) void k_cache_forward(FwdArgs a) {
    __shared__ __align__(16) uint16_t xs[4 * 64 * kRow];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, g4 = lane >> 4;
    const Range rg = batch_range(a.count, a.n, a.cap, a.batch, a.batches);
    uint16_t* tile_lds = xs + wave * 64 * kRow;
    for (int tile = blockIdx.x; tile * 256 < rg.m; tile += gridDim.x) {
        const int li = tile * 256 + wave * 64 + lane;
        {
            uint16_t x[64];
            if (li < rg.m) {
                cache_features(a.g, a.in + (size_t)(rg.begin + li) * a.in_stride, x);
            } else {
#pragma unroll
                for (int k = 0; k < 64; ++k) x[k] = 0;
            }
            store_row(tile_lds + lane * kRow, x);
            if (MODE == 2 && li < rg.m) store_row(a.act + (size_t)li * 64, x);
        }
        __syncthreads();
        for (int l = 0; l < kCacheHiddenLayers; ++l) {
            bf16x8 B[2][4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
                for (int c = 0; c < 2; ++c) B[c][s] = ld8(tile_lds + (16 * s + r16) * kRow + 32 * c + 8 * g4);
            }
            __syncthreads();
            const uint16_t* W = a.pk.fwd[l];
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const bf16x8 A0 = ld8(W + (16 * o + r16) * 64 + 8 * g4);
                const bf16x8 A1 = ld8(W + (16 * o + r16) * 64 + 32 + 8 * g4);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B[0][s], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B[1][s], acc, 0, 0, 0);
                    // D[row = channel 16 o + 4 g4 + e][col = sample 16 s + r16]
                    *reinterpret_cast<uint2*>(tile_lds + (16 * s + r16) * kRow + 16 * o + 4 * g4) =
                        pack4(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f), fmaxf(acc[2], 0.0f),
                            fmaxf(acc[3], 0.0f));
                }
            }
            __syncthreads();
            if (MODE == 2 && li < rg.m)
                copy_row(a.act + ((size_t)(l + 1) * a.cap + li) * 64, tile_lds + lane * kRow);
        }
        {
            bf16x8 B[2][4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
                for (int c = 0; c < 2; ++c) B[c][s] = ld8(tile_lds + (16 * s + r16) * kRow + 32 * c + 8 * g4);
            }
            const uint16_t* W = a.pk.fwd[kCacheHiddenLayers];
            const bf16x8 A0 = ld8(W + r16 * 64 + 8 * g4);
            const bf16x8 A1 = ld8(W + r16 * 64 + 32 + 8 * g4);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B[0][s], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B[1][s], acc, 0, 0, 0);
                const int ls = tile * 256 + wave * 64 + 16 * s + r16;
                if (g4 == 0 && ls < rg.m) {
                    if (MODE == 1) {
                        const float4 q = a.in[(size_t)(rg.begin + ls) * a.in_stride];
                        const uint32_t slot = __float_as_uint(q.w);
                        const float4 t = a.thr[slot];
                        float4 rad = a.radiance[slot];
                        rad.x = rad.x + t.x * fmaxf(acc[0], 0.0f);
                        rad.y = rad.y + t.y * fmaxf(acc[1], 0.0f);
                        rad.z = rad.z + t.z * fmaxf(acc[2], 0.0f);
                        a.radiance[slot] = rad;
                    } else {
                        a.out[ls] = make_float4(acc[0], acc[1], acc[2], 0.0f);
                    }
                }
            }
        }
        __syncthreads();
    }
}

struct BwdArgs {
    CacheGridView g;
    CachePacked pk;
    const float4* in;
    int in_stride;
    const float4* target;
    int t_stride;
    const uint32_t* count;
    int cap, batch, batches;
    const float4* pred;
    const uint16_t* act;   // [5][cap][64]
    uint16_t* dz;          // [5][cap][64]: dz1..dz4, dy
    float* loss_partial;   // per 256-sample tile
};

// loss, output gradient, back-propagation through the four hidden layers
// (stores dy, dz4..dz1) and the hash-grid gradient of the first 32 inputs
__global__ __launch_bounds__(256) void k_cache_backward(BwdArgs a) {
    __shared__ __align__(16) uint16_t xs[4 * 64 * kRow];
    __shared__ float red[256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, g4 = lane >> 4;
    const Range rg = batch_range(a.count, 0, a.cap, a.batch, a.batches);
    const float scale = rg.m > 0 ? 1.0f / (3.0f * (float)rg.m) : 0.0f;
    uint16_t* tile_lds = xs + wave * 64 * kRow;
    const size_t plane = (size_t)a.cap * 64;
    for (int tile = blockIdx.x; tile * 256 < rg.m; tile += gridDim.x) {
        const int li = tile * 256 + wave * 64 + lane;
        float lsum = 0.0f;
        {
            uint16_t x[64];
#pragma unroll
            for (int k = 0; k < 64; ++k) x[k] = 0;
            if (li < rg.m) {
                const float4 y = a.pred[li];
                const float4 t = a.target[(size_t)(rg.begin + li) * a.t_stride];
                const float yv[3] = {y.x, y.y, y.z}, tv[3] = {t.x, t.y, t.z};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float d = yv[c] - tv[c];
                    const float den = yv[c] * yv[c] + kCacheLossEps;
                    lsum = lsum + (d * d) / den;
                    x[c] = to_bf16(((2.0f * d) / den) * scale);
                }
                store_row(a.dz + 4 * plane + (size_t)li * 64, x);
            }
            store_row(tile_lds + lane * kRow, x);
        }
        red[threadIdx.x] = lsum * scale;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.loss_partial[tile] = red[0];
        // output layer: dh4 = W4^T dy (K = 32, dy in columns 0..2)
        {
            bf16x8 B[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) B[s] = ld8(tile_lds + (16 * s + r16) * kRow + 8 * g4);
            __syncthreads();
            const uint16_t* WT = a.pk.bwd[kCacheHiddenLayers];
            const uint16_t* h = a.act + 4 * plane;
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const bf16x8 A = ld8(WT + (16 * o + r16) * 32 + 8 * g4);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[s], acc, 0, 0, 0);
                    const int ls = tile * 256 + wave * 64 + 16 * s + r16;
                    uint2 hm = make_uint2(0u, 0u);
                    if (ls < rg.m) hm = *reinterpret_cast<const uint2*>(h + (size_t)ls * 64 + 16 * o + 4 * g4);
                    const float m0 = from_bf16((uint16_t)(hm.x & 0xFFFFu)), m1 = from_bf16((uint16_t)(hm.x >> 16));
                    const float m2 = from_bf16((uint16_t)(hm.y & 0xFFFFu)), m3 = from_bf16((uint16_t)(hm.y >> 16));
                    *reinterpret_cast<uint2*>(tile_lds + (16 * s + r16) * kRow + 16 * o + 4 * g4) =
                        pack4(m0 > 0.0f ? acc[0] : 0.0f, m1 > 0.0f ? acc[1] : 0.0f,
                            m2 > 0.0f ? acc[2] : 0.0f, m3 > 0.0f ? acc[3] : 0.0f);
                }
            }
            __syncthreads();
            if (li < rg.m) copy_row(a.dz + 3 * plane + (size_t)li * 64, tile_lds + lane * kRow);
        }
        // hidden layers l = 3, 2, 1: dh_l = W_l^T dz_{l+1}, masked by h_l > 0
        for (int l = kCacheHiddenLayers - 1; l >= 1; --l) {
            bf16x8 B[2][4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
                for (int c = 0; c < 2; ++c) B[c][s] = ld8(tile_lds + (16 * s + r16) * kRow + 32 * c + 8 * g4);
            }
            __syncthreads();
            const uint16_t* WT = a.pk.bwd[l];
            const uint16_t* h = a.act + (size_t)l * plane;
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const bf16x8 A0 = ld8(WT + (16 * o + r16) * 64 + 8 * g4);
                const bf16x8 A1 = ld8(WT + (16 * o + r16) * 64 + 32 + 8 * g4);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B[0][s], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B[1][s], acc, 0, 0, 0);
                    const int ls = tile * 256 + wave * 64 + 16 * s + r16;
                    uint2 hm = make_uint2(0u, 0u);
                    if (ls < rg.m) hm = *reinterpret_cast<const uint2*>(h + (size_t)ls * 64 + 16 * o + 4 * g4);
                    const float m0 = from_bf16((uint16_t)(hm.x & 0xFFFFu)), m1 = from_bf16((uint16_t)(hm.x >> 16));
                    const float m2 = from_bf16((uint16_t)(hm.y & 0xFFFFu)), m3 = from_bf16((uint16_t)(hm.y >> 16));
                    *reinterpret_cast<uint2*>(tile_lds + (16 * s + r16) * kRow + 16 * o + 4 * g4) =
                        pack4(m0 > 0.0f ? acc[0] : 0.0f, m1 > 0.0f ? acc[1] : 0.0f,
                            m2 > 0.0f ? acc[2] : 0.0f, m3 > 0.0f ? acc[3] : 0.0f);
                }
            }
            __syncthreads();
            if (li < rg.m) copy_row(a.dz + (size_t)(l - 1) * plane + (size_t)li * 64, tile_lds + lane * kRow);
        }
        // input gradient of the 32 hash features: dx = W0^T dz1, scattered to
        // the grid cells with the trilinear weights
        {
            bf16x8 B[2][4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
                for (int c = 0; c < 2; ++c) B[c][s] = ld8(tile_lds + (16 * s + r16) * kRow + 32 * c + 8 * g4);
            }
            const uint16_t* WT = a.pk.bwd[0];
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                const bf16x8 A0 = ld8(WT + (16 * o + r16) * 64 + 8 * g4);
                const bf16x8 A1 = ld8(WT + (16 * o + r16) * 64 + 32 + 8 * g4);
#pragma unroll 1
                for (int s = 0; s < 4; ++s) {
                    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B[0][s], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B[1][s], acc, 0, 0, 0);
                    const int ls = tile * 256 + wave * 64 + 16 * s + r16;
                    if (ls >= rg.m) continue;
                    // channels 16 o + 4 g4 + e = levels 8 o + 2 g4 + (e >> 1), feature e & 1
                    float p[3];
                    unit_position(a.g, a.in[(size_t)(rg.begin + ls) * a.in_stride], p);
#pragma unroll
                    for (int half = 0; half < 2; ++half) {
                        const int lv = 8 * o + 2 * g4 + half;
                        if (lv >= a.g.levels) continue;
                        uint32_t idx[8];
                        float w[8];
                        grid_corners(a.g, lv, p[0], p[1], p[2], idx, w);
                        float* gt = a.g.grad + ((size_t)lv << a.g.log2_table) * 2;
                        const float d0 = acc[2 * half], d1 = acc[2 * half + 1];
                        uint8_t* tl = a.g.touched + ((size_t)lv << a.g.log2_table);
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            atomicAdd(gt + (size_t)idx[k] * 2, w[k] * d0);
                            atomicAdd(gt + (size_t)idx[k] * 2 + 1, w[k] * d1);
                            tl[idx[k]] = 1;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
}

// Weight gradient of layer l as a GEMM over samples:
//   G[o][i] = sum_s dz[s][o] * x[s][i]
// (dz = dz_{l+1}, or dy for the output layer; x = x or h_l).  Block b covers
// local samples [b * kCacheGradSamples, ...); wave w owns input tile w and
// output tiles 0 .. out_tiles - 1; G goes to partial[b] ([64][64]).
__global__ __launch_bounds__(256) void k_cache_wgrad(const uint16_t* dz, const uint16_t* x,
    const uint32_t* count, int cap, int batch, int batches, int out_tiles, float* partial) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, g4 = lane >> 4;
    const Range rg = batch_range(count, 0, cap, batch, batches);
    const int p0 = blockIdx.x * kCacheGradSamples;
    if (p0 >= rg.m) return;
    const int p1 = min(rg.m, p0 + kCacheGradSamples);
    f32x4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int it = wave;
    for (int s0 = p0; s0 < p1; s0 += 32) {
        uint16_t bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int sidx = s0 + 8 * g4 + j;
            bv[j] = sidx < p1 ? x[(size_t)sidx * 64 + 16 * it + r16] : (uint16_t)0;
        }
        uint4 rb;
        rb.x = (uint32_t)bv[0] | ((uint32_t)bv[1] << 16);
        rb.y = (uint32_t)bv[2] | ((uint32_t)bv[3] << 16);
        rb.z = (uint32_t)bv[4] | ((uint32_t)bv[5] << 16);
        rb.w = (uint32_t)bv[6] | ((uint32_t)bv[7] << 16);
        const bf16x8 B = __builtin_bit_cast(bf16x8, rb);
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
            if (ot >= out_tiles) break;
            uint16_t av[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int sidx = s0 + 8 * g4 + j;
                av[j] = sidx < p1 ? dz[(size_t)sidx * 64 + 16 * ot + r16] : (uint16_t)0;
            }
            uint4 ra;
            ra.x = (uint32_t)av[0] | ((uint32_t)av[1] << 16);
            ra.y = (uint32_t)av[2] | ((uint32_t)av[3] << 16);
            ra.z = (uint32_t)av[4] | ((uint32_t)av[5] << 16);
            ra.w = (uint32_t)av[6] | ((uint32_t)av[7] << 16);
            const bf16x8 A = __builtin_bit_cast(bf16x8, ra);
            acc[ot] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[ot], 0, 0, 0);
        }
    }
    float* dst = partial + (size_t)blockIdx.x * 64 * 64;
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        if (ot >= out_tiles) break;
#pragma unroll
        for (int e = 0; e < 4; ++e) dst[(size_t)(16 * ot + 4 * g4 + e) * 64 + 16 * it + r16] = acc[ot][e];
    }
}

// sum the partials of the blocks that held samples, in block order
__global__ void k_cache_wgrad_reduce(const float* partial, const uint32_t* count, int cap, int batch,
    int batches, int outs, float* grad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= outs * 64) return;
    const Range rg = batch_range(count, 0, cap, batch, batches);
    const int nb = cdiv(rg.m, kCacheGradSamples);
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += partial[(size_t)b * 64 * 64 + i];
    grad[i] = s;
}

__global__ void k_cache_loss_reduce(const float* partial, const uint32_t* count, int cap, int batch,
    int batches, float* loss) {
    if (threadIdx.x != 0) return;
    const Range rg = batch_range(count, 0, cap, batch, batches);
    const int nt = cdiv(rg.m, 256);
    float s = 0.0f;
    for (int t = 0; t < nt; ++t) s += partial[t];
    *loss = s;
}

__global__ void k_cache_adam(int64_t n, float* params, const float* grad, float* m, float* v,
    float lr, const CacheState* state, const uint32_t* count) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || (count && *count == 0u)) return;
    const float b1 = 0.9f, b2 = 0.999f;
    const float t = (float)(state->steps + 1);
    const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// sparse Adam over the grid: 4 cells per thread, only the touched ones
// (their 2 features), whose gradients and flags are cleared afterwards
__global__ void k_cache_adam_grid(int64_t cells, float* params, float* grad, float* m, float* v,
    uint8_t* touched, float lr, const CacheState* state, const uint32_t* count) {
    const int64_t c0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (c0 >= cells || (count && *count == 0u)) return;
    const uint32_t flags = *reinterpret_cast<const uint32_t*>(touched + c0);
    if (flags == 0u) return;
    const float b1 = 0.9f, b2 = 0.999f;
    const float t = (float)(state->steps + 1);
    const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (((flags >> (8 * k)) & 0xFFu) == 0u) continue;
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            const int64_t i = (c0 + k) * 2 + f;
            const float gi = grad[i];
            const float mi = b1 * m[i] + (1.0f - b1) * gi;
            const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
            m[i] = mi;
            v[i] = vi;
            params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
            grad[i] = 0.0f;
        }
    }
    *reinterpret_cast<uint32_t*>(touched + c0) = 0u;
}

__global__ void k_cache_step(CacheState* state, const uint32_t* count) {
    if (threadIdx.x == 0 && !(count && *count == 0u)) state->steps += 1;
}

__global__ void k_cache_update_state(CacheState* state, const float* first_loss,
    const uint32_t* count, int warmup, float max_rel_loss, int adapt, int target) {
    if (threadIdx.x != 0) return;
    const uint32_t n = *count;
    if (adapt) {
        const float want = 1.1f * (float)target;
        const float ratio = n > 0u ? want / (float)n : 2.0f;
        const float f = state->train_fraction * fminf(fmaxf(ratio, 0.5f), 2.0f);
        state->train_fraction = fminf(fmaxf(f, 1e-5f), 1.0f);
    }
    if (n == 0u) return;
    const float l = *first_loss;
    state->last_loss = l;
    state->loss_ema = state->frames == 0 ? l : 0.8f * state->loss_ema + 0.2f * l;
    state->frames += 1;
    state->usable = (state->frames >= warmup && state->loss_ema <= max_rel_loss) ? 1 : 0;
}

__global__ void k_cache_fill(float* p, size_t n, float value) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = value;
}

// forward A operands W [out][64] (W4 padded to 16 rows) and backward A
// operands W^T [64][out] (W4^T padded to 32 columns), bf16
__global__ void k_cache_pack(const float* params, CachePacked pk) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 64 * 64) return;
    const int r = i / 64, c = i % 64;
#pragma unroll
    for (int l = 0; l < kCacheHiddenLayers; ++l) {
        const float* W = params + cache_layer_offset(l);
        pk.fwd[l][i] = to_bf16(W[i]);              // [o = r][i = c]
        pk.bwd[l][i] = to_bf16(W[c * 64 + r]);     // [i = r][o = c] = W[c][r]
    }
    const float* W4 = params + cache_layer_offset(kCacheHiddenLayers);
    if (i < 16 * 64) pk.fwd[kCacheHiddenLayers][i] = r < kCacheOut ? to_bf16(W4[i]) : (uint16_t)0;
    if (c < 32) {
        pk.bwd[kCacheHiddenLayers][r * 32 + c] = c < kCacheOut ? to_bf16(W4[c * 64 + r]) : (uint16_t)0;
    }
}

}  // namespace

hipError_t cache_pack(const float* params, const CachePacked& pk, hipStream_t st) {
    k_cache_pack<<<16, 256, 0, st>>>(params, pk);
    return hipGetLastError();
}

hipError_t cache_infer(const CacheGridView& g, const CachePacked& pk, const CacheInput* in, int n,
    float4* out, int grid, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    FwdArgs a{};
    a.g = g;
    a.pk = pk;
    a.in = reinterpret_cast<const float4*>(in);
    a.in_stride = 4;
    a.count = nullptr;
    a.n = n;
    a.cap = n;
    a.batch = 0;
    a.batches = 1;
    a.out = out;
    k_cache_forward<0><<<std::max(1, std::min(grid, cdiv(n, 256))), 256, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t cache_infer_paths(const CacheGridView& g, const CachePacked& pk, const CacheInput* in,
    const uint32_t* count, int cap, const float4* throughput, float4* radiance, int grid,
    hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    FwdArgs a{};
    a.g = g;
    a.pk = pk;
    a.in = reinterpret_cast<const float4*>(in);
    a.in_stride = 4;
    a.count = count;
    a.n = 0;
    a.cap = cap;
    a.batch = 0;
    a.batches = 1;
    a.thr = throughput;
    a.radiance = radiance;
    k_cache_forward<1><<<std::max(1, std::min(grid, cdiv(cap, 256))), 256, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t cache_train_pass(const CacheGridView& g, const CachePacked& pk,
    const CacheTrainScratch& w, const float4* in, int in_stride, const float4* target,
    int t_stride, const uint32_t* count, int cap, int batch, int batches, float* grad, float* loss,
    int grid, hipStream_t st) {
    if (cap <= 0) return hipSuccess;
    const int blocks = std::max(1, std::min(grid, cdiv(cap, 256)));
    FwdArgs f{};
    f.g = g;
    f.pk = pk;
    f.in = in;
    f.in_stride = in_stride;
    f.count = count;
    f.n = 0;
    f.cap = cap;
    f.batch = batch;
    f.batches = batches;
    f.out = w.pred;
    f.act = w.act;
    k_cache_forward<2><<<blocks, 256, 0, st>>>(f);
    BwdArgs b{};
    b.g = g;
    b.pk = pk;
    b.in = in;
    b.in_stride = in_stride;
    b.target = target;
    b.t_stride = t_stride;
    b.count = count;
    b.cap = cap;
    b.batch = batch;
    b.batches = batches;
    b.pred = w.pred;
    b.act = w.act;
    b.dz = w.dz;
    b.loss_partial = w.loss_partial;
    k_cache_backward<<<blocks, 256, 0, st>>>(b);
    const size_t plane = (size_t)cap * 64;
    const int gblocks = cdiv(cap, kCacheGradSamples);
    for (int l = 0; l <= kCacheHiddenLayers; ++l) {
        const bool out_layer = l == kCacheHiddenLayers;
        // dW_l pairs dz_{l+1} (dz plane l; dy is plane 4) with the layer input (act plane l)
        k_cache_wgrad<<<gblocks, 256, 0, st>>>(w.dz + (size_t)l * plane, w.act + (size_t)l * plane,
            count, cap, batch, batches, out_layer ? 1 : 4, w.grad_partial);
        const int outs = out_layer ? kCacheOut : kCacheHidden;
        k_cache_wgrad_reduce<<<cdiv(outs * 64, 256), 256, 0, st>>>(w.grad_partial, count, cap,
            batch, batches, outs, grad + cache_layer_offset(l));
    }
    k_cache_loss_reduce<<<1, 64, 0, st>>>(w.loss_partial, count, cap, batch, batches, loss);
    return hipGetLastError();
}

hipError_t cache_adam(int64_t n_mlp, float* params, const float* grad, float* m, float* v, float lr,
    const CacheGridView& g, CacheState* state, const uint32_t* count, hipStream_t st) {
    if (n_mlp > 0)
        k_cache_adam<<<(unsigned)((n_mlp + 255) / 256), 256, 0, st>>>(n_mlp, params, grad, m, v, lr,
            state, count);
    const int64_t cells = (int64_t)g.levels << g.log2_table;   // a multiple of 4 (log2_table >= 8)
    if (cells > 0)
        k_cache_adam_grid<<<(unsigned)((cells / 4 + 255) / 256), 256, 0, st>>>(cells, params + n_mlp,
            g.grad, m + n_mlp, v + n_mlp, g.touched, lr, state, count);
    k_cache_step<<<1, 64, 0, st>>>(state, count);
    return hipGetLastError();
}

hipError_t cache_fill(float* p, size_t n, float value, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_cache_fill<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n, value);
    return hipGetLastError();
}

hipError_t cache_update_state(CacheState* state, const float* first_loss, const uint32_t* count,
    int warmup, float max_rel_loss, int adapt, int target, hipStream_t st) {
    k_cache_update_state<<<1, 64, 0, st>>>(state, first_loss, count, warmup, max_rel_loss, adapt,
        target);
    return hipGetLastError();
}

}  // namespace rtk
