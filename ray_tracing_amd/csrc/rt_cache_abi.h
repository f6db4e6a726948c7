// rt_cache_abi.h -- the radiance-cache object shared by its C ABI and the
// renderer that trains and queries it.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>

#include "../../include/rt/rt_hip.h"
#include "rt_alloc.h"
#include "rt_cache.h"

template <typename T>
struct CacheBuf {
    T* p = nullptr;
    size_t n = 0;
    CacheBuf() = default;
    CacheBuf(const CacheBuf&) = delete;
    CacheBuf& operator=(const CacheBuf&) = delete;
    ~CacheBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t reserve(size_t c) {
        if (c <= n && p) return hipSuccess;
        release();
        if (c == 0) c = 1;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * c);
        if (e == hipSuccess) n = c;
        rt_alloc_bump();
        return e;
    }
};

struct rt_cache {
    std::atomic<int> refs{1};  // the caller's handle plus each renderer it is attached to
    int ordinal = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    rt_cache_config cfg{};
    int64_t n_mlp = 0, n_grid = 0, n_total = 0;
    CacheBuf<float> params, grad, m, v;    // MLP then grid
    CacheBuf<uint8_t> touched;             // grid cells with a gradient in the current step
    CacheBuf<uint16_t> fwd[rtk::kCacheHiddenLayers + 1], bwd[rtk::kCacheHiddenLayers + 1];
    CacheBuf<rtk::CacheState> state;
    CacheBuf<float> loss;                  // one per training batch of a frame
    // training scratch (capacity samples)
    int cap = 0;
    CacheBuf<uint16_t> act, dz;
    CacheBuf<float4> pred;
    CacheBuf<float> loss_partial, grad_partial;
    // host-API staging
    CacheBuf<float4> stage_in, stage_tgt, stage_out;
    CacheBuf<uint32_t> stage_count;
    // encoding constants
    float res[rtk::kCacheMaxLevels] = {};
    int32_t dense[rtk::kCacheMaxLevels] = {};
    float lo[3] = {0.0f, 0.0f, 0.0f}, inv_extent[3] = {1.0f, 1.0f, 1.0f};
    // last frame of the renderer it is attached to
    int32_t last_records = 0, last_queries = 0;
    bool fraction_set = false;  // CacheState::train_fraction holds a seeded value
    float fraction_seed = 0.0f; // host source of that seed copy

    rtk::CacheGridView grid_view() const {
        rtk::CacheGridView g{};
        g.table = params.p + n_mlp;
        g.grad = grad.p + n_mlp;
        g.touched = touched.p;
        g.levels = cfg.levels;
        g.tmask = (1u << cfg.log2_table) - 1u;
        g.log2_table = cfg.log2_table;
        for (int l = 0; l < rtk::kCacheMaxLevels; ++l) {
            g.res[l] = res[l];
            g.dense[l] = dense[l];
        }
        for (int a = 0; a < 3; ++a) {
            g.lo[a] = lo[a];
            g.inv_extent[a] = inv_extent[a];
        }
        return g;
    }

    rtk::CachePacked packed() const {
        rtk::CachePacked pk{};
        for (int l = 0; l <= rtk::kCacheHiddenLayers; ++l) {
            pk.fwd[l] = fwd[l].p;
            pk.bwd[l] = bwd[l].p;
        }
        return pk;
    }

    hipError_t ensure(int c) {
        if (c <= cap) return hipSuccess;
        const size_t planes = (size_t)(rtk::kCacheHiddenLayers + 1) * (size_t)c * 64;
        hipError_t e = act.reserve(planes);
        if (e == hipSuccess) e = dz.reserve(planes);
        if (e == hipSuccess) e = pred.reserve((size_t)c);
        if (e == hipSuccess) e = loss_partial.reserve((size_t)(c + 255) / 256);
        if (e == hipSuccess)
            e = grad_partial.reserve((size_t)((c + rtk::kCacheGradSamples - 1) / rtk::kCacheGradSamples) * 64 * 64);
        if (e == hipSuccess) cap = c;
        return e;
    }

    rtk::CacheTrainScratch scratch() const {
        rtk::CacheTrainScratch w{};
        w.act = act.p;
        w.dz = dz.p;
        w.pred = pred.p;
        w.loss_partial = loss_partial.p;
        w.grad_partial = grad_partial.p;
        return w;
    }

    int grid() const { return (cu_count > 0 ? cu_count : 1) * 4; }
};

void rt_cache_retain(rt_cache* c);
void rt_cache_release(rt_cache* c);

// One frame's training on the renderer's stream: `train_steps` batches of
// the first min(*count, cap) records (each: forward, loss, backward, Adam on
// the MLP and the grid, repack), then the state update (fallback flag).
hipError_t rt_cache_train_frame(rt_cache* c, const rtk::CacheRecord* records,
    const uint32_t* count, int cap, hipStream_t st);
