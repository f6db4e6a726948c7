// rt_cache_api.hip -- C ABI of the radiance cache: configuration, parameters,
// host-buffer inference and training steps, and the per-frame training used
// by the renderer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rt_cache_abi.h"
#include "rt_nn_abi.h"

namespace {

rt_status cache_fail(rt_status s, const std::string& msg) {
    rt_set_error(msg.c_str());
    return s;
}

#define CACHE_HIP(expr)                                                                          \
    do {                                                                                         \
        const hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                                    \
            return cache_fail(RT_ERROR_HIP, std::string("HIP runtime failure at ") + #expr + ": " \
                                                + hipGetErrorString(e_));                        \
    } while (0)

#define CACHE_REQUIRE(cond, status, msg) \
    do {                                 \
        if (!(cond)) return cache_fail(status, msg); \
    } while (0)

rt_status check_config(const rt_cache_config* c) {
    CACHE_REQUIRE(c != nullptr, RT_ERROR_INVALID_ARGUMENT, "null cache configuration");
    CACHE_REQUIRE(c->levels >= 1 && c->levels <= rtk::kCacheMaxLevels, RT_ERROR_INVALID_ARGUMENT,
        "levels must be in [1, 16]");
    CACHE_REQUIRE(c->log2_table >= 8 && c->log2_table <= 22, RT_ERROR_INVALID_ARGUMENT,
        "log2_table must be in [8, 22]");
    CACHE_REQUIRE(c->base_resolution >= 2 && c->max_resolution >= c->base_resolution
                      && c->max_resolution <= (1 << 20),
        RT_ERROR_INVALID_ARGUMENT, "resolutions must satisfy 2 <= base <= max <= 2^20");
    CACHE_REQUIRE(c->train_samples >= 256 && c->train_samples <= (1 << 20), RT_ERROR_INVALID_ARGUMENT,
        "train_samples must be in [256, 2^20]");
    CACHE_REQUIRE(c->train_steps >= 1 && c->train_steps <= 64, RT_ERROR_INVALID_ARGUMENT,
        "train_steps must be in [1, 64]");
    CACHE_REQUIRE(c->query_bounce >= 1, RT_ERROR_INVALID_ARGUMENT, "query_bounce must be >= 1");
    CACHE_REQUIRE(c->warmup_frames >= 0, RT_ERROR_INVALID_ARGUMENT, "warmup_frames must be >= 0");
    CACHE_REQUIRE(c->learning_rate > 0.0f && std::isfinite(c->learning_rate), RT_ERROR_INVALID_ARGUMENT,
        "learning_rate must be positive");
    CACHE_REQUIRE(c->train_fraction >= 0.0f && c->train_fraction <= 1.0f, RT_ERROR_INVALID_ARGUMENT,
        "train_fraction must be in [0, 1] (0: automatic)");
    return RT_OK;
}

// res_l = floor(base * b^l + 1e-6), b = (max / base)^(1 / (L - 1)); dense
// storage where the (res + 1)^3 grid vertices fit the table
void resolutions(const rt_cache_config* c, float res[16], int32_t dense[16]) {
    const double base = c->base_resolution, mx = c->max_resolution;
    const double b = c->levels > 1 ? std::exp((std::log(mx) - std::log(base)) / (double)(c->levels - 1)) : 1.0;
    const uint64_t T = 1ull << c->log2_table;
    for (int l = 0; l < rtk::kCacheMaxLevels; ++l) {
        if (l < c->levels) {
            const double r = std::floor(base * std::pow(b, (double)l) + 1e-6);
            res[l] = (float)r;
            const uint64_t v = (uint64_t)r + 1ull;
            dense[l] = v * v * v <= T ? 1 : 0;
        } else {
            res[l] = 0.0f;
            dense[l] = 0;
        }
    }
}

// PCG32 (O'Neill 2014), XSH-RR output
struct Pcg {
    uint64_t state = 0, inc = 0;
    explicit Pcg(uint64_t seed) {
        inc = (54ull << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// He-uniform MLP weights, grid entries U(-1e-4, 1e-4)
void init_params(const rt_cache_config* c, int64_t n_mlp, int64_t n_total, std::vector<float>& p) {
    p.assign((size_t)n_total, 0.0f);
    Pcg r(c->seed);
    const float a = std::sqrt(6.0f / (float)rtk::kCacheIn);
    for (int64_t i = 0; i < n_mlp; ++i) p[(size_t)i] = (2.0f * r.uniform() - 1.0f) * a;
    for (int64_t i = n_mlp; i < n_total; ++i) p[(size_t)i] = (2.0f * r.uniform() - 1.0f) * 1e-4f;
}

rt_status reset_training(rt_cache* c) {
    const size_t bytes = sizeof(float) * (size_t)c->n_total;
    CACHE_HIP(hipMemsetAsync(c->m.p, 0, bytes, c->stream));
    CACHE_HIP(hipMemsetAsync(c->v.p, 0, bytes, c->stream));
    CACHE_HIP(hipMemsetAsync(c->grad.p, 0, bytes, c->stream));
    CACHE_HIP(hipMemsetAsync(c->touched.p, 0, (size_t)(c->n_grid / 2), c->stream));
    CACHE_HIP(hipMemsetAsync(c->state.p, 0, sizeof(rtk::CacheState), c->stream));
    c->fraction_set = false;  // the renderer seeds the training share on its next frame
    CACHE_HIP(rtk::cache_pack(c->params.p, c->packed(), c->stream));
    return RT_OK;
}

}  // namespace

void rt_cache_retain(rt_cache* c) { c->refs.fetch_add(1); }

void rt_cache_release(rt_cache* c) {
    if (c->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(c->ordinal);
    (void)hipStreamSynchronize(c->stream);
    hipStream_t st = c->stream;
    delete c;
    if (st) (void)hipStreamDestroy(st);
}

hipError_t rt_cache_train_frame(rt_cache* c, const rtk::CacheRecord* records,
    const uint32_t* count, int cap, hipStream_t st) {
    const int steps = c->cfg.train_steps;
    const rtk::CacheGridView g = c->grid_view();
    const rtk::CachePacked pk = c->packed();
    const rtk::CacheTrainScratch w = c->scratch();
    const float4* in = reinterpret_cast<const float4*>(records);
    const float4* tgt = &records[0].target;
    constexpr int stride = (int)(sizeof(rtk::CacheRecord) / sizeof(float4));
    hipError_t e = hipSuccess;
    // the grid gradient is zero on entry to each step: the sparse Adam of the
    // previous step cleared the cells it touched
    for (int k = 0; k < steps && e == hipSuccess; ++k) {
        e = rtk::cache_train_pass(g, pk, w, in, stride, tgt, stride, count, cap, k, steps,
            c->grad.p, c->loss.p + k, c->grid(), st);
        if (e == hipSuccess)
            e = rtk::cache_adam(c->n_mlp, c->params.p, c->grad.p, c->m.p, c->v.p,
                c->cfg.learning_rate, g, c->state.p, count, st);
        if (e == hipSuccess) e = rtk::cache_pack(c->params.p, pk, st);
    }
    if (e == hipSuccess)
        e = rtk::cache_update_state(c->state.p, c->loss.p, count, c->cfg.warmup_frames,
            c->cfg.max_rel_loss, c->cfg.train_fraction > 0.0f ? 0 : 1, c->cfg.train_samples, st);
    return e;
}

extern "C" {

void rt_cache_default_config(rt_cache_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->levels = 16;
    c->log2_table = 19;
    c->base_resolution = 16;
    c->max_resolution = 2048;
    c->train_samples = 65536;
    c->train_steps = 4;
    c->query_bounce = 2;
    c->warmup_frames = 4;
    c->learning_rate = 1e-2f;
    c->max_rel_loss = 8.0f;
    c->train_fraction = 0.0f;
    c->seed = 1u;
}

int32_t rt_cache_mlp_param_count(void) { return rtk::cache_mlp_params(); }

int64_t rt_cache_param_count(const rt_cache_config* c) {
    if (check_config(c) != RT_OK) return -1;
    return (int64_t)rtk::cache_mlp_params() + ((int64_t)c->levels << c->log2_table) * 2;
}

rt_status rt_cache_create(int ordinal, const rt_cache_config* cfg, rt_cache** out) {
    CACHE_REQUIRE(out != nullptr, RT_ERROR_INVALID_ARGUMENT, "null output");
    const rt_status cs = check_config(cfg);
    if (cs != RT_OK) return cs;
    int count = 0;
    CACHE_HIP(hipGetDeviceCount(&count));
    CACHE_REQUIRE(ordinal >= 0 && ordinal < count, RT_ERROR_INVALID_ARGUMENT, "device ordinal out of range");
    CACHE_HIP(hipSetDevice(ordinal));
    auto* c = new rt_cache();
    c->ordinal = ordinal;
    c->cfg = *cfg;
    c->n_mlp = rtk::cache_mlp_params();
    c->n_grid = ((int64_t)cfg->levels << cfg->log2_table) * 2;
    c->n_total = c->n_mlp + c->n_grid;
    resolutions(cfg, c->res, c->dense);
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, ordinal);
    if (e == hipSuccess) {
        c->cu_count = prop.multiProcessorCount;
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    }
    if (e == hipSuccess) e = c->touched.reserve((size_t)(c->n_grid / 2));
    for (CacheBuf<float>* b : {&c->params, &c->grad, &c->m, &c->v}) {
        if (e == hipSuccess) e = b->reserve((size_t)c->n_total);
    }
    for (int l = 0; l <= rtk::kCacheHiddenLayers && e == hipSuccess; ++l) {
        const bool last = l == rtk::kCacheHiddenLayers;
        e = c->fwd[l].reserve(last ? 16 * 64 : 64 * 64);
        if (e == hipSuccess) e = c->bwd[l].reserve(last ? 64 * 32 : 64 * 64);
    }
    if (e == hipSuccess) e = c->state.reserve(1);
    if (e == hipSuccess) e = c->loss.reserve((size_t)cfg->train_steps);
    if (e == hipSuccess) e = c->ensure(cfg->train_samples);
    std::vector<float> init;
    if (e == hipSuccess) {
        init_params(cfg, c->n_mlp, c->n_total, init);
        e = hipMemcpyAsync(c->params.p, init.data(), sizeof(float) * init.size(),
            hipMemcpyHostToDevice, c->stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(c->loss.p, 0, sizeof(float) * (size_t)cfg->train_steps, c->stream);
    rt_status s = RT_OK;
    if (e == hipSuccess) s = reset_training(c);
    if (e == hipSuccess && s == RT_OK) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess || s != RT_OK) {
        const std::string msg = e != hipSuccess ? std::string("HIP runtime failure at cache allocation: ")
                                                      + hipGetErrorString(e)
                                                : std::string(rt_last_error());
        if (c->stream) (void)hipStreamDestroy(c->stream);
        c->stream = nullptr;
        delete c;
        return cache_fail(RT_ERROR_HIP, msg);
    }
    *out = c;
    return RT_OK;
}

rt_status rt_cache_destroy(rt_cache* c) {
    if (c) rt_cache_release(c);
    return RT_OK;
}

rt_status rt_cache_set_params(rt_cache* c, const float* params, int64_t n) {
    CACHE_REQUIRE(c && params, RT_ERROR_INVALID_ARGUMENT, "null cache or parameters");
    CACHE_REQUIRE(n == c->n_total, RT_ERROR_INVALID_ARGUMENT,
        "expected " + std::to_string(c->n_total) + " parameters");
    CACHE_HIP(hipSetDevice(c->ordinal));
    CACHE_HIP(hipMemcpyAsync(c->params.p, params, sizeof(float) * (size_t)n, hipMemcpyHostToDevice,
        c->stream));
    const rt_status s = reset_training(c);
    if (s != RT_OK) return s;
    CACHE_HIP(hipStreamSynchronize(c->stream));
    return RT_OK;
}

rt_status rt_cache_get_params(rt_cache* c, float* params, int64_t n) {
    CACHE_REQUIRE(c && params, RT_ERROR_INVALID_ARGUMENT, "null cache or parameters");
    CACHE_REQUIRE(n == c->n_total, RT_ERROR_INVALID_ARGUMENT,
        "expected " + std::to_string(c->n_total) + " parameters");
    CACHE_HIP(hipSetDevice(c->ordinal));
    CACHE_HIP(hipMemcpyAsync(params, c->params.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost,
        c->stream));
    CACHE_HIP(hipStreamSynchronize(c->stream));
    return RT_OK;
}

rt_status rt_cache_set_bounds(rt_cache* c, const float lo[3], const float hi[3]) {
    CACHE_REQUIRE(c && lo && hi, RT_ERROR_INVALID_ARGUMENT, "null cache or bounds");
    for (int a = 0; a < 3; ++a) {
        CACHE_REQUIRE(std::isfinite(lo[a]) && std::isfinite(hi[a]), RT_ERROR_INVALID_ARGUMENT,
            "bounds must be finite");
        const double ext = (double)hi[a] - (double)lo[a];
        c->lo[a] = lo[a];
        c->inv_extent[a] = ext > 0.0 ? (float)(1.0 / ext) : 1.0f;
    }
    return RT_OK;
}

rt_status rt_cache_grid_info(rt_cache* c, float res[16], int32_t dense[16], float lo[3],
    float inv_extent[3]) {
    CACHE_REQUIRE(c != nullptr, RT_ERROR_INVALID_ARGUMENT, "null cache");
    for (int l = 0; l < rtk::kCacheMaxLevels; ++l) {
        if (res) res[l] = c->res[l];
        if (dense) dense[l] = c->dense[l];
    }
    for (int a = 0; a < 3; ++a) {
        if (lo) lo[a] = c->lo[a];
        if (inv_extent) inv_extent[a] = c->inv_extent[a];
    }
    return RT_OK;
}

rt_status rt_cache_infer_host(rt_cache* c, const float* inputs, int32_t n, float* out) {
    CACHE_REQUIRE(c && inputs && out, RT_ERROR_INVALID_ARGUMENT, "null cache or buffers");
    CACHE_REQUIRE(n >= 0 && n <= (1 << 24), RT_ERROR_INVALID_ARGUMENT, "n must be in [0, 2^24]");
    if (n == 0) return RT_OK;
    CACHE_HIP(hipSetDevice(c->ordinal));
    CACHE_HIP(c->stage_in.reserve((size_t)n * 4));
    CACHE_HIP(c->stage_out.reserve((size_t)n));
    CACHE_HIP(hipMemcpyAsync(c->stage_in.p, inputs, sizeof(float) * 16 * (size_t)n,
        hipMemcpyHostToDevice, c->stream));
    CACHE_HIP(rtk::cache_infer(c->grid_view(), c->packed(),
        reinterpret_cast<const rtk::CacheInput*>(c->stage_in.p), n, c->stage_out.p, c->grid(),
        c->stream));
    CACHE_HIP(hipMemcpyAsync(out, c->stage_out.p, sizeof(float) * 4 * (size_t)n,
        hipMemcpyDeviceToHost, c->stream));
    CACHE_HIP(hipStreamSynchronize(c->stream));
    return RT_OK;
}

rt_status rt_cache_train_step_host(rt_cache* c, const float* inputs, const float* targets,
    int32_t n, float* loss, float* grad) {
    CACHE_REQUIRE(c && inputs && targets, RT_ERROR_INVALID_ARGUMENT, "null cache or buffers");
    CACHE_REQUIRE(n >= 1 && n <= (1 << 20), RT_ERROR_INVALID_ARGUMENT, "n must be in [1, 2^20]");
    CACHE_HIP(hipSetDevice(c->ordinal));
    CACHE_HIP(c->ensure(n));
    CACHE_HIP(c->stage_in.reserve((size_t)n * 4));
    CACHE_HIP(c->stage_tgt.reserve((size_t)n));
    CACHE_HIP(c->stage_count.reserve(1));
    const uint32_t un = (uint32_t)n;
    CACHE_HIP(hipMemcpyAsync(c->stage_in.p, inputs, sizeof(float) * 16 * (size_t)n,
        hipMemcpyHostToDevice, c->stream));
    CACHE_HIP(hipMemcpyAsync(c->stage_tgt.p, targets, sizeof(float) * 4 * (size_t)n,
        hipMemcpyHostToDevice, c->stream));
    CACHE_HIP(hipMemcpyAsync(c->stage_count.p, &un, sizeof(un), hipMemcpyHostToDevice, c->stream));
    CACHE_HIP(rtk::cache_fill(c->grad.p + c->n_mlp, (size_t)c->n_grid, 0.0f, c->stream));
    CACHE_HIP(rtk::cache_train_pass(c->grid_view(), c->packed(), c->scratch(), c->stage_in.p, 4,
        c->stage_tgt.p, 1, c->stage_count.p, n, 0, 1, c->grad.p, c->loss.p, c->grid(), c->stream));
    if (grad) {
        CACHE_HIP(hipMemcpyAsync(grad, c->grad.p, sizeof(float) * (size_t)c->n_total,
            hipMemcpyDeviceToHost, c->stream));
    }
    CACHE_HIP(rtk::cache_adam(c->n_mlp, c->params.p, c->grad.p, c->m.p, c->v.p,
        c->cfg.learning_rate, c->grid_view(), c->state.p, c->stage_count.p, c->stream));
    CACHE_HIP(rtk::cache_pack(c->params.p, c->packed(), c->stream));
    float l = 0.0f;
    CACHE_HIP(hipMemcpyAsync(&l, c->loss.p, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CACHE_HIP(hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    return RT_OK;
}

rt_status rt_cache_get_stats(rt_cache* c, rt_cache_stats* out) {
    CACHE_REQUIRE(c && out, RT_ERROR_INVALID_ARGUMENT, "null cache or output");
    CACHE_HIP(hipSetDevice(c->ordinal));
    rtk::CacheState st{};
    CACHE_HIP(hipMemcpy(&st, c->state.p, sizeof(st), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->loss_ema = st.loss_ema;
    out->last_loss = st.last_loss;
    out->frames = st.frames;
    out->usable = st.usable;
    out->steps = st.steps;
    out->records = c->last_records;
    out->queries = c->last_queries;
    out->train_fraction = st.train_fraction;
    return RT_OK;
}

}  // extern "C"
