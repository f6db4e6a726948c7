// rt_nn_api.hip -- C ABI of the learned denoiser: parameters, inference,
// one training step (forward, backward, Adam), all on device buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <vector>

#include "rt_alloc.h"
#include "rt_nn.h"
#include "rt_nn_abi.h"

namespace {

thread_local std::string g_nn_error;

rt_status nn_fail(rt_status s, const std::string& msg) {
    g_nn_error = msg;
    rt_set_error(msg.c_str());
    return s;
}

#define NN_HIP(expr)                                                                       \
    do {                                                                                   \
        const hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                              \
            return nn_fail(RT_ERROR_HIP, std::string("HIP runtime failure at ") + #expr + ": " \
                                             + hipGetErrorString(e_));                     \
    } while (0)

#define NN_REQUIRE(cond, status, msg) \
    do {                              \
        if (!(cond)) return nn_fail(status, msg); \
    } while (0)

template <typename T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { release(); }
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

}  // namespace

struct rt_nn {
    std::atomic<int> refs{1};  // the caller's handle plus each renderer it is attached to
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    int step = 0;
    Buf<float> params, grad, m, v, loss;
    Buf<uint16_t> fwd[rtk::kNnConvs], bwd[rtk::kNnConvs];
    Buf<float> bias[rtk::kNnConvs];
    Buf<float> head;
    // per-batch scratch, sized for the largest batch seen at each level
    size_t lv_cap[rtk::kNnLevels] = {0, 0, 0, 0};
    size_t lv_cap_train[rtk::kNnLevels] = {0, 0, 0, 0};
    Buf<uint16_t> feat, act[rtk::kNnConvs], pool[rtk::kNnLevels], dz[rtk::kNnConvs], r[rtk::kNnConvs];
    Buf<float> head_partial, grad_partial;

    rtk::NnPacked packed() {
        rtk::NnPacked pk{};
        for (int l = 0; l < rtk::kNnConvs; ++l) {
            pk.fwd[l] = fwd[l].p;
            pk.bwd[l] = bwd[l].p;
            pk.bias[l] = bias[l].p;
        }
        pk.head = head.p;
        return pk;
    }

    // activations for inference; with `train` also the gradient buffers
    hipError_t ensure(const rtk::NnLevels& lv, bool train) {
        size_t P[rtk::kNnLevels];
        bool grow = false, grow_train = false;
        for (int l = 0; l < rtk::kNnLevels; ++l) {
            P[l] = (size_t)lv.lv[l].offset[lv.lv[l].count];
            grow = grow || P[l] > lv_cap[l];
            grow_train = grow_train || (train && P[l] > lv_cap_train[l]);
        }
        hipError_t e = hipSuccess;
        if (grow) {
            e = feat.reserve(P[0] * rtk::kNnFeat);
            for (int l = 0; l < rtk::kNnConvs && e == hipSuccess; ++l) {
                const rtk::NnConv c = rtk::nn_conv(l);
                e = act[l].reserve(P[c.level] * (size_t)c.cout);
            }
            for (int l = 1; l < rtk::kNnLevels && e == hipSuccess; ++l)
                e = pool[l].reserve(P[l] * (size_t)rtk::nn_conv(2 * l - 1).cout);
            if (e == hipSuccess) e = head_partial.reserve(((P[0] + 255) / 256) * (rtk::kNnHeadParams + 1));
            if (e == hipSuccess)
                for (int l = 0; l < rtk::kNnLevels; ++l) lv_cap[l] = std::max(lv_cap[l], P[l]);
        }
        if (e == hipSuccess && grow_train) {
            for (int l = 0; l < rtk::kNnConvs && e == hipSuccess; ++l) {
                const rtk::NnConv c = rtk::nn_conv(l);
                e = dz[l].reserve(P[c.level] * (size_t)c.cout);
                if (e == hipSuccess && l > 0) e = r[l].reserve(P[c.level] * (size_t)rtk::nn_cin(l));
            }
            if (e == hipSuccess) e = grad_partial.reserve(rtk::nn_grad_partial_floats(lv));
            if (e == hipSuccess)
                for (int l = 0; l < rtk::kNnLevels; ++l) lv_cap_train[l] = std::max(lv_cap_train[l], P[l]);
        }
        return e;
    }

    rtk::NnScratch scratch() {
        rtk::NnScratch w{};
        w.feat = feat.p;
        for (int l = 0; l < rtk::kNnConvs; ++l) {
            w.act[l] = act[l].p;
            w.dz[l] = dz[l].p;
            w.r[l] = r[l].p;
        }
        for (int l = 0; l < rtk::kNnLevels; ++l) w.pool[l] = pool[l].p;
        w.head_partial = head_partial.p;
        w.grad_partial = grad_partial.p;
        return w;
    }

    rtk::NnTiling tiling{};   // convolution tile rows per level (0: by image width)
};

namespace {

rt_status check_images(const rt_nn_images* im, int* P) {
    NN_REQUIRE(im != nullptr, RT_ERROR_INVALID_ARGUMENT, "null image table");
    NN_REQUIRE(im->count >= 1 && im->count <= rtk::kNnMaxImages, RT_ERROR_INVALID_ARGUMENT,
        "image count must be in [1, " + std::to_string(rtk::kNnMaxImages) + "]");
    NN_REQUIRE(im->offset[0] == 0, RT_ERROR_INVALID_ARGUMENT, "images must start at pixel 0");
    for (int i = 0; i < im->count; ++i) {
        const int n = im->offset[i + 1] - im->offset[i];
        NN_REQUIRE(im->width[i] > 0 && n > 0 && n % im->width[i] == 0, RT_ERROR_INVALID_ARGUMENT,
            "image " + std::to_string(i) + " has an invalid extent");
    }
    *P = im->offset[im->count];
    return RT_OK;
}

rtk::NnImages to_images(const rt_nn_images* im) {
    rtk::NnImages r{};
    r.count = im->count;
    for (int i = 0; i <= im->count; ++i) r.offset[i] = im->offset[i];
    for (int i = 0; i < im->count; ++i) r.width[i] = im->width[i];
    return r;
}

}  // namespace

hipError_t rt_nn_infer_on_stream(rt_nn* nn, const rtk::NnImages& images, int P,
    const float4* beauty, const float4* normal, const float4* albedo, const float* depth,
    float4* out, hipStream_t st) {
    (void)P;
    const rtk::NnLevels lv = rtk::nn_levels(images);
    const hipError_t e = nn->ensure(lv, false);
    if (e != hipSuccess) return e;
    return rtk::nn_forward(lv, nn->packed(), nn->scratch(), beauty, normal, albedo, depth, nullptr,
        out, nn->tiling, st);
}

int rt_nn_device(const rt_nn* nn) { return nn->ordinal; }

hipError_t rt_nn_reserve(rt_nn* nn, const rtk::NnImages& images) {
    return nn->ensure(rtk::nn_levels(images), false);
}

void rt_nn_retain(rt_nn* nn) { nn->refs.fetch_add(1); }

void rt_nn_release(rt_nn* nn) {
    if (nn->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(nn->ordinal);
    (void)hipStreamSynchronize(nn->stream);
    hipStream_t st = nn->stream;
    delete nn;
    if (st) (void)hipStreamDestroy(st);
}

extern "C" {

int32_t rt_nn_param_count(void) { return rtk::nn_param_count(); }

rt_status rt_nn_create(int ordinal, rt_nn** out) {
    NN_REQUIRE(out != nullptr, RT_ERROR_INVALID_ARGUMENT, "null output");
    int count = 0;
    NN_HIP(hipGetDeviceCount(&count));
    NN_REQUIRE(ordinal >= 0 && ordinal < count, RT_ERROR_INVALID_ARGUMENT, "device ordinal out of range");
    NN_HIP(hipSetDevice(ordinal));
    auto* nn = new rt_nn();
    nn->ordinal = ordinal;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, ordinal);
    if (e == hipSuccess) {
        nn->cu_count = prop.multiProcessorCount;
        e = hipStreamCreateWithFlags(&nn->stream, hipStreamNonBlocking);
    }
    const size_t np = (size_t)rtk::nn_param_count();
    for (Buf<float>* b : {&nn->params, &nn->grad, &nn->m, &nn->v}) {
        if (e == hipSuccess) e = b->reserve(np);
    }
    if (e == hipSuccess) e = nn->loss.reserve(1);
    for (int l = 0; l < rtk::kNnConvs && e == hipSuccess; ++l) {
        const rtk::NnConv c = rtk::nn_conv(l);
        e = nn->fwd[l].reserve((size_t)c.cout * rtk::nn_kp_fwd(l));
        if (e == hipSuccess) e = nn->bwd[l].reserve((size_t)rtk::nn_cin(l) * rtk::nn_kp_bwd(l));
        if (e == hipSuccess) e = nn->bias[l].reserve((size_t)c.cout);
    }
    if (e == hipSuccess) e = nn->head.reserve(rtk::kNnHeadParams);
    if (e == hipSuccess) e = hipMemsetAsync(nn->params.p, 0, np * sizeof(float), nn->stream);
    if (e == hipSuccess) e = hipMemsetAsync(nn->m.p, 0, np * sizeof(float), nn->stream);
    if (e == hipSuccess) e = hipMemsetAsync(nn->v.p, 0, np * sizeof(float), nn->stream);
    if (e == hipSuccess) e = rtk::nn_pack(nn->params.p, nn->packed(), nn->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(nn->stream);
    if (e != hipSuccess) {
        if (nn->stream) (void)hipStreamDestroy(nn->stream);
        delete nn;
        return nn_fail(RT_ERROR_HIP, std::string("HIP runtime failure at denoiser allocation: ")
                                        + hipGetErrorString(e));
    }
    *out = nn;
    return RT_OK;
}

rt_status rt_nn_destroy(rt_nn* nn) {
    if (!nn) return RT_OK;
    rt_nn_release(nn);
    return RT_OK;
}


rt_status rt_nn_set_params(rt_nn* nn, const float* params, int32_t n) {
    NN_REQUIRE(nn && params, RT_ERROR_INVALID_ARGUMENT, "null network or parameters");
    NN_REQUIRE(n == rtk::nn_param_count(), RT_ERROR_INVALID_ARGUMENT,
        "expected " + std::to_string(rtk::nn_param_count()) + " parameters");
    NN_HIP(hipSetDevice(nn->ordinal));
    const size_t bytes = sizeof(float) * (size_t)n;
    NN_HIP(hipMemcpyAsync(nn->params.p, params, bytes, hipMemcpyHostToDevice, nn->stream));
    NN_HIP(hipMemsetAsync(nn->m.p, 0, bytes, nn->stream));
    NN_HIP(hipMemsetAsync(nn->v.p, 0, bytes, nn->stream));
    NN_HIP(rtk::nn_pack(nn->params.p, nn->packed(), nn->stream));
    NN_HIP(hipStreamSynchronize(nn->stream));
    nn->step = 0;
    return RT_OK;
}

rt_status rt_nn_get_params(rt_nn* nn, float* params, int32_t n) {
    NN_REQUIRE(nn && params, RT_ERROR_INVALID_ARGUMENT, "null network or parameters");
    NN_REQUIRE(n == rtk::nn_param_count(), RT_ERROR_INVALID_ARGUMENT,
        "expected " + std::to_string(rtk::nn_param_count()) + " parameters");
    NN_HIP(hipSetDevice(nn->ordinal));
    NN_HIP(hipMemcpyAsync(params, nn->params.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, nn->stream));
    NN_HIP(hipStreamSynchronize(nn->stream));
    return RT_OK;
}

rt_status rt_nn_infer(rt_nn* nn, const rt_nn_images* images, const float* beauty,
    const float* normal, const float* albedo, const float* depth, float* out) {
    NN_REQUIRE(nn != nullptr, RT_ERROR_INVALID_ARGUMENT, "null network");
    int P = 0;
    const rt_status s = check_images(images, &P);
    if (s != RT_OK) return s;
    NN_REQUIRE(beauty && normal && albedo && depth && out, RT_ERROR_INVALID_ARGUMENT, "null frame buffer");
    NN_HIP(hipSetDevice(nn->ordinal));
    const rtk::NnLevels lv = rtk::nn_levels(to_images(images));
    NN_HIP(nn->ensure(lv, false));
    NN_HIP(rtk::nn_forward(lv, nn->packed(), nn->scratch(), reinterpret_cast<const float4*>(beauty),
        reinterpret_cast<const float4*>(normal), reinterpret_cast<const float4*>(albedo), depth,
        nullptr, reinterpret_cast<float4*>(out), nn->tiling, nn->stream));
    NN_HIP(hipStreamSynchronize(nn->stream));
    return RT_OK;
}

rt_status rt_nn_train_step(rt_nn* nn, const rt_nn_images* images, const float* beauty,
    const float* normal, const float* albedo, const float* depth, const float* target, float lr,
    float* loss, float* grad) {
    NN_REQUIRE(nn != nullptr, RT_ERROR_INVALID_ARGUMENT, "null network");
    int P = 0;
    const rt_status s = check_images(images, &P);
    if (s != RT_OK) return s;
    NN_REQUIRE(beauty && normal && albedo && depth && target, RT_ERROR_INVALID_ARGUMENT,
        "null frame or target buffer");
    NN_REQUIRE(lr >= 0.0f, RT_ERROR_INVALID_ARGUMENT, "negative learning rate");
    NN_HIP(hipSetDevice(nn->ordinal));
    const rtk::NnLevels lv = rtk::nn_levels(to_images(images));
    NN_HIP(nn->ensure(lv, true));
    const rtk::NnPacked pk = nn->packed();
    const rtk::NnScratch w = nn->scratch();
    NN_HIP(rtk::nn_forward(lv, pk, w, reinterpret_cast<const float4*>(beauty),
        reinterpret_cast<const float4*>(normal), reinterpret_cast<const float4*>(albedo), depth,
        reinterpret_cast<const float4*>(target), nullptr, nn->tiling, nn->stream));
    NN_HIP(rtk::nn_backward(lv, pk, w, nn->grad.p, nn->loss.p, nn->tiling, nn->stream));
    const int np = rtk::nn_param_count();
    if (grad) NN_HIP(hipMemcpyAsync(grad, nn->grad.p, sizeof(float) * (size_t)np, hipMemcpyDeviceToHost, nn->stream));
    if (loss) NN_HIP(hipMemcpyAsync(loss, nn->loss.p, sizeof(float), hipMemcpyDeviceToHost, nn->stream));
    if (lr > 0.0f) {
        ++nn->step;
        NN_HIP(rtk::nn_adam(np, nn->params.p, nn->grad.p, nn->m.p, nn->v.p, lr, nn->step, nn->stream));
        NN_HIP(rtk::nn_pack(nn->params.p, pk, nn->stream));
    }
    NN_HIP(hipStreamSynchronize(nn->stream));
    return RT_OK;
}

rt_status rt_nn_set_tiling(rt_nn* nn, const rt_nn_tiling* t) {
    NN_REQUIRE(nn && t, RT_ERROR_INVALID_ARGUMENT, "null network or tiling");
    for (int l = 0; l < rtk::kNnLevels; ++l)
        NN_REQUIRE(t->conv_rows[l] == 0 || t->conv_rows[l] == 4 || t->conv_rows[l] == 8,
            RT_ERROR_INVALID_ARGUMENT, "convolution tile rows must be 0, 4 or 8");
    for (int l = 0; l < rtk::kNnLevels; ++l) nn->tiling.rows[l] = t->conv_rows[l];
    rt_alloc_bump();   // frame graphs captured with the old launch shapes are recaptured
    return RT_OK;
}

rt_status rt_nn_get_tiling(rt_nn* nn, rt_nn_tiling* t) {
    NN_REQUIRE(nn && t, RT_ERROR_INVALID_ARGUMENT, "null network or tiling");
    for (int l = 0; l < rtk::kNnLevels; ++l) t->conv_rows[l] = nn->tiling.rows[l];
    return RT_OK;
}

// Coordinate descent over the convolution tile rows of each level (4, 8),
// each candidate timed by `reps` inferences of this image batch on device
// buffers (median, after one warm-up); applies the fastest tiling.
rt_status rt_nn_autotune(rt_nn* nn, const rt_nn_images* images, int32_t reps,
    rt_nn_autotune_result* out) {
    NN_REQUIRE(nn && out, RT_ERROR_INVALID_ARGUMENT, "null network or result");
    NN_REQUIRE(reps >= 1 && reps <= 64, RT_ERROR_INVALID_ARGUMENT, "reps must be in [1, 64]");
    int P = 0;
    const rt_status s = check_images(images, &P);
    if (s != RT_OK) return s;
    std::memset(out, 0, sizeof(*out));
    NN_HIP(hipSetDevice(nn->ordinal));
    const rtk::NnLevels lv = rtk::nn_levels(to_images(images));
    NN_HIP(nn->ensure(lv, false));
    Buf<float4> b, n, a, o;
    Buf<float> d;
    NN_HIP(b.reserve((size_t)P));
    NN_HIP(n.reserve((size_t)P));
    NN_HIP(a.reserve((size_t)P));
    NN_HIP(o.reserve((size_t)P));
    NN_HIP(d.reserve((size_t)P));
    NN_HIP(hipMemsetAsync(b.p, 0, (size_t)P * 16, nn->stream));
    NN_HIP(hipMemsetAsync(n.p, 0, (size_t)P * 16, nn->stream));
    NN_HIP(hipMemsetAsync(a.p, 0, (size_t)P * 16, nn->stream));
    NN_HIP(hipMemsetAsync(d.p, 0, (size_t)P * 4, nn->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    NN_HIP(hipEventCreate(&e0));
    hipError_t herr = hipEventCreate(&e1);
    if (herr != hipSuccess) {
        (void)hipEventDestroy(e0);
        NN_HIP(herr);
    }
    int trials = 0;
    auto measure = [&](const rtk::NnTiling& t, double* ms) -> hipError_t {
        std::vector<double> v;
        for (int i = 0; i <= reps; ++i) {
            hipError_t e = hipEventRecord(e0, nn->stream);
            if (e == hipSuccess)
                e = rtk::nn_forward(lv, nn->packed(), nn->scratch(), b.p, n.p, a.p, d.p, nullptr, o.p, t,
                    nn->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, nn->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float f = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
            if (e != hipSuccess) return e;
            if (i > 0) v.push_back(f);
        }
        std::sort(v.begin(), v.end());
        *ms = v[v.size() / 2];
        ++trials;
        return hipSuccess;
    };
    // resolve the default rule into explicit rows, then try the other value per level
    rtk::NnTiling best{};
    for (int l = 0; l < rtk::kNnLevels; ++l) best.rows[l] = rtk::nn_tile_rows(nn->tiling, lv.lv[l], l);
    double best_ms = 0.0;
    herr = measure(best, &best_ms);
    out->baseline_ms = best_ms;
    for (int l = 0; l < rtk::kNnLevels && herr == hipSuccess; ++l) {
        rtk::NnTiling t = best;
        t.rows[l] = best.rows[l] == 8 ? 4 : 8;
        double ms = 0.0;
        herr = measure(t, &ms);
        if (herr == hipSuccess && ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    NN_HIP(herr);
    nn->tiling = best;
    rt_alloc_bump();
    for (int l = 0; l < rtk::kNnLevels; ++l) out->best.conv_rows[l] = best.rows[l];
    out->best_ms = best_ms;
    out->trials = trials;
    return RT_OK;
}

// host-array forms (tests, small batches): upload, run, download
rt_status rt_nn_infer_host(rt_nn* nn, const rt_nn_images* images, const float* beauty,
    const float* normal, const float* albedo, const float* depth, float* out) {
    int P = 0;
    const rt_status s = check_images(images, &P);
    if (s != RT_OK) return s;
    NN_REQUIRE(nn && beauty && normal && albedo && depth && out, RT_ERROR_INVALID_ARGUMENT, "null argument");
    NN_HIP(hipSetDevice(nn->ordinal));
    Buf<float> b, n, a, d, o;
    NN_HIP(b.reserve((size_t)P * 4));
    NN_HIP(n.reserve((size_t)P * 4));
    NN_HIP(a.reserve((size_t)P * 4));
    NN_HIP(d.reserve((size_t)P));
    NN_HIP(o.reserve((size_t)P * 4));
    NN_HIP(hipMemcpyAsync(b.p, beauty, (size_t)P * 16, hipMemcpyHostToDevice, nn->stream));
    NN_HIP(hipMemcpyAsync(n.p, normal, (size_t)P * 16, hipMemcpyHostToDevice, nn->stream));
    NN_HIP(hipMemcpyAsync(a.p, albedo, (size_t)P * 16, hipMemcpyHostToDevice, nn->stream));
    NN_HIP(hipMemcpyAsync(d.p, depth, (size_t)P * 4, hipMemcpyHostToDevice, nn->stream));
    const rt_status r = rt_nn_infer(nn, images, b.p, n.p, a.p, d.p, o.p);
    if (r != RT_OK) return r;
    NN_HIP(hipMemcpy(out, o.p, (size_t)P * 16, hipMemcpyDeviceToHost));
    return RT_OK;
}

rt_status rt_nn_train_step_host(rt_nn* nn, const rt_nn_images* images, const float* beauty,
    const float* normal, const float* albedo, const float* depth, const float* target, float lr,
    float* loss, float* grad) {
    int P = 0;
    const rt_status s = check_images(images, &P);
    if (s != RT_OK) return s;
    NN_REQUIRE(nn && beauty && normal && albedo && depth && target, RT_ERROR_INVALID_ARGUMENT, "null argument");
    NN_HIP(hipSetDevice(nn->ordinal));
    Buf<float> b, n, a, d, t;
    NN_HIP(b.reserve((size_t)P * 4));
    NN_HIP(n.reserve((size_t)P * 4));
    NN_HIP(a.reserve((size_t)P * 4));
    NN_HIP(d.reserve((size_t)P));
    NN_HIP(t.reserve((size_t)P * 4));
    NN_HIP(hipMemcpyAsync(b.p, beauty, (size_t)P * 16, hipMemcpyHostToDevice, nn->stream));
    NN_HIP(hipMemcpyAsync(n.p, normal, (size_t)P * 16, hipMemcpyHostToDevice, nn->stream));
    NN_HIP(hipMemcpyAsync(a.p, albedo, (size_t)P * 16, hipMemcpyHostToDevice, nn->stream));
    NN_HIP(hipMemcpyAsync(d.p, depth, (size_t)P * 4, hipMemcpyHostToDevice, nn->stream));
    NN_HIP(hipMemcpyAsync(t.p, target, (size_t)P * 16, hipMemcpyHostToDevice, nn->stream));
    return rt_nn_train_step(nn, images, b.p, n.p, a.p, d.p, t.p, lr, loss, grad);
}

}  // extern "C"
