// rt_api.hip -- C ABI of the gfx950 path tracer (declared in include/rt/rt_hip.h).
//
// Host orchestration: device and stream set-up, scene upload with light CDF
// construction, LBVH build, trace test hooks, and the per-frame wavefront
// launch sequence.  Errors are returned as rt_status with a thread-local
// message (the reference throws std::runtime_error / std::invalid_argument
// with the same kind of text, src/realtime/gpu/optix_renderer.cpp:31-59).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <chrono>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "rt_alloc.h"
#include "rt_host.h"
#include "rt_cache_abi.h"
#include "rt_nn_abi.h"
#include "rt_kernels.h"
#include "rt_openpbr.h"

namespace {

thread_local std::string g_error;

rt_status fail(rt_status s, const std::string& msg) {
    g_error = msg;
    return s;
}

}  // namespace

void rt_set_error(const char* msg) { g_error = msg; }

static std::atomic<uint64_t> g_alloc_epoch{0};
uint64_t rt_alloc_epoch() { return g_alloc_epoch.load(); }
void rt_alloc_bump() { g_alloc_epoch.fetch_add(1); }

namespace {

#define RT_HIP(expr)                                                                        \
    do {                                                                                    \
        const hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                               \
            return fail(RT_ERROR_HIP, std::string("HIP runtime failure at ") + #expr + ": " \
                                          + hipGetErrorString(e_));                        \
    } while (0)

#define RT_REQUIRE(cond, status, msg) \
    do {                              \
        if (!(cond)) return fail(status, msg); \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t reserve(size_t count) {
        if (count <= n && p) return hipSuccess;
        release();
        if (count == 0) count = 1;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count);
        if (e == hipSuccess) n = count;
        rt_alloc_bump();
        return e;
    }
};

template <typename T>
hipError_t upload(DevBuf<T>& b, const T* src, size_t count, hipStream_t st) {
    hipError_t e = b.reserve(count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, st);
}

float lum_clamped(const float* c) {
    const float r = std::max(c[0], 0.0f), g = std::max(c[1], 0.0f), b = std::max(c[2], 0.0f);
    return std::max(0.0f, 0.2126f * r + 0.7152f * g + 0.0722f * b);
}

struct V3 {
    float x, y, z;
};
V3 v3(const float* p) { return V3{p[0], p[1], p[2]}; }
V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

}  // namespace

// Handles are reference counted: a scene holds its device and a renderer
// holds its scene, so destroying them in any order frees each object only
// after the last handle that uses it (the reference shares the prepared
// scene through shared_ptr<SharedGpuSceneState>, optix_renderer.h:117).
struct rt_device {
    std::atomic<int> refs{1};
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    std::string name;
    size_t total_mem = 0;
};

// the viewer presenter: an RGBA8 pixel buffer of a fixed resolution
// (CudaGlTexturePresenter's pixel-unpack buffer, cuda_gl_presenter.cu:53-67)
struct rt_presenter {
    rt_device* dev = nullptr;
    int32_t width = 0, height = 0;
    DevBuf<uint32_t> pixels;
};

struct rt_scene {
    std::atomic<int> refs{1};
    rt_device* dev = nullptr;
    int32_t sphere_count = 0, quad_count = 0, triangle_count = 0, material_count = 0;
    int32_t openpbr_count = 0, analytic_count = 0;
    // shade passes: 0 one full pass (OpenPBR only), 1 the legacy-material
    // pass alone (no OpenPBR), 2 both (legacy pass, then its OpenPBR hits)
    int32_t shade_mode = 0;
    bool pbr_full = false;   // an OpenPBR material uses lobes beyond the opaque core
    int32_t texture_count = 0, texel_count = 0;
    bool has_albedo_textures = false;
    bool has_triangle_uvs = false, has_triangle_normals = false;
    DevBuf<float> tri_uv, tri_n;
    DevBuf<rt_texture> textures;
    DevBuf<float> texels;
    DevBuf<int32_t> mat_albedo_tex;
    bool has_emission_textures = false, has_pbr_textures = false;
    DevBuf<int32_t> mat_emission_tex;
    DevBuf<rt_openpbr_textures> pbr_tex;
    int32_t background_mode = 0;
    float bg[3] = {0, 0, 0}, bg_top[3] = {0, 0, 0};
    DevBuf<rt_sphere> spheres;
    DevBuf<rt_quad> quads;
    DevBuf<rt_triangle> triangles;
    DevBuf<rt_material> materials;
    DevBuf<rt_openpbr_material> openpbr;
    DevBuf<rt_analytic_light> analytic;
    std::vector<rtk::LightRec> host_lights;
    DevBuf<rtk::LightRec> lights;
    DevBuf<float> a_sel, a_cdf;
    DevBuf<int32_t> a_infinite;   // dome / distant lights in index order
    int32_t a_infinite_count = 0;
    DevBuf<rt_medium> media;
    int32_t medium_count = 0;
    DevBuf<float4> a_nodes;       // BVH over the finite analytic lights
    DevBuf<int32_t> a_order;
    int32_t a_node_count = 0;
    DevBuf<int32_t> light_of_prim;
    int32_t env_light = -1;
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};  // primitive bounds
    // BVH
    bool bvh_ready = false;
    int32_t prim_count = 0;
    int32_t node_count = 0;
    int32_t max_depth = 0;
    int32_t wide_count = 0;
    int32_t wide_levels = 0;
    DevBuf<float4> nodes4;
    DevBuf<float4> nodes;
    DevBuf<float4> leaves;
    DevBuf<uint32_t> sorted;
    DevBuf<uint32_t> morton;
    DevBuf<uint32_t> leaf_parent;  // binary LBVH parent of each sorted leaf (kept for refits)
    // acceleration lifecycle (GpuSceneAcceleration::update,
    // src/realtime/gpu/gpu_scene_acceleration.cpp:247-291)
    uint64_t geometry_sig = 0, scene_sig = 0, generation = 0;
    rt_acceleration_update last_update{};

    rtk::SceneView view() const {
        rtk::SceneView v{};
        v.nodes = nodes.p;
        v.leaves = leaves.p;
        v.node_count = node_count;
        v.prim_count = bvh_ready ? prim_count : 0;
        v.spheres = spheres.p;
        v.quads = quads.p;
        v.triangles = triangles.p;
        v.sphere_count = sphere_count;
        v.quad_count = quad_count;
        v.triangle_count = triangle_count;
        v.material_count = material_count;
        v.materials = materials.p;
        v.openpbr = openpbr.p;
        v.openpbr_count = openpbr_count;
        v.pbr_full = pbr_full ? 1 : 0;
        v.light_count = static_cast<int32_t>(host_lights.size());
        v.lights = lights.p;
        v.light_of_prim = light_of_prim.p;
        v.env_light = env_light;
        v.background_mode = background_mode;
        for (int i = 0; i < 3; ++i) {
            v.bg[i] = bg[i];
            v.bg_top[i] = bg_top[i];
        }
        v.analytic = analytic.p;
        v.analytic_count = analytic_count;
        v.nodes4 = nodes4.p;
        v.node4_count = bvh_ready ? wide_count : 0;
        v.stack4_entries = std::max(3 * wide_levels, 1);
        v.a_sel = a_sel.p;
        v.a_cdf = a_cdf.p;
        v.a_infinite = a_infinite.p;
        v.a_infinite_count = a_infinite_count;
        v.a_nodes = a_nodes.p;
        v.a_order = a_order.p;
        v.a_node_count = a_node_count;
        v.media = media.p;
        v.medium_count = medium_count;
        v.textures = textures.p;
        v.texels = texels.p;
        v.mat_albedo_tex = has_albedo_textures ? mat_albedo_tex.p : nullptr;
        v.tri_uv = has_triangle_uvs ? tri_uv.p : nullptr;
        v.tri_n = has_triangle_normals ? tri_n.p : nullptr;
        v.texture_count = texture_count;
        v.texel_count = texel_count;
        v.mat_emission_tex = has_emission_textures ? mat_emission_tex.p : nullptr;
        v.pbr_tex = has_pbr_textures ? pbr_tex.p : nullptr;
        return v;
    }
};

// Everything a captured graph of the frame bakes in: the sequence and shape
// of launches and every buffer pointer they receive.  A frame whose key
// differs from the graph's is recaptured.
struct GraphKey {
    int32_t spp, max_b, rr, jitter, layers, total, count, restir, denoise, block, grid, refill,
        count_traversal, fs, res_cur;
    int32_t restir_params[10];
    int32_t widths[4], heights[4];
    const void* ptrs[10];
    uint64_t alloc_epoch;  // device allocations of the library when captured
};

// Budgets of the learned denoiser's guard: the network's output may move
// the mean log(1 + luminance) of a camera's surface pixels by at most 80 %
// (an output collapsed to zero moves it by 100 %; a 1-spp frame whose few
// bright samples the network spreads out can move it by half), and must
// remove at least a tenth of the input's high-frequency energy (a network
// reduced to the identity removes none; a trained one removes most of it
// from a 1-spp frame).
constexpr float kNnMaxMeanShift = 0.8f;
constexpr float kNnMaxResidualNoise = 0.9f;

struct rt_renderer {
    rt_device* dev = nullptr;
    rt_scene* scene = nullptr;
    rt_tuning tuning{};
    uint32_t sequence_seed = 0;
    uint64_t parameter_uploads = 0;  // frames launched (one FrameDyn upload each)
    // rig
    rt_camera* host_cams = nullptr;  // pinned, two frame slots of 8 cameras
    DevBuf<rt_camera> cams;
    int camera_count = 0;
    int pixel_offset[5] = {0, 0, 0, 0, 0};
    int widths[4] = {0, 0, 0, 0};
    int heights[4] = {0, 0, 0, 0};
    // frame outputs: two sets of beauty / normal / depth, the last frame's
    // set is the next frame's reprojection history
    // (DeviceFrameBufferSet::copy_frame_to_history, device_frame_buffers.cpp:151-162,
    // without the copy)
    DevBuf<float4> beauty_set[2], normal_set[2];
    DevBuf<float> depth_set[2];
    DevBuf<float4> albedo;
    DevBuf<float2> flow;
    DevBuf<float> trust;
    DevBuf<float4> denoised, denoise_scratch;
    int frame_set = 0;
    int last_set = 0;
    bool last_denoised = false;
    rt_nn* nn = nullptr;  // learned denoiser (replaces the a-trous filter when set)
    // its guard: per camera, a frame whose network output fails the budgets
    // is filtered by the a-trous filter instead (decided on the device)
    DevBuf<rtk::NnGuard> nn_guard;
    rt_cache* cache = nullptr;  // radiance cache (trained and queried when set)
    DevBuf<rtk::CacheInput> cache_queries;   // one per path slot
    DevBuf<rtk::CacheRecord> cache_records;  // training records of a frame
    // per-frame device values and captured frame graphs
    DevBuf<rtk::FrameDyn> dyn;
    rtk::FrameDyn* host_dyn = nullptr;  // pinned, two frame slots
    // Host staging is double-buffered so that an asynchronous frame returns
    // once enqueued: a frame writes slot `slot` and a call waits only for the
    // frame two calls back (slot_ev) before it rewrites that slot.  The
    // frame's statistics words land in host_stats[slot] and are folded into
    // the radiance cache's record / query counts once the slot is waited on.
    uint32_t* host_stats = nullptr;  // pinned, two slots of kStatWords
    hipEvent_t slot_ev[2] = {nullptr, nullptr};
    bool slot_pending[2] = {false, false};
    rt_cache* slot_cache[2] = {nullptr, nullptr};
    int slot = 0;
    bool async_frames = false;

    void absorb_stats(int k) {
        rt_cache* c = slot_cache[k];
        const uint32_t* w = host_stats + (size_t)k * rtk::kStatWords;
        if (c != nullptr) {
            c->last_records = (int32_t)std::min<uint32_t>(w[rtk::kStatCacheRecords], (uint32_t)c->cfg.train_samples);
            c->last_queries = (int32_t)w[rtk::kStatCacheQueries];
        }
        slot_pending[k] = false;
        slot_cache[k] = nullptr;
    }
    // waits for every frame enqueued on the stream and folds in their statistics
    hipError_t drain() {
        const hipError_t e = hipStreamSynchronize(dev->stream);
        if (e != hipSuccess) return e;
        for (int j = 1; j <= 2; ++j) {  // older slot first
            const int k = (slot + j) & 1;
            if (slot_pending[k]) absorb_stats(k);
        }
        return hipSuccess;
    }
    struct Graph {
        GraphKey key;
        hipGraphExec_t exec = nullptr;
    };
    Graph graphs[4];
    int graph_next = 0;
    int graph_captures = 0;
    hipEvent_t post_ev[2] = {nullptr, nullptr};
    // wavefront state
    DevBuf<float4> ray_o, ray_d, hit, throughput, radiance, prev_pos, prev_n, aux_normal, aux_albedo;
    DevBuf<int32_t> prev_flags;
    DevBuf<uint32_t> queue0, queue1;
    DevBuf<uint32_t> pbr_queue;  // OpenPBR hits handed from the legacy shade pass
    DevBuf<float4> shadow_o, shadow_d, shadow_c, shadow2_o, shadow2_d, shadow2_c;
    DevBuf<uint32_t> restir_queue;
    DevBuf<float4> restir_o, restir_d, restir_thr;
    // ReSTIR reservoirs (ping-pong, global pixel index) and per-camera history
    DevBuf<rtk::Reservoir> reservoirs[2];
    int res_cur = 0;
    size_t reservoir_pixels = 0;
    rt_camera prev_cams[4];
    int history_length[4] = {0, 0, 0, 0};
    int prev_valid[4] = {0, 0, 0, 0};
    int prev_width[4] = {0, 0, 0, 0};
    int prev_height[4] = {0, 0, 0, 0};
    int prev_offset[4] = {0, 0, 0, 0};
    bool restir_last = false;
    int last_reservoirs = -1;
    DevBuf<uint32_t> counters;
    DevBuf<uint2> spill;  // traversal stack entries beyond the LDS part
    DevBuf<float4> acc_beauty, acc_normal, acc_albedo;
    size_t path_capacity = 0;
    size_t pixel_capacity = 0;
    std::vector<hipEvent_t> events;
    bool rendered = false;

    ~rt_renderer() {
        for (Graph& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (host_dyn) (void)hipHostFree(host_dyn);
        if (host_stats) (void)hipHostFree(host_stats);
        for (hipEvent_t e : slot_ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (hipEvent_t e : post_ev)
            if (e) (void)hipEventDestroy(e);
        if (host_cams) (void)hipHostFree(host_cams);
    }

    rtk::WavefrontBuffers buffers() {
        rtk::WavefrontBuffers b{};
        b.ray_o = ray_o.p;
        b.ray_d = ray_d.p;
        b.hit = hit.p;
        b.throughput = throughput.p;
        b.radiance = radiance.p;
        b.prev_pos = prev_pos.p;
        b.prev_n = prev_n.p;
        b.prev_flags = prev_flags.p;
        b.aux_normal = aux_normal.p;
        b.aux_albedo = aux_albedo.p;
        b.queue[0] = queue0.p;
        b.queue[1] = queue1.p;
        b.pbr_queue = pbr_queue.p;
        b.shadow_o = shadow_o.p;
        b.shadow_d = shadow_d.p;
        b.shadow_c = shadow_c.p;
        b.shadow2_o = shadow2_o.p;
        b.shadow2_d = shadow2_d.p;
        b.shadow2_c = shadow2_c.p;
        b.restir_queue = restir_queue.p;
        b.restir_o = restir_o.p;
        b.restir_d = restir_d.p;
        b.restir_thr = restir_thr.p;
        b.counters = counters.p;
        b.spill = spill.p;
        b.acc_beauty = acc_beauty.p;
        b.acc_normal = acc_normal.p;
        b.acc_albedo = acc_albedo.p;
        return b;
    }

    hipError_t ensure_paths(size_t q) {
        if (q <= path_capacity) return hipSuccess;
        hipError_t e;
        for (DevBuf<float4>* b : {&ray_o, &ray_d, &hit, &throughput, &radiance, &prev_pos, &prev_n,
                 &aux_normal, &aux_albedo, &shadow_o, &shadow_d, &shadow_c, &shadow2_o, &shadow2_d,
                 &shadow2_c, &restir_o, &restir_d, &restir_thr}) {
            if ((e = b->reserve(q)) != hipSuccess) return e;
        }
        if ((e = prev_flags.reserve(q)) != hipSuccess) return e;
        if ((e = queue0.reserve(q)) != hipSuccess) return e;
        if ((e = queue1.reserve(q)) != hipSuccess) return e;
        if ((e = pbr_queue.reserve(q)) != hipSuccess) return e;
        if ((e = restir_queue.reserve(q)) != hipSuccess) return e;
        path_capacity = q;
        return hipSuccess;
    }

    hipError_t ensure_pixels(size_t p) {
        if (p <= pixel_capacity) return hipSuccess;
        hipError_t e;
        for (DevBuf<float4>* b : {&beauty_set[0], &beauty_set[1], &normal_set[0], &normal_set[1],
                 &albedo, &denoised, &denoise_scratch, &acc_beauty, &acc_normal, &acc_albedo}) {
            if ((e = b->reserve(p)) != hipSuccess) return e;
        }
        if ((e = depth_set[0].reserve(p)) != hipSuccess) return e;
        if ((e = depth_set[1].reserve(p)) != hipSuccess) return e;
        if ((e = flow.reserve(p)) != hipSuccess) return e;
        if ((e = trust.reserve(p)) != hipSuccess) return e;
        pixel_capacity = p;
        return hipSuccess;
    }
};

namespace {

// device address of the guard's per-camera fallback flags
const int32_t* nn_fallback_flags(const rt_renderer* r) {
    return reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(r->nn_guard.p)
        + offsetof(rtk::NnGuard, fallback));
}

void device_release(rt_device* d) {
    if (d->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(d->ordinal);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

void scene_release(rt_scene* s) {
    if (s->refs.fetch_sub(1) != 1) return;
    rt_device* d = s->dev;
    (void)hipSetDevice(d->ordinal);
    (void)hipStreamSynchronize(d->stream);
    delete s;
    device_release(d);
}

}  // namespace

namespace {

// light distribution: power x area CDF over emissive primitives plus the
// environment (src/realtime/gpu/packed_scene_preparation.cpp:250-302)
// average_texture (packed_scene_preparation.cpp:195-222): a texture's mean
// colour, the emission a textured light is selected by. Checkers average
// their two children (8 levels at most), images their texels, noise is 0.5,
// a bad index or a missing image 0.
static void average_texture(const rt_scene_desc* d, int index, int depth, float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    if (index < 0 || index >= d->texture_count || d->textures == nullptr || depth >= 8) return;
    const rt_texture& t = d->textures[index];
    if (t.type == RT_TEX_CONSTANT) {
        for (int k = 0; k < 3; ++k) out[k] = std::max(t.color[k], 0.0f);
    } else if (t.type == RT_TEX_CHECKER) {
        float a[3], b[3];
        average_texture(d, t.even_texture, depth + 1, a);
        average_texture(d, t.odd_texture, depth + 1, b);
        for (int k = 0; k < 3; ++k) out[k] = 0.5f * (a[k] + b[k]);
    } else if (t.type == RT_TEX_IMAGE && t.image_width > 0 && t.image_height > 0) {
        const int64_t count = (int64_t)t.image_width * t.image_height;
        if (t.image_offset < 0 || t.image_offset + count > d->image_texel_count) return;
        float sum[3] = {0.0f, 0.0f, 0.0f};
        const float* px = d->image_texels + 3 * (int64_t)t.image_offset;
        for (int64_t i = 0; i < count; ++i)
            for (int k = 0; k < 3; ++k) sum[k] += px[3 * i + k];
        for (int k = 0; k < 3; ++k) out[k] = std::max(sum[k] / (float)count, 0.0f);
    } else if (t.type == RT_TEX_NOISE) {
        out[0] = out[1] = out[2] = 0.5f;
    }
}

static void build_lights(const rt_scene_desc* d, std::vector<rtk::LightRec>& out) {
    const float kPi = 3.14159265358979323846f;
    out.clear();
    // material_emission_luminance (packed_scene_preparation.cpp:224-248)
    auto emission_lum = [&](int m) -> float {
        if (m < 0 || m >= d->material_count) return 0.0f;
        const rt_material& mt = d->materials[m];
        if (mt.type == RT_MAT_DIFFUSE_LIGHT) {
            if (d->material_emission_texture != nullptr && d->material_emission_texture[m] >= 0) {
                float avg[3];
                average_texture(d, d->material_emission_texture[m], 0, avg);
                return lum_clamped(avg);
            }
            return lum_clamped(mt.emission);
        }
        if (mt.type != RT_MAT_OPENPBR || mt.openpbr_index < 0 || mt.openpbr_index >= d->openpbr_count)
            return 0.0f;
        rt_openpbr_material pm = d->openpbr[mt.openpbr_index];
        if (d->openpbr_textures != nullptr && d->openpbr_textures[mt.openpbr_index].emission_color >= 0) {
            const rt_openpbr_textures& tb = d->openpbr_textures[mt.openpbr_index];
            float avg[3];
            average_texture(d, tb.emission_color, 0, avg);
            const rtk::f3 c = rtk::source_to_linear(rtk::f3{avg[0], avg[1], avg[2]}, tb.emission_color_space);
            pm.emission_color[0] = c.x;
            pm.emission_color[1] = c.y;
            pm.emission_color[2] = c.z;
        }
        const rtk::f3 e = rtk::openpbr_emission(pm);
        return 0.2126f * std::max(0.0f, e.x) + 0.7152f * std::max(0.0f, e.y)
               + 0.0722f * std::max(0.0f, e.z);
    };
    auto append = [&](int type, int prim, float w) {
        if (std::isfinite(w) && w > 1e-12f) out.push_back(rtk::LightRec{type, prim, w, 0.0f});
    };
    for (int i = 0; i < d->sphere_count; ++i) {
        const float area = 4.0f * kPi * d->spheres[i].radius * d->spheres[i].radius;
        append(0, i, area * emission_lum(d->spheres[i].material));
    }
    for (int i = 0; i < d->quad_count; ++i) {
        const float area = length(cross(v3(d->quads[i].edge_u), v3(d->quads[i].edge_v)));
        append(1, i, area * emission_lum(d->quads[i].material));
    }
    for (int i = 0; i < d->triangle_count; ++i) {
        const rt_triangle& t = d->triangles[i];
        const float area = 0.5f * length(cross(sub(v3(t.p1), v3(t.p0)), sub(v3(t.p2), v3(t.p0))));
        append(2, i, area * emission_lum(t.material));
    }
    float mean[3];
    if (d->background_mode == RT_BG_SKY_GRADIENT) {
        for (int i = 0; i < 3; ++i) mean[i] = (d->background[i] + d->background_top[i]) * 0.5f;
    } else {
        for (int i = 0; i < 3; ++i) mean[i] = d->background[i];
    }
    append(3, -1, 4.0f * kPi * lum_clamped(mean));
    float total = 0.0f;
    for (const auto& l : out) total += l.selection_pdf;
    float cdf = 0.0f;
    for (auto& l : out) {
        l.selection_pdf /= total;
        cdf += l.selection_pdf;
        l.cdf = cdf;
    }
    if (!out.empty()) out.back().cdf = 1.0f;
}

// the analytic lights' selection tables, read from their compiled fields
// (sample_packed_analytic_light, light_sampling.h:104-116, walks the cdf),
// and the dome / distant lights a missed ray collects
// (analytic_infinite_radiance, programs.cu:871-890), in index order
static void analytic_tables(const rt_scene_desc* d, std::vector<float>& sel, std::vector<float>& cdf,
    std::vector<int32_t>& infinite) {
    const int n = d->analytic_light_count;
    sel.resize((size_t)n);
    cdf.resize((size_t)n);
    infinite.clear();
    for (int i = 0; i < n; ++i) {
        const rt_analytic_light& L = d->analytic_lights[i];
        sel[(size_t)i] = L.selection_pdf;
        cdf[(size_t)i] = L.cdf;
        if (L.type == RT_ALIGHT_DISTANT || L.type == RT_ALIGHT_DOME) infinite.push_back(i);
    }
}

// A binary BVH over the finite analytic lights that rays can hit (sphere,
// disk, rect, cylinder; not points), so the analytic hit test visits the
// lights near a ray instead of all of them.  Each light's box is its exact
// extent (sphere c +- r |bx|; disk r sqrt(bx_a^2 + by_a^2); rect
// w/2 |bx_a| + h/2 |by_a|; cylinder the disk term plus l/2 |bz_a|), padded
// so the box test in float never culls a light the exact test would hit.
// Median splits on the widest centroid axis, at most 4 lights per leaf.
static void analytic_light_bvh(const rt_scene_desc* d, std::vector<float4>& nodes,
    std::vector<int32_t>& order) {
    nodes.clear();
    order.clear();
    auto bits = [](int32_t v) {
        float f;
        std::memcpy(&f, &v, sizeof(f));
        return f;
    };
    struct Item { double lo[3], hi[3], c[3]; int32_t index; };
    std::vector<Item> items;
    for (int i = 0; i < d->analytic_light_count; ++i) {
        const rt_analytic_light& L = d->analytic_lights[i];
        if (L.type > RT_ALIGHT_CYLINDER || L.treat_as_point != 0) continue;
        double ext[3];
        for (int a = 0; a < 3; ++a) {
            const double bx = L.basis_x[a], by = L.basis_y[a], bz = L.basis_z[a];
            const double nx = std::sqrt((double)L.basis_x[0] * L.basis_x[0] + (double)L.basis_x[1] * L.basis_x[1]
                                        + (double)L.basis_x[2] * L.basis_x[2]);
            switch (L.type) {
                case RT_ALIGHT_SPHERE: ext[a] = (double)L.radius * nx; break;
                case RT_ALIGHT_DISK: ext[a] = (double)L.radius * std::sqrt(bx * bx + by * by); break;
                case RT_ALIGHT_RECT:
                    ext[a] = 0.5 * (double)L.width * std::fabs(bx) + 0.5 * (double)L.height * std::fabs(by);
                    break;
                default:
                    ext[a] = (double)L.radius * std::sqrt(bx * bx + by * by) + 0.5 * (double)L.length * std::fabs(bz);
                    break;
            }
        }
        Item it;
        for (int a = 0; a < 3; ++a) {
            const double c = L.position[a];
            const double pad = 1e-3 * (std::fabs(c) + ext[a]) + 1e-4;
            it.lo[a] = c - ext[a] - pad;
            it.hi[a] = c + ext[a] + pad;
            it.c[a] = c;
        }
        it.index = i;
        items.push_back(it);
    }
    if (items.empty()) return;
    order.resize(items.size());
    // depth-first emission: a node's left child follows it, its skip index
    // (lo.w) is the first node after its subtree; hi.w holds a leaf's light
    // count in the top 4 bits (0 for an inner node) and its first slot in
    // the low 28
    std::function<void(int, int)> emit = [&](int begin, int end) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
        for (int k = begin; k < end; ++k)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], items[(size_t)k].lo[a]);
                hi[a] = std::max(hi[a], items[(size_t)k].hi[a]);
                clo[a] = std::min(clo[a], items[(size_t)k].c[a]);
                chi[a] = std::max(chi[a], items[(size_t)k].c[a]);
            }
        float flo[3], fhi[3];   // rounded outward
        for (int a = 0; a < 3; ++a) {
            flo[a] = std::nextafter((float)lo[a], -INFINITY);
            fhi[a] = std::nextafter((float)hi[a], INFINITY);
        }
        const size_t node = nodes.size() / 2;
        nodes.push_back(make_float4(flo[0], flo[1], flo[2], 0.0f));
        nodes.push_back(make_float4(fhi[0], fhi[1], fhi[2], 0.0f));
        const int n = end - begin;
        if (n <= 4) {
            for (int k = begin; k < end; ++k) order[(size_t)k] = items[(size_t)k].index;
            nodes[2 * node + 1].w = bits((int32_t)(((uint32_t)n << 28) | (uint32_t)begin));
        } else {
            int axis = 0;
            for (int a = 1; a < 3; ++a)
                if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
            const int mid = begin + n / 2;
            std::nth_element(items.begin() + begin, items.begin() + mid, items.begin() + end,
                [axis](const Item& x, const Item& y) {
                    return x.c[axis] < y.c[axis] || (x.c[axis] == y.c[axis] && x.index < y.index);
                });
            emit(begin, mid);
            emit(mid, end);
        }
        nodes[2 * node].w = bits((int32_t)(nodes.size() / 2));   // skip: past the subtree
    };
    emit(0, (int)items.size());
}

// scene arrays and counts (rt_scene_upload / rt_scene_update)
static rt_status scene_validate(const rt_scene_desc* d) {
    RT_REQUIRE(d->sphere_count >= 0 && d->quad_count >= 0 && d->triangle_count >= 0
                   && d->material_count >= 0 && d->openpbr_count >= 0 && d->analytic_light_count >= 0,
        RT_ERROR_INVALID_ARGUMENT, "scene counts must be non-negative");
    RT_REQUIRE((d->sphere_count == 0 || d->spheres) && (d->quad_count == 0 || d->quads)
                   && (d->triangle_count == 0 || d->triangles)
                   && (d->material_count == 0 || d->materials)
                   && (d->openpbr_count == 0 || d->openpbr)
                   && (d->analytic_light_count == 0 || d->analytic_lights),
        RT_ERROR_INVALID_ARGUMENT, "scene arrays must be non-null when their count is positive");
    const int64_t prims = (int64_t)d->sphere_count + d->quad_count + d->triangle_count;
    // leaf ranges address 28-bit sorted slots
    RT_REQUIRE(prims < (int64_t(1) << 28), RT_ERROR_INVALID_ARGUMENT, "scene primitive count exceeds 2^28");
    // a path carries the OpenPBR index of the subsurface medium it is inside in
    // 19 bits of its vertex flags
    RT_REQUIRE(d->openpbr_count < (1 << 19) - 1, RT_ERROR_INVALID_ARGUMENT,
        "scene OpenPBR material count exceeds 2^19 - 2");
    float prev_cdf = 0.0f;
    for (int i = 0; i < d->analytic_light_count; ++i) {
        const rt_analytic_light& L = d->analytic_lights[i];
        RT_REQUIRE(L.type >= RT_ALIGHT_SPHERE && L.type <= RT_ALIGHT_DOME, RT_ERROR_INVALID_ARGUMENT,
            "analytic light " + std::to_string(i) + " has an unknown type " + std::to_string(L.type));
        const float* f = &L.position[0];
        bool finite = true;
        for (int k = 0; k < 15 + 8; ++k) finite = finite && std::isfinite(f[k]);
        RT_REQUIRE(finite, RT_ERROR_INVALID_ARGUMENT,
            "analytic light " + std::to_string(i) + " has a non-finite field");
        RT_REQUIRE(L.selection_pdf >= 0.0f && L.cdf >= prev_cdf && L.cdf <= 1.0f, RT_ERROR_INVALID_ARGUMENT,
            "analytic light " + std::to_string(i)
                + " has a selection pdf below 0 or a cdf outside [previous cdf, 1] (rt_analytic_lights_compile)");
        prev_cdf = L.cdf;
    }
    RT_REQUIRE(d->medium_count >= 0 && d->medium_count <= 60 && (d->medium_count == 0 || d->media),
        RT_ERROR_INVALID_ARGUMENT, "medium_count must be in [0, 60] with a media array");
    for (int i = 0; i < d->medium_count; ++i) {
        const rt_medium& m = d->media[i];
        RT_REQUIRE(m.boundary_type == RT_MEDIUM_SPHERE || m.boundary_type == RT_MEDIUM_BOX,
            RT_ERROR_INVALID_ARGUMENT, "medium " + std::to_string(i) + " has an unknown boundary type");
        RT_REQUIRE(m.material_index >= 0 && m.material_index < d->material_count, RT_ERROR_INVALID_ARGUMENT,
            "medium " + std::to_string(i) + " names material " + std::to_string(m.material_index)
                + " outside the material table");
        const float* f = &m.local_center_or_min[0];
        bool finite = true;
        for (int k = 0; k < 23; ++k) {
            if (k == 11 || k == 15 || k == 19) continue;   // material_index, boundary_type, pad0
            finite = finite && std::isfinite(f[k]);
        }
        RT_REQUIRE(finite && m.density >= 0.0f, RT_ERROR_INVALID_ARGUMENT,
            "medium " + std::to_string(i) + " has a negative density or a non-finite field");
    }
    RT_REQUIRE(d->texture_count >= 0 && d->image_texel_count >= 0, RT_ERROR_INVALID_ARGUMENT,
        "texture counts must be non-negative");
    RT_REQUIRE((d->texture_count == 0 || d->textures) && (d->image_texel_count == 0 || d->image_texels),
        RT_ERROR_INVALID_ARGUMENT, "texture arrays must be non-null when their count is positive");
    RT_REQUIRE(d->image_texel_count <= (INT32_MAX / 3), RT_ERROR_INVALID_ARGUMENT,
        "image texel count exceeds 2^31 / 3");
    return RT_OK;
}

// host-side tables and device copies of a scene description (everything but
// the BVH); the device is current
static rt_status scene_fill(rt_scene* s, const rt_scene_desc* d) {
    const int64_t prims = (int64_t)d->sphere_count + d->quad_count + d->triangle_count;
    s->env_light = -1;
    s->host_lights.clear();
    {
        // bounding box of the primitives (the radiance cache's grid domain)
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        auto grow = [&](const float* p, double pad) {
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], (double)p[a] - pad);
                hi[a] = std::max(hi[a], (double)p[a] + pad);
            }
        };
        for (int i = 0; i < d->sphere_count; ++i) grow(d->spheres[i].center, std::fabs((double)d->spheres[i].radius));
        for (int i = 0; i < d->quad_count; ++i) {
            const rt_quad& q = d->quads[i];
            for (int k = 0; k < 4; ++k) {
                float c[3];
                for (int a = 0; a < 3; ++a)
                    c[a] = q.origin[a] + ((k & 1) ? q.edge_u[a] : 0.0f) + ((k & 2) ? q.edge_v[a] : 0.0f);
                grow(c, 0.0);
            }
        }
        for (int i = 0; i < d->triangle_count; ++i) {
            grow(d->triangles[i].p0, 0.0);
            grow(d->triangles[i].p1, 0.0);
            grow(d->triangles[i].p2, 0.0);
        }
        for (int a = 0; a < 3; ++a) {
            if (lo[a] > hi[a]) lo[a] = hi[a] = 0.0;
            s->lo[a] = (float)lo[a];
            s->hi[a] = (float)hi[a];
        }
    }
    s->sphere_count = d->sphere_count;
    s->quad_count = d->quad_count;
    s->triangle_count = d->triangle_count;
    s->material_count = d->material_count;
    s->openpbr_count = d->openpbr_count;
    s->analytic_count = d->analytic_light_count;
    {
        bool pbr = false, legacy = false;
        for (int i = 0; i < d->material_count; ++i) {
            const rt_material& m = d->materials[i];
            if (m.type == RT_MAT_OPENPBR && m.openpbr_index >= 0 && m.openpbr_index < d->openpbr_count)
                pbr = true;
            else
                legacy = true;
        }
        s->shade_mode = !pbr ? 1 : (legacy ? 2 : 0);
        s->pbr_full = false;
        for (int i = 0; i < d->openpbr_count; ++i) s->pbr_full = s->pbr_full || rtk::opbr::needs_full(d->openpbr[i]);
    }
    s->background_mode = d->background_mode;
    for (int i = 0; i < 3; ++i) {
        s->bg[i] = d->background[i];
        s->bg_top[i] = d->background_top[i];
    }
    s->prim_count = static_cast<int32_t>(prims);
    build_lights(d, s->host_lights);
    std::vector<int32_t> lop(static_cast<size_t>(std::max<int64_t>(prims, 1)), -1);
    for (size_t i = 0; i < s->host_lights.size(); ++i) {
        const rtk::LightRec& l = s->host_lights[i];
        if (l.type == 3) {
            s->env_light = static_cast<int32_t>(i);
            continue;
        }
        const int base = l.type == 0 ? 0 : (l.type == 1 ? d->sphere_count : d->sphere_count + d->quad_count);
        lop[static_cast<size_t>(base + l.prim)] = static_cast<int32_t>(i);
    }
    hipStream_t st = s->dev->stream;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = upload(s->spheres, d->spheres, (size_t)d->sphere_count, st);
    if (e == hipSuccess) e = upload(s->quads, d->quads, (size_t)d->quad_count, st);
    if (e == hipSuccess) e = upload(s->triangles, d->triangles, (size_t)d->triangle_count, st);
    if (e == hipSuccess) e = upload(s->materials, d->materials, (size_t)d->material_count, st);
    if (e == hipSuccess) e = upload(s->openpbr, d->openpbr, (size_t)d->openpbr_count, st);
    if (e == hipSuccess) e = upload(s->analytic, d->analytic_lights, (size_t)d->analytic_light_count, st);
    s->medium_count = d->medium_count;
    if (e == hipSuccess) e = upload(s->media, d->media, (size_t)d->medium_count, st);
    if (e == hipSuccess) e = upload(s->lights, s->host_lights.data(), s->host_lights.size(), st);
    if (e == hipSuccess) e = upload(s->light_of_prim, lop.data(), lop.size(), st);
    std::vector<float> asel, acdf;
    std::vector<int32_t> ainf;
    analytic_tables(d, asel, acdf, ainf);
    if (e == hipSuccess) e = upload(s->a_sel, asel.data(), asel.size(), st);
    if (e == hipSuccess) e = upload(s->a_cdf, acdf.data(), acdf.size(), st);
    if (e == hipSuccess) e = upload(s->a_infinite, ainf.data(), ainf.size(), st);
    s->a_infinite_count = (int32_t)ainf.size();
    std::vector<float4> anodes;
    std::vector<int32_t> aorder;
    analytic_light_bvh(d, anodes, aorder);
    if (e == hipSuccess) e = upload(s->a_nodes, anodes.data(), anodes.size(), st);
    if (e == hipSuccess) e = upload(s->a_order, aorder.data(), aorder.size(), st);
    s->a_node_count = (int32_t)(anodes.size() / 2);
    s->texture_count = d->texture_count;
    s->texel_count = d->image_texel_count;
    s->has_albedo_textures = d->material_albedo_texture != nullptr && d->material_count > 0;
    if (e == hipSuccess) e = upload(s->textures, d->textures, (size_t)d->texture_count, st);
    if (e == hipSuccess) e = upload(s->texels, d->image_texels, (size_t)d->image_texel_count * 3, st);
    if (e == hipSuccess && s->has_albedo_textures)
        e = upload(s->mat_albedo_tex, d->material_albedo_texture, (size_t)d->material_count, st);
    // bindings that are all -1 leave the scene on the untextured kernels
    s->has_emission_textures = false;
    if (d->material_emission_texture != nullptr)
        for (int i = 0; i < d->material_count; ++i)
            s->has_emission_textures = s->has_emission_textures || d->material_emission_texture[i] >= 0;
    if (e == hipSuccess && s->has_emission_textures)
        e = upload(s->mat_emission_tex, d->material_emission_texture, (size_t)d->material_count, st);
    s->has_pbr_textures = false;
    if (d->openpbr_textures != nullptr)
        for (int i = 0; i < d->openpbr_count; ++i) {
            const rt_openpbr_textures& t = d->openpbr_textures[i];
            s->has_pbr_textures = s->has_pbr_textures || t.base_color >= 0 || t.specular_color >= 0
                                  || t.transmission_color >= 0 || t.emission_color >= 0
                                  || t.base_metalness >= 0 || t.specular_roughness >= 0;
        }
    if (e == hipSuccess && s->has_pbr_textures)
        e = upload(s->pbr_tex, d->openpbr_textures, (size_t)d->openpbr_count, st);
    s->has_triangle_uvs = d->triangle_uvs != nullptr && d->triangle_count > 0;
    if (e == hipSuccess && s->has_triangle_uvs)
        e = upload(s->tri_uv, d->triangle_uvs, (size_t)d->triangle_count * 6, st);
    s->has_triangle_normals = d->triangle_normals != nullptr && d->triangle_count > 0;
    if (e == hipSuccess && s->has_triangle_normals)
        e = upload(s->tri_n, d->triangle_normals, (size_t)d->triangle_count * 9, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return fail(RT_ERROR_HIP, std::string("HIP runtime failure at scene upload: ") + hipGetErrorString(e));
    return RT_OK;
}

// Content signature of a byte range: FNV-1a 64 (the signatures of
// gpu_scene_acceleration.cpp:93-135) taken over 64-bit words in four
// interleaved lanes, folded together at the end, so a million-triangle
// scene hashes in milliseconds; equal contents give equal signatures.
struct Signature {
    uint64_t lane[4] = {14695981039346656037ull, 14695981039346656037ull ^ 1ull,
        14695981039346656037ull ^ 2ull, 14695981039346656037ull ^ 3ull};
    static constexpr uint64_t kPrime = 1099511628211ull;

    void bytes(const void* p, size_t n) {
        const auto* b = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 32 <= n; i += 32) {
            uint64_t w[4];
            std::memcpy(w, b + i, 32);
            for (int k = 0; k < 4; ++k) lane[k] = (lane[k] ^ w[k]) * kPrime;
        }
        for (; i < n; ++i) lane[0] = (lane[0] ^ b[i]) * kPrime;
    }
    template <class T>
    void array(const T* p, int64_t n) {
        bytes(&n, sizeof(n));
        if (n > 0 && p) bytes(p, sizeof(T) * (size_t)n);
    }
    uint64_t value() const {
        uint64_t h = 14695981039346656037ull;
        for (int k = 0; k < 4; ++k) h = (h ^ lane[k]) * kPrime;
        return h;
    }
};

// the primitive positions the BVH bounds depend on (geometry), then
// everything else a frame reads: materials, lights, background, textures,
// per-vertex attributes (scene, which covers the geometry too)
static void scene_signatures(const rt_scene_desc* d, uint64_t* geometry, uint64_t* scene) {
    Signature g;
    g.array(d->spheres, d->sphere_count);
    g.array(d->quads, d->quad_count);
    g.array(d->triangles, d->triangle_count);
    *geometry = g.value();
    Signature h = g;
    h.bytes(&d->background_mode, sizeof(d->background_mode));
    h.bytes(d->background, sizeof(d->background));
    h.bytes(d->background_top, sizeof(d->background_top));
    h.array(d->materials, d->material_count);
    h.array(d->openpbr, d->openpbr_count);
    h.array(d->analytic_lights, d->analytic_light_count);
    h.array(d->textures, d->texture_count);
    h.array(d->image_texels, (int64_t)d->image_texel_count * 3);
    h.array(d->material_albedo_texture, d->material_albedo_texture ? d->material_count : 0);
    h.array(d->material_emission_texture, d->material_emission_texture ? d->material_count : 0);
    h.array(d->openpbr_textures, d->openpbr_textures ? d->openpbr_count : 0);
    h.array(d->triangle_uvs, d->triangle_uvs ? (int64_t)d->triangle_count * 6 : 0);
    h.array(d->triangle_normals, d->triangle_normals ? (int64_t)d->triangle_count * 9 : 0);
    h.array(d->media, d->medium_count);
    *scene = h.value();
}

}  // namespace

extern "C" {

int rt_abi_version(void) { return RT_ABI_VERSION; }

const char* rt_last_error(void) { return g_error.c_str(); }

rt_status rt_device_create(int ordinal, rt_device** out) {
    RT_REQUIRE(out != nullptr, RT_ERROR_INVALID_ARGUMENT, "rt_device_create requires an output handle");
    int count = 0;
    RT_HIP(hipGetDeviceCount(&count));
    RT_REQUIRE(ordinal >= 0 && ordinal < count, RT_ERROR_INVALID_ARGUMENT,
        "device ordinal " + std::to_string(ordinal) + " out of range [0, " + std::to_string(count) + ")");
    RT_HIP(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    RT_HIP(hipGetDeviceProperties(&prop, ordinal));
    auto* d = new rt_device();
    d->ordinal = ordinal;
    d->cu_count = prop.multiProcessorCount;
    d->name = prop.gcnArchName;
    d->total_mem = prop.totalGlobalMem;
    const hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete d;
        return fail(RT_ERROR_HIP, std::string("HIP runtime failure at hipStreamCreate: ") + hipGetErrorString(e));
    }
    // the allocation-free trace entry point's scratch on this device
    const hipError_t se = rtk::trace_scratch_init(ordinal);
    if (se != hipSuccess) {
        (void)hipStreamDestroy(d->stream);
        delete d;
        return fail(RT_ERROR_HIP, std::string("HIP runtime failure at hipGetSymbolAddress: ") + hipGetErrorString(se));
    }
    *out = d;
    return RT_OK;
}

rt_status rt_device_destroy(rt_device* dev) {
    if (!dev) return RT_OK;
    device_release(dev);
    return RT_OK;
}

rt_status rt_device_synchronize(rt_device* dev) {
    RT_REQUIRE(dev != nullptr, RT_ERROR_INVALID_ARGUMENT, "null device");
    RT_HIP(hipStreamSynchronize(dev->stream));
    return RT_OK;
}

rt_status rt_device_stream(rt_device* dev, void** stream) {
    RT_REQUIRE(dev != nullptr && stream != nullptr, RT_ERROR_INVALID_ARGUMENT, "null device or output");
    *stream = static_cast<void*>(dev->stream);
    return RT_OK;
}

// AnalyticLightDesc compilation (analytic_light_compiler.cpp:114-128,
// 178-240) and finalize_analytic_light_distribution (:15-39), over the
// packed fields in FP64
rt_status rt_analytic_lights_compile(rt_analytic_light* lights, int32_t count) {
    RT_REQUIRE(count >= 0 && (count == 0 || lights != nullptr), RT_ERROR_INVALID_ARGUMENT,
        "rt_analytic_lights_compile requires count >= 0 and a light array");
    const double kPiD = 3.14159265358979323846;
    std::vector<double> weight((size_t)count, 0.0);
    double total = 0.0;
    for (int i = 0; i < count; ++i) {
        rt_analytic_light& L = lights[i];
        RT_REQUIRE(L.type >= RT_ALIGHT_SPHERE && L.type <= RT_ALIGHT_DOME, RT_ERROR_INVALID_ARGUMENT,
            "analytic light " + std::to_string(i) + " has an unknown type " + std::to_string(L.type));
        double B[3][3];   // columns basis_x, basis_y, basis_z
        for (int r = 0; r < 3; ++r) {
            B[r][0] = L.basis_x[r];
            B[r][1] = L.basis_y[r];
            B[r][2] = L.basis_z[r];
        }
        auto col_cross_norm = [&]() {
            const double cx = B[1][0] * B[2][1] - B[2][0] * B[1][1];
            const double cy = B[2][0] * B[0][1] - B[0][0] * B[2][1];
            const double cz = B[0][0] * B[1][1] - B[1][0] * B[0][1];
            return std::sqrt(cx * cx + cy * cy + cz * cz);
        };
        // require_similarity_scale (:114-128): B^T B = s^2 I within 1e-6 max(1, s^2)
        auto similarity_scale = [&](const char* what, double* scale) -> rt_status {
            double G[3][3];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) G[a][b] = B[0][a] * B[0][b] + B[1][a] * B[1][b] + B[2][a] * B[2][b];
            const double s2 = (G[0][0] + G[1][1] + G[2][2]) / 3.0;
            RT_REQUIRE(std::isfinite(s2) && s2 > 0.0, RT_ERROR_INVALID_ARGUMENT,
                std::string(what) + " light has a singular transform: analytic light " + std::to_string(i));
            const double tol = 1e-6 * std::max(1.0, s2);
            double dev = 0.0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) dev = std::max(dev, std::fabs(G[a][b] - (a == b ? s2 : 0.0)));
            RT_REQUIRE(dev <= tol, RT_ERROR_INVALID_ARGUMENT,
                std::string(what) + " light requires a similarity transform in this backend: analytic light "
                    + std::to_string(i));
            *scale = std::sqrt(s2);
            return RT_OK;
        };
        double area = 0.0, scale = 1.0;
        switch (L.type) {
            case RT_ALIGHT_SPHERE: {
                const rt_status st = similarity_scale("sphere", &scale);
                if (st != RT_OK) return st;
                area = 4.0 * kPiD * std::pow((double)L.radius * scale, 2.0);
                break;
            }
            case RT_ALIGHT_DISK:
                area = kPiD * (double)L.radius * (double)L.radius * col_cross_norm();
                break;
            case RT_ALIGHT_RECT:
                area = (double)L.width * (double)L.height * col_cross_norm();
                break;
            case RT_ALIGHT_CYLINDER: {
                const rt_status st = similarity_scale("cylinder", &scale);
                if (st != RT_OK) return st;
                area = 2.0 * kPiD * ((double)L.radius * scale) * ((double)L.length * scale);
                break;
            }
            default: break;
        }
        if (L.type <= RT_ALIGHT_CYLINDER) L.world_area = (float)area;
        if (L.type == RT_ALIGHT_DISTANT) L.delta = L.cos_theta_max >= 1.0f ? 1 : 0;
        const double lum = std::max(0.0, 0.2126 * (double)L.radiance[0] + 0.7152 * (double)L.radiance[1]
                                             + 0.0722 * (double)L.radiance[2]);
        double w;
        if (L.type == RT_ALIGHT_DOME) w = 4.0 * kPiD * lum;
        else if (L.type == RT_ALIGHT_DISTANT)
            w = (L.delta ? 1.0 : 2.0 * kPiD * (1.0 - (double)L.cos_theta_max)) * lum;
        else w = area * lum;
        w = std::isfinite(w) ? std::max(0.0, w) : 0.0;
        weight[(size_t)i] = w;
        total += w;
    }
    double acc = 0.0;
    int last = -1;
    for (int i = 0; i < count; ++i) {
        const double p = total > 0.0 ? weight[(size_t)i] / total : 0.0;
        acc += p;
        lights[i].selection_pdf = (float)p;
        lights[i].cdf = (float)acc;
        if (p > 0.0) last = i;
    }
    if (total > 0.0)
        for (int i = last; i < count; ++i) lights[i].cdf = 1.0f;
    return RT_OK;
}

rt_status rt_device_info(rt_device* dev, char* name, size_t name_len, int* cu_count,
    uint64_t* total_bytes, uint64_t* free_bytes) {
    RT_REQUIRE(dev != nullptr, RT_ERROR_INVALID_ARGUMENT, "null device");
    if (name && name_len > 0) {
        std::strncpy(name, dev->name.c_str(), name_len - 1);
        name[name_len - 1] = '\0';
    }
    if (cu_count) *cu_count = dev->cu_count;
    if (total_bytes) *total_bytes = dev->total_mem;
    if (free_bytes) {
        size_t f = 0, t = 0;
        RT_HIP(hipMemGetInfo(&f, &t));
        *free_bytes = f;
    }
    return RT_OK;
}

rt_status rt_scene_upload(rt_device* dev, const rt_scene_desc* d, rt_scene** out) {
    RT_REQUIRE(dev && d && out, RT_ERROR_INVALID_ARGUMENT, "rt_scene_upload requires device, description and output");
    const rt_status vs = scene_validate(d);
    if (vs != RT_OK) return vs;
    RT_HIP(hipSetDevice(dev->ordinal));
    auto* s = new rt_scene();
    s->dev = dev;
    dev->refs.fetch_add(1);
    const rt_status fs = scene_fill(s, d);
    if (fs != RT_OK) {
        scene_release(s);
        return fs;
    }
    scene_signatures(d, &s->geometry_sig, &s->scene_sig);
    *out = s;
    return RT_OK;
}

rt_status rt_scene_destroy(rt_scene* scene) {
    if (!scene) return RT_OK;
    scene_release(scene);
    return RT_OK;
}

// LBVH build (refit = false) or refit over the last build's topology, then
// the 4-wide collapse; the device is current
static rt_status bvh_work(rt_scene* s, bool refit, rt_bvh_stats* stats) {
    const int n = s->prim_count;
    hipStream_t st = s->dev->stream;
    s->bvh_ready = false;
    if (!refit) {
        s->node_count = n > 1 ? n - 1 : (n == 1 ? 1 : 0);
        s->max_depth = 0;
    }
    double build_ms = 0.0;
    if (n > 0) {
        RT_HIP(s->nodes.reserve((size_t)s->node_count * 4));
        RT_HIP(s->leaves.reserve((size_t)n * 3));
        RT_HIP(s->sorted.reserve((size_t)n));
        RT_HIP(s->morton.reserve((size_t)n));
        RT_HIP(s->leaf_parent.reserve((size_t)n));
        DevBuf<float4> pbox;
        DevBuf<uint32_t> k0, k1, v0, v1, counts, flags, cb;
        const size_t nblocks = ((size_t)n + 1023) / 1024;
        RT_HIP(pbox.reserve((size_t)n * 2));
        RT_HIP(k0.reserve((size_t)n));
        RT_HIP(k1.reserve((size_t)n));
        RT_HIP(v0.reserve((size_t)n));
        RT_HIP(v1.reserve((size_t)n));
        RT_HIP(counts.reserve(16 * nblocks));
        RT_HIP(flags.reserve((size_t)n));
        RT_HIP(cb.reserve(8));
        rtk::BvhBuildScratch w{};
        w.pbox = pbox.p;
        w.keys[0] = k0.p;
        w.keys[1] = k1.p;
        w.vals[0] = v0.p;
        w.vals[1] = v1.p;
        w.counts = counts.p;
        w.leaf_parent = s->leaf_parent.p;
        w.flags = flags.p;
        w.cb = cb.p;
        hipEvent_t e0, e1;
        RT_HIP(hipEventCreate(&e0));
        RT_HIP(hipEventCreate(&e1));
        rtk::SceneView v = s->view();
        RT_HIP(hipEventRecord(e0, st));
        if (refit) {
            RT_HIP(rtk::bvh_refit(v, n, w, s->nodes.p, s->leaves.p, s->sorted.p, s->leaf_parent.p, st));
        } else {
            RT_HIP(rtk::bvh_build(v, n, w, s->nodes.p, s->leaves.p, s->sorted.p, s->morton.p, st));
        }
        RT_HIP(hipEventRecord(e1, st));
        uint32_t hcb[8] = {0};
        RT_HIP(hipMemcpyAsync(hcb, cb.p, sizeof(hcb), hipMemcpyDeviceToHost, st));
        RT_HIP(hipStreamSynchronize(st));
        float ms = 0.0f;
        RT_HIP(hipEventElapsedTime(&ms, e0, e1));
        build_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (!refit) s->max_depth = static_cast<int32_t>(hcb[6]);
        RT_REQUIRE(s->max_depth >= 1 && s->max_depth <= 64, RT_ERROR_HIP,
            "LBVH build produced an invalid depth " + std::to_string(s->max_depth));
        // collapse to the 4-wide quantized tree every ray traverses
        const int nbin = s->node_count;
        DevBuf<uint32_t> is_wide, index, f0, f1, fc, tiles;
        RT_HIP(tiles.reserve((size_t)rtk::scan_tiles(nbin)));
        RT_HIP(is_wide.reserve((size_t)nbin));
        RT_HIP(index.reserve((size_t)nbin));
        RT_HIP(f0.reserve((size_t)nbin));
        RT_HIP(f1.reserve((size_t)nbin));
        RT_HIP(fc.reserve(3));
        RT_HIP(s->nodes4.reserve((size_t)nbin * 4));
        rtk::WideScratch ws{};
        ws.is_wide = is_wide.p;
        ws.index = index.p;
        ws.front[0] = f0.p;
        ws.front[1] = f1.p;
        ws.counts = fc.p;
        ws.tiles = tiles.p;
        hipEvent_t w0, w1;
        RT_HIP(hipEventCreate(&w0));
        RT_HIP(hipEventCreate(&w1));
        RT_HIP(hipEventRecord(w0, st));
        int wc = 0, wl = 0;
        RT_HIP(rtk::bvh_collapse_wide(s->nodes.p, nbin, ws, s->nodes4.p, &wc, &wl, st));
        RT_HIP(hipEventRecord(w1, st));
        RT_HIP(hipStreamSynchronize(st));
        float wms = 0.0f;
        RT_HIP(hipEventElapsedTime(&wms, w0, w1));
        (void)hipEventDestroy(w0);
        (void)hipEventDestroy(w1);
        build_ms += wms;
        s->wide_count = wc;
        s->wide_levels = wl;
        RT_REQUIRE(wc >= 1 && wl >= 1 && wl <= 64, RT_ERROR_HIP,
            "4-wide collapse produced an invalid tree (" + std::to_string(wc) + " nodes, "
                + std::to_string(wl) + " levels)");
    }
    s->bvh_ready = true;
    if (stats) {
        stats->prim_count = n;
        stats->node_count = s->node_count;
        stats->max_depth = s->max_depth;
        stats->wide_node_count = s->wide_count;
        stats->wide_levels = s->wide_levels;
        stats->pad = 0;
        stats->build_ms = build_ms;
    }
    return RT_OK;
}

static void record_update(rt_scene* s, int kind, double elapsed_ms, double bvh_ms) {
    if (kind != RT_ACCEL_REUSE) ++s->generation;
    rt_acceleration_update& u = s->last_update;
    u.kind = kind;
    u.node_count = s->node_count;
    u.primitive_reference_count = s->prim_count;
    u.wide_node_count = s->wide_count;
    u.elapsed_ms = elapsed_ms;
    u.bvh_ms = bvh_ms;
    u.generation = s->generation;
}

rt_status rt_bvh_build(rt_scene* s, rt_bvh_stats* stats) {
    RT_REQUIRE(s != nullptr, RT_ERROR_INVALID_ARGUMENT, "null scene");
    RT_HIP(hipSetDevice(s->dev->ordinal));
    const auto t0 = std::chrono::steady_clock::now();
    rt_bvh_stats local{};
    const rt_status st = bvh_work(s, false, &local);
    if (st != RT_OK) return st;
    if (stats) *stats = local;
    record_update(s, RT_ACCEL_REBUILD,
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
        local.build_ms);
    return RT_OK;
}

rt_status rt_scene_update(rt_scene* s, const rt_scene_desc* d, rt_acceleration_update* out) {
    RT_REQUIRE(s != nullptr && d != nullptr, RT_ERROR_INVALID_ARGUMENT,
        "rt_scene_update requires a scene and a description");
    const rt_status vs = scene_validate(d);
    if (vs != RT_OK) return vs;
    RT_HIP(hipSetDevice(s->dev->ordinal));
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t gsig = 0, ssig = 0;
    scene_signatures(d, &gsig, &ssig);
    const bool topology_changed = s->sphere_count != d->sphere_count
                                  || s->quad_count != d->quad_count
                                  || s->triangle_count != d->triangle_count;
    int kind = RT_ACCEL_REUSE;
    if (topology_changed || !s->bvh_ready || s->prim_count <= 1) kind = RT_ACCEL_REBUILD;
    else if (gsig != s->geometry_sig) kind = RT_ACCEL_REFIT;
    else if (ssig != s->scene_sig) kind = RT_ACCEL_UPDATE;
    double bvh_ms = 0.0;
    if (kind != RT_ACCEL_REUSE) {
        // frames in flight read the old contents until here
        RT_HIP(hipStreamSynchronize(s->dev->stream));
        const rt_status fs = scene_fill(s, d);
        if (fs != RT_OK) return fs;
        if (kind == RT_ACCEL_REBUILD || kind == RT_ACCEL_REFIT) {
            rt_bvh_stats bs{};
            const rt_status bw = bvh_work(s, kind == RT_ACCEL_REFIT, &bs);
            if (bw != RT_OK) return bw;
            bvh_ms = bs.build_ms;
        }
        rt_alloc_bump();  // captured frame graphs may hold the replaced scene buffers
        s->geometry_sig = gsig;
        s->scene_sig = ssig;
    }
    record_update(s, kind,
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
        bvh_ms);
    if (out) *out = s->last_update;
    return RT_OK;
}

rt_status rt_scene_last_update(rt_scene* s, rt_acceleration_update* out) {
    RT_REQUIRE(s != nullptr && out != nullptr, RT_ERROR_INVALID_ARGUMENT,
        "rt_scene_last_update requires a scene and an output");
    *out = s->last_update;
    return RT_OK;
}

rt_status rt_bvh_download(rt_scene* s, rt_bvh_node* nodes, uint32_t* sorted_prims, uint32_t* morton) {
    RT_REQUIRE(s != nullptr, RT_ERROR_INVALID_ARGUMENT, "null scene");
    RT_REQUIRE(s->bvh_ready, RT_ERROR_NOT_READY, "rt_bvh_download requires rt_bvh_build() first");
    RT_HIP(hipSetDevice(s->dev->ordinal));
    hipStream_t st = s->dev->stream;
    if (s->prim_count > 0) {
        if (nodes)
            RT_HIP(hipMemcpyAsync(nodes, s->nodes.p, sizeof(rt_bvh_node) * (size_t)s->node_count,
                hipMemcpyDeviceToHost, st));
        if (sorted_prims)
            RT_HIP(hipMemcpyAsync(sorted_prims, s->sorted.p, sizeof(uint32_t) * (size_t)s->prim_count,
                hipMemcpyDeviceToHost, st));
        if (morton)
            RT_HIP(hipMemcpyAsync(morton, s->morton.p, sizeof(uint32_t) * (size_t)s->prim_count,
                hipMemcpyDeviceToHost, st));
    }
    RT_HIP(hipStreamSynchronize(st));
    return RT_OK;
}

rt_status rt_bvh_download_wide(rt_scene* s, void* nodes4) {
    RT_REQUIRE(s != nullptr && nodes4 != nullptr, RT_ERROR_INVALID_ARGUMENT, "null scene or output");
    RT_REQUIRE(s->bvh_ready, RT_ERROR_NOT_READY, "rt_bvh_download_wide requires rt_bvh_build() first");
    RT_HIP(hipSetDevice(s->dev->ordinal));
    if (s->prim_count > 0) {
        RT_HIP(hipMemcpyAsync(nodes4, s->nodes4.p, 64u * (size_t)s->wide_count, hipMemcpyDeviceToHost,
            s->dev->stream));
    }
    RT_HIP(hipStreamSynchronize(s->dev->stream));
    return RT_OK;
}

namespace {

// fetch cursor + stack spill of the test-hook trace kernels (one launch)
struct TraceScratch {
    DevBuf<uint32_t> fetch;
    DevBuf<uint2> spill;
    int grid = 1;
    hipError_t prepare(const rt_scene* s, int64_t n, hipStream_t st) {
        const int64_t blocks = (n + rtk::kTraceBlock - 1) / rtk::kTraceBlock;
        grid = (int)std::max<int64_t>(1, std::min<int64_t>(blocks, (int64_t)s->dev->cu_count * 8));
        const rtk::SceneView v = s->view();
        const size_t entries = (size_t)std::max(v.stack4_entries - rtk::kStackLds, 1);
        hipError_t e = fetch.reserve(1);
        if (e == hipSuccess) e = spill.reserve(entries * (size_t)grid * rtk::kTraceBlock);
        if (e == hipSuccess) e = hipMemsetAsync(fetch.p, 0, sizeof(uint32_t), st);
        return e;
    }
};

}  // namespace

static rt_status trace_host(rt_scene* s, const rt_ray* rays, int64_t n, void* out, bool any) {
    RT_REQUIRE(s != nullptr, RT_ERROR_INVALID_ARGUMENT, "null scene");
    RT_REQUIRE(n >= 0 && (n == 0 || (rays && out)), RT_ERROR_INVALID_ARGUMENT, "invalid ray arrays");
    RT_REQUIRE(s->bvh_ready, RT_ERROR_NOT_READY, "tracing requires rt_bvh_build() first");
    if (n == 0) return RT_OK;
    RT_HIP(hipSetDevice(s->dev->ordinal));
    hipStream_t st = s->dev->stream;
    RT_REQUIRE(n < (int64_t(1) << 31), RT_ERROR_INVALID_ARGUMENT, "at most 2^31 - 1 rays per call");
    DevBuf<rt_ray> dr;
    RT_HIP(upload(dr, rays, (size_t)n, st));
    const rtk::SceneView v = s->view();
    TraceScratch ts;
    RT_HIP(ts.prepare(s, n, st));
    if (any) {
        DevBuf<int32_t> dh;
        RT_HIP(dh.reserve((size_t)n));
        RT_HIP(rtk::trace_any_launch(v, ts.grid, dr.p, n, dh.p, ts.fetch.p, ts.spill.p, st));
        RT_HIP(hipMemcpyAsync(out, dh.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, st));
        RT_HIP(hipStreamSynchronize(st));
    } else {
        DevBuf<rt_hit> dh;
        RT_HIP(dh.reserve((size_t)n));
        RT_HIP(rtk::trace_closest_launch(v, ts.grid, dr.p, n, dh.p, ts.fetch.p, ts.spill.p, st));
        RT_HIP(hipMemcpyAsync(out, dh.p, sizeof(rt_hit) * (size_t)n, hipMemcpyDeviceToHost, st));
        RT_HIP(hipStreamSynchronize(st));
    }
    return RT_OK;
}

rt_status rt_trace_closest(rt_scene* s, const rt_ray* rays, int64_t n, rt_hit* hits) {
    return trace_host(s, rays, n, hits, false);
}

rt_status rt_trace_any(rt_scene* s, const rt_ray* rays, int64_t n, int32_t* occluded) {
    return trace_host(s, rays, n, occluded, true);
}

rt_status rt_trace_closest_device_noalloc(rt_scene* s, const rt_ray* d_rays, int64_t n,
    rt_hit* d_hits) {
    RT_REQUIRE(s != nullptr, RT_ERROR_INVALID_ARGUMENT, "null scene");
    RT_REQUIRE(s->bvh_ready, RT_ERROR_NOT_READY, "tracing requires rt_bvh_build() first");
    RT_REQUIRE(n >= 0 && (n == 0 || (d_rays && d_hits)), RT_ERROR_INVALID_ARGUMENT, "invalid ray arrays");
    RT_REQUIRE(n < (int64_t(1) << 31), RT_ERROR_INVALID_ARGUMENT, "at most 2^31 - 1 rays per call");
    if (n == 0) return RT_OK;
    RT_HIP(hipSetDevice(s->dev->ordinal));
    hipError_t e = hipSuccess;
    const int rc = rtk::trace_closest_noalloc(s->view(), s->dev->ordinal, s->dev->cu_count, d_rays,
        n, d_hits, s->dev->stream, &e);
    RT_REQUIRE(rc != 1, RT_ERROR_NOT_READY,
        "rt_trace_closest_device_noalloc: scratch not resolved on device " + std::to_string(s->dev->ordinal));
    RT_REQUIRE(rc != 2, RT_ERROR_INVALID_ARGUMENT,
        "rt_trace_closest_device_noalloc: the scene's traversal stack needs more than the " +
        std::to_string(rtk::kNoAllocScratchBytes - 256u) + "-byte scratch holds for one block");
    RT_HIP(e);
    return RT_OK;
}

rt_status rt_trace_closest_device(rt_scene* s, const rt_ray* d_rays, int64_t n, rt_hit* d_hits,
    double* kernel_ms) {
    RT_REQUIRE(s != nullptr, RT_ERROR_INVALID_ARGUMENT, "null scene");
    RT_REQUIRE(s->bvh_ready, RT_ERROR_NOT_READY, "tracing requires rt_bvh_build() first");
    RT_REQUIRE(n >= 0 && (n == 0 || (d_rays && d_hits)), RT_ERROR_INVALID_ARGUMENT, "invalid ray arrays");
    RT_HIP(hipSetDevice(s->dev->ordinal));
    hipStream_t st = s->dev->stream;
    hipEvent_t e0, e1;
    RT_HIP(hipEventCreate(&e0));
    RT_HIP(hipEventCreate(&e1));
    RT_REQUIRE(n < (int64_t(1) << 31), RT_ERROR_INVALID_ARGUMENT, "at most 2^31 - 1 rays per call");
    const rtk::SceneView v = s->view();
    TraceScratch ts;
    RT_HIP(ts.prepare(s, n, st));
    RT_HIP(hipEventRecord(e0, st));
    RT_HIP(rtk::trace_closest_launch(v, ts.grid, d_rays, n, d_hits, ts.fetch.p, ts.spill.p, st));
    RT_HIP(hipEventRecord(e1, st));
    RT_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    RT_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (kernel_ms) *kernel_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return RT_OK;
}

rt_status rt_renderer_create(rt_device* dev, rt_scene* scene, rt_renderer** out) {
    RT_REQUIRE(dev && scene && out, RT_ERROR_INVALID_ARGUMENT, "rt_renderer_create requires device, scene and output");
    RT_REQUIRE(scene->dev == dev, RT_ERROR_INVALID_ARGUMENT, "the scene was uploaded to another device");
    RT_HIP(hipSetDevice(dev->ordinal));
    auto* r = new rt_renderer();
    r->dev = dev;
    r->scene = scene;
    scene->refs.fetch_add(1);
    // the autotuner's choice on the 4-camera Cornell mesh workload (MI355X):
    // room for every sample of a 4 x 640x480 x 16-spp frame in one wavefront,
    // so each bounce is one traversal launch (a launch costs about the
    // traversal latency of its slowest ray, whatever its size)
    r->tuning.wavefront_paths = 1 << 25;
    r->tuning.trace_block = 256;
    r->tuning.trace_blocks_per_cu = 8;
    r->tuning.trace_refill = 4;
    r->tuning.use_graph = 0;
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&r->host_cams), sizeof(rt_camera) * 16,
        hipHostMallocDefault);
    if (e == hipSuccess) e = r->cams.reserve(8);
    if (e == hipSuccess) e = r->dyn.reserve(1);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&r->host_dyn), sizeof(rtk::FrameDyn) * 2, hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&r->host_stats), sizeof(uint32_t) * rtk::kStatWords * 2,
            hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->slot_ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->slot_ev[1], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&r->post_ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&r->post_ev[1]);
    if (e == hipSuccess) e = r->counters.reserve(rtk::kCounterAlloc);
    if (e == hipSuccess) e = r->nn_guard.reserve(1);
    if (e == hipSuccess) {
        rtk::NnGuard g{};
        g.max_mean_shift = kNnMaxMeanShift;
        g.max_residual_noise = kNnMaxResidualNoise;
        e = hipMemcpy(r->nn_guard.p, &g, sizeof(g), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        delete r;
        scene_release(scene);
        return fail(RT_ERROR_HIP, std::string("HIP runtime failure at renderer allocation: ") + hipGetErrorString(e));
    }
    *out = r;
    return RT_OK;
}

rt_status rt_renderer_destroy(rt_renderer* r) {
    if (!r) return RT_OK;
    rt_scene* scene = r->scene;
    rt_nn* nn = r->nn;
    rt_cache* cache = r->cache;
    (void)hipSetDevice(r->dev->ordinal);
    (void)hipStreamSynchronize(r->dev->stream);
    delete r;
    if (nn) rt_nn_release(nn);
    if (cache) rt_cache_release(cache);
    scene_release(scene);
    return RT_OK;
}

rt_status rt_renderer_set_tuning(rt_renderer* r, const rt_tuning* t) {
    RT_REQUIRE(r && t, RT_ERROR_INVALID_ARGUMENT, "null renderer or tuning");
    if (t->wavefront_paths > 0) r->tuning.wavefront_paths = t->wavefront_paths;
    if (t->trace_block > 0) {
        RT_REQUIRE(t->trace_block == 64 || t->trace_block == 128 || t->trace_block == 256,
            RT_ERROR_INVALID_ARGUMENT, "trace_block must be 64, 128 or 256");
        r->tuning.trace_block = t->trace_block;
    }
    if (t->trace_blocks_per_cu > 0) r->tuning.trace_blocks_per_cu = t->trace_blocks_per_cu;
    if (t->use_graph >= 0) r->tuning.use_graph = t->use_graph;
    if (t->count_traversal >= 0) r->tuning.count_traversal = t->count_traversal;
    if (t->trace_refill > 0) {
        RT_REQUIRE(t->trace_refill <= 64, RT_ERROR_INVALID_ARGUMENT, "trace_refill must be in [1, 64]");
        r->tuning.trace_refill = t->trace_refill;
    }
    return RT_OK;
}

rt_status rt_renderer_autotune(rt_renderer* r, const rt_camera_rig* rig, const rt_profile* p,
    int32_t frames_per_trial, rt_autotune_result* out) {
    RT_REQUIRE(r && rig && p && out, RT_ERROR_INVALID_ARGUMENT, "null renderer, rig, profile or result");
    RT_REQUIRE(frames_per_trial >= 1 && frames_per_trial <= 64, RT_ERROR_INVALID_ARGUMENT,
        "frames_per_trial must be in [1, 64]");
    std::memset(out, 0, sizeof(*out));
    int trials = 0;
    uint32_t stream = 0x7A5E0000u;
    // median device time of a tuning (one warm-up frame first)
    auto measure = [&](const rt_tuning& t, double* ms) -> rt_status {
        r->tuning = t;
        rt_frame_timing tm{};
        rt_status st = rt_render_frame(r, rig, p, stream++, nullptr);
        if (st != RT_OK) return st;
        std::vector<double> v;
        for (int i = 0; i < frames_per_trial; ++i) {
            st = rt_render_frame(r, rig, p, stream++, &tm);
            if (st != RT_OK) return st;
            v.push_back(tm.render_ms);
        }
        std::sort(v.begin(), v.end());
        *ms = v[v.size() / 2];
        ++trials;
        return RT_OK;
    };
    rt_tuning best = r->tuning;
    best.count_traversal = 0;
    double best_ms = 0.0;
    rt_status st = measure(best, &best_ms);
    if (st != RT_OK) return st;
    out->baseline_ms = best_ms;
    // candidates per knob; paths in flight 4..32 Mi (allocated only up to
    // the frame's pixels x samples)
    const int blocks[] = {64, 128, 256};
    const int per_cu[] = {4, 6, 8, 10, 12, 16};
    const int refill[] = {4, 16, 64};
    const int paths[] = {1 << 22, 1 << 23, 1 << 24, 1 << 25};
    const int graph[] = {0, 1};
    for (int pass = 0; pass < 2; ++pass) {
        const rt_tuning start = best;
        auto sweep = [&](const int* vals, int n, int32_t rt_tuning::*field) -> rt_status {
            for (int i = 0; i < n; ++i) {
                if (best.*field == vals[i]) continue;
                rt_tuning t = best;
                t.*field = vals[i];
                double ms = 0.0;
                const rt_status s2 = measure(t, &ms);
                if (s2 != RT_OK) return s2;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = t;
                }
            }
            return RT_OK;
        };
        if ((st = sweep(blocks, 3, &rt_tuning::trace_block)) != RT_OK) return st;
        if ((st = sweep(per_cu, 6, &rt_tuning::trace_blocks_per_cu)) != RT_OK) return st;
        if ((st = sweep(refill, 3, &rt_tuning::trace_refill)) != RT_OK) return st;
        if ((st = sweep(paths, 4, &rt_tuning::wavefront_paths)) != RT_OK) return st;
        if ((st = sweep(graph, 2, &rt_tuning::use_graph)) != RT_OK) return st;
        if (std::memcmp(&start, &best, sizeof(best)) == 0) break;
    }
    r->tuning = best;
    out->best = best;
    out->best_ms = best_ms;
    out->trials = trials;
    if (r->nn != nullptr && p->enable_denoise != 0) {
        // the network's tile shapes for this rig's camera images, back to back
        rt_nn_images im{};
        im.count = rig->active_count;
        int off = 0;
        for (int c = 0; c < rig->active_count; ++c) {
            im.offset[c] = off;
            im.width[c] = rig->cameras[c].width;
            off += rig->cameras[c].width * rig->cameras[c].height;
        }
        im.offset[rig->active_count] = off;
        if ((st = rt_nn_autotune(r->nn, &im, frames_per_trial, &out->denoiser)) != RT_OK) return st;
        out->denoiser_tuned = 1;
    }
    return rt_reset_accumulation(r);
}

rt_status rt_renderer_get_tuning(rt_renderer* r, rt_tuning* t) {
    RT_REQUIRE(r && t, RT_ERROR_INVALID_ARGUMENT, "null renderer or tuning");
    *t = r->tuning;
    return RT_OK;
}

// render_request_validation.cpp:15-66
static rt_status validate_request(const rt_camera_rig* rig, const rt_profile* p) {
    RT_REQUIRE(rig && p, RT_ERROR_INVALID_ARGUMENT, "render request requires a camera rig and a profile");
    RT_REQUIRE(rig->active_count >= 1 && rig->active_count <= 4, RT_ERROR_INVALID_ARGUMENT,
        "camera rig active_count must be in [1, 4]");
    for (int c = 0; c < rig->active_count; ++c) {
        const rt_camera& cam = rig->cameras[c];
        RT_REQUIRE(cam.enabled != 0, RT_ERROR_INVALID_ARGUMENT,
            "camera " + std::to_string(c) + " is disabled");
        RT_REQUIRE(cam.width > 0 && cam.height > 0, RT_ERROR_INVALID_ARGUMENT,
            "camera " + std::to_string(c) + " has invalid resolution");
        RT_REQUIRE(cam.model == RT_CAM_PINHOLE32 || cam.model == RT_CAM_EQUI62_LUT1D,
            RT_ERROR_INVALID_ARGUMENT, "camera " + std::to_string(c) + " has an unknown model");
        RT_REQUIRE((int64_t)cam.width * cam.height <= (int64_t(1) << 24), RT_ERROR_INVALID_ARGUMENT,
            "camera " + std::to_string(c) + " exceeds 2^24 pixels");
    }
    RT_REQUIRE(p->max_bounces <= 64, RT_ERROR_INVALID_ARGUMENT, "max_bounces must be <= 64");
    RT_REQUIRE(p->samples_per_pixel <= (1 << 20), RT_ERROR_INVALID_ARGUMENT, "samples_per_pixel must be <= 2^20");
    return RT_OK;
}

// one frame of the rig's active cameras; slot c draws its random numbers
// under camera key keys[c] from sample stream streams[c]
static rt_status render_frame_impl(rt_renderer* r, const rt_camera_rig* rig, const rt_profile* p,
    const uint32_t streams[4], const uint32_t keys[4], rt_frame_timing* timing) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    const rt_status vs = validate_request(rig, p);
    if (vs != RT_OK) return vs;
    rt_scene* s = r->scene;
    RT_REQUIRE(s->bvh_ready, RT_ERROR_NOT_READY, "render_prepared_device requires prepare_scene() first");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    hipStream_t st = r->dev->stream;

    // rig
    r->camera_count = rig->active_count;
    int total = 0;
    for (int c = 0; c < 4; ++c) {
        r->pixel_offset[c] = total;
        if (c < rig->active_count) {
            r->widths[c] = rig->cameras[c].width;
            r->heights[c] = rig->cameras[c].height;
            total += rig->cameras[c].width * rig->cameras[c].height;
        } else {
            r->widths[c] = 0;
            r->heights[c] = 0;
        }
    }
    r->pixel_offset[4] = total;
    // per-camera history: a resized or moved-in-the-layout camera starts over
    // (DeviceFrameBufferSet::resize_history, device_frame_buffers.cpp:88-98)
    for (int c = 0; c < 4; ++c) {
        if (c >= rig->active_count || r->prev_width[c] != r->widths[c]
            || r->prev_height[c] != r->heights[c] || r->prev_offset[c] != r->pixel_offset[c]) {
            r->history_length[c] = 0;
            r->prev_valid[c] = 0;
        }
    }
    // this slot's staging may still feed the copies of the frame two calls back
    const int slot = r->slot;
    if (r->slot_pending[slot]) {
        RT_HIP(hipEventSynchronize(r->slot_ev[slot]));
        r->absorb_stats(slot);
    }
    rt_camera* host_cams = r->host_cams + 8 * slot;
    std::memcpy(host_cams, rig->cameras, sizeof(rt_camera) * 4);
    std::memcpy(host_cams + 4, r->prev_cams, sizeof(rt_camera) * 4);
    RT_HIP(hipMemcpyAsync(r->cams.p, host_cams, sizeof(rt_camera) * 8, hipMemcpyHostToDevice, st));

    const int spp = p->samples_per_pixel > 0 ? p->samples_per_pixel : 1;
    const int max_b = p->max_bounces > 0 ? p->max_bounces : 1;
    const size_t P = (size_t)total;
    const size_t cap_target = (size_t)std::max(r->tuning.wavefront_paths, 1);
    const int layers = (int)std::max<size_t>(1, std::min<size_t>((size_t)spp, cap_target / P));
    RT_HIP(r->ensure_pixels(P));
    RT_HIP(r->ensure_paths(P * (size_t)layers));

    rtk::CameraView cv{};
    cv.cams = r->cams.p;
    cv.count = rig->active_count;
    for (int c = 0; c < 5; ++c) cv.pixel_offset[c] = r->pixel_offset[c];
    for (int c = 0; c < 4; ++c) cv.width[c] = r->widths[c];

    const int min_lights = p->restir_min_analytic_lights > 0 ? p->restir_min_analytic_lights : 1;
    const int restir_on = p->enable_restir_di != 0 && s->analytic_count >= min_lights ? 1 : 0;
    if (restir_on && r->reservoir_pixels < P) {
        RT_HIP(r->reservoirs[0].reserve(P));
        RT_HIP(r->reservoirs[1].reserve(P));
        r->reservoir_pixels = P;
        for (int c = 0; c < 4; ++c) r->history_length[c] = 0;
    }

    // per-frame values, read by the kernels from device memory
    rtk::FrameDyn* dyn = r->host_dyn + slot;
    for (int c = 0; c < 4; ++c) {
        dyn->stream[c] = streams[c];
        dyn->camera_key[c] = keys[c];
    }
    dyn->restir_history = r->restir_last ? 1 : 0;
    for (int c = 0; c < 4; ++c) {
        dyn->history_length[c] = r->history_length[c];
        dyn->prev_valid[c] = r->prev_valid[c];
    }
    RT_HIP(hipMemcpyAsync(r->dyn.p, dyn, sizeof(rtk::FrameDyn), hipMemcpyHostToDevice, st));
    ++r->parameter_uploads;

    rtk::FrameParams fp{};
    fp.spp = spp;
    fp.max_bounces = max_b;
    fp.rr_start_bounce = p->rr_start_bounce;
    fp.jitter = p->pixel_jitter;
    fp.dyn = r->dyn.p;
    fp.total_pixels = total;
    rtk::RestirView rv{};
    rv.enabled = restir_on;
    fp.restir = rv.enabled;
    if (rv.enabled) {
        rv.initial_candidates = p->restir_initial_candidates;
        rv.temporal_reuse = p->restir_temporal_reuse;
        rv.max_history_age = p->restir_max_history_age;
        rv.max_temporal_candidates = p->restir_max_temporal_candidates;
        rv.spatial_neighbors = p->restir_spatial_neighbors;
        rv.max_spatial_candidates = p->restir_max_spatial_candidates;
        rv.bias_correction = p->restir_bias_correction;
        rv.spatial_radius = p->restir_spatial_radius;
        rv.spp = spp;
        rv.current = r->reservoirs[r->res_cur].p;
        rv.history = r->reservoirs[r->res_cur ^ 1].p;
        rv.prev_cams = r->cams.p + 4;
        rv.dyn = r->dyn.p;
    }

    const rtk::SceneView sv = s->view();
    rtk::WavefrontBuffers b = r->buffers();
    const int block = r->tuning.trace_block;
    const rtk::StageLaunch trace_l{r->dev->cu_count * r->tuning.trace_blocks_per_cu, block,
        r->tuning.trace_refill};
    const rtk::StageLaunch shade_l{r->dev->cu_count * rtk::wf_shade_blocks_per_cu(), 256};
    const rtk::StageLaunch shade_lite{r->dev->cu_count * rtk::wf_shade_blocks_per_cu(true), 256};
    const rtk::StageLaunch restir_l{std::min(trace_l.grid * block / 128,
                                        r->dev->cu_count * rtk::wf_restir_blocks_per_cu()),
        128};
    {
        const size_t entries = (size_t)std::max(sv.stack4_entries - rtk::kStackLds, 1);
        RT_HIP(r->spill.reserve(entries * (size_t)trace_l.grid * (size_t)block));
        b.spill = r->spill.p;
    }
    const bool denoise = p->enable_denoise != 0;
    // the denoiser's image table: the cameras back to back
    rtk::NnImages nn_images{};
    nn_images.count = rig->active_count;
    for (int c = 0; c <= rig->active_count; ++c) nn_images.offset[c] = r->pixel_offset[c];
    for (int c = 0; c < rig->active_count; ++c) nn_images.width[c] = r->widths[c];
    if (denoise && r->nn != nullptr) RT_HIP(rt_nn_reserve(r->nn, nn_images));
    // radiance cache: queries (one per path slot) and the frame's training records
    rt_cache* cache = r->cache;
    fp.cache_bounce = -1;
    if (cache != nullptr) {
        const rt_cache_config& cc = cache->cfg;
        RT_HIP(r->cache_queries.reserve(P * (size_t)layers));
        RT_HIP(r->cache_records.reserve((size_t)cc.train_samples));
        RT_HIP(cache->ensure(cc.train_samples));
        fp.cache_bounce = cc.query_bounce;
        if (!cache->fraction_set) {
            // the training share starts at the configured value, or (adaptive)
            // at 1.5 x train_samples / paths of this frame
            const double paths_per_frame = (double)P * (double)spp;
            cache->fraction_seed = cc.train_fraction > 0.0f
                ? cc.train_fraction
                : (float)std::min(1.0, 1.5 * (double)cc.train_samples / std::max(paths_per_frame, 1.0));
            RT_HIP(hipMemcpyAsync(reinterpret_cast<char*>(cache->state.p) + offsetof(rtk::CacheState, train_fraction),
                &cache->fraction_seed, sizeof(float), hipMemcpyHostToDevice, st));
            cache->fraction_set = true;
        }
        fp.cache_state = cache->state.p;
        fp.cache_queries = r->cache_queries.p;
        fp.cache_records = r->cache_records.p;
        fp.cache_record_cap = cc.train_samples;
    }

    const int chunks = (spp + layers - 1) / layers;
    const size_t stage_events = (size_t)chunks * (size_t)(max_b + 1) * 4 + 4;
    const bool timed = timing != nullptr;
    const bool use_graph = r->tuning.use_graph != 0;
    const bool stage_timing = timed && !use_graph;
    if (timed) {
        while (r->events.size() < stage_events) {
            hipEvent_t e;
            RT_HIP(hipEventCreate(&e));
            r->events.push_back(e);
        }
    }
    const int fs = r->frame_set, hs = fs ^ 1;
    size_t ev = 2;
    // every launch of the frame, in order (captured once per key in graph mode)
    auto enqueue = [&](bool marks) -> hipError_t {
        hipError_t e;
        auto mark = [&]() -> hipError_t { return marks ? hipEventRecord(r->events[ev++], st) : hipSuccess; };
#define ENQ(x) \
    if ((e = (x)) != hipSuccess) return e
        ENQ(rtk::wf_reset_counters(b.counters, st));
        ENQ(rtk::wf_clear_accum(total, b, st));
        // the frame's reservoirs start cleared: pixels without a surface keep
        // an empty reservoir (optix_renderer.cpp:256-257 clears them per launch)
        if (rv.enabled) ENQ(rtk::wf_zero(rv.current, P * sizeof(rtk::Reservoir), st));
        rtk::FrameParams f = fp;
        for (int s0 = 0; s0 < spp; s0 += layers) {
            f.layer_count = std::min(layers, spp - s0);
            f.first_sample = s0;
            const int paths = f.layer_count * total;
            ENQ(rtk::wf_raygen(cv, f, b, paths, st));
            const bool count = r->tuning.count_traversal != 0;
            // per bounce one traversal launch: its extension rays together
            // with the shadow rays of the bounce before; the last bounce's
            // shadow rays get a launch of their own
            for (int bounce = 0; bounce < max_b; ++bounce) {
                ENQ(mark());
                ENQ(rtk::wf_trace(sv, cv, f, b, bounce, true, count, trace_l, shade_l.grid, st));
                ENQ(mark());
                ENQ(rtk::wf_shade(sv, cv, f, b, bounce, shade_l, shade_lite, s->shade_mode, st));
                if (f.restir != 0 && bounce == 0) ENQ(rtk::wf_restir(sv, cv, f, rv, b, restir_l, st));
                ENQ(rtk::wf_end_bounce(b, bounce, st));
                ENQ(mark());
                ENQ(mark());
            }
            ENQ(mark());
            ENQ(mark());
            ENQ(mark());
            ENQ(rtk::wf_trace(sv, cv, f, b, max_b, false, count, trace_l, shade_l.grid, st));
            ENQ(mark());
            if (cache != nullptr) {
                ENQ(rtk::wf_cache_chunk(cache->grid_view(), cache->packed(), f, b, paths,
                    cache->grid(), st));
            }
            ENQ(rtk::wf_accumulate(f, b, st));
        }
        ENQ(rtk::wf_finalize(total, spp, b, r->beauty_set[fs].p, r->normal_set[fs].p, r->albedo.p,
            r->depth_set[fs].p, st));
        // launch_resolve_kernel (optix_renderer.cpp:263): cameras without
        // history only clear their flow / trust
        rtk::ResolveParams rp{};
        rp.dyn = r->dyn.p;
        rp.restir_enabled = rv.enabled;
        ENQ(rtk::wf_resolve(cv, rp, r->beauty_set[fs].p, r->normal_set[fs].p, r->depth_set[fs].p,
            r->beauty_set[hs].p, r->normal_set[hs].p, r->depth_set[hs].p, r->flow.p, r->trust.p,
            total, st));
        if (denoise) {
            if (marks) ENQ(hipEventRecord(r->post_ev[0], st));
            if (r->nn != nullptr) {
                ENQ(rt_nn_infer_on_stream(r->nn, nn_images, total, r->beauty_set[fs].p,
                    r->normal_set[fs].p, r->albedo.p, r->depth_set[fs].p, r->denoised.p, st));
                // the learned part's fallback: cameras whose network output is
                // not finite, shifts the mean or keeps the noise are filtered
                ENQ(rtk::wf_nn_guard(cv, r->beauty_set[fs].p, r->denoised.p, r->depth_set[fs].p,
                    r->nn_guard.p, st));
                ENQ(rtk::wf_denoise_filter(cv, rtk::kDenoiseIterations, r->beauty_set[fs].p,
                    r->normal_set[fs].p, r->albedo.p, r->depth_set[fs].p, r->denoised.p,
                    r->denoise_scratch.p, total, st, nn_fallback_flags(r)));
            } else {
                ENQ(rtk::wf_denoise_filter(cv, rtk::kDenoiseIterations, r->beauty_set[fs].p,
                    r->normal_set[fs].p, r->albedo.p, r->depth_set[fs].p, r->denoised.p,
                    r->denoise_scratch.p, total, st));
            }
            if (marks) ENQ(hipEventRecord(r->post_ev[1], st));
        }
        // the cache learns from this frame's training paths (used from the next frame)
        if (cache != nullptr) {
            ENQ(rt_cache_train_frame(cache, r->cache_records.p, rtk::wf_cache_record_count(b),
                cache->cfg.train_samples, st));
        }
#undef ENQ
        return hipSuccess;
    };

    if (timed) RT_HIP(hipEventRecord(r->events[0], st));
    if (use_graph) {
        GraphKey key;
        std::memset(&key, 0, sizeof(key));
        key.spp = spp;
        key.max_b = max_b;
        key.rr = fp.rr_start_bounce;
        key.jitter = fp.jitter;
        key.layers = layers;
        key.total = total;
        key.count = rig->active_count;
        key.restir = rv.enabled;
        key.denoise = denoise ? 1 : 0;
        key.block = block;
        key.grid = trace_l.grid;
        key.refill = trace_l.refill_min;
        key.count_traversal = r->tuning.count_traversal;
        key.fs = fs;
        key.res_cur = r->res_cur;
        const int32_t rp10[10] = {rv.initial_candidates, rv.temporal_reuse, rv.max_history_age,
            rv.max_temporal_candidates, rv.spatial_neighbors, rv.max_spatial_candidates,
            rv.bias_correction, rv.spatial_radius, rv.spp, s->analytic_count};
        std::memcpy(key.restir_params, rp10, sizeof(rp10));
        for (int c = 0; c < 4; ++c) {
            key.widths[c] = r->widths[c];
            key.heights[c] = r->heights[c];
        }
        const void* ptrs[10] = {b.ray_o, b.acc_beauty, r->beauty_set[0].p, r->reservoirs[0].p,
            b.spill, r->nn, sv.nodes4, r->denoised.p, cache, r->cache_queries.p};
        std::memcpy(key.ptrs, ptrs, sizeof(ptrs));
        key.alloc_epoch = rt_alloc_epoch();
        rt_renderer::Graph* g = nullptr;
        for (auto& gr : r->graphs) {
            if (gr.exec && std::memcmp(&gr.key, &key, sizeof(key)) == 0) g = &gr;
        }
        if (g == nullptr) {
            // evict the oldest of the cached graphs
            g = &r->graphs[r->graph_next];
            r->graph_next = (r->graph_next + 1) % (int)(sizeof(r->graphs) / sizeof(r->graphs[0]));
            if (g->exec) {
                (void)hipGraphExecDestroy(g->exec);
                g->exec = nullptr;
            }
            hipGraph_t graph = nullptr;
            RT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const hipError_t ce = enqueue(false);
            const hipError_t ee = hipStreamEndCapture(st, &graph);
            RT_HIP(ce);
            RT_HIP(ee);
            const hipError_t ie = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            RT_HIP(ie);
            g->key = key;
            ++r->graph_captures;
        }
        RT_HIP(hipGraphLaunch(g->exec, st));
    } else {
        RT_HIP(enqueue(stage_timing));
    }
    if (timed) RT_HIP(hipEventRecord(r->events[1], st));
    uint32_t* stats = r->host_stats + (size_t)slot * rtk::kStatWords;
    RT_HIP(hipMemcpyAsync(stats, b.counters, sizeof(uint32_t) * rtk::kStatWords, hipMemcpyDeviceToHost, st));
    RT_HIP(hipEventRecord(r->slot_ev[slot], st));
    r->slot_pending[slot] = true;
    r->slot_cache[slot] = cache;
    r->slot = slot ^ 1;
    if (timed || !r->async_frames) {
        // timing reads the frame's events and counters; a synchronous frame
        // returns with its outputs complete
        RT_HIP(hipEventSynchronize(r->slot_ev[slot]));
        r->absorb_stats(slot);
    }
    r->rendered = true;
    // capture_launch_history (radiance_launch_setup.cpp:122-136)
    for (int c = 0; c < rig->active_count; ++c) {
        r->prev_cams[c] = rig->cameras[c];
        r->history_length[c] = r->history_length[c] == 0 ? 1 : r->history_length[c] + 1;
        r->prev_valid[c] = 1;
        r->prev_width[c] = r->widths[c];
        r->prev_height[c] = r->heights[c];
        r->prev_offset[c] = r->pixel_offset[c];
    }
    r->last_set = fs;
    r->frame_set = hs;
    r->last_denoised = denoise;
    r->last_reservoirs = rv.enabled ? r->res_cur : -1;
    if (rv.enabled) r->res_cur ^= 1;
    r->restir_last = rv.enabled != 0;
    if (timed) {
        std::memset(timing, 0, sizeof(*timing));
        float ms = 0.0f;
        RT_HIP(hipEventElapsedTime(&ms, r->events[0], r->events[1]));
        timing->render_ms = ms;
        double ext = 0.0, shd = 0.0, con = 0.0;
        for (size_t i = 2; stage_timing && i + 3 < ev; i += 4) {
            float a = 0.0f, bb = 0.0f, c = 0.0f;
            RT_HIP(hipEventElapsedTime(&a, r->events[i], r->events[i + 1]));
            RT_HIP(hipEventElapsedTime(&bb, r->events[i + 1], r->events[i + 2]));
            RT_HIP(hipEventElapsedTime(&c, r->events[i + 2], r->events[i + 3]));
            ext += a;
            shd += bb;
            con += c;
        }
        timing->extend_ms = ext;
        timing->shade_ms = shd;
        timing->connect_ms = con;
        float dn = 0.0f;
        if (denoise && stage_timing) RT_HIP(hipEventElapsedTime(&dn, r->post_ev[0], r->post_ev[1]));
        timing->denoise_ms = dn;
        timing->other_ms = stage_timing ? std::max(0.0, timing->render_ms - ext - shd - con - (double)dn) : 0.0;
        uint64_t ex = 0, sh = 0;
        std::memcpy(&ex, stats + 4, 8);
        std::memcpy(&sh, stats + 6, 8);
        timing->paths = (uint64_t)P * (uint64_t)spp;
        timing->extend_rays = ex;
        timing->shadow_rays = sh;
        timing->extend_launches = (uint64_t)chunks * (uint64_t)max_b;
        std::memcpy(&timing->node_visits, stats + 8, 8);
        std::memcpy(&timing->prim_tests, stats + 10, 8);
        std::memcpy(&timing->sphere_tests, stats + 12, 8);
        std::memcpy(&timing->shadow_node_visits, stats + 24, 8);
        std::memcpy(&timing->shadow_prim_tests, stats + 26, 8);
        std::memcpy(&timing->shadow_sphere_tests, stats + 28, 8);
    }
    return RT_OK;
}

// render_direction_debug_frame (direction_debug_renderer.cpp:51-66)
rt_status rt_render_direction_debug(rt_device* dev, const rt_camera_rig* rig, int camera_index,
    uint8_t* rgba) {
    RT_REQUIRE(dev && rig && rgba, RT_ERROR_INVALID_ARGUMENT, "null device, rig or output");
    RT_REQUIRE(rig->active_count >= 1 && rig->active_count <= 4, RT_ERROR_INVALID_ARGUMENT,
        "rig.active_count must be in [1, 4]");
    RT_REQUIRE(camera_index >= 0 && camera_index < rig->active_count, RT_ERROR_INVALID_ARGUMENT,
        "camera_index " + std::to_string(camera_index) + " out of range");
    const rt_camera& cam = rig->cameras[camera_index];
    RT_REQUIRE(cam.enabled != 0 && cam.width > 0 && cam.height > 0, RT_ERROR_INVALID_ARGUMENT,
        "camera " + std::to_string(camera_index) + " is disabled or has no pixels");
    RT_REQUIRE(cam.model == RT_CAM_PINHOLE32 || cam.model == RT_CAM_EQUI62_LUT1D, RT_ERROR_INVALID_ARGUMENT,
        "camera " + std::to_string(camera_index) + " has an unknown model");
    RT_REQUIRE((int64_t)cam.width * cam.height <= (int64_t(1) << 28), RT_ERROR_INVALID_ARGUMENT,
        "camera " + std::to_string(camera_index) + " exceeds 2^28 pixels");
    RT_HIP(hipSetDevice(dev->ordinal));
    hipStream_t st = dev->stream;
    const size_t n = (size_t)cam.width * (size_t)cam.height;
    DevBuf<rt_camera> dcam;
    DevBuf<uint32_t> out;
    RT_HIP(dcam.reserve(1));
    RT_HIP(out.reserve(n));
    RT_HIP(hipMemcpyAsync(dcam.p, &cam, sizeof(rt_camera), hipMemcpyHostToDevice, st));
    RT_HIP(rtk::wf_direction_debug(dcam.p, out.p, cam.width, cam.height, st));
    RT_HIP(hipMemcpyAsync(rgba, out.p, n * 4, hipMemcpyDeviceToHost, st));
    RT_HIP(hipStreamSynchronize(st));
    return RT_OK;
}

rt_status rt_render_frame(rt_renderer* r, const rt_camera_rig* rig, const rt_profile* p,
    uint32_t sample_stream, rt_frame_timing* timing) {
    const uint32_t streams[4] = {sample_stream, sample_stream, sample_stream, sample_stream};
    const uint32_t keys[4] = {0u, 1u, 2u, 3u};
    return render_frame_impl(r, rig, p, streams, keys, timing);
}

rt_status rt_render_frame_streams(rt_renderer* r, const rt_camera_rig* rig, const rt_profile* p,
    const uint32_t* sample_streams, const uint32_t* camera_keys, rt_frame_timing* timing) {
    RT_REQUIRE(sample_streams != nullptr, RT_ERROR_INVALID_ARGUMENT,
        "rt_render_frame_streams requires four sample streams");
    uint32_t keys[4] = {0u, 1u, 2u, 3u};
    if (camera_keys != nullptr) {
        for (int c = 0; c < 4; ++c) {
            RT_REQUIRE(camera_keys[c] < 4u, RT_ERROR_INVALID_ARGUMENT,
                "camera key " + std::to_string(camera_keys[c]) + " of slot " + std::to_string(c)
                    + " out of range [0, 4)");
            keys[c] = camera_keys[c];
        }
    }
    return render_frame_impl(r, rig, p, sample_streams, keys, timing);
}

rt_status rt_renderer_set_async(rt_renderer* r, int32_t enable) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    if (!enable) RT_HIP(r->drain());
    r->async_frames = enable != 0;
    return RT_OK;
}

rt_status rt_renderer_synchronize(rt_renderer* r) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    RT_HIP(r->drain());
    return RT_OK;
}

// LaunchParameterDiagnostics (optix_renderer.cpp:130-138,217-219,257): the
// per-frame parameter block is allocated once with the renderer and
// uploaded once per frame
rt_status rt_renderer_launch_diagnostics(rt_renderer* r, uint64_t* allocation_count,
    uint64_t* upload_count) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    if (allocation_count) *allocation_count = 1u;
    if (upload_count) *upload_count = r->parameter_uploads;
    return RT_OK;
}

rt_status rt_frame_download(rt_renderer* r, int camera, float* beauty, float* normal, float* albedo,
    float* depth) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    RT_REQUIRE(r->rendered, RT_ERROR_NOT_READY, "rt_frame_download requires a rendered frame");
    RT_REQUIRE(camera >= 0 && camera < r->camera_count, RT_ERROR_INVALID_ARGUMENT,
        "camera index " + std::to_string(camera) + " out of range");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    hipStream_t st = r->dev->stream;
    const size_t off = (size_t)r->pixel_offset[camera];
    const size_t n = (size_t)r->widths[camera] * (size_t)r->heights[camera];
    const float4* bsrc = r->last_denoised ? r->denoised.p : r->beauty_set[r->last_set].p;
    if (beauty) RT_HIP(hipMemcpyAsync(beauty, bsrc + off, n * 16, hipMemcpyDeviceToHost, st));
    if (normal) RT_HIP(hipMemcpyAsync(normal, r->normal_set[r->last_set].p + off, n * 16, hipMemcpyDeviceToHost, st));
    if (albedo) RT_HIP(hipMemcpyAsync(albedo, r->albedo.p + off, n * 16, hipMemcpyDeviceToHost, st));
    if (depth) RT_HIP(hipMemcpyAsync(depth, r->depth_set[r->last_set].p + off, n * 4, hipMemcpyDeviceToHost, st));
    RT_HIP(hipStreamSynchronize(st));
    return RT_OK;
}

rt_status rt_frame_device_view(rt_renderer* r, int camera, rt_device_frame_view* view) {
    RT_REQUIRE(r && view, RT_ERROR_INVALID_ARGUMENT, "null renderer or view");
    RT_REQUIRE(r->rendered, RT_ERROR_NOT_READY, "rt_frame_device_view requires a rendered frame");
    RT_REQUIRE(camera >= 0 && camera < r->camera_count, RT_ERROR_INVALID_ARGUMENT,
        "camera index " + std::to_string(camera) + " out of range");
    const size_t off = (size_t)r->pixel_offset[camera];
    view->width = r->widths[camera];
    view->height = r->heights[camera];
    const float4* bsrc = r->last_denoised ? r->denoised.p : r->beauty_set[r->last_set].p;
    view->beauty = reinterpret_cast<const float*>(bsrc + off);
    view->normal = reinterpret_cast<const float*>(r->normal_set[r->last_set].p + off);
    view->albedo = reinterpret_cast<const float*>(r->albedo.p + off);
    view->depth = r->depth_set[r->last_set].p + off;
    view->resolved_beauty = reinterpret_cast<const float*>(r->beauty_set[r->last_set].p + off);
    view->flow = reinterpret_cast<const float*>(r->flow.p + off);
    view->trust = r->trust.p + off;
    return RT_OK;
}

rt_status rt_renderer_set_denoiser(rt_renderer* r, rt_nn* nn) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    RT_REQUIRE(nn == nullptr || rt_nn_device(nn) == r->dev->ordinal, RT_ERROR_INVALID_ARGUMENT,
        "the denoiser lives on another device");
    if (nn) rt_nn_retain(nn);
    if (r->nn) {
        (void)hipStreamSynchronize(r->dev->stream);
        rt_nn_release(r->nn);
    }
    r->nn = nn;
    rt_alloc_bump();  // captured frame graphs hold the previous network's buffers
    return RT_OK;
}

rt_status rt_renderer_set_denoiser_budget(rt_renderer* r, float max_mean_shift,
    float max_residual_noise) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    RT_REQUIRE(max_mean_shift >= 0.0f && max_residual_noise >= 0.0f, RT_ERROR_INVALID_ARGUMENT,
        "denoiser budgets must be non-negative");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    RT_HIP(hipStreamSynchronize(r->dev->stream));
    const float b[2] = {max_mean_shift, max_residual_noise};
    RT_HIP(hipMemcpy(reinterpret_cast<char*>(r->nn_guard.p) + offsetof(rtk::NnGuard, max_mean_shift),
        b, sizeof(b), hipMemcpyHostToDevice));
    return RT_OK;
}

rt_status rt_renderer_denoiser_guard(rt_renderer* r, rt_denoiser_guard* out) {
    RT_REQUIRE(r != nullptr && out != nullptr, RT_ERROR_INVALID_ARGUMENT,
        "rt_renderer_denoiser_guard requires a renderer and an output");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    RT_HIP(hipStreamSynchronize(r->dev->stream));
    rtk::NnGuard g{};
    RT_HIP(hipMemcpy(&g, r->nn_guard.p, sizeof(g), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    for (int c = 0; c < 4; ++c) {
        out->fallback[c] = g.fallback[c];
        out->nonfinite[c] = g.metric[c][0];
        out->mean_shift[c] = g.metric[c][1];
        out->residual_noise[c] = g.metric[c][2];
    }
    out->frames_checked = g.frames_checked;
    out->frames_fallback = g.frames_fallback;
    out->max_mean_shift = g.max_mean_shift;
    out->max_residual_noise = g.max_residual_noise;
    return RT_OK;
}

rt_status rt_renderer_set_radiance_cache(rt_renderer* r, rt_cache* cache) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    RT_REQUIRE(cache == nullptr || cache->ordinal == r->dev->ordinal, RT_ERROR_INVALID_ARGUMENT,
        "the radiance cache lives on another device");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    if (cache) {
        const rt_status bs = rt_cache_set_bounds(cache, r->scene->lo, r->scene->hi);
        if (bs != RT_OK) return bs;
        rt_cache_retain(cache);
    }
    if (r->cache) {
        RT_HIP(r->drain());
        rt_cache_release(r->cache);
    }
    r->cache = cache;
    rt_alloc_bump();  // captured frame graphs hold the previous cache's buffers and bounds
    return RT_OK;
}

rt_status rt_frame_download_guides(rt_renderer* r, int camera, float* flow_xy, float* trust,
    float* resolved_beauty_rgba) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    RT_REQUIRE(r->rendered, RT_ERROR_NOT_READY, "rt_frame_download_guides requires a rendered frame");
    RT_REQUIRE(camera >= 0 && camera < r->camera_count, RT_ERROR_INVALID_ARGUMENT,
        "camera index " + std::to_string(camera) + " out of range");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    hipStream_t st = r->dev->stream;
    const size_t off = (size_t)r->pixel_offset[camera];
    const size_t n = (size_t)r->widths[camera] * (size_t)r->heights[camera];
    if (flow_xy) RT_HIP(hipMemcpyAsync(flow_xy, r->flow.p + off, n * 8, hipMemcpyDeviceToHost, st));
    if (trust) RT_HIP(hipMemcpyAsync(trust, r->trust.p + off, n * 4, hipMemcpyDeviceToHost, st));
    if (resolved_beauty_rgba)
        RT_HIP(hipMemcpyAsync(resolved_beauty_rgba, r->beauty_set[r->last_set].p + off, n * 16,
            hipMemcpyDeviceToHost, st));
    RT_HIP(hipStreamSynchronize(st));
    return RT_OK;
}

// DeviceFrameBufferSet::reset_accumulation (device_frame_buffers.cpp:146-149)
rt_status rt_reset_accumulation(rt_renderer* r) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    for (int c = 0; c < 4; ++c) {
        r->history_length[c] = 0;
        r->prev_valid[c] = 0;
    }
    return RT_OK;
}

rt_status rt_reset_sequence(rt_renderer* r, uint32_t seed) {
    RT_REQUIRE(r != nullptr, RT_ERROR_INVALID_ARGUMENT, "null renderer");
    r->sequence_seed = seed;
    return rt_reset_accumulation(r);
}

// RestirDiagnostics over the reservoirs of the last frame
// (OptixRenderer::restir_diagnostics, optix_renderer.cpp:183-215)
rt_status rt_restir_diagnostics(rt_renderer* r, int camera, rt_restir_stats* out) {
    RT_REQUIRE(r && out, RT_ERROR_INVALID_ARGUMENT, "null renderer or output");
    std::memset(out, 0, sizeof(*out));
    RT_REQUIRE(camera >= 0 && camera < r->camera_count, RT_ERROR_INVALID_ARGUMENT,
        "camera index " + std::to_string(camera) + " out of range");
    if (r->last_reservoirs < 0) return RT_OK;
    RT_HIP(hipSetDevice(r->dev->ordinal));
    const size_t n = (size_t)r->widths[camera] * (size_t)r->heights[camera];
    std::vector<rtk::Reservoir> host(n);
    RT_HIP(hipMemcpyAsync(host.data(), r->reservoirs[r->last_reservoirs].p + r->pixel_offset[camera],
        n * sizeof(rtk::Reservoir), hipMemcpyDeviceToHost, r->dev->stream));
    RT_HIP(hipStreamSynchronize(r->dev->stream));
    out->pixel_count = (int32_t)n;
    for (const rtk::Reservoir& v : host) {
        if (v.valid == 0) continue;
        ++out->active_reservoir_count;
        if (v.temporal_candidate_count > 0) ++out->temporal_reuse_count;
        if (v.spatial_candidate_count > 0) ++out->spatial_reuse_count;
        if (v.bias_correction_source_count > 0) ++out->bias_corrected_count;
        out->max_candidate_count = std::max(out->max_candidate_count, v.candidate_count);
        out->max_age = std::max(out->max_age, v.age);
    }
    return RT_OK;
}

rt_status rt_restir_download(rt_renderer* r, int camera, void* reservoirs) {
    RT_REQUIRE(r && reservoirs, RT_ERROR_INVALID_ARGUMENT, "null renderer or output");
    RT_REQUIRE(camera >= 0 && camera < r->camera_count, RT_ERROR_INVALID_ARGUMENT,
        "camera index " + std::to_string(camera) + " out of range");
    RT_REQUIRE(r->last_reservoirs >= 0, RT_ERROR_NOT_READY, "the last frame did not run ReSTIR DI");
    RT_HIP(hipSetDevice(r->dev->ordinal));
    const size_t n = (size_t)r->widths[camera] * (size_t)r->heights[camera];
    RT_HIP(hipMemcpyAsync(reservoirs, r->reservoirs[r->last_reservoirs].p + r->pixel_offset[camera],
        n * sizeof(rtk::Reservoir), hipMemcpyDeviceToHost, r->dev->stream));
    RT_HIP(hipStreamSynchronize(r->dev->stream));
    return RT_OK;
}

rt_status rt_scene_generate(int config, uint64_t seed, rt_generated_scene** out) {
    RT_REQUIRE(out != nullptr, RT_ERROR_INVALID_ARGUMENT, "null output");
    auto* g = new rt_generated_scene();
    if (!rth::generate_scene(config, seed, g)) {
        delete g;
        return fail(RT_ERROR_INVALID_ARGUMENT, "unknown scene config " + std::to_string(config));
    }
    *out = g;
    return RT_OK;
}

rt_status rt_scene_generate_rig(int config, uint64_t seed, int frame, int frames,
    rt_camera_rig* out) {
    RT_REQUIRE(out != nullptr, RT_ERROR_INVALID_ARGUMENT, "null output");
    if (!rth::generate_rig(config, seed, frame, frames, out))
        return fail(RT_ERROR_INVALID_ARGUMENT, "unknown scene config " + std::to_string(config));
    return RT_OK;
}

rt_status rt_camera_build_lut(rt_camera* c) {
    RT_REQUIRE(c != nullptr, RT_ERROR_INVALID_ARGUMENT, "null camera");
    RT_REQUIRE(c->model == RT_CAM_EQUI62_LUT1D, RT_ERROR_INVALID_ARGUMENT,
        "only equi62_lut1d cameras have a theta LUT");
    RT_REQUIRE(c->width > 0 && c->height > 0 && c->fx > 0.0 && c->fy > 0.0, RT_ERROR_INVALID_ARGUMENT,
        "camera needs a positive size and focal lengths");
    rth::make_equi_lut(c);
    return RT_OK;
}

rt_status rt_scene_generated_free(rt_generated_scene* g) {
    if (!g) return RT_OK;
    rth::free_generated(g);
    delete g;
    return RT_OK;
}

rt_status rt_presenter_create(rt_device* dev, int32_t width, int32_t height, rt_presenter** out) {
    RT_REQUIRE(dev != nullptr && out != nullptr, RT_ERROR_INVALID_ARGUMENT, "null device or output handle");
    RT_REQUIRE(width > 0 && height > 0, RT_ERROR_INVALID_ARGUMENT,
        "rt_presenter_create requires a positive resolution");
    RT_HIP(hipSetDevice(dev->ordinal));
    auto* p = new rt_presenter();
    p->width = width;
    p->height = height;
    const hipError_t e = p->pixels.reserve((size_t)width * (size_t)height);
    if (e != hipSuccess) {
        delete p;
        return fail(e == hipErrorOutOfMemory ? RT_ERROR_OUT_OF_MEMORY : RT_ERROR_HIP,
            std::string("HIP runtime failure at presenter buffer allocation: ") + hipGetErrorString(e));
    }
    dev->refs.fetch_add(1);
    p->dev = dev;
    *out = p;
    return RT_OK;
}

rt_status rt_presenter_destroy(rt_presenter* p) {
    if (!p) return RT_OK;
    rt_device* d = p->dev;
    (void)hipSetDevice(d->ordinal);
    (void)hipStreamSynchronize(d->stream);
    delete p;
    device_release(d);
    return RT_OK;
}

rt_status rt_presenter_present(rt_presenter* p, const rt_device_frame_view* view,
    int32_t* present_status) {
    RT_REQUIRE(p != nullptr && present_status != nullptr, RT_ERROR_INVALID_ARGUMENT,
        "null presenter or status");
    if (view == nullptr || view->beauty == nullptr) {
        *present_status = RT_PRESENT_INVALID_DEVICE_VIEW;
        return RT_OK;
    }
    if (view->width != p->width || view->height != p->height) {
        *present_status = RT_PRESENT_RESOLUTION_MISMATCH;
        return RT_OK;
    }
    RT_HIP(hipSetDevice(p->dev->ordinal));
    RT_HIP(rtk::wf_tone_map_rgba8(reinterpret_cast<const float4*>(view->beauty), p->pixels.p,
        (int64_t)p->width * p->height, p->dev->stream));
    RT_HIP(hipStreamSynchronize(p->dev->stream));
    *present_status = RT_PRESENT_OK;
    return RT_OK;
}

rt_status rt_presenter_pixels(rt_presenter* p, uint8_t** d_rgba8) {
    RT_REQUIRE(p != nullptr && d_rgba8 != nullptr, RT_ERROR_INVALID_ARGUMENT, "null presenter or output");
    *d_rgba8 = reinterpret_cast<uint8_t*>(p->pixels.p);
    return RT_OK;
}

rt_status rt_presenter_download(rt_presenter* p, uint8_t* rgba8) {
    RT_REQUIRE(p != nullptr && rgba8 != nullptr, RT_ERROR_INVALID_ARGUMENT, "null presenter or output");
    RT_HIP(hipSetDevice(p->dev->ordinal));
    RT_HIP(hipMemcpyAsync(rgba8, p->pixels.p, (size_t)p->width * p->height * 4,
        hipMemcpyDeviceToHost, p->dev->stream));
    RT_HIP(hipStreamSynchronize(p->dev->stream));
    return RT_OK;
}

rt_status rt_tone_map_rgba8(rt_device* dev, const float* d_beauty_rgba, uint8_t* d_rgba8,
    int64_t pixel_count) {
    RT_REQUIRE(dev != nullptr, RT_ERROR_INVALID_ARGUMENT, "null device");
    RT_REQUIRE(pixel_count >= 0, RT_ERROR_INVALID_ARGUMENT, "negative pixel count");
    if (pixel_count == 0) return RT_OK;
    RT_REQUIRE(d_beauty_rgba != nullptr && d_rgba8 != nullptr, RT_ERROR_INVALID_ARGUMENT,
        "null source or destination");
    RT_REQUIRE(((uintptr_t)d_beauty_rgba & 15) == 0 && ((uintptr_t)d_rgba8 & 3) == 0,
        RT_ERROR_INVALID_ARGUMENT, "the source must be 16-byte and the destination 4-byte aligned");
    RT_HIP(hipSetDevice(dev->ordinal));
    RT_HIP(rtk::wf_tone_map_rgba8(reinterpret_cast<const float4*>(d_beauty_rgba),
        reinterpret_cast<uint32_t*>(d_rgba8), pixel_count, dev->stream));
    return RT_OK;
}

}  // extern "C"
