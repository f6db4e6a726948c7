// rt_host.cpp -- implementation of the C++ host API (see rt_host.hpp).
#include "rt_host.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace rt_amd {

RenderProfile RenderProfile::quality() {
    RenderProfile p;
    p.samples_per_pixel = 4;
    p.max_bounces = 8;
    p.enable_denoise = false;
    p.rr_start_bounce = 6;
    p.accumulation_reset_rotation_deg = 0.5;
    p.accumulation_reset_translation = 0.01;
    p.enable_restir_di = false;
    return p;
}

RenderProfile RenderProfile::balanced() {
    RenderProfile p;
    p.samples_per_pixel = 2;
    p.max_bounces = 4;
    p.enable_denoise = true;
    p.rr_start_bounce = 3;
    p.accumulation_reset_rotation_deg = 1.0;
    p.accumulation_reset_translation = 0.02;
    p.enable_restir_di = true;
    return p;
}

RenderProfile RenderProfile::realtime() {
    RenderProfile p;
    p.samples_per_pixel = 1;
    p.max_bounces = 2;
    p.enable_denoise = true;
    p.rr_start_bounce = 2;
    p.accumulation_reset_rotation_deg = 2.0;
    p.accumulation_reset_translation = 0.05;
    p.enable_restir_di = true;
    return p;
}

namespace {

bool same(const RenderProfile& a, const RenderProfile& b) {
    return a.samples_per_pixel == b.samples_per_pixel && a.max_bounces == b.max_bounces
           && a.enable_denoise == b.enable_denoise && a.rr_start_bounce == b.rr_start_bounce
           && a.accumulation_reset_rotation_deg == b.accumulation_reset_rotation_deg
           && a.accumulation_reset_translation == b.accumulation_reset_translation
           && a.enable_restir_di == b.enable_restir_di
           && a.restir_initial_candidates == b.restir_initial_candidates
           && a.restir_temporal_reuse == b.restir_temporal_reuse
           && a.restir_max_history_age == b.restir_max_history_age
           && a.restir_max_temporal_candidates == b.restir_max_temporal_candidates
           && a.restir_spatial_neighbors == b.restir_spatial_neighbors
           && a.restir_max_spatial_candidates == b.restir_max_spatial_candidates
           && a.restir_bias_correction == b.restir_bias_correction
           && a.restir_min_analytic_lights == b.restir_min_analytic_lights
           && a.pixel_jitter == b.pixel_jitter && a.restir_spatial_radius == b.restir_spatial_radius;
}

double average_luminance(const std::vector<float>& rgba) {
    // radiance_frame_assembly.cpp:19-28: mean Rec.709 luminance of the beauty
    const size_t n = rgba.size() / 4;
    if (n == 0) return 0.0;
    double s = 0.0;
    for (size_t i = 0; i < n; ++i)
        s += 0.2126 * rgba[4 * i] + 0.7152 * rgba[4 * i + 1] + 0.0722 * rgba[4 * i + 2];
    return s / (double)n;
}

}  // namespace

std::optional<RenderProfile> render_profile_from_name(const std::string& name) {
    if (name == "quality") return RenderProfile::quality();
    if (name == "balanced") return RenderProfile::balanced();
    if (name == "realtime") return RenderProfile::realtime();
    return std::nullopt;
}

std::string render_profile_name(const RenderProfile& profile) {
    if (same(profile, RenderProfile::quality())) return "quality";
    if (same(profile, RenderProfile::balanced())) return "balanced";
    if (same(profile, RenderProfile::realtime())) return "realtime";
    return "default";
}

rt_profile to_abi(const RenderProfile& p) {
    rt_profile r{};
    r.samples_per_pixel = p.samples_per_pixel;
    r.max_bounces = p.max_bounces;
    r.enable_denoise = p.enable_denoise ? 1 : 0;
    r.rr_start_bounce = p.rr_start_bounce;
    r.accumulation_reset_rotation_deg = p.accumulation_reset_rotation_deg;
    r.accumulation_reset_translation = p.accumulation_reset_translation;
    r.enable_restir_di = p.enable_restir_di ? 1 : 0;
    r.restir_initial_candidates = p.restir_initial_candidates;
    r.restir_temporal_reuse = p.restir_temporal_reuse ? 1 : 0;
    r.restir_max_history_age = p.restir_max_history_age;
    r.restir_max_temporal_candidates = p.restir_max_temporal_candidates;
    r.restir_spatial_neighbors = p.restir_spatial_neighbors;
    r.restir_max_spatial_candidates = p.restir_max_spatial_candidates;
    r.restir_bias_correction = static_cast<int32_t>(p.restir_bias_correction);
    r.restir_min_analytic_lights = p.restir_min_analytic_lights;
    r.pixel_jitter = p.pixel_jitter;
    r.restir_spatial_radius = p.restir_spatial_radius;
    return r;
}

RenderProfile from_abi(const rt_profile& r) {
    RenderProfile p;
    p.samples_per_pixel = r.samples_per_pixel;
    p.max_bounces = r.max_bounces;
    p.enable_denoise = r.enable_denoise != 0;
    p.rr_start_bounce = r.rr_start_bounce;
    p.accumulation_reset_rotation_deg = r.accumulation_reset_rotation_deg;
    p.accumulation_reset_translation = r.accumulation_reset_translation;
    p.enable_restir_di = r.enable_restir_di != 0;
    p.restir_initial_candidates = r.restir_initial_candidates;
    p.restir_temporal_reuse = r.restir_temporal_reuse != 0;
    p.restir_max_history_age = r.restir_max_history_age;
    p.restir_max_temporal_candidates = r.restir_max_temporal_candidates;
    p.restir_spatial_neighbors = r.restir_spatial_neighbors;
    p.restir_max_spatial_candidates = r.restir_max_spatial_candidates;
    p.restir_bias_correction = static_cast<RestirBiasCorrectionMode>(r.restir_bias_correction);
    p.restir_min_analytic_lights = r.restir_min_analytic_lights;
    p.pixel_jitter = r.pixel_jitter;
    p.restir_spatial_radius = r.restir_spatial_radius;
    return p;
}

void throw_on_error(rt_status s) {
    if (s == RT_OK) return;
    const std::string msg = rt_last_error();
    if (s == RT_ERROR_INVALID_ARGUMENT || s == RT_ERROR_UNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// ------------------------------------------------------- request validation
void validate_render_camera_request(const rt_camera_rig& rig, int camera_index,
    std::string_view caller) {
    constexpr int kMaxCameraSlots = 4;
    const std::string name(caller);
    if (rig.active_count < 1 || rig.active_count > kMaxCameraSlots) {
        throw std::runtime_error(name + " requires rig.active_count in [1, 4], got "
            + std::to_string(rig.active_count));
    }
    if (camera_index < 0 || camera_index >= rig.active_count) {
        throw std::runtime_error(name + " camera_index out of range: camera_index="
            + std::to_string(camera_index) + ", active_count=" + std::to_string(rig.active_count));
    }
    const rt_camera& camera = rig.cameras[camera_index];
    if (camera.enabled == 0) {
        throw std::runtime_error(name + " camera slot is disabled at index " + std::to_string(camera_index));
    }
    if (camera.width <= 0 || camera.height <= 0) {
        throw std::runtime_error(name + " camera slot has invalid resolution at index "
            + std::to_string(camera_index) + ": width=" + std::to_string(camera.width)
            + ", height=" + std::to_string(camera.height));
    }
}

void validate_render_pool_request(const rt_camera_rig& rig, int active_cameras, int renderer_count,
    std::string_view caller) {
    const std::string name(caller);
    if (active_cameras < 1 || active_cameras > renderer_count)
        throw std::runtime_error(name + " active_cameras out of range");
    if (rig.active_count < 1 || rig.active_count > 4)
        throw std::runtime_error(name + " rig.active_count out of range");
    if (active_cameras > rig.active_count)
        throw std::runtime_error(name + " active_cameras exceeds rig.active_count");
    for (int c = 0; c < active_cameras; ++c) {
        const rt_camera& camera = rig.cameras[c];
        if (camera.enabled == 0)
            throw std::runtime_error(name + " leading camera slot disabled at index " + std::to_string(c));
        if (camera.width <= 0 || camera.height <= 0)
            throw std::runtime_error(name + " leading camera slot has invalid resolution at index "
                + std::to_string(c));
    }
}

std::string acceleration_update_kind_name(AccelerationUpdateKind kind) {
    switch (kind) {
        case AccelerationUpdateKind::reuse: return "reuse";
        case AccelerationUpdateKind::update: return "update";
        case AccelerationUpdateKind::refit: return "refit";
        case AccelerationUpdateKind::rebuild: return "rebuild";
    }
    return "rebuild";
}

namespace {

AccelerationUpdateStats to_stats(const rt_acceleration_update& u) {
    AccelerationUpdateStats a;
    switch (u.kind) {
        case RT_ACCEL_REUSE: a.kind = AccelerationUpdateKind::reuse; break;
        case RT_ACCEL_UPDATE: a.kind = AccelerationUpdateKind::update; break;
        case RT_ACCEL_REFIT: a.kind = AccelerationUpdateKind::refit; break;
        default: a.kind = AccelerationUpdateKind::rebuild; break;
    }
    a.elapsed_ms = u.elapsed_ms;
    a.node_count = u.node_count;
    a.primitive_reference_count = u.primitive_reference_count;
    a.generation = u.generation;
    return a;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// ----------------------------------------------------------------- Renderer
Renderer::Renderer(int ordinal) { throw_on_error(rt_device_create(ordinal, &device_)); }

Renderer::~Renderer() {
    if (renderer_) rt_renderer_destroy(renderer_);
    if (scene_) rt_scene_destroy(scene_);
    if (device_) rt_device_destroy(device_);
}

DirectionDebugFrame Renderer::render_direction_debug(const rt_camera_rig& rig, int camera_index) {
    validate_render_camera_request(rig, camera_index, "render_radiance");
    DirectionDebugFrame f;
    f.width = rig.cameras[camera_index].width;
    f.height = rig.cameras[camera_index].height;
    f.rgba.assign((size_t)f.width * (size_t)f.height * 4, 0);
    throw_on_error(rt_render_direction_debug(device_, &rig, camera_index, f.rgba.data()));
    return f;
}

// OptixRenderer::upload_scene / SharedGpuSceneState::prepare
// (optix_renderer.cpp:76-90,154-158): the first scene is uploaded and built,
// later ones go through the acceleration lifecycle
void Renderer::upload_scene(const rt_scene_desc& scene) {
    scene_prepared_ = false;
    if (scene.sphere_count + scene.quad_count + scene.triangle_count <= 0)
        throw std::runtime_error("render_radiance requires at least one surface primitive");
    rt_acceleration_update u{};
    if (scene_ == nullptr) {
        throw_on_error(rt_scene_upload(device_, &scene, &scene_));
        throw_on_error(rt_bvh_build(scene_, &bvh_));
        throw_on_error(rt_scene_last_update(scene_, &u));
    } else {
        throw_on_error(rt_scene_update(scene_, &scene, &u));
        bvh_.prim_count = u.primitive_reference_count;
        bvh_.node_count = u.node_count;
        bvh_.wide_node_count = u.wide_node_count;
        bvh_.build_ms = u.bvh_ms;
    }
    if (renderer_ == nullptr) throw_on_error(rt_renderer_create(device_, scene_, &renderer_));
    acceleration_ = to_stats(u);
}

void Renderer::prepare_scene(const rt_scene_desc& scene) {
    upload_scene(scene);
    use_prepared_scene(scene);
}

void Renderer::use_prepared_scene(const rt_scene_desc&) {
    if (scene_ == nullptr || renderer_ == nullptr)
        throw std::runtime_error("use_prepared_scene requires a prepared scene");
    scene_prepared_ = true;
    launch_sample_stream_ = 0;
    reset_accumulation();
}

void Renderer::reset_accumulation() {
    if (renderer_) throw_on_error(rt_reset_accumulation(renderer_));
}

void Renderer::reset_sequence(std::uint32_t sample_stream) {
    launch_sample_stream_ = sample_stream;
    reset_accumulation();
}

rt_frame_timing Renderer::render_rig(const rt_camera_rig& rig, const RenderProfile& profile,
    const std::uint32_t sample_streams[4], const std::uint32_t* camera_keys) {
    if (!scene_prepared_) throw std::runtime_error("render_prepared_device requires prepare_scene() first");
    const rt_profile p = to_abi(profile);
    rt_frame_timing t{};
    throw_on_error(rt_render_frame_streams(renderer_, &rig, &p, sample_streams, camera_keys, &t));
    return t;
}

// one camera in slot 0 under its own camera key, with the next sample stream
rt_frame_timing Renderer::launch_camera(const rt_camera_rig& rig, const RenderProfile& profile,
    int camera_index) {
    rt_camera_rig one{};
    one.active_count = 1;
    one.cameras[0] = rig.cameras[camera_index];
    const std::uint32_t s = launch_sample_stream_++;
    const std::uint32_t streams[4] = {s, s, s, s};
    const std::uint32_t keys[4] = {(std::uint32_t)camera_index, 1u, 2u, 3u};
    const rt_profile p = to_abi(profile);
    rt_frame_timing t{};
    throw_on_error(rt_render_frame_streams(renderer_, &one, &p, streams, keys, &t));
    last_slot_ = 0;
    return t;
}

RadianceFrame Renderer::download(int slot) const { return download_profiled(slot, nullptr); }

RadianceFrame Renderer::download_profiled(int slot, RadianceTiming* timing) const {
    if (renderer_ == nullptr) throw std::runtime_error("no frame has been rendered");
    const auto t0 = std::chrono::steady_clock::now();
    rt_device_frame_view v{};
    throw_on_error(rt_frame_device_view(renderer_, slot, &v));
    RadianceFrame f;
    f.width = v.width;
    f.height = v.height;
    const size_t n = (size_t)v.width * (size_t)v.height;
    f.beauty_rgba.resize(4 * n);
    f.normal_rgba.resize(4 * n);
    f.albedo_rgba.resize(4 * n);
    f.depth.resize(n);
    throw_on_error(rt_frame_download(renderer_, slot, f.beauty_rgba.data(), f.normal_rgba.data(),
        f.albedo_rgba.data(), f.depth.data()));
    f.average_luminance = average_luminance(f.beauty_rgba);
    if (timing) timing->download_ms = (float)ms_since(t0);
    return f;
}

DeviceRadianceFrameView Renderer::device_view(int slot) const {
    if (renderer_ == nullptr) throw std::runtime_error("no frame has been rendered");
    rt_device_frame_view v{};
    throw_on_error(rt_frame_device_view(renderer_, slot, &v));
    DeviceRadianceFrameView out;
    out.beauty_rgba = v.beauty;
    out.width = v.width;
    out.height = v.height;
    throw_on_error(rt_device_stream(device_, &out.stream));
    return out;
}

RestirDiagnostics Renderer::restir_diagnostics(int slot) const {
    RestirDiagnostics d;
    if (renderer_ == nullptr) return d;
    rt_restir_stats s{};
    throw_on_error(rt_restir_diagnostics(renderer_, slot, &s));
    d.pixel_count = s.pixel_count;
    d.active_reservoir_count = s.active_reservoir_count;
    d.temporal_reuse_count = s.temporal_reuse_count;
    d.spatial_reuse_count = s.spatial_reuse_count;
    d.bias_corrected_count = s.bias_corrected_count;
    d.max_candidate_count = s.max_candidate_count;
    d.max_age = s.max_age;
    return d;
}

RestirDiagnostics Renderer::restir_diagnostics() const { return restir_diagnostics(last_slot_); }

LaunchParameterDiagnostics Renderer::launch_parameter_diagnostics() const {
    LaunchParameterDiagnostics d;
    d.allocation_count = 1;  // the frame parameter block lives as long as the renderer
    if (renderer_) throw_on_error(rt_renderer_launch_diagnostics(renderer_, &d.allocation_count, &d.upload_count));
    return d;
}

RadianceFrame Renderer::render_radiance(const rt_scene_desc& scene, const rt_camera_rig& rig,
    const RenderProfile& profile, int camera_index) {
    validate_render_camera_request(rig, camera_index, "render_radiance");
    reset_accumulation();
    upload_scene(scene);
    launch_camera(rig, profile, camera_index);
    return download_profiled(0, nullptr);
}

ProfiledDeviceRadianceFrame Renderer::render_prepared_device(const rt_camera_rig& rig,
    const RenderProfile& profile, int camera_index) {
    validate_render_camera_request(rig, camera_index, "render_radiance");
    if (!scene_prepared_) throw std::runtime_error("render_prepared_device requires prepare_scene() first");
    const rt_frame_timing t = launch_camera(rig, profile, camera_index);
    ProfiledDeviceRadianceFrame out;
    out.frame = device_view(0);
    out.timing.render_ms = (float)(t.render_ms - t.denoise_ms);
    out.timing.denoise_ms = (float)t.denoise_ms;
    return out;
}

ProfiledRadianceFrame Renderer::render_prepared_radiance(const rt_camera_rig& rig,
    const RenderProfile& profile, int camera_index) {
    const ProfiledDeviceRadianceFrame device = render_prepared_device(rig, profile, camera_index);
    ProfiledRadianceFrame out;
    out.timing = device.timing;
    out.frame = download_profiled(0, &out.timing);
    return out;
}

ProfiledRadianceFrame Renderer::render_radiance_profiled(const rt_scene_desc& scene,
    const rt_camera_rig& rig, const RenderProfile& profile, int camera_index) {
    validate_render_camera_request(rig, camera_index, "render_radiance");
    prepare_scene(scene);
    return render_prepared_radiance(rig, profile, camera_index);
}

// ------------------------------------------------------------- RendererPool
namespace {
int checked_pool_size(int n) {
    if (n < 1 || n > 4) throw std::runtime_error("RendererPool requires renderer_count in [1, 4]");
    return n;
}
}  // namespace

// the count is validated before the device is opened
RendererPool::RendererPool(int renderer_count, int ordinal)
    : count_(checked_pool_size(renderer_count)), renderer_(std::make_unique<Renderer>(ordinal)) {}

RendererPool::~RendererPool() = default;

void RendererPool::prepare_scene(const rt_scene_desc& scene) {
    std::lock_guard<std::mutex> lock(mutex_);
    renderer_->prepare_scene(scene);
    for (auto& s : streams_) s = 0u;
    submissions_ += (std::uint64_t)count_;
}

void RendererPool::reset_accumulation() {
    std::lock_guard<std::mutex> lock(mutex_);
    renderer_->reset_accumulation();
    submissions_ += (std::uint64_t)count_;
}

void RendererPool::reset_sequence(std::uint32_t sample_stream) {
    std::lock_guard<std::mutex> lock(mutex_);
    renderer_->reset_accumulation();
    for (auto& s : streams_) s = sample_stream;
    submissions_ += (std::uint64_t)count_;
}

// the leading active_cameras cameras in one launch, each from its slot's
// sample stream; the mutex is held by the caller
rt_frame_timing RendererPool::launch(const rt_camera_rig& rig, const RenderProfile& profile,
    int active_cameras) {
    validate_render_pool_request(rig, active_cameras, count_, "RendererPool");
    rt_camera_rig lead = rig;
    lead.active_count = active_cameras;
    std::uint32_t streams[4];
    for (int c = 0; c < 4; ++c) streams[c] = streams_[c];
    const rt_frame_timing t = renderer_->render_rig(lead, profile, streams);
    for (int c = 0; c < active_cameras; ++c) ++streams_[c];
    submissions_ += (std::uint64_t)active_cameras;
    return t;
}

std::vector<CameraDeviceRenderResult> RendererPool::render_device_frame(const rt_camera_rig& rig,
    const RenderProfile& profile, int active_cameras) {
    std::lock_guard<std::mutex> lock(mutex_);
    const rt_frame_timing t = launch(rig, profile, active_cameras);
    std::vector<CameraDeviceRenderResult> out;
    out.reserve((size_t)active_cameras);
    for (int c = 0; c < active_cameras; ++c) {
        CameraDeviceRenderResult r;
        r.camera_index = c;
        r.profiled.frame = renderer_->device_view(c);
        r.profiled.timing.render_ms = (float)(t.render_ms - t.denoise_ms);
        r.profiled.timing.denoise_ms = (float)t.denoise_ms;
        out.push_back(r);
    }
    return out;
}

std::vector<CameraRenderResult> RendererPool::render_frame(const rt_camera_rig& rig,
    const RenderProfile& profile, int active_cameras) {
    std::lock_guard<std::mutex> lock(mutex_);
    const rt_frame_timing t = launch(rig, profile, active_cameras);
    std::vector<CameraRenderResult> out;
    out.reserve((size_t)active_cameras);
    for (int c = 0; c < active_cameras; ++c) {
        CameraRenderResult r;
        r.camera_index = c;
        r.profiled.timing.render_ms = (float)(t.render_ms - t.denoise_ms);
        r.profiled.timing.denoise_ms = (float)t.denoise_ms;
        const auto d0 = std::chrono::steady_clock::now();
        r.profiled.frame = renderer_->download(c);
        r.profiled.timing.download_ms = (float)ms_since(d0);
        out.push_back(std::move(r));
    }
    return out;
}

RendererPoolDiagnostics RendererPool::diagnostics() const {
    std::lock_guard<std::mutex> lock(mutex_);
    RendererPoolDiagnostics d;
    d.worker_count = count_;
    d.worker_start_count = (std::uint64_t)count_;
    d.task_submission_count = submissions_;
    const LaunchParameterDiagnostics l = renderer_->launch_parameter_diagnostics();
    d.launch_parameter_allocation_count = l.allocation_count;
    d.launch_parameter_upload_count = l.upload_count;
    d.acceleration = renderer_->acceleration_diagnostics();
    return d;
}

// ---------------------------------------------------------------- profiling
namespace {

enum BundleTag : std::uint32_t {
    kBundleBackground = 1, kBundleSpheres, kBundleQuads, kBundleTriangles, kBundleMaterials,
    kBundleOpenpbr, kBundleAnalytic, kBundleTextures, kBundleTexels, kBundleAlbedoTex,
    kBundleTriUvs, kBundleTriNormals, kBundleMedia, kBundleEmissionTex, kBundlePbrTex,
    kBundleCameras, kBundleProfile, kBundleName
};

}  // namespace

SceneBundle load_scene_bundle(const std::filesystem::path& path) {
    const std::string where = "scene bundle " + path.string();
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error(where + ": cannot open");
    std::vector<unsigned char> data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    auto rd32 = [&](size_t off) {
        std::uint32_t v;
        std::memcpy(&v, data.data() + off, 4);
        return v;
    };
    if (data.size() < 12 || std::memcmp(data.data(), "RTSB", 4) != 0)
        throw std::runtime_error(where + ": not a scene bundle");
    if (rd32(4) != 1u) throw std::runtime_error(where + ": unsupported version " + std::to_string(rd32(4)));
    const std::uint32_t n = rd32(8);
    SceneBundle b;
    struct Sec {
        std::uint32_t size = 0;
        std::uint64_t count = 0;
        const unsigned char* bytes = nullptr;
    };
    Sec sec[kBundleName + 1];
    size_t off = 12;
    for (std::uint32_t i = 0; i < n; ++i) {
        if (off + 16 > data.size()) throw std::runtime_error(where + ": truncated section header");
        const std::uint32_t tag = rd32(off), size = rd32(off + 4);
        std::uint64_t count;
        std::memcpy(&count, data.data() + off + 8, 8);
        off += 16;
        if (size == 0 || count > (data.size() - off) / size)
            throw std::runtime_error(where + ": truncated section " + std::to_string(tag));
        if (tag < kBundleBackground || tag > kBundleName)
            throw std::runtime_error(where + ": unknown section " + std::to_string(tag));
        sec[tag] = Sec{size, count, data.data() + off};
        off += (size_t)(size * count);
    }
    if (off != data.size()) throw std::runtime_error(where + ": trailing bytes");
    // copies section `tag` into owned storage after checking its element size
    auto take = [&](std::uint32_t tag, size_t elem, const char* what) -> const void* {
        const Sec& s = sec[tag];
        if (s.bytes == nullptr) return nullptr;
        if (s.size != elem)
            throw std::runtime_error(where + ": " + what + " elements are " + std::to_string(s.size)
                + " bytes, this build's are " + std::to_string(elem));
        b.storage.emplace_back(s.bytes, s.bytes + (size_t)s.size * s.count);
        return b.storage.back().data();
    };
    auto count_of = [&](std::uint32_t tag) { return (int32_t)sec[tag].count; };
    if (sec[kBundleBackground].bytes == nullptr || sec[kBundleBackground].size != 28)
        throw std::runtime_error(where + ": missing background");
    std::memcpy(&b.desc.background_mode, sec[kBundleBackground].bytes, 4);
    std::memcpy(&b.desc.background, sec[kBundleBackground].bytes + 4, 12);
    std::memcpy(&b.desc.background_top, sec[kBundleBackground].bytes + 16, 12);
    rt_scene_desc& d = b.desc;
    d.spheres = static_cast<const rt_sphere*>(take(kBundleSpheres, sizeof(rt_sphere), "sphere"));
    d.sphere_count = d.spheres ? count_of(kBundleSpheres) : 0;
    d.quads = static_cast<const rt_quad*>(take(kBundleQuads, sizeof(rt_quad), "quad"));
    d.quad_count = d.quads ? count_of(kBundleQuads) : 0;
    d.triangles = static_cast<const rt_triangle*>(take(kBundleTriangles, sizeof(rt_triangle), "triangle"));
    d.triangle_count = d.triangles ? count_of(kBundleTriangles) : 0;
    d.materials = static_cast<const rt_material*>(take(kBundleMaterials, sizeof(rt_material), "material"));
    d.material_count = d.materials ? count_of(kBundleMaterials) : 0;
    d.openpbr = static_cast<const rt_openpbr_material*>(
        take(kBundleOpenpbr, sizeof(rt_openpbr_material), "OpenPBR material"));
    d.openpbr_count = d.openpbr ? count_of(kBundleOpenpbr) : 0;
    d.analytic_lights = static_cast<const rt_analytic_light*>(
        take(kBundleAnalytic, sizeof(rt_analytic_light), "analytic light"));
    d.analytic_light_count = d.analytic_lights ? count_of(kBundleAnalytic) : 0;
    d.textures = static_cast<const rt_texture*>(take(kBundleTextures, sizeof(rt_texture), "texture"));
    d.texture_count = d.textures ? count_of(kBundleTextures) : 0;
    d.image_texels = static_cast<const float*>(take(kBundleTexels, 3 * sizeof(float), "texel"));
    d.image_texel_count = d.image_texels ? count_of(kBundleTexels) : 0;
    d.media = static_cast<const rt_medium*>(take(kBundleMedia, sizeof(rt_medium), "medium"));
    d.medium_count = d.media ? count_of(kBundleMedia) : 0;
    // per-material, per-triangle and per-record side arrays must match their owners' counts
    auto side = [&](std::uint32_t tag, size_t elem, int32_t owner, const char* what) -> const void* {
        if (sec[tag].bytes != nullptr && (int64_t)sec[tag].count != (int64_t)owner)
            throw std::runtime_error(where + ": " + what + " has " + std::to_string(sec[tag].count)
                + " entries for " + std::to_string(owner));
        return take(tag, elem, what);
    };
    d.material_albedo_texture = static_cast<const int32_t*>(
        side(kBundleAlbedoTex, 4, d.material_count, "material albedo textures"));
    d.material_emission_texture = static_cast<const int32_t*>(
        side(kBundleEmissionTex, 4, d.material_count, "material emission textures"));
    d.triangle_uvs = static_cast<const float*>(side(kBundleTriUvs, 24, d.triangle_count, "triangle uvs"));
    d.triangle_normals = static_cast<const float*>(
        side(kBundleTriNormals, 36, d.triangle_count, "triangle normals"));
    d.openpbr_textures = static_cast<const rt_openpbr_textures*>(
        side(kBundlePbrTex, sizeof(rt_openpbr_textures), d.openpbr_count, "OpenPBR texture bindings"));
    const Sec& cams = sec[kBundleCameras];
    if (cams.bytes == nullptr || cams.size != sizeof(rt_camera) || cams.count < 1 || cams.count > 4)
        throw std::runtime_error(where + ": needs 1-4 cameras of " + std::to_string(sizeof(rt_camera)) + " bytes");
    b.rig.active_count = (int32_t)cams.count;
    std::memcpy(b.rig.cameras, cams.bytes, sizeof(rt_camera) * cams.count);
    const Sec& prof = sec[kBundleProfile];
    if (prof.bytes == nullptr || prof.size != sizeof(rt_profile) || prof.count != 1)
        throw std::runtime_error(where + ": needs one profile of " + std::to_string(sizeof(rt_profile)) + " bytes");
    std::memcpy(&b.profile, prof.bytes, sizeof(rt_profile));
    if (sec[kBundleName].bytes != nullptr)
        b.name.assign(reinterpret_cast<const char*>(sec[kBundleName].bytes), (size_t)sec[kBundleName].count);
    return b;
}

namespace profiling {
namespace {

const char* bool_text(bool v) { return v ? "true" : "false"; }

std::string csv_field(const std::string& v) {
    if (v.find_first_of(",\"\n\r") == std::string::npos) return v;
    std::string e = "\"";
    for (char ch : v) e += ch == '"' ? std::string("\"\"") : std::string(1, ch);
    return e + "\"";
}

std::string json_string(const std::string& v) {
    std::string e;
    static const char* hex = "0123456789abcdef";
    for (unsigned char ch : v) {
        switch (ch) {
            case '"': e += "\\\""; break;
            case '\\': e += "\\\\"; break;
            case '\b': e += "\\b"; break;
            case '\f': e += "\\f"; break;
            case '\n': e += "\\n"; break;
            case '\r': e += "\\r"; break;
            case '\t': e += "\\t"; break;
            default:
                if (ch < 0x20) {
                    e += "\\u00";
                    e += hex[ch >> 4];
                    e += hex[ch & 15];
                } else {
                    e += (char)ch;
                }
        }
    }
    return e;
}

std::ofstream open_or_throw(const std::filesystem::path& path, const char* what) {
    std::ofstream out(path);
    if (!out.is_open())
        throw std::runtime_error(std::string("failed to open ") + what + " output file: " + path.string());
    return out;
}

void close_or_throw(std::ofstream& out, const std::filesystem::path& path, const char* what) {
    out.flush();
    if (!out.good())
        throw std::runtime_error(std::string("failed to write ") + what + " output file: " + path.string());
}

void stats_line(std::ostream& out, const char* name, const AggregateStats& s, bool comma) {
    out << "    \"" << name << "\": {\"avg\": " << s.avg << ", \"p50\": " << s.p50
        << ", \"p95\": " << s.p95 << ", \"p99\": " << s.p99 << ", \"max\": " << s.max << "}"
        << (comma ? "," : "") << "\n";
}

}  // namespace

AggregateStats compute_stats(std::vector<double> values) {
    AggregateStats s;
    if (values.empty()) return s;
    std::sort(values.begin(), values.end());
    const size_t n = values.size();
    auto at = [&](double q) {
        const double rank = std::ceil(q * (double)n);
        const size_t i = rank <= 1.0 ? 0 : (size_t)(rank - 1.0);
        return values[std::min(i, n - 1)];
    };
    double sum = 0.0;
    for (double v : values) sum += v;
    s.avg = sum / (double)n;
    s.p50 = at(0.50);
    s.p95 = at(0.95);
    s.p99 = at(0.99);
    s.max = values.back();
    return s;
}

RunAggregate compute_aggregate(const std::vector<FrameStageSample>& frames) {
    auto col = [&](double FrameStageSample::*m) {
        std::vector<double> v;
        for (const auto& f : frames) v.push_back(f.*m);
        return compute_stats(v);
    };
    RunAggregate a;
    a.frame_ms = col(&FrameStageSample::frame_ms);
    a.pipeline_ms = col(&FrameStageSample::pipeline_ms);
    a.render_ms = col(&FrameStageSample::render_ms);
    a.denoise_ms = col(&FrameStageSample::denoise_ms);
    a.download_ms = col(&FrameStageSample::download_ms);
    a.render_work_ms = col(&FrameStageSample::render_work_ms);
    a.denoise_work_ms = col(&FrameStageSample::denoise_work_ms);
    a.download_work_ms = col(&FrameStageSample::download_work_ms);
    a.image_write_ms = col(&FrameStageSample::image_write_ms);
    a.host_overhead_ms = col(&FrameStageSample::host_overhead_ms);
    return a;
}

void write_csv(const RunReport& r, const std::filesystem::path& path) {
    std::ofstream out = open_or_throw(path, "csv");
    out << "frame_index,sample_stream,camera_count,profile,width,height,samples_per_pixel,"
           "max_bounces,denoise_enabled,frame_ms,pipeline_ms,render_ms,denoise_ms,download_ms,"
           "render_work_ms,denoise_work_ms,download_work_ms,image_write_ms,host_overhead_ms,fps\n";
    for (const FrameStageSample& f : r.frames) {
        out << f.frame_index << "," << f.sample_stream << "," << f.camera_count << ","
            << csv_field(f.profile) << "," << f.width << "," << f.height << "," << f.samples_per_pixel
            << "," << f.max_bounces << "," << bool_text(f.denoise_enabled) << "," << f.frame_ms << ","
            << f.pipeline_ms << "," << f.render_ms << "," << f.denoise_ms << "," << f.download_ms << ","
            << f.render_work_ms << "," << f.denoise_work_ms << "," << f.download_work_ms << ","
            << f.image_write_ms << "," << f.host_overhead_ms << "," << f.fps << "\n";
    }
    close_or_throw(out, path, "csv");
}

void write_json(const RunReport& r, const std::filesystem::path& path) {
    std::ofstream out = open_or_throw(path, "json");
    const auto& m = r;
    out << "{\n  \"schema_version\": " << r.schema_version << ",\n  \"metadata\": {\n";
    out << "    \"scene\": \"" << json_string(m.scene) << "\",\n";
    out << "    \"profile\": \"" << json_string(m.profile) << "\",\n";
    out << "    \"camera_count\": " << m.camera_count << ",\n";
    out << "    \"width\": " << m.width << ",\n    \"height\": " << m.height << ",\n";
    out << "    \"frames_requested\": " << m.frames_requested << ",\n";
    out << "    \"warmup_frames\": " << m.warmup_frames << ",\n";
    out << "    \"random_seed\": " << m.random_seed << ",\n";
    out << "    \"measurement_sample_stream_start\": "
        << (uint64_t)m.random_seed + (uint64_t)m.warmup_frames << ",\n";
    out << "    \"rng_algorithm\": \"pcg4d_pixel_sample_stream_v1\",\n";
    out << "    \"samples_per_pixel\": " << m.samples_per_pixel << ",\n";
    out << "    \"max_bounces\": " << m.max_bounces << ",\n";
    out << "    \"denoise_enabled\": " << bool_text(m.denoise_enabled) << ",\n";
    out << "    \"image_write_enabled\": " << bool_text(m.image_write_enabled) << "\n  },\n";
    const RunProvenance& p = r.provenance;
    out << "  \"provenance\": {\n";
    out << "    \"captured_at_utc\": \"" << json_string(p.captured_at_utc) << "\",\n";
    out << "    \"project_version\": \"" << json_string(p.project_version) << "\",\n";
    out << "    \"source_revision\": \"" << json_string(p.source_revision) << "\",\n";
    out << "    \"source_dirty\": " << bool_text(p.source_dirty) << ",\n";
    out << "    \"source_scope\": \"" << json_string(p.source_scope) << "\",\n";
    out << "    \"source_state_sha256\": \"" << json_string(p.source_state_sha256) << "\",\n";
    out << "    \"build_configuration\": \"" << json_string(p.build_configuration) << "\",\n";
    out << "    \"cxx_compiler\": \"" << json_string(p.cxx_compiler) << "\"\n  },\n";
    const RunEnvironment& e = r.environment;
    out << "  \"environment\": {\n";
    out << "    \"operating_system\": \"" << json_string(e.operating_system) << "\",\n";
    out << "    \"architecture\": \"" << json_string(e.architecture) << "\",\n";
    out << "    \"gpu_device\": " << e.gpu_device << ",\n";
    out << "    \"gpu_name\": \"" << json_string(e.gpu_name) << "\",\n";
    out << "    \"compute_capability\": \"" << json_string(e.compute_capability) << "\",\n";
    out << "    \"nvidia_driver_version\": \"" << json_string(e.nvidia_driver_version) << "\",\n";
    out << "    \"cuda_driver_api_version\": " << e.cuda_driver_api_version << ",\n";
    out << "    \"cuda_driver_api_version_text\": \"" << json_string(e.cuda_driver_api_version_text) << "\",\n";
    out << "    \"cuda_runtime_version\": " << e.cuda_runtime_version << ",\n";
    out << "    \"cuda_runtime_version_text\": \"" << json_string(e.cuda_runtime_version_text) << "\",\n";
    out << "    \"optix_version\": " << e.optix_version << ",\n";
    out << "    \"optix_version_text\": \"" << json_string(e.optix_version_text) << "\"\n  },\n";
    const GpuMemoryReport& g = r.gpu_memory;
    out << "  \"gpu_memory\": {\n    \"scope\": \"" << json_string(g.scope) << "\",\n";
    out << "    \"total_bytes\": " << g.total_bytes << ",\n";
    out << "    \"baseline_used_bytes\": " << g.baseline_used_bytes << ",\n";
    out << "    \"prepared_used_bytes\": " << g.prepared_used_bytes << ",\n";
    out << "    \"peak_used_bytes\": " << g.peak_used_bytes << ",\n";
    out << "    \"peak_delta_bytes\": " << g.peak_delta_bytes << ",\n";
    out << "    \"final_used_bytes\": " << g.final_used_bytes << "\n  },\n";
    const GpuSchedulingReport& s = r.gpu_scheduling;
    out << "  \"gpu_scheduling\": {\n";
    out << "    \"persistent_worker_count\": " << s.persistent_worker_count << ",\n";
    out << "    \"worker_start_count\": " << s.worker_start_count << ",\n";
    out << "    \"task_submission_count\": " << s.task_submission_count << ",\n";
    out << "    \"launch_parameter_allocation_count\": " << s.launch_parameter_allocation_count << ",\n";
    out << "    \"launch_parameter_upload_count\": " << s.launch_parameter_upload_count << ",\n";
    out << "    \"acceleration_update_kind\": \"" << json_string(s.acceleration_update_kind) << "\",\n";
    out << "    \"acceleration_update_ms\": " << s.acceleration_update_ms << ",\n";
    out << "    \"acceleration_node_count\": " << s.acceleration_node_count << ",\n";
    out << "    \"acceleration_reference_count\": " << s.acceleration_reference_count << ",\n";
    out << "    \"acceleration_prototype_count\": " << s.acceleration_prototype_count << ",\n";
    out << "    \"acceleration_instance_count\": " << s.acceleration_instance_count << ",\n";
    out << "    \"acceleration_instanced_primitive_count\": " << s.acceleration_instanced_primitive_count << ",\n";
    out << "    \"acceleration_generation\": " << s.acceleration_generation << "\n  },\n";
    out << "  \"aggregate\": {\n";
    stats_line(out, "frame_ms", r.aggregate.frame_ms, true);
    stats_line(out, "pipeline_ms", r.aggregate.pipeline_ms, true);
    stats_line(out, "render_ms", r.aggregate.render_ms, true);
    stats_line(out, "denoise_ms", r.aggregate.denoise_ms, true);
    stats_line(out, "download_ms", r.aggregate.download_ms, true);
    stats_line(out, "render_work_ms", r.aggregate.render_work_ms, true);
    stats_line(out, "denoise_work_ms", r.aggregate.denoise_work_ms, true);
    stats_line(out, "download_work_ms", r.aggregate.download_work_ms, true);
    stats_line(out, "image_write_ms", r.aggregate.image_write_ms, true);
    stats_line(out, "host_overhead_ms", r.aggregate.host_overhead_ms, false);
    out << "  },\n  \"frames\": [\n";
    for (size_t i = 0; i < r.frames.size(); ++i) {
        const FrameStageSample& f = r.frames[i];
        out << "    {\"frame_index\": " << f.frame_index << ", \"sample_stream\": " << f.sample_stream
            << ", \"camera_count\": " << f.camera_count << ", \"profile\": \"" << json_string(f.profile)
            << "\", \"width\": " << f.width << ", \"height\": " << f.height
            << ", \"samples_per_pixel\": " << f.samples_per_pixel << ", \"max_bounces\": " << f.max_bounces
            << ", \"denoise_enabled\": " << bool_text(f.denoise_enabled) << ", \"frame_ms\": " << f.frame_ms
            << ", \"pipeline_ms\": " << f.pipeline_ms << ", \"render_ms\": " << f.render_ms
            << ", \"denoise_ms\": " << f.denoise_ms << ", \"download_ms\": " << f.download_ms
            << ", \"render_work_ms\": " << f.render_work_ms << ", \"denoise_work_ms\": " << f.denoise_work_ms
            << ", \"download_work_ms\": " << f.download_work_ms << ", \"image_write_ms\": " << f.image_write_ms
            << ", \"host_overhead_ms\": " << f.host_overhead_ms << ", \"fps\": " << f.fps << "}"
            << (i + 1 != r.frames.size() ? "," : "") << "\n";
    }
    out << "  ],\n  \"per-camera\": [\n";
    bool first = true;
    for (const FrameStageSample& f : r.frames) {
        for (const CameraStageSample& c : f.cameras) {
            if (!first) out << ",\n";
            first = false;
            out << "    {\"frame_index\": " << f.frame_index << ", \"camera_index\": " << c.camera_index
                << ", \"render_ms\": " << c.render_ms << ", \"denoise_ms\": " << c.denoise_ms
                << ", \"download_ms\": " << c.download_ms << ", \"average_luminance\": "
                << c.average_luminance << "}";
        }
    }
    out << "\n  ]\n}\n";
    close_or_throw(out, path, "json");
}

std::string fnv1a64_hex(const std::string& bytes) {
    uint64_t h = 14695981039346656037ULL;
    for (unsigned char c : bytes) {
        h ^= c;
        h *= 1099511628211ULL;
    }
    std::ostringstream s;
    s << std::hex << std::setfill('0') << std::setw(16) << h;
    return s.str();
}

void write_artifact_manifest(const RunReport& r, const std::vector<std::filesystem::path>& artifacts,
    const std::filesystem::path& path) {
    if (artifacts.empty())
        throw std::invalid_argument("benchmark artifact manifest requires at least one artifact");
    struct Rec {
        std::string name, kind, digest;
        uintmax_t size;
    };
    std::vector<Rec> recs;
    for (const auto& a : artifacts) {
        if (!std::filesystem::is_regular_file(a))
            throw std::runtime_error("benchmark artifact is not a regular file: " + a.string());
        std::ifstream in(a, std::ios::binary);
        std::string bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        recs.push_back({a.filename().string(),
            a.extension() == ".csv" ? "raw_frame_samples" : "summary_and_provenance",
            fnv1a64_hex(bytes), std::filesystem::file_size(a)});
    }
    size_t per_camera = 0;
    for (const auto& f : r.frames) per_camera += f.cameras.size();
    std::ofstream out = open_or_throw(path, "artifact manifest");
    out << "{\n  \"schema_version\": 1,\n  \"report_schema_version\": " << r.schema_version << ",\n";
    out << "  \"measured_frame_count\": " << r.frames.size() << ",\n";
    out << "  \"per_camera_record_count\": " << per_camera << ",\n  \"artifacts\": [\n";
    for (size_t i = 0; i < recs.size(); ++i) {
        out << "    {\"filename\": \"" << json_string(recs[i].name) << "\", \"kind\": \"" << recs[i].kind
            << "\", \"size_bytes\": " << recs[i].size
            << ", \"digest_algorithm\": \"fnv1a64\", \"digest\": \"" << recs[i].digest << "\"}"
            << (i + 1 != recs.size() ? "," : "") << "\n";
    }
    out << "  ]\n}\n";
    close_or_throw(out, path, "artifact manifest");
}

}  // namespace profiling
}  // namespace rt_amd
