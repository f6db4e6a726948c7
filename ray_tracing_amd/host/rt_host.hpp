// rt_host.hpp -- C++ host API over the C ABI (include/rt/rt_hip.h), shaped
// like the reference's realtime host API so its callers can switch over:
//   RenderProfile + presets        (src/realtime/render_profile.h:10-64)
//   Renderer                       (OptixRenderer, src/realtime/gpu/optix_renderer.h:69-128)
//   RendererPool                   (src/realtime/gpu/renderer_pool.h:30-53)
//   profiling::RunReport + writers (src/realtime/profiling/benchmark_report.h:10-147)
// Request-validation failures are thrown as std::runtime_error with the
// reference's messages; library errors as std::invalid_argument (bad
// arguments) or std::runtime_error with the library's message.
#pragma once

#include <cstdint>
#include <filesystem>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/rt/rt_hip.h"

namespace rt_amd {

enum class RestirBiasCorrectionMode : int { none = 0, basic = 1 };

struct RenderProfile {
    int samples_per_pixel = 1;
    int max_bounces = 4;
    bool enable_denoise = true;
    int rr_start_bounce = 3;
    double accumulation_reset_rotation_deg = 2.0;
    double accumulation_reset_translation = 0.05;
    bool enable_restir_di = false;
    int restir_initial_candidates = 4;
    bool restir_temporal_reuse = true;
    int restir_max_history_age = 20;
    int restir_max_temporal_candidates = 64;
    int restir_spatial_neighbors = 1;
    int restir_max_spatial_candidates = 4;
    RestirBiasCorrectionMode restir_bias_correction = RestirBiasCorrectionMode::basic;
    int restir_min_analytic_lights = 16;
    // extensions of this library (rt_profile): primary-ray jitter inside the
    // pixel, and the ReSTIR spatial radius in pixels (0: the reference's)
    int pixel_jitter = 0;
    int restir_spatial_radius = 0;

    static RenderProfile quality();
    static RenderProfile balanced();
    static RenderProfile realtime();
    static RenderProfile realtime_default() { return balanced(); }
};

std::optional<RenderProfile> render_profile_from_name(const std::string& name);
std::string render_profile_name(const RenderProfile& profile);
rt_profile to_abi(const RenderProfile& p);
RenderProfile from_abi(const rt_profile& p);

void throw_on_error(rt_status s);

// RenderRequestValidation (src/realtime/gpu/render_request_validation.cpp:
// 15-66): the reference's checks and messages, thrown as std::runtime_error
void validate_render_camera_request(const rt_camera_rig& rig, int camera_index,
    std::string_view caller);
void validate_render_pool_request(const rt_camera_rig& rig, int active_cameras, int renderer_count,
    std::string_view caller);

// frames of one camera (RadianceFrame, src/realtime/gpu/frame_types.h:14-21)
struct RadianceFrame {
    int width = 0;
    int height = 0;
    std::vector<float> beauty_rgba;
    std::vector<float> normal_rgba;
    std::vector<float> albedo_rgba;
    std::vector<float> depth;
    double average_luminance = 0.0;
};

// DirectionDebugFrame (src/realtime/gpu/frame_types.h:8-12)
struct DirectionDebugFrame {
    int width = 0;
    int height = 0;
    std::vector<std::uint8_t> rgba;
};

// RadianceTiming / ProfiledRadianceFrame / ProfiledDeviceRadianceFrame
// (src/realtime/gpu/optix_renderer.h:26-38)
struct RadianceTiming {
    float render_ms = 0.0f;
    float denoise_ms = 0.0f;
    float download_ms = 0.0f;
};

struct ProfiledRadianceFrame {
    RadianceFrame frame;
    RadianceTiming timing;
};

// DeviceRadianceFrameView (src/realtime/gpu/device_radiance_frame_view.h):
// beauty_rgba is a device float4 per pixel; stream is the device's
// hipStream_t, on which the frame was produced
struct DeviceRadianceFrameView {
    const float* beauty_rgba = nullptr;
    int width = 0;
    int height = 0;
    void* stream = nullptr;
};

struct ProfiledDeviceRadianceFrame {
    DeviceRadianceFrameView frame;
    RadianceTiming timing;
};

// RestirDiagnostics (optix_renderer.h:40-48)
struct RestirDiagnostics {
    int pixel_count = 0;
    int active_reservoir_count = 0;
    int temporal_reuse_count = 0;
    int spatial_reuse_count = 0;
    int bias_corrected_count = 0;
    int max_candidate_count = 0;
    int max_age = 0;
};

// LaunchParameterDiagnostics (optix_renderer.h:50-53)
struct LaunchParameterDiagnostics {
    std::uint64_t allocation_count = 0;
    std::uint64_t upload_count = 0;
};

// AccelerationUpdateKind / AccelerationUpdateStats
// (src/realtime/gpu/gpu_scene_acceleration.h:12-30).  Scenes reach the
// library flattened, so the prototype and instance counts are 0.
enum class AccelerationUpdateKind { reuse, update, refit, rebuild };

struct AccelerationUpdateStats {
    AccelerationUpdateKind kind = AccelerationUpdateKind::rebuild;
    double elapsed_ms = 0.0;
    int node_count = 0;
    int primitive_reference_count = 0;
    int prototype_count = 0;
    int instance_count = 0;
    int instanced_primitive_count = 0;
    std::uint64_t generation = 0;
};

std::string acceleration_update_kind_name(AccelerationUpdateKind kind);

// OptixRenderer (src/realtime/gpu/optix_renderer.h:69-128) on one device.
// The reference renders one camera per launch; here a launch traces any
// number of a rig's cameras, and the single-camera calls render the asked
// camera in slot 0 under its own camera key, so its samples are those of
// the same camera in a whole-rig frame (rt_render_frame_streams).  Every
// launch takes the renderer's sample stream and advances it by one
// (launch_sample_stream_, optix_renderer.cpp:237); prepare_scene and
// use_prepared_scene restart it at 0, reset_sequence sets it.
class Renderer {
public:
    explicit Renderer(int ordinal = 0);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    Renderer(Renderer&&) = delete;
    Renderer& operator=(Renderer&&) = delete;

    // the camera-model debug image of one camera (optix_renderer.cpp:146-152);
    // needs no scene
    DirectionDebugFrame render_direction_debug(const rt_camera_rig& rig, int camera_index = 0);
    // upload + acceleration update (reuse / update / refit / rebuild), then
    // use_prepared_scene
    void prepare_scene(const rt_scene_desc& scene);
    // adopts the scene prepared last (optix_renderer.cpp:412-417)
    void use_prepared_scene(const rt_scene_desc& scene);
    void reset_accumulation();
    void reset_sequence(std::uint32_t sample_stream);
    RestirDiagnostics restir_diagnostics() const;
    LaunchParameterDiagnostics launch_parameter_diagnostics() const;
    const AccelerationUpdateStats& acceleration_diagnostics() const { return acceleration_; }
    // uploads `scene` and renders one camera; leaves the renderer unprepared
    // (optix_renderer.cpp:419-432)
    RadianceFrame render_radiance(const rt_scene_desc& scene, const rt_camera_rig& rig,
        const RenderProfile& profile, int camera_index);
    // the view remains valid until this renderer launches another frame
    ProfiledDeviceRadianceFrame render_prepared_device(const rt_camera_rig& rig,
        const RenderProfile& profile, int camera_index);
    ProfiledRadianceFrame render_prepared_radiance(const rt_camera_rig& rig,
        const RenderProfile& profile, int camera_index);
    ProfiledRadianceFrame render_radiance_profiled(const rt_scene_desc& scene,
        const rt_camera_rig& rig, const RenderProfile& profile, int camera_index);

    // Whole-rig launch (the pool's path): slot c of `rig` draws
    // sample_streams[c] under camera key camera_keys[c] (nullptr: the slot
    // index); requires a prepared scene.  Does not touch the renderer's own
    // sample stream.
    rt_frame_timing render_rig(const rt_camera_rig& rig, const RenderProfile& profile,
        const std::uint32_t sample_streams[4], const std::uint32_t* camera_keys = nullptr);
    RadianceFrame download(int slot) const;
    DeviceRadianceFrameView device_view(int slot) const;
    RestirDiagnostics restir_diagnostics(int slot) const;
    std::uint32_t next_sample_stream() const { return launch_sample_stream_; }
    bool scene_prepared() const { return scene_prepared_; }
    rt_bvh_stats acceleration_stats() const { return bvh_; }
    rt_renderer* handle() const { return renderer_; }
    rt_device* device() const { return device_; }

private:
    void upload_scene(const rt_scene_desc& scene);
    rt_frame_timing launch_camera(const rt_camera_rig& rig, const RenderProfile& profile,
        int camera_index);
    RadianceFrame download_profiled(int slot, RadianceTiming* timing) const;

    rt_device* device_ = nullptr;
    rt_scene* scene_ = nullptr;
    rt_renderer* renderer_ = nullptr;
    rt_bvh_stats bvh_{};
    AccelerationUpdateStats acceleration_{};
    bool scene_prepared_ = false;
    std::uint32_t launch_sample_stream_ = 0;
    int last_slot_ = 0;
};

// CameraRenderResult / CameraDeviceRenderResult / RendererPoolDiagnostics
// (src/realtime/gpu/renderer_pool.h:11-28)
struct CameraRenderResult {
    int camera_index = 0;
    ProfiledRadianceFrame profiled;
};

struct CameraDeviceRenderResult {
    int camera_index = 0;
    ProfiledDeviceRadianceFrame profiled;
};

struct RendererPoolDiagnostics {
    int worker_count = 0;
    std::uint64_t worker_start_count = 0;
    std::uint64_t task_submission_count = 0;
    std::uint64_t launch_parameter_allocation_count = 0;
    std::uint64_t launch_parameter_upload_count = 0;
    AccelerationUpdateStats acceleration;
};

// RendererPool (src/realtime/gpu/renderer_pool.h:30-53).  The reference
// keeps one renderer, worker thread and sample-stream counter per camera;
// here one renderer traces the pool's leading active_cameras cameras in one
// launch, each camera slot keeping its own sample-stream counter (advanced
// by every frame that renders it, restarted by prepare_scene, set by
// reset_sequence), so camera i's frame equals what a lone Renderer makes of
// camera i.  Frame results carry the shared launch's render and denoise
// times.  Diagnostics count the reference's units: worker_count and
// worker_start_count are the pool's camera slots (started once, at
// construction), task_submission_count counts one task per slot for
// prepare / reset calls and one per rendered camera, and the launch
// parameter counts are the one renderer's (one upload per frame).
class RendererPool {
public:
    explicit RendererPool(int renderer_count, int ordinal = 0);
    ~RendererPool();
    RendererPool(const RendererPool&) = delete;
    RendererPool& operator=(const RendererPool&) = delete;
    RendererPool(RendererPool&&) = delete;
    RendererPool& operator=(RendererPool&&) = delete;

    void prepare_scene(const rt_scene_desc& scene);
    void reset_accumulation();
    void reset_sequence(std::uint32_t sample_stream);
    // device views are consumed before the next pool operation
    std::vector<CameraDeviceRenderResult> render_device_frame(const rt_camera_rig& rig,
        const RenderProfile& profile, int active_cameras);
    std::vector<CameraRenderResult> render_frame(const rt_camera_rig& rig,
        const RenderProfile& profile, int active_cameras);
    RendererPoolDiagnostics diagnostics() const;

    int renderer_count() const { return count_; }
    Renderer& renderer() { return *renderer_; }

private:
    rt_frame_timing launch(const rt_camera_rig& rig, const RenderProfile& profile,
        int active_cameras);

    int count_;
    std::unique_ptr<Renderer> renderer_;
    std::uint32_t streams_[4] = {0u, 0u, 0u, 0u};
    std::uint64_t submissions_ = 0;
    mutable std::mutex mutex_;
};

// A scene bundle written by ray_tracing_amd.scene_bundle (the package's
// YAML ingest flattened into every array of an rt_scene_desc, the scene's
// camera rig and a render profile).  Layout: magic "RTSB", u32 version 1,
// u32 section count, then per section u32 tag, u32 element size, u64
// element count and the elements; element sizes must equal this build's
// struct sizes.  `desc` points into the bundle's own storage.
struct SceneBundle {
    std::string name;
    rt_scene_desc desc{};
    rt_camera_rig rig{};
    rt_profile profile{};
    std::vector<std::vector<unsigned char>> storage;
};
// throws std::runtime_error naming the file and the defect
SceneBundle load_scene_bundle(const std::filesystem::path& path);

namespace profiling {

struct CameraStageSample {
    int camera_index = 0;
    double render_ms = 0.0;
    double denoise_ms = 0.0;
    double download_ms = 0.0;
    double average_luminance = 0.0;
};

struct FrameStageSample {
    int frame_index = 0;
    std::uint32_t sample_stream = 0;
    int camera_count = 0;
    std::string profile;
    int width = 0;
    int height = 0;
    int samples_per_pixel = 0;
    int max_bounces = 0;
    bool denoise_enabled = false;
    double frame_ms = 0.0;
    double pipeline_ms = 0.0;
    double render_ms = 0.0;
    double denoise_ms = 0.0;
    double download_ms = 0.0;
    double render_work_ms = 0.0;
    double denoise_work_ms = 0.0;
    double download_work_ms = 0.0;
    double image_write_ms = 0.0;
    double host_overhead_ms = 0.0;
    double fps = 0.0;
    std::vector<CameraStageSample> cameras;
};

struct AggregateStats {
    double avg = 0.0, p50 = 0.0, p95 = 0.0, p99 = 0.0, max = 0.0;
};

struct RunAggregate {
    AggregateStats frame_ms, pipeline_ms, render_ms, denoise_ms, download_ms, render_work_ms,
        denoise_work_ms, download_work_ms, image_write_ms, host_overhead_ms;
};

struct RunProvenance {
    std::string captured_at_utc, project_version, source_revision;
    bool source_dirty = false;
    std::string source_scope, source_state_sha256, build_configuration, cxx_compiler;
};

// The reference's environment keys are kept (so its consumers parse the
// report); they carry the HIP / ROCm equivalents on this platform.
struct RunEnvironment {
    std::string operating_system, architecture;
    int gpu_device = 0;
    std::string gpu_name, compute_capability, nvidia_driver_version;
    int cuda_driver_api_version = 0;
    std::string cuda_driver_api_version_text;
    int cuda_runtime_version = 0;
    std::string cuda_runtime_version_text;
    int optix_version = 0;
    std::string optix_version_text;
};

struct GpuMemoryReport {
    std::string scope = "hip_device_global";
    std::uint64_t total_bytes = 0, baseline_used_bytes = 0, prepared_used_bytes = 0,
                  peak_used_bytes = 0, peak_delta_bytes = 0, final_used_bytes = 0;
};

struct GpuSchedulingReport {
    int persistent_worker_count = 0;
    std::uint64_t worker_start_count = 0, task_submission_count = 0,
                  launch_parameter_allocation_count = 0, launch_parameter_upload_count = 0;
    std::string acceleration_update_kind;
    double acceleration_update_ms = 0.0;
    int acceleration_node_count = 0, acceleration_reference_count = 0,
        acceleration_prototype_count = 0, acceleration_instance_count = 0,
        acceleration_instanced_primitive_count = 0;
    std::uint64_t acceleration_generation = 0;
};

struct RunReport {
    int schema_version = 3;
    RunProvenance provenance;
    RunEnvironment environment;
    GpuMemoryReport gpu_memory;
    GpuSchedulingReport gpu_scheduling;
    std::string scene, profile;
    int camera_count = 0, width = 0, height = 0, frames_requested = 0, warmup_frames = 0;
    std::uint32_t random_seed = 0;
    int samples_per_pixel = 0, max_bounces = 0;
    bool denoise_enabled = false, image_write_enabled = true;
    std::vector<FrameStageSample> frames;
    RunAggregate aggregate;
};

// nearest-rank statistics (p = values[ceil(p n) - 1] of the sorted values)
AggregateStats compute_stats(std::vector<double> values);
RunAggregate compute_aggregate(const std::vector<FrameStageSample>& frames);
void write_csv(const RunReport& report, const std::filesystem::path& path);
void write_json(const RunReport& report, const std::filesystem::path& path);
void write_artifact_manifest(const RunReport& report,
    const std::vector<std::filesystem::path>& artifacts, const std::filesystem::path& path);
std::string fnv1a64_hex(const std::string& bytes);

}  // namespace profiling
}  // namespace rt_amd
