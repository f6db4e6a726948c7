// rt_host.hpp -- C++ host API over the C ABI (include/rt/rt_hip.h), shaped
// like the reference's realtime host API so its callers can switch over:
//   RenderProfile + presets        (src/realtime/render_profile.h:10-64)
//   Renderer                       (OptixRenderer, src/realtime/gpu/optix_renderer.h:69-128)
//   RendererPool                   (src/realtime/gpu/renderer_pool.h:30-53)
//   profiling::RunReport + writers (src/realtime/profiling/benchmark_report.h:10-147)
// Errors are thrown as std::invalid_argument / std::runtime_error with the
// library's message, as the reference throws them.
#pragma once

#include <cstdint>
#include <filesystem>
#include <optional>
#include <stdexcept>
#include <string>
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

    static RenderProfile quality();
    static RenderProfile balanced();
    static RenderProfile realtime();
    static RenderProfile realtime_default() { return balanced(); }
};

std::optional<RenderProfile> render_profile_from_name(const std::string& name);
std::string render_profile_name(const RenderProfile& profile);
rt_profile to_abi(const RenderProfile& p);

void throw_on_error(rt_status s);

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

struct CameraRenderResult {
    int camera_index = 0;
    RadianceFrame frame;
    double render_ms = 0.0;
    double denoise_ms = 0.0;
    double download_ms = 0.0;
};

struct RestirDiagnostics {
    uint64_t valid_reservoirs = 0;
    uint64_t temporal_candidates = 0;
    uint64_t spatial_candidates = 0;
    uint64_t bias_correction_sources = 0;
};

// One device, one prepared scene, one renderer tracing all cameras of a rig
// in a single launch sequence.
class Renderer {
public:
    explicit Renderer(int ordinal = 0);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    void prepare_scene(const rt_scene_desc& scene);  // upload + BVH build
    bool has_scene() const { return renderer_ != nullptr; }
    void reset_accumulation();
    void reset_sequence(uint32_t seed);
    // renders every active camera of `rig`; returns the device-timed frame
    rt_frame_timing render_frame(const rt_camera_rig& rig, const RenderProfile& profile,
        uint32_t sample_stream);
    RadianceFrame download(int camera_index) const;
    RestirDiagnostics restir_diagnostics(int camera_index) const;
    rt_bvh_stats acceleration_stats() const { return bvh_; }
    rt_renderer* handle() const { return renderer_; }
    rt_device* device() const { return device_; }

private:
    rt_device* device_ = nullptr;
    rt_scene* scene_ = nullptr;
    rt_renderer* renderer_ = nullptr;
    rt_bvh_stats bvh_{};
};

// RendererPool(int 1..4): the reference keeps one renderer and worker per
// camera; here one renderer traces the whole rig, and the pool keeps the
// reference's interface and its camera-count validation.
class RendererPool {
public:
    explicit RendererPool(int renderer_count, int ordinal = 0);
    void prepare_scene(const rt_scene_desc& scene) { renderer_.prepare_scene(scene); }
    void reset_accumulation() { renderer_.reset_accumulation(); }
    void reset_sequence(uint32_t seed) { renderer_.reset_sequence(seed); }
    // results sorted by camera_index; render_ms is the shared frame time
    std::vector<CameraRenderResult> render_frame(const rt_camera_rig& rig,
        const RenderProfile& profile, uint32_t sample_stream);
    int renderer_count() const { return count_; }
    Renderer& renderer() { return renderer_; }

private:
    int count_;
    Renderer renderer_;
};

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
