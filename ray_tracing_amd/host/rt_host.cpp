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
           && a.restir_min_analytic_lights == b.restir_min_analytic_lights;
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
    r.pixel_jitter = 0;
    r.restir_spatial_radius = 0;
    return r;
}

void throw_on_error(rt_status s) {
    if (s == RT_OK) return;
    const std::string msg = rt_last_error();
    if (s == RT_ERROR_INVALID_ARGUMENT || s == RT_ERROR_UNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// ----------------------------------------------------------------- Renderer
Renderer::Renderer(int ordinal) { throw_on_error(rt_device_create(ordinal, &device_)); }

Renderer::~Renderer() {
    if (renderer_) rt_renderer_destroy(renderer_);
    if (scene_) rt_scene_destroy(scene_);
    if (device_) rt_device_destroy(device_);
}

void Renderer::prepare_scene(const rt_scene_desc& scene) {
    if (renderer_) {
        rt_renderer_destroy(renderer_);
        renderer_ = nullptr;
    }
    if (scene_) {
        rt_scene_destroy(scene_);
        scene_ = nullptr;
    }
    throw_on_error(rt_scene_upload(device_, &scene, &scene_));
    throw_on_error(rt_bvh_build(scene_, &bvh_));
    throw_on_error(rt_renderer_create(device_, scene_, &renderer_));
}

void Renderer::reset_accumulation() {
    if (renderer_) throw_on_error(rt_reset_accumulation(renderer_));
}

void Renderer::reset_sequence(uint32_t seed) {
    if (renderer_) throw_on_error(rt_reset_sequence(renderer_, seed));
}

rt_frame_timing Renderer::render_frame(const rt_camera_rig& rig, const RenderProfile& profile,
    uint32_t sample_stream) {
    if (!renderer_) throw std::runtime_error("render_prepared_device requires prepare_scene() first");
    const rt_profile p = to_abi(profile);
    rt_frame_timing t{};
    throw_on_error(rt_render_frame(renderer_, &rig, &p, sample_stream, &t));
    return t;
}

RadianceFrame Renderer::download(int camera_index) const {
    rt_device_frame_view v{};
    throw_on_error(rt_frame_device_view(renderer_, camera_index, &v));
    RadianceFrame f;
    f.width = v.width;
    f.height = v.height;
    const size_t n = (size_t)v.width * (size_t)v.height;
    f.beauty_rgba.resize(4 * n);
    f.normal_rgba.resize(4 * n);
    f.albedo_rgba.resize(4 * n);
    f.depth.resize(n);
    throw_on_error(rt_frame_download(renderer_, camera_index, f.beauty_rgba.data(),
        f.normal_rgba.data(), f.albedo_rgba.data(), f.depth.data()));
    f.average_luminance = average_luminance(f.beauty_rgba);
    return f;
}

RestirDiagnostics Renderer::restir_diagnostics(int camera_index) const {
    rt_restir_stats s{};
    throw_on_error(rt_restir_diagnostics(renderer_, camera_index, &s));
    return RestirDiagnostics{s.valid_reservoirs, s.temporal_candidates, s.spatial_candidates,
        s.bias_correction_sources};
}

namespace {
int checked_pool_size(int n) {
    if (n < 1 || n > 4) throw std::invalid_argument("RendererPool requires renderer_count in [1, 4]");
    return n;
}
}  // namespace

// the count is validated before the device is opened
RendererPool::RendererPool(int renderer_count, int ordinal)
    : count_(checked_pool_size(renderer_count)), renderer_(ordinal) {}

std::vector<CameraRenderResult> RendererPool::render_frame(const rt_camera_rig& rig,
    const RenderProfile& profile, uint32_t sample_stream) {
    if (rig.active_count > count_)
        throw std::invalid_argument("camera count exceeds the renderer pool size");
    const rt_frame_timing t = renderer_.render_frame(rig, profile, sample_stream);
    std::vector<CameraRenderResult> out;
    for (int c = 0; c < rig.active_count; ++c) {
        CameraRenderResult r;
        r.camera_index = c;
        const auto d0 = std::chrono::steady_clock::now();
        r.frame = renderer_.download(c);
        r.download_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - d0).count();
        r.render_ms = t.render_ms - t.denoise_ms;
        r.denoise_ms = t.denoise_ms;
        out.push_back(std::move(r));
    }
    return out;
}

// ---------------------------------------------------------------- profiling
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
