// test_report -- known answers of the benchmark report (the reference's
// tests/test_realtime_benchmark_report.cpp fixture values): nearest-rank
// aggregates, CSV quoting, JSON keys, manifest counts, FNV-1a64("hello").
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iterator>

#include "rt_host.hpp"

using namespace rt_amd;

static int failures = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}
static std::string slurp(const std::filesystem::path& p) {
    std::ifstream in(p);
    return std::string((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

int main(int argc, char** argv) {
    const std::filesystem::path dir = argc > 1 ? argv[1] : std::filesystem::temp_directory_path() / "rt_amd_report_test";
    std::filesystem::remove_all(dir);
    std::filesystem::create_directories(dir);
    const std::string tricky = "rt,\"x\"\\path";
    profiling::RunReport r;
    r.scene = "final_room";
    r.profile = tricky;
    r.camera_count = 2;
    r.width = 640;
    r.height = 480;
    r.frames_requested = 2;
    r.warmup_frames = 3;
    r.random_seed = 42;
    r.samples_per_pixel = 1;
    r.max_bounces = 2;
    r.denoise_enabled = true;
    r.image_write_enabled = false;
    r.gpu_memory.peak_used_bytes = 300;
    r.gpu_memory.peak_delta_bytes = 200;
    r.gpu_scheduling.persistent_worker_count = 2;
    r.gpu_scheduling.acceleration_update_kind = "rebuild";
    profiling::FrameStageSample f0;
    f0.frame_index = 0; f0.sample_stream = 45; f0.camera_count = 2; f0.profile = tricky;
    f0.width = 640; f0.height = 480; f0.samples_per_pixel = 1; f0.max_bounces = 2; f0.denoise_enabled = true;
    f0.frame_ms = 10.0; f0.pipeline_ms = 9.0; f0.render_ms = 3.0; f0.denoise_ms = 1.0; f0.download_ms = 0.5;
    f0.render_work_ms = 6.0; f0.denoise_work_ms = 2.0; f0.download_work_ms = 1.0; f0.image_write_ms = 0.5;
    f0.host_overhead_ms = 0.5; f0.fps = 100.0;
    f0.cameras = {{0, 3.0, 1.0, 0.5, 0.12}, {1, 3.0, 1.0, 0.5, 0.13}};
    profiling::FrameStageSample f1 = f0;
    f1.frame_index = 1; f1.sample_stream = 46; f1.frame_ms = 14.0; f1.pipeline_ms = 13.0; f1.render_ms = 4.0;
    f1.denoise_ms = 1.5; f1.download_ms = 0.75; f1.render_work_ms = 8.0; f1.denoise_work_ms = 3.0;
    f1.download_work_ms = 1.5; f1.image_write_ms = 1.0; f1.host_overhead_ms = 0.0; f1.fps = 71.428571;
    f1.cameras = {{0, 4.0, 1.5, 0.75, 0.14}, {1, 4.0, 1.5, 0.75, 0.15}};
    r.frames = {f0, f1};
    r.aggregate = profiling::compute_aggregate(r.frames);
    expect(std::fabs(r.aggregate.frame_ms.avg - 12.0) < 1e-12, "frame avg 12");
    expect(std::fabs(r.aggregate.frame_ms.p50 - 10.0) < 1e-12, "frame p50 10");
    expect(std::fabs(r.aggregate.frame_ms.p95 - 14.0) < 1e-12, "frame p95 14");
    expect(std::fabs(r.aggregate.frame_ms.p99 - 14.0) < 1e-12, "frame p99 14");
    expect(std::fabs(r.aggregate.pipeline_ms.avg - 11.0) < 1e-12, "pipeline avg 11");
    expect(std::fabs(r.aggregate.denoise_ms.max - 1.5) < 1e-12, "denoise max 1.5");
    expect(std::fabs(r.aggregate.denoise_work_ms.max - 3.0) < 1e-12, "denoise work max 3");
    expect(std::fabs(r.aggregate.host_overhead_ms.avg - 0.25) < 1e-12, "host overhead avg 0.25");
    const auto csv = dir / "benchmark_frames.csv", json = dir / "benchmark_summary.json",
               man = dir / "benchmark_manifest.json";
    profiling::write_csv(r, csv);
    profiling::write_json(r, json);
    profiling::write_artifact_manifest(r, {csv, json}, man);
    const std::string c = slurp(csv), j = slurp(json), m = slurp(man);
    expect(c.find("frame_index,sample_stream,camera_count,profile,width,height") != std::string::npos, "csv header");
    expect(c.find("0,45,2,\"rt,\"\"x\"\"\\path\",640,480") != std::string::npos, "csv first row quoting");
    expect(j.find("\"schema_version\": 3") != std::string::npos, "json schema 3");
    expect(j.find("\"scene\": \"final_room\"") != std::string::npos, "json scene");
    expect(j.find("\"profile\": \"rt,\\\"x\\\"\\\\path\"") != std::string::npos, "json escaping");
    expect(j.find("\"p99\": 14") != std::string::npos, "json p99");
    expect(j.find("\"warmup_frames\": 3") != std::string::npos, "json warmup");
    expect(j.find("\"random_seed\": 42") != std::string::npos, "json seed");
    expect(j.find("\"sample_stream\": 45") != std::string::npos, "json sample stream");
    expect(j.find("\"peak_used_bytes\": 300") != std::string::npos, "json peak memory");
    expect(j.find("\"peak_delta_bytes\": 200") != std::string::npos, "json memory delta");
    expect(j.find("\"persistent_worker_count\": 2") != std::string::npos, "json workers");
    expect(j.find("\"acceleration_update_kind\": \"rebuild\"") != std::string::npos, "json acceleration kind");
    expect(j.find("\"host_overhead_ms\": 0.5") != std::string::npos, "json host overhead");
    expect(j.find("\"fps\": 100") != std::string::npos, "json fps");
    expect(j.find("\"per-camera\"") != std::string::npos && j.find("\"camera_index\": 1") != std::string::npos,
        "json per-camera");
    expect(m.find("\"report_schema_version\": 3") != std::string::npos, "manifest schema");
    expect(m.find("\"measured_frame_count\": 2") != std::string::npos, "manifest frames");
    expect(m.find("\"per_camera_record_count\": 4") != std::string::npos, "manifest per-camera count");
    expect(m.find("\"filename\": \"benchmark_frames.csv\"") != std::string::npos, "manifest csv");
    expect(m.find("\"digest_algorithm\": \"fnv1a64\"") != std::string::npos, "manifest digest algorithm");
    expect(profiling::fnv1a64_hex("hello") == "a430d84680aabd0b", "FNV-1a64(hello)");
    bool threw = false;
    try {
        profiling::write_csv(r, dir / "missing-parent" / "x.csv");
    } catch (...) { threw = true; }
    expect(threw, "csv write failure throws");
    threw = false;
    try {
        profiling::write_artifact_manifest(r, {csv, dir / "missing.json"}, dir / "m2.json");
    } catch (...) { threw = true; }
    expect(threw, "manifest of a missing artifact throws");
    threw = false;
    try {
        RendererPool pool(5);
    } catch (const std::runtime_error& e) {
        threw = std::string(e.what()) == "RendererPool requires renderer_count in [1, 4]";
    } catch (...) {}
    expect(threw, "RendererPool(5) throws the reference's message");
    expect(render_profile_name(RenderProfile::realtime()) == "realtime", "profile name realtime");
    expect(render_profile_name(RenderProfile::realtime_default()) == "balanced", "default is balanced");
    expect(!render_profile_from_name("fast").has_value(), "unknown profile name");
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
