// render_realtime -- benchmark CLI of the MI355X path-tracing core, with the
// flags, stdout lines and artifacts of the reference's utils/render_realtime.cpp
// (benchmark_frames.csv, benchmark_summary.json schema 3, benchmark_manifest.json).
//
//   render_realtime --scene cornell_mesh --profile scene --frames 20 --warmup-frames 3
//                   --camera-count 4 --output-dir out --skip-image-write
// Extensions: --denoiser-weights <params.npy> (learned denoiser), --graph 0|1,
// --autotune.  Scenes: the built-in ids below, or a YAML scene file: --scene
// <file.yaml>, --scene <id> with --scene-dir <dir> (or RT_SCENE_DIR) holding
// <id>/scene.yaml or <id>.yaml, or --scene-bundle <file> (written by
// `python -m ray_tracing_amd.scene_yaml <file.yaml> --bundle <file>`).  A YAML
// file is flattened by that same ingest, run as a child process before the
// device is touched, into a temporary bundle.
#include <spawn.h>
#include <sys/utsname.h>
#include <sys/wait.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "rt_host.hpp"

namespace {

const char* kVersion = "ray_tracing_amd 0.1";
#ifndef RT_SOURCE_REVISION
#define RT_SOURCE_REVISION "unversioned"
#endif
#ifndef RT_SOURCE_STATE_SHA256
#define RT_SOURCE_STATE_SHA256 "0"
#endif

struct Args {
    int camera_count = 4;
    int frames = 10;
    int warmup_frames = 2;
    unsigned seed = 0;
    std::string output_dir = "render_realtime_out";
    std::string scene = "smoke";
    std::string profile;
    bool skip_image_write = false;
    std::string denoiser_weights;
    int graph = -1;
    bool autotune = false;
    std::string scene_dir;
    std::string scene_bundle;
    bool print_scene = false;  // describe the resolved scene and exit before the device is used
};

// the YAML scene file `--scene` names: a path to a .yaml / .yml file, or an
// id found in the scene directory; empty when there is none
std::filesystem::path scene_yaml_path(const Args& a) {
    namespace fs = std::filesystem;
    const fs::path direct(a.scene);
    const std::string ext = direct.extension().string();
    if ((ext == ".yaml" || ext == ".yml") && fs::is_regular_file(direct)) return direct;
    std::string dir = a.scene_dir;
    if (dir.empty()) {
        const char* env = std::getenv("RT_SCENE_DIR");
        if (env != nullptr) dir = env;
    }
    if (dir.empty() || a.scene.empty() || a.scene.find('/') != std::string::npos) return {};
    for (const fs::path& p : {fs::path(dir) / a.scene / "scene.yaml", fs::path(dir) / (a.scene + ".yaml")})
        if (fs::is_regular_file(p)) return p;
    return {};
}

// runs the package's YAML ingest (python3 -m ray_tracing_amd.scene_yaml) as a
// child process that writes a scene bundle; the package root is the parent of
// the directory that holds this executable
void convert_yaml_scene(const std::filesystem::path& yaml, const std::filesystem::path& bundle) {
    namespace fs = std::filesystem;
    const fs::path root = fs::read_symlink("/proc/self/exe").parent_path().parent_path();
    if (!fs::is_regular_file(root / "ray_tracing_amd" / "scene_yaml.py"))
        throw std::runtime_error("the YAML scene ingest (ray_tracing_amd/scene_yaml.py) is not beside " + root.string());
    const char* old = std::getenv("PYTHONPATH");
    const std::string pp = root.string() + (old != nullptr && *old ? ":" + std::string(old) : "");
    setenv("PYTHONPATH", pp.c_str(), 1);
    const std::string ys = yaml.string(), bs = bundle.string();
    const char* argv[] = {"python3", "-m", "ray_tracing_amd.scene_yaml", ys.c_str(), "--bundle", bs.c_str(), nullptr};
    pid_t pid = 0;
    const int rc = posix_spawnp(&pid, "python3", nullptr, nullptr, const_cast<char* const*>(argv), environ);
    if (rc != 0) throw std::runtime_error("cannot start python3 for the YAML scene ingest: " + std::string(std::strerror(rc)));
    int status = 0;
    if (waitpid(pid, &status, 0) != pid || !WIFEXITED(status) || WEXITSTATUS(status) != 0)
        throw std::runtime_error("the YAML scene ingest failed on " + ys);
}

int scene_config(const std::string& name) {
    if (name == "random_spheres") return 0;
    if (name == "cornell_mesh") return 1;
    if (name == "instanced_mesh_restir") return 2;
    if (name == "radiance_cache_realtime") return 3;
    if (name == "smoke") return 5;
    if (name == "final_room") return 6;
    return -1;
}

bool parse(int argc, char** argv, Args& a) {
    std::map<std::string, std::string> kv;
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        if (k == "--skip-image-write" || k == "--autotune" || k == "--print-scene") {
            kv[k] = "1";
            continue;
        }
        if (k.rfind("--", 0) != 0 || i + 1 >= argc) {
            std::fprintf(stderr, "unknown or incomplete argument: %s\n", k.c_str());
            return false;
        }
        kv[k] = argv[++i];
    }
    try {
        for (const auto& [k, v] : kv) {
            if (k == "--camera-count") a.camera_count = std::stoi(v);
            else if (k == "--frames") a.frames = std::stoi(v);
            else if (k == "--warmup-frames") a.warmup_frames = std::stoi(v);
            else if (k == "--seed") a.seed = (unsigned)std::stoul(v);
            else if (k == "--output-dir") a.output_dir = v;
            else if (k == "--scene") a.scene = v;
            else if (k == "--scene-dir") a.scene_dir = v;
            else if (k == "--scene-bundle") a.scene_bundle = v;
            else if (k == "--profile") a.profile = v;
            else if (k == "--skip-image-write") a.skip_image_write = true;
            else if (k == "--denoiser-weights") a.denoiser_weights = v;
            else if (k == "--graph") a.graph = std::stoi(v);
            else if (k == "--autotune") a.autotune = true;
            else if (k == "--print-scene") a.print_scene = true;
            else {
                std::fprintf(stderr, "unknown argument: %s\n", k.c_str());
                return false;
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return false;
    }
    return true;
}

// float32 1-D .npy (little endian) as written by numpy.save
std::vector<float> load_npy(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    char magic[6];
    in.read(magic, 6);
    if (std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error(path + " is not a .npy file");
    unsigned char ver[2];
    in.read(reinterpret_cast<char*>(ver), 2);
    uint32_t hlen = 0;
    if (ver[0] == 1) {
        uint16_t h16 = 0;
        in.read(reinterpret_cast<char*>(&h16), 2);
        hlen = h16;
    } else {
        in.read(reinterpret_cast<char*>(&hlen), 4);
    }
    std::string header(hlen, ' ');
    in.read(header.data(), hlen);
    if (header.find("'<f4'") == std::string::npos || header.find("'fortran_order': False") == std::string::npos)
        throw std::runtime_error(path + ": expected a C-ordered float32 array");
    std::vector<float> v;
    float x;
    while (in.read(reinterpret_cast<char*>(&x), 4)) v.push_back(x);
    return v;
}

// PNG (RFC 2083) of the display-encoded beauty image: 8-bit RGB, filter 0
// per row, zlib stream of stored (uncompressed) deflate blocks, CRC-32 per
// chunk and Adler-32 over the raw rows.
uint32_t crc32_update(uint32_t crc, const unsigned char* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return crc;
}

void put_be32(std::string& s, uint32_t v) {
    for (int k = 3; k >= 0; --k) s.push_back((char)((v >> (8 * k)) & 0xFFu));
}

void png_chunk(std::ofstream& out, const char* type, const std::string& data) {
    std::string len;
    put_be32(len, (uint32_t)data.size());
    out << len;
    std::string body = std::string(type, 4) + data;
    out << body;
    std::string crc;
    put_be32(crc, crc32_update(0xFFFFFFFFu, reinterpret_cast<const unsigned char*>(body.data()), body.size())
        ^ 0xFFFFFFFFu);
    out << crc;
}

void write_png(const std::string& path, const rt_amd::RadianceFrame& f) {
    std::string raw;
    raw.reserve((size_t)f.height * (1 + 3 * (size_t)f.width));
    for (int y = 0; y < f.height; ++y) {
        raw.push_back(0);
        for (int x = 0; x < f.width; ++x) {
            const size_t i = (size_t)y * (size_t)f.width + (size_t)x;
            for (int c = 0; c < 3; ++c) {
                float v = f.beauty_rgba[4 * i + c];
                v = v <= 0.0f ? 0.0f : (v >= 1.0f ? 1.0f : v);
                raw.push_back((char)(unsigned char)(255.0f * std::pow(v, 1.0f / 2.2f) + 0.5f));
            }
        }
    }
    std::string z = {(char)0x78, (char)0x01};
    uint32_t a = 1, b = 0;
    for (unsigned char ch : raw) {
        a = (a + ch) % 65521u;
        b = (b + a) % 65521u;
    }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + n >= raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((char)(n & 0xFFu));
        z.push_back((char)(n >> 8));
        z.push_back((char)(~n & 0xFFu));
        z.push_back((char)((~n >> 8) & 0xFFu));
        z.append(raw, off, n);
        if (last) break;
    }
    put_be32(z, (b << 16) | a);
    std::string ihdr;
    put_be32(ihdr, (uint32_t)f.width);
    put_be32(ihdr, (uint32_t)f.height);
    ihdr += std::string("\x08\x02\x00\x00\x00", 5);
    std::ofstream out(path, std::ios::binary);
    out << std::string("\x89PNG\r\n\x1a\n", 8);
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", "");
}

std::string utc_now() {
    char buf[64];
    std::time_t t = std::time(nullptr);
    std::strftime(buf, sizeof(buf), "%Y-%m-%dT%H:%M:%SZ", std::gmtime(&t));
    return buf;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    if (!parse(argc, argv, a)) return EXIT_FAILURE;
    if (a.camera_count < 1 || a.camera_count > 4) {
        std::fprintf(stderr, "--camera-count must be in [1, 4]\n");
        return EXIT_FAILURE;
    }
    if (a.frames < 1) {
        std::fprintf(stderr, "--frames must be >= 1\n");
        return EXIT_FAILURE;
    }
    if (a.warmup_frames < 0) {
        std::fprintf(stderr, "--warmup-frames must be >= 0\n");
        return EXIT_FAILURE;
    }
    if ((uint64_t)a.seed + (uint64_t)a.warmup_frames + (uint64_t)a.frames - 1u
        > std::numeric_limits<uint32_t>::max()) {
        std::fprintf(stderr, "--seed + --warmup-frames + --frames exceeds uint32 sample streams\n");
        return EXIT_FAILURE;
    }
    const int config = a.scene_bundle.empty() ? scene_config(a.scene) : -1;
    rt_generated_scene* gen = nullptr;
    rt_amd::SceneBundle bundle;
    if (config >= 0) {
        rt_amd::throw_on_error(rt_scene_generate(config, 0, &gen));
    } else {
        std::filesystem::path bundle_path = a.scene_bundle;
        bool temporary = false;
        if (bundle_path.empty()) {
            const std::filesystem::path yaml = scene_yaml_path(a);
            if (yaml.empty()) {
                std::fprintf(stderr, "--scene must reference a registered realtime scene\n");
                return EXIT_FAILURE;
            }
            bundle_path = std::filesystem::temp_directory_path()
                / ("rt_scene_" + std::to_string((long)getpid()) + ".rtsb");
            temporary = true;
            try {
                convert_yaml_scene(yaml, bundle_path);
            } catch (const std::exception& e) {
                std::fprintf(stderr, "%s\n", e.what());
                return EXIT_FAILURE;
            }
        }
        try {
            bundle = rt_amd::load_scene_bundle(bundle_path);
        } catch (const std::exception& e) {
            if (temporary) std::filesystem::remove(bundle_path);
            std::fprintf(stderr, "%s\n", e.what());
            return EXIT_FAILURE;
        }
        if (temporary) std::filesystem::remove(bundle_path);
        if (!bundle.name.empty()) a.scene = bundle.name;  // the scene id, not the file path
        // the generated-scene shape the rest of the run reads
        gen = new rt_generated_scene{};
        gen->scene = bundle.desc;
        gen->camera_count = bundle.rig.active_count;
        for (int c = 0; c < bundle.rig.active_count; ++c) gen->cameras[c] = bundle.rig.cameras[c];
        gen->profile = bundle.profile;
    }
    auto free_scene = [&]() {
        if (config >= 0) rt_scene_generated_free(gen);
        else delete gen;
    };
    if (a.print_scene) {
        const rt_scene_desc& d = gen->scene;
        std::printf("scene %s: %d spheres, %d quads, %d triangles, %d materials, %d openpbr, "
                    "%d analytic lights, %d textures, %d texels, %d media, %d cameras",
            a.scene.c_str(), d.sphere_count, d.quad_count, d.triangle_count, d.material_count,
            d.openpbr_count, d.analytic_light_count, d.texture_count, d.image_texel_count, d.medium_count,
            gen->camera_count);
        for (int c = 0; c < gen->camera_count; ++c)
            std::printf("%s%dx%d", c ? ", " : " ", gen->cameras[c].width, gen->cameras[c].height);
        std::printf("; spp %d, depth %d\n", gen->profile.samples_per_pixel, gen->profile.max_bounces);
        free_scene();
        return EXIT_SUCCESS;
    }
    rt_amd::RenderProfile profile = rt_amd::RenderProfile::realtime_default();
    std::string profile_name = rt_amd::render_profile_name(profile);
    rt_profile abi_profile{};
    bool scene_profile = false;
    if (!a.profile.empty()) {
        if (a.profile == "scene") {
            scene_profile = true;
            profile_name = "scene";
        } else {
            const auto p = rt_amd::render_profile_from_name(a.profile);
            if (!p) {
                std::fprintf(stderr, "--profile must be one of: quality, balanced, realtime\n");
                return EXIT_FAILURE;
            }
            profile = *p;
            profile_name = a.profile;
        }
    }
    abi_profile = scene_profile ? gen->profile : rt_amd::to_abi(profile);
    std::filesystem::create_directories(a.output_dir);
    try {
        rt_camera_rig rig{};
        rig.active_count = std::min(a.camera_count, gen->camera_count);
        for (int c = 0; c < rig.active_count; ++c) rig.cameras[c] = gen->cameras[c];
        const rt_amd::RenderProfile run_profile = rt_amd::from_abi(abi_profile);
        // one pool slot per camera, as the reference's CLI sizes its pool
        // (render_realtime.cpp:289-291)
        rt_amd::RendererPool pool(rig.active_count, 0);
        rt_amd::Renderer& renderer = pool.renderer();
        char name[128] = {0};
        int cus = 0;
        uint64_t total_b = 0, free_b = 0;
        rt_amd::throw_on_error(rt_device_info(renderer.device(), name, sizeof(name), &cus, &total_b, &free_b));
        const uint64_t baseline_used = total_b - free_b;
        pool.prepare_scene(gen->scene);
        rt_nn* nn = nullptr;
        if (!a.denoiser_weights.empty()) {
            const std::vector<float> w = load_npy(a.denoiser_weights);
            rt_amd::throw_on_error(rt_nn_create(0, &nn));
            rt_amd::throw_on_error(rt_nn_set_params(nn, w.data(), (int32_t)w.size()));
            rt_amd::throw_on_error(rt_renderer_set_denoiser(renderer.handle(), nn));
        }
        if (a.graph >= 0) {
            rt_tuning t{};
            t.use_graph = a.graph;
            t.count_traversal = -1;
            rt_amd::throw_on_error(rt_renderer_set_tuning(renderer.handle(), &t));
        }
        if (a.autotune) {
            rt_autotune_result ar{};
            rt_amd::throw_on_error(rt_renderer_autotune(renderer.handle(), &rig, &abi_profile, 2, &ar));
            std::printf("autotune trials=%d baseline_ms=%.3f best_ms=%.3f block=%d blocks_per_cu=%d "
                        "refill=%d paths=%d graph=%d\n",
                ar.trials, ar.baseline_ms, ar.best_ms, ar.best.trace_block, ar.best.trace_blocks_per_cu,
                ar.best.trace_refill, ar.best.wavefront_paths, ar.best.use_graph);
        }
        // the run's frames start at the seed's sample stream (after any
        // autotuning trial frames)
        pool.reset_sequence(a.seed);
        rt_device_info(renderer.device(), name, sizeof(name), &cus, &total_b, &free_b);
        const uint64_t prepared_used = total_b - free_b;
        uint64_t peak_used = std::max(baseline_used, prepared_used);

        rt_amd::profiling::RunReport report;
        report.provenance.captured_at_utc = utc_now();
        report.provenance.project_version = kVersion;
        report.provenance.source_scope = "ray_tracing_amd,include,oracle";
        report.provenance.source_revision = RT_SOURCE_REVISION;
        report.provenance.source_state_sha256 = RT_SOURCE_STATE_SHA256;
        report.provenance.build_configuration = "Release";
        report.provenance.cxx_compiler = __VERSION__;
        struct utsname un;
        if (uname(&un) == 0) {
            report.environment.operating_system = std::string(un.sysname) + " " + un.release;
            report.environment.architecture = un.machine;
        }
        report.environment.gpu_name = name;
        report.environment.compute_capability = name;
        report.gpu_memory.total_bytes = total_b;
        report.gpu_memory.baseline_used_bytes = baseline_used;
        report.gpu_memory.prepared_used_bytes = prepared_used;
        const rt_bvh_stats bs = renderer.acceleration_stats();
        report.scene = a.scene;
        report.profile = profile_name;
        report.camera_count = rig.active_count;
        report.width = rig.cameras[0].width;
        report.height = rig.cameras[0].height;
        report.frames_requested = a.frames;
        report.warmup_frames = a.warmup_frames;
        report.random_seed = a.seed;
        report.samples_per_pixel = abi_profile.samples_per_pixel;
        report.max_bounces = abi_profile.max_bounces;
        report.denoise_enabled = abi_profile.enable_denoise != 0;
        report.image_write_enabled = !a.skip_image_write;

        for (int w = 0; w < a.warmup_frames; ++w) {
            const auto w0 = std::chrono::steady_clock::now();
            const std::vector<rt_amd::CameraRenderResult> res = pool.render_frame(rig, run_profile, rig.active_count);
            const double wms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
            std::printf("warmup=%d sample_stream=%u cameras=%zu pipeline_ms=%.3f\n", w, a.seed + w,
                res.size(), wms);
        }
        for (int f = 0; f < a.frames; ++f) {
            rt_amd::profiling::FrameStageSample s;
            s.frame_index = f;
            s.sample_stream = a.seed + (uint32_t)a.warmup_frames + (uint32_t)f;
            s.camera_count = rig.active_count;
            s.profile = profile_name;
            s.width = report.width;
            s.height = report.height;
            s.samples_per_pixel = report.samples_per_pixel;
            s.max_bounces = report.max_bounces;
            s.denoise_enabled = report.denoise_enabled;
            const auto t0 = std::chrono::steady_clock::now();
            const std::vector<rt_amd::CameraRenderResult> res = pool.render_frame(rig, run_profile, rig.active_count);
            const auto t1 = std::chrono::steady_clock::now();
            double lum_sum = 0.0, download_ms = 0.0;
            for (const rt_amd::CameraRenderResult& cr : res) {
                const rt_amd::RadianceFrame& fr = cr.profiled.frame;
                download_ms += cr.profiled.timing.download_ms;
                rt_amd::profiling::CameraStageSample cs;
                cs.camera_index = cr.camera_index;
                cs.render_ms = cr.profiled.timing.render_ms;
                cs.denoise_ms = cr.profiled.timing.denoise_ms;
                cs.download_ms = cr.profiled.timing.download_ms;
                cs.average_luminance = fr.average_luminance;
                lum_sum += fr.average_luminance;
                s.cameras.push_back(cs);
                if (!a.skip_image_write) {
                    const auto w0 = std::chrono::steady_clock::now();
                    write_png(a.output_dir + "/frame_" + std::to_string(f) + "_cam" + std::to_string(cr.camera_index) + ".png", fr);
                    s.image_write_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
                }
            }
            const auto t2 = std::chrono::steady_clock::now();
            // all cameras in one launch sequence: the shared render / denoise times
            s.render_ms = res.empty() ? 0.0 : res[0].profiled.timing.render_ms;
            s.denoise_ms = res.empty() ? 0.0 : res[0].profiled.timing.denoise_ms;
            s.download_ms = download_ms;
            s.render_work_ms = s.render_ms;
            s.denoise_work_ms = s.denoise_ms;
            s.download_work_ms = download_ms;
            s.pipeline_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            s.frame_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
            s.host_overhead_ms = std::max(0.0, s.frame_ms - s.pipeline_ms - s.image_write_ms);
            s.fps = s.frame_ms > 0.0 ? 1000.0 / s.frame_ms : 0.0;
            rt_device_info(renderer.device(), name, sizeof(name), &cus, &total_b, &free_b);
            peak_used = std::max<uint64_t>(peak_used, total_b - free_b);
            std::printf("frame=%d sample_stream=%u cameras=%d avg_luminance=%.6f pipeline_ms=%.3f "
                        "render_ms=%.3f denoise_ms=%.3f download_ms=%.3f render_work_ms=%.3f "
                        "denoise_work_ms=%.3f download_work_ms=%.3f image_write_ms=%.3f "
                        "host_overhead_ms=%.3f frame_ms=%.3f\n",
                f, s.sample_stream, rig.active_count, lum_sum / rig.active_count, s.pipeline_ms,
                s.render_ms, s.denoise_ms, s.download_ms, s.render_work_ms, s.denoise_work_ms,
                s.download_work_ms, s.image_write_ms, s.host_overhead_ms, s.frame_ms);
            report.frames.push_back(std::move(s));
        }
        // GpuSchedulingReport from the pool's diagnostics (render_realtime.cpp:444-460)
        const rt_amd::RendererPoolDiagnostics pd = pool.diagnostics();
        report.gpu_scheduling.persistent_worker_count = pd.worker_count;
        report.gpu_scheduling.worker_start_count = pd.worker_start_count;
        report.gpu_scheduling.task_submission_count = pd.task_submission_count;
        report.gpu_scheduling.launch_parameter_allocation_count = pd.launch_parameter_allocation_count;
        report.gpu_scheduling.launch_parameter_upload_count = pd.launch_parameter_upload_count;
        report.gpu_scheduling.acceleration_update_kind = rt_amd::acceleration_update_kind_name(pd.acceleration.kind);
        report.gpu_scheduling.acceleration_update_ms = pd.acceleration.elapsed_ms;
        report.gpu_scheduling.acceleration_node_count = pd.acceleration.node_count;
        report.gpu_scheduling.acceleration_reference_count = pd.acceleration.primitive_reference_count;
        report.gpu_scheduling.acceleration_prototype_count = pd.acceleration.prototype_count;
        report.gpu_scheduling.acceleration_instance_count = pd.acceleration.instance_count;
        report.gpu_scheduling.acceleration_instanced_primitive_count = pd.acceleration.instanced_primitive_count;
        report.gpu_scheduling.acceleration_generation = pd.acceleration.generation;
        rt_device_info(renderer.device(), name, sizeof(name), &cus, &total_b, &free_b);
        report.gpu_memory.final_used_bytes = total_b - free_b;
        report.gpu_memory.peak_used_bytes = std::max<uint64_t>(peak_used, report.gpu_memory.final_used_bytes);
        report.gpu_memory.peak_delta_bytes =
            report.gpu_memory.peak_used_bytes >= baseline_used ? report.gpu_memory.peak_used_bytes - baseline_used : 0;
        report.aggregate = rt_amd::profiling::compute_aggregate(report.frames);
        const std::filesystem::path out = a.output_dir;
        rt_amd::profiling::write_csv(report, out / "benchmark_frames.csv");
        rt_amd::profiling::write_json(report, out / "benchmark_summary.json");
        rt_amd::profiling::write_artifact_manifest(report,
            {out / "benchmark_frames.csv", out / "benchmark_summary.json"}, out / "benchmark_manifest.json");
        const auto& ag = report.aggregate;
        std::printf("summary profile=%s warmup_frames=%d frames=%d seed=%u cameras=%d resolution=%dx%d spp=%d "
                    "max_bounces=%d denoise=%s avg_frame_ms=%.3f avg_denoise_ms=%.3f p95_frame_ms=%.3f "
                    "p99_frame_ms=%.3f fps=%.2f peak_gpu_memory_mib=%.2f peak_gpu_memory_delta_mib=%.2f "
                    "workers=%d/%llu launch_param_allocations=%llu launch_param_uploads=%llu "
                    "as_update=%s as_update_ms=%.3f as_nodes=%d as_refs=%d as_prototypes=%d "
                    "as_instances=%d instanced_primitives=%d "
                    "artifacts=benchmark_frames.csv,benchmark_summary.json,benchmark_manifest.json "
                    "output_dir=%s\n",
            profile_name.c_str(), a.warmup_frames, a.frames, a.seed, rig.active_count, report.width,
            report.height, report.samples_per_pixel, report.max_bounces,
            report.denoise_enabled ? "true" : "false", ag.frame_ms.avg, ag.denoise_ms.avg, ag.frame_ms.p95,
            ag.frame_ms.p99, 1000.0 / ag.frame_ms.avg,
            (double)report.gpu_memory.peak_used_bytes / (1024.0 * 1024.0),
            (double)report.gpu_memory.peak_delta_bytes / (1024.0 * 1024.0), pd.worker_count,
            (unsigned long long)pd.worker_start_count,
            (unsigned long long)pd.launch_parameter_allocation_count,
            (unsigned long long)pd.launch_parameter_upload_count,
            rt_amd::acceleration_update_kind_name(pd.acceleration.kind).c_str(), pd.acceleration.elapsed_ms,
            pd.acceleration.node_count, pd.acceleration.primitive_reference_count,
            pd.acceleration.prototype_count, pd.acceleration.instance_count,
            pd.acceleration.instanced_primitive_count, a.output_dir.c_str());
        (void)bs;
        if (nn) {
            rt_renderer_set_denoiser(renderer.handle(), nullptr);
            rt_nn_destroy(nn);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        free_scene();
        return EXIT_FAILURE;
    }
    free_scene();
    return EXIT_SUCCESS;
}
