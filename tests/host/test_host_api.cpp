// test_host_api -- the C++ host API's request validation and call shapes,
// without a GPU: the checks and messages of the reference's
// tests/test_render_request_validation.cpp:42-110, RendererPool's size
// check (tests/test_renderer_pool.cpp:83-84) and the signatures of
// RendererPool (renderer_pool.h:30-53) and OptixRenderer
// (optix_renderer.h:69-128) that a caller of the reference compiles against.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "rt_host.hpp"

using namespace rt_amd;

// RendererPool and Renderer are neither copyable nor movable (renderer_pool.h:35-38)
static_assert(!std::is_copy_constructible_v<RendererPool>);
static_assert(!std::is_copy_assignable_v<RendererPool>);
static_assert(!std::is_move_constructible_v<RendererPool>);
static_assert(!std::is_move_assignable_v<RendererPool>);
static_assert(!std::is_copy_constructible_v<Renderer>);
static_assert(!std::is_move_constructible_v<Renderer>);

// the reference's call shapes: a caller passing active_cameras gets the
// camera count, not a sample stream
static_assert(std::is_same_v<decltype(&RendererPool::render_frame),
    std::vector<CameraRenderResult> (RendererPool::*)(const rt_camera_rig&, const RenderProfile&, int)>);
static_assert(std::is_same_v<decltype(&RendererPool::render_device_frame),
    std::vector<CameraDeviceRenderResult> (RendererPool::*)(const rt_camera_rig&, const RenderProfile&, int)>);
static_assert(std::is_same_v<decltype(&RendererPool::diagnostics),
    RendererPoolDiagnostics (RendererPool::*)() const>);
static_assert(std::is_same_v<decltype(&RendererPool::reset_sequence), void (RendererPool::*)(std::uint32_t)>);
static_assert(std::is_same_v<decltype(&Renderer::render_prepared_radiance),
    ProfiledRadianceFrame (Renderer::*)(const rt_camera_rig&, const RenderProfile&, int)>);
static_assert(std::is_same_v<decltype(&Renderer::render_prepared_device),
    ProfiledDeviceRadianceFrame (Renderer::*)(const rt_camera_rig&, const RenderProfile&, int)>);
static_assert(std::is_same_v<decltype(&Renderer::render_radiance_profiled),
    ProfiledRadianceFrame (Renderer::*)(const rt_scene_desc&, const rt_camera_rig&, const RenderProfile&, int)>);
static_assert(std::is_same_v<decltype(&Renderer::render_radiance),
    RadianceFrame (Renderer::*)(const rt_scene_desc&, const rt_camera_rig&, const RenderProfile&, int)>);
static_assert(std::is_same_v<decltype(&Renderer::render_direction_debug),
    DirectionDebugFrame (Renderer::*)(const rt_camera_rig&, int)>);
static_assert(std::is_same_v<decltype(&Renderer::launch_parameter_diagnostics),
    LaunchParameterDiagnostics (Renderer::*)() const>);
static_assert(std::is_same_v<decltype(&Renderer::acceleration_diagnostics),
    const AccelerationUpdateStats& (Renderer::*)() const>);

static int failures = 0;
static void expect(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what.c_str());
        ++failures;
    }
}

// the call throws std::runtime_error whose message contains `message`
template <typename Fn>
static void expect_runtime_error(Fn&& fn, const std::string& message, const std::string& label) {
    bool threw = false, matched = false;
    try {
        fn();
    } catch (const std::runtime_error& e) {
        threw = true;
        matched = std::string(e.what()).find(message) != std::string::npos;
        if (!matched) std::printf("  %s: message was \"%s\"\n", label.c_str(), e.what());
    } catch (...) {
        threw = true;
    }
    expect(threw, label + " threw std::runtime_error");
    expect(matched, label + " message");
}

static rt_camera_rig make_rig(int active_count) {
    rt_camera_rig rig{};
    rig.active_count = active_count;
    for (int i = 0; i < active_count && i < 4; ++i) {
        rig.cameras[i].enabled = 1;
        rig.cameras[i].width = 64;
        rig.cameras[i].height = 48;
    }
    return rig;
}

int main() {
    const rt_camera_rig rig = make_rig(2);
    bool ok = true;
    try {
        validate_render_camera_request(rig, 0, "render_radiance");
        validate_render_camera_request(rig, 1, "render_radiance");
        validate_render_pool_request(rig, 2, 4, "RendererPool");
    } catch (...) {
        ok = false;
    }
    expect(ok, "valid requests pass");

    expect_runtime_error([] { validate_render_camera_request(make_rig(0), 0, "render_radiance"); },
        "render_radiance requires rig.active_count in [1, 4], got 0", "reject zero active rig");
    expect_runtime_error([&] { validate_render_camera_request(rig, 2, "render_radiance"); },
        "render_radiance camera_index out of range: camera_index=2, active_count=2",
        "reject camera index outside active slots");
    expect_runtime_error(
        [&] {
            rt_camera_rig disabled = rig;
            disabled.cameras[1].enabled = 0;
            validate_render_camera_request(disabled, 1, "render_radiance");
        },
        "render_radiance camera slot is disabled at index 1", "reject disabled camera");
    expect_runtime_error(
        [&] {
            rt_camera_rig bad = rig;
            bad.cameras[1].width = 0;
            validate_render_camera_request(bad, 1, "render_radiance");
        },
        "render_radiance camera slot has invalid resolution at index 1: width=0, height=48",
        "reject invalid camera resolution");
    expect_runtime_error([&] { validate_render_pool_request(rig, 0, 4, "RendererPool"); },
        "RendererPool active_cameras out of range", "reject zero active cameras");
    expect_runtime_error([&] { validate_render_pool_request(rig, 5, 4, "RendererPool"); },
        "RendererPool active_cameras out of range", "reject active cameras above renderer count");
    expect_runtime_error([&] { validate_render_pool_request(rig, 3, 4, "RendererPool"); },
        "RendererPool active_cameras exceeds rig.active_count",
        "reject pool active cameras above rig active count");
    expect_runtime_error(
        [] {
            rt_camera_rig disabled = make_rig(2);
            disabled.cameras[0].enabled = 0;
            validate_render_pool_request(disabled, 2, 4, "RendererPool");
        },
        "RendererPool leading camera slot disabled at index 0", "reject disabled leading camera");
    expect_runtime_error(
        [] {
            rt_camera_rig bad = make_rig(2);
            bad.cameras[1].height = -1;
            validate_render_pool_request(bad, 2, 4, "RendererPool");
        },
        "RendererPool leading camera slot has invalid resolution at index 1",
        "reject invalid leading camera resolution");
    expect_runtime_error([] { validate_render_pool_request(make_rig(5), 1, 4, "RendererPool"); },
        "RendererPool rig.active_count out of range", "reject rig with five cameras");

    // pool sizes are checked before any device is opened
    expect_runtime_error([] { RendererPool p(0); }, "RendererPool requires renderer_count in [1, 4]",
        "reject renderer_count below range");
    expect_runtime_error([] { RendererPool p(5); }, "RendererPool requires renderer_count in [1, 4]",
        "reject renderer_count above range");

    // extension fields of the profile survive the C ABI round trip
    RenderProfile p = RenderProfile::realtime();
    p.pixel_jitter = 1;
    p.restir_spatial_radius = 4;
    const RenderProfile q = from_abi(to_abi(p));
    expect(render_profile_name(q) == "default" && q.pixel_jitter == 1 && q.restir_spatial_radius == 4
               && q.samples_per_pixel == 1 && q.enable_restir_di,
        "profile round trip");
    expect(render_profile_name(from_abi(to_abi(RenderProfile::quality()))) == "quality",
        "quality round trip");
    expect(acceleration_update_kind_name(AccelerationUpdateKind::refit) == "refit", "kind names");

    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
