"""ctypes mirror of the POD structs and entry points of include/rt/rt_hip.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

f3 = C.c_float * 3
d3 = C.c_double * 3


class Sphere(C.Structure):
    _fields_ = [("center", f3), ("radius", C.c_float), ("material", C.c_int32)]


class Quad(C.Structure):
    _fields_ = [("origin", f3), ("edge_u", f3), ("edge_v", f3), ("material", C.c_int32)]


class Triangle(C.Structure):
    _fields_ = [("p0", f3), ("p1", f3), ("p2", f3), ("material", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("albedo", f3), ("emission", f3), ("fuzz", C.c_float),
                ("ior", C.c_float), ("openpbr_index", C.c_int32)]


class OpenPbr(C.Structure):
    _fields_ = [
        ("base_weight", C.c_float), ("base_color", f3), ("base_diffuse_roughness", C.c_float),
        ("base_metalness", C.c_float), ("specular_weight", C.c_float), ("specular_color", f3),
        ("specular_roughness", C.c_float), ("specular_ior", C.c_float),
        ("specular_roughness_anisotropy", C.c_float), ("transmission_weight", C.c_float),
        ("transmission_color", f3), ("transmission_depth", C.c_float),
        ("transmission_dispersion_scale", C.c_float),
        ("transmission_dispersion_abbe_number", C.c_float), ("subsurface_weight", C.c_float),
        ("subsurface_color", f3), ("subsurface_radius", C.c_float),
        ("subsurface_radius_scale", f3), ("subsurface_scatter_anisotropy", C.c_float),
        ("fuzz_weight", C.c_float), ("fuzz_color", f3), ("fuzz_roughness", C.c_float),
        ("coat_weight", C.c_float), ("coat_color", f3), ("coat_roughness", C.c_float),
        ("coat_roughness_anisotropy", C.c_float), ("coat_ior", C.c_float),
        ("coat_darkening", C.c_float), ("thin_film_weight", C.c_float),
        ("thin_film_thickness", C.c_float), ("thin_film_ior", C.c_float),
        ("emission_luminance", C.c_float), ("emission_color", f3),
        ("geometry_opacity", C.c_float), ("geometry_thin_walled", C.c_int32),
    ]


class AnalyticLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", f3), ("basis_x", f3), ("basis_y", f3),
                ("basis_z", f3), ("radiance", f3), ("radius", C.c_float), ("width", C.c_float),
                ("height", C.c_float), ("length", C.c_float), ("world_area", C.c_float),
                ("cos_theta_max", C.c_float), ("selection_pdf", C.c_float), ("cdf", C.c_float),
                ("delta", C.c_int32), ("treat_as_point", C.c_int32), ("treat_as_line", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("sphere_count", C.c_int32), ("quad_count", C.c_int32), ("triangle_count", C.c_int32),
        ("material_count", C.c_int32), ("openpbr_count", C.c_int32),
        ("analytic_light_count", C.c_int32), ("background_mode", C.c_int32),
        ("background", f3), ("background_top", f3),
        ("spheres", C.c_void_p), ("quads", C.c_void_p), ("triangles", C.c_void_p),
        ("materials", C.c_void_p), ("openpbr", C.c_void_p), ("analytic_lights", C.c_void_p),
        ("texture_count", C.c_int32), ("image_texel_count", C.c_int32),
        ("textures", C.c_void_p), ("image_texels", C.c_void_p),
        ("material_albedo_texture", C.c_void_p), ("triangle_uvs", C.c_void_p),
        ("triangle_normals", C.c_void_p),
        ("medium_count", C.c_int32), ("pad_media", C.c_int32), ("media", C.c_void_p),
        ("material_emission_texture", C.c_void_p), ("openpbr_textures", C.c_void_p),
    ]


class OpenPbrTextures(C.Structure):
    """rt_openpbr_textures: texture index per bound OpenPBR input (-1 none) and the
    colour space of each colour binding (RT_COLOR_*)."""
    _fields_ = [(n, C.c_int32) for n in (
        "base_color", "base_color_space", "specular_color", "specular_color_space",
        "transmission_color", "transmission_color_space", "emission_color", "emission_color_space",
        "base_metalness", "specular_roughness")]


class Medium(C.Structure):
    _fields_ = [("local_center_or_min", f3), ("radius", C.c_float), ("local_max", f3),
                ("density", C.c_float), ("rotation_row0", f3), ("material_index", C.c_int32),
                ("rotation_row1", f3), ("boundary_type", C.c_int32), ("rotation_row2", f3),
                ("pad0", C.c_float), ("translation", f3), ("pad1", C.c_float)]


class Texture(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("even_texture", C.c_int32), ("odd_texture", C.c_int32),
        ("image_offset", C.c_int32), ("image_width", C.c_int32), ("image_height", C.c_int32),
        ("u_address_mode", C.c_int32), ("v_address_mode", C.c_int32), ("filter_type", C.c_int32),
        ("scale", C.c_float), ("color", f3),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("model", C.c_int32),
        ("origin", d3), ("basis_x", d3), ("basis_y", d3), ("basis_z", d3),
        ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("k1", C.c_double), ("k2", C.c_double), ("k3", C.c_double), ("p1", C.c_double),
        ("p2", C.c_double), ("radial", C.c_double * 6), ("tangential", C.c_double * 2),
        ("lut_step", C.c_double), ("lut", C.c_double * 1024),
    ]


class CameraRig(C.Structure):
    _fields_ = [("active_count", C.c_int32), ("pad", C.c_int32), ("cameras", Camera * 4)]


class Profile(C.Structure):
    _fields_ = [
        ("samples_per_pixel", C.c_int32), ("max_bounces", C.c_int32),
        ("enable_denoise", C.c_int32), ("rr_start_bounce", C.c_int32),
        ("accumulation_reset_rotation_deg", C.c_double),
        ("accumulation_reset_translation", C.c_double),
        ("enable_restir_di", C.c_int32), ("restir_initial_candidates", C.c_int32),
        ("restir_temporal_reuse", C.c_int32), ("restir_max_history_age", C.c_int32),
        ("restir_max_temporal_candidates", C.c_int32), ("restir_spatial_neighbors", C.c_int32),
        ("restir_max_spatial_candidates", C.c_int32), ("restir_bias_correction", C.c_int32),
        ("restir_min_analytic_lights", C.c_int32), ("pixel_jitter", C.c_int32),
        ("restir_spatial_radius", C.c_int32), ("pad0", C.c_int32),
    ]


class BvhStats(C.Structure):
    _fields_ = [("prim_count", C.c_int32), ("node_count", C.c_int32), ("max_depth", C.c_int32),
                ("wide_node_count", C.c_int32), ("build_ms", C.c_double),
                ("wide_levels", C.c_int32), ("pad", C.c_int32)]


WIDE_NODE_DTYPE = np.dtype([("origin", "<f4", 3), ("meta", "<u4"), ("child", "<u4", 4),
                            ("qlo", "<u4", 3), ("qhi", "<u4", 3), ("pad", "<u4", 2)])


class FrameTiming(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("denoise_ms", C.c_double),
                ("download_ms", C.c_double), ("extend_ms", C.c_double),
                ("shade_ms", C.c_double), ("connect_ms", C.c_double), ("other_ms", C.c_double),
                ("paths", C.c_uint64), ("extend_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("extend_launches", C.c_uint64), ("node_visits", C.c_uint64),
                ("prim_tests", C.c_uint64), ("sphere_tests", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_prim_tests", C.c_uint64),
                ("shadow_sphere_tests", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("wavefront_paths", C.c_int32), ("trace_block", C.c_int32),
                ("trace_blocks_per_cu", C.c_int32), ("use_graph", C.c_int32),
                ("count_traversal", C.c_int32), ("trace_refill", C.c_int32)]


class RestirStats(C.Structure):
    _fields_ = [("pixel_count", C.c_int32), ("active_reservoir_count", C.c_int32),
                ("temporal_reuse_count", C.c_int32), ("spatial_reuse_count", C.c_int32),
                ("bias_corrected_count", C.c_int32), ("max_candidate_count", C.c_int32),
                ("max_age", C.c_int32), ("pad", C.c_int32)]


class NnTiling(C.Structure):
    _fields_ = [("conv_rows", C.c_int32 * 4)]


class NnAutotuneResult(C.Structure):
    _fields_ = [("best", NnTiling), ("best_ms", C.c_double), ("baseline_ms", C.c_double),
                ("trials", C.c_int32), ("pad", C.c_int32)]


class AutotuneResult(C.Structure):
    _fields_ = [("best", Tuning), ("best_ms", C.c_double), ("baseline_ms", C.c_double),
                ("trials", C.c_int32), ("denoiser_tuned", C.c_int32),
                ("denoiser", NnAutotuneResult)]


class NnImages(C.Structure):
    _fields_ = [("count", C.c_int32), ("offset", C.c_int32 * 17), ("width", C.c_int32 * 16)]


def nn_images(sizes) -> "NnImages":
    """Image table for [(width, height), ...] laid back to back."""
    im = NnImages()
    im.count = len(sizes)
    o = 0
    for i, (w, h) in enumerate(sizes):
        im.offset[i] = o
        im.width[i] = w
        o += w * h
    im.offset[len(sizes)] = o
    return im


class AccelerationUpdate(C.Structure):
    _fields_ = [("kind", C.c_int32), ("node_count", C.c_int32),
                ("primitive_reference_count", C.c_int32), ("wide_node_count", C.c_int32),
                ("elapsed_ms", C.c_double), ("bvh_ms", C.c_double), ("generation", C.c_uint64)]


ACCEL_KINDS = {0: "reuse", 1: "update", 2: "refit", 3: "rebuild"}


class DenoiserGuard(C.Structure):
    _fields_ = [("fallback", C.c_int32 * 4), ("nonfinite", C.c_float * 4),
                ("mean_shift", C.c_float * 4), ("residual_noise", C.c_float * 4),
                ("frames_checked", C.c_uint32), ("frames_fallback", C.c_uint32),
                ("max_mean_shift", C.c_float), ("max_residual_noise", C.c_float)]


class CacheConfig(C.Structure):
    _fields_ = [("levels", C.c_int32), ("log2_table", C.c_int32),
                ("base_resolution", C.c_int32), ("max_resolution", C.c_int32),
                ("train_samples", C.c_int32), ("train_steps", C.c_int32),
                ("query_bounce", C.c_int32), ("warmup_frames", C.c_int32),
                ("learning_rate", C.c_float), ("max_rel_loss", C.c_float),
                ("train_fraction", C.c_float), ("seed", C.c_uint32)]


class CacheStats(C.Structure):
    _fields_ = [("loss_ema", C.c_float), ("last_loss", C.c_float), ("frames", C.c_int32),
                ("usable", C.c_int32), ("steps", C.c_int32), ("records", C.c_int32),
                ("queries", C.c_int32), ("train_fraction", C.c_float)]


class DeviceFrameView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("beauty", C.c_void_p),
                ("normal", C.c_void_p), ("albedo", C.c_void_p), ("depth", C.c_void_p),
                ("resolved_beauty", C.c_void_p), ("flow", C.c_void_p), ("trust", C.c_void_p)]


class GeneratedScene(C.Structure):
    _fields_ = [("scene", SceneDesc), ("camera_count", C.c_int32), ("pad", C.c_int32),
                ("cameras", Camera * 4), ("profile", Profile)]


RESERVOIR_DTYPE = np.dtype([("light_index", "<i4"), ("sample_u0", "<f4"), ("sample_u1", "<f4"),
                            ("position", "<f4", 3), ("normal", "<f4", 3),
                            ("material_type", "<i4"), ("primitive_type", "<i4"),
                            ("primitive_index", "<i4"), ("weight_sum", "<f4"),
                            ("selected_target", "<f4"), ("estimator_weight", "<f4"),
                            ("candidate_count", "<i4"), ("temporal_candidate_count", "<i4"),
                            ("spatial_candidate_count", "<i4"),
                            ("bias_correction_source_count", "<i4"), ("age", "<i4"),
                            ("valid", "<i4")])

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<i4"), ("u", "<f4"), ("v", "<f4")])
NODE_DTYPE = np.dtype([("lo0", "<f4", 3), ("hi0", "<f4", 3), ("lo1", "<f4", 3),
                       ("hi1", "<f4", 3), ("child", "<u4", 2), ("parent", "<u4"),
                       ("pad", "<u4")])


def declare(L):
    P = C.POINTER
    vp = C.c_void_p
    L.rt_abi_version.restype = C.c_int
    L.rt_last_error.restype = C.c_char_p
    L.rt_device_create.argtypes = [C.c_int, P(vp)]
    L.rt_device_destroy.argtypes = [vp]
    L.rt_device_synchronize.argtypes = [vp]
    L.rt_device_info.argtypes = [vp, C.c_char_p, C.c_size_t, P(C.c_int), P(C.c_uint64),
                                 P(C.c_uint64)]
    L.rt_scene_upload.argtypes = [vp, P(SceneDesc), P(vp)]
    L.rt_scene_destroy.argtypes = [vp]
    L.rt_bvh_build.argtypes = [vp, P(BvhStats)]
    L.rt_bvh_download.argtypes = [vp, vp, vp, vp]
    L.rt_bvh_download_wide.argtypes = [vp, vp]
    L.rt_trace_closest.argtypes = [vp, vp, C.c_int64, vp]
    L.rt_trace_any.argtypes = [vp, vp, C.c_int64, vp]
    L.rt_trace_closest_device.argtypes = [vp, vp, C.c_int64, vp, P(C.c_double)]
    L.rt_trace_closest_device_noalloc.argtypes = [vp, vp, C.c_int64, vp]
    L.rt_scene_update.argtypes = [vp, P(SceneDesc), P(AccelerationUpdate)]
    L.rt_scene_last_update.argtypes = [vp, P(AccelerationUpdate)]
    L.rt_renderer_create.argtypes = [vp, vp, P(vp)]
    L.rt_renderer_destroy.argtypes = [vp]
    L.rt_render_frame.argtypes = [vp, P(CameraRig), P(Profile), C.c_uint32, P(FrameTiming)]
    L.rt_render_frame_streams.argtypes = [vp, P(CameraRig), P(Profile), P(C.c_uint32),
                                          P(C.c_uint32), P(FrameTiming)]
    L.rt_renderer_launch_diagnostics.argtypes = [vp, P(C.c_uint64), P(C.c_uint64)]
    L.rt_renderer_set_async.argtypes = [vp, C.c_int32]
    L.rt_renderer_synchronize.argtypes = [vp]
    L.rt_analytic_lights_compile.argtypes = [vp, C.c_int32]
    L.rt_render_direction_debug.argtypes = [vp, P(CameraRig), C.c_int, vp]
    L.rt_frame_download.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    L.rt_frame_device_view.argtypes = [vp, C.c_int, P(DeviceFrameView)]
    L.rt_presenter_create.argtypes = [vp, C.c_int32, C.c_int32, P(vp)]
    L.rt_presenter_destroy.argtypes = [vp]
    L.rt_presenter_present.argtypes = [vp, P(DeviceFrameView), P(C.c_int32)]
    L.rt_presenter_pixels.argtypes = [vp, P(vp)]
    L.rt_presenter_download.argtypes = [vp, vp]
    L.rt_tone_map_rgba8.argtypes = [vp, vp, vp, C.c_int64]
    L.rt_reset_accumulation.argtypes = [vp]
    L.rt_reset_sequence.argtypes = [vp, C.c_uint32]
    L.rt_restir_diagnostics.argtypes = [vp, C.c_int, P(RestirStats)]
    L.rt_renderer_set_tuning.argtypes = [vp, P(Tuning)]
    L.rt_renderer_get_tuning.argtypes = [vp, P(Tuning)]
    L.rt_scene_generate.argtypes = [C.c_int, C.c_uint64, P(P(GeneratedScene))]
    L.rt_scene_generated_free.argtypes = [P(GeneratedScene)]
    L.rt_restir_download.argtypes = [vp, C.c_int, vp]
    L.rt_nn_param_count.argtypes = []
    L.rt_nn_param_count.restype = C.c_int32
    L.rt_nn_create.argtypes = [C.c_int, P(vp)]
    L.rt_nn_destroy.argtypes = [vp]
    L.rt_nn_set_params.argtypes = [vp, vp, C.c_int32]
    L.rt_nn_get_params.argtypes = [vp, vp, C.c_int32]
    L.rt_nn_infer.argtypes = [vp, P(NnImages), vp, vp, vp, vp, vp]
    L.rt_nn_train_step.argtypes = [vp, P(NnImages), vp, vp, vp, vp, vp, C.c_float, vp, vp]
    L.rt_nn_infer_host.argtypes = [vp, P(NnImages), vp, vp, vp, vp, vp]
    L.rt_nn_train_step_host.argtypes = [vp, P(NnImages), vp, vp, vp, vp, vp, C.c_float, vp, vp]
    L.rt_nn_set_tiling.argtypes = [vp, P(NnTiling)]
    L.rt_nn_get_tiling.argtypes = [vp, P(NnTiling)]
    L.rt_nn_autotune.argtypes = [vp, P(NnImages), C.c_int32, P(NnAutotuneResult)]
    L.rt_renderer_set_denoiser.argtypes = [vp, vp]
    L.rt_renderer_set_denoiser_budget.argtypes = [vp, C.c_float, C.c_float]
    L.rt_renderer_denoiser_guard.argtypes = [vp, P(DenoiserGuard)]
    L.rt_renderer_autotune.argtypes = [vp, P(CameraRig), P(Profile), C.c_int32, P(AutotuneResult)]
    L.rt_frame_download_guides.argtypes = [vp, C.c_int, vp, vp, vp]
    L.rt_scene_generate_rig.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_int, P(CameraRig)]
    L.rt_camera_build_lut.argtypes = [P(Camera)]
    L.rt_cache_default_config.argtypes = [P(CacheConfig)]
    L.rt_cache_default_config.restype = None
    L.rt_cache_mlp_param_count.argtypes = []
    L.rt_cache_mlp_param_count.restype = C.c_int32
    L.rt_cache_param_count.argtypes = [P(CacheConfig)]
    L.rt_cache_param_count.restype = C.c_int64
    L.rt_cache_create.argtypes = [C.c_int, P(CacheConfig), P(vp)]
    L.rt_cache_destroy.argtypes = [vp]
    L.rt_cache_set_params.argtypes = [vp, vp, C.c_int64]
    L.rt_cache_get_params.argtypes = [vp, vp, C.c_int64]
    L.rt_cache_set_bounds.argtypes = [vp, vp, vp]
    L.rt_cache_grid_info.argtypes = [vp, vp, vp, vp, vp]
    L.rt_cache_infer_host.argtypes = [vp, vp, C.c_int32, vp]
    L.rt_cache_train_step_host.argtypes = [vp, vp, vp, C.c_int32, vp, vp]
    L.rt_cache_get_stats.argtypes = [vp, P(CacheStats)]
    L.rt_renderer_set_radiance_cache.argtypes = [vp, vp]
