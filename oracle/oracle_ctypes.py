"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The struct
layouts mirror oracle/rt_oracle.h, which in turn mirrors the C-ABI structs in
include/rt/rt_hip.h byte for byte.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

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


OPENPBR_FIELDS = [
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


class OpenPbr(C.Structure):
    _fields_ = OPENPBR_FIELDS


class AnalyticLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", f3), ("basis_x", f3), ("basis_y", f3),
                ("basis_z", f3), ("radiance", f3), ("radius", C.c_float), ("width", C.c_float),
                ("height", C.c_float), ("length", C.c_float), ("world_area", C.c_float),
                ("cos_theta_max", C.c_float), ("selection_pdf", C.c_float), ("cdf", C.c_float),
                ("delta", C.c_int32), ("treat_as_point", C.c_int32), ("treat_as_line", C.c_int32)]


class Texture(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("even_texture", C.c_int32), ("odd_texture", C.c_int32),
        ("image_offset", C.c_int32), ("image_width", C.c_int32), ("image_height", C.c_int32),
        ("u_address_mode", C.c_int32), ("v_address_mode", C.c_int32), ("filter_type", C.c_int32),
        ("scale", C.c_float), ("color", f3),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("sphere_count", C.c_int32), ("quad_count", C.c_int32), ("triangle_count", C.c_int32),
        ("material_count", C.c_int32), ("openpbr_count", C.c_int32),
        ("analytic_light_count", C.c_int32), ("background_mode", C.c_int32),
        ("background", f3), ("background_top", f3),
        ("spheres", C.POINTER(Sphere)), ("quads", C.POINTER(Quad)),
        ("triangles", C.POINTER(Triangle)), ("materials", C.POINTER(Material)),
        ("openpbr", C.POINTER(OpenPbr)), ("analytic_lights", C.POINTER(AnalyticLight)),
        ("texture_count", C.c_int32), ("image_texel_count", C.c_int32),
        ("textures", C.c_void_p), ("image_texels", C.c_void_p),
        ("material_albedo_texture", C.c_void_p), ("triangle_uvs", C.c_void_p),
        ("triangle_normals", C.c_void_p),
        ("medium_count", C.c_int32), ("pad_media", C.c_int32), ("media", C.c_void_p),
        ("material_emission_texture", C.c_void_p), ("openpbr_textures", C.c_void_p),
    ]


class OpenPbrTextures(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "base_color", "base_color_space", "specular_color", "specular_color_space",
        "transmission_color", "transmission_color_space", "emission_color", "emission_color_space",
        "base_metalness", "specular_roughness")]


class Medium(C.Structure):
    _fields_ = [("local_center_or_min", f3), ("radius", C.c_float), ("local_max", f3),
                ("density", C.c_float), ("rotation_row0", f3), ("material_index", C.c_int32),
                ("rotation_row1", f3), ("boundary_type", C.c_int32), ("rotation_row2", f3),
                ("pad0", C.c_float), ("translation", f3), ("pad1", C.c_float)]


class Camera(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("model", C.c_int32),
        ("origin", d3), ("basis_x", d3), ("basis_y", d3), ("basis_z", d3),
        ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("k1", C.c_double), ("k2", C.c_double), ("k3", C.c_double), ("p1", C.c_double),
        ("p2", C.c_double), ("radial", C.c_double * 6), ("tangential", C.c_double * 2),
        ("lut_step", C.c_double), ("lut", C.c_double * 1024),
    ]


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


class BvhNode(C.Structure):
    _fields_ = [("lo0", f3), ("hi0", f3), ("lo1", f3), ("hi1", f3),
                ("child", C.c_uint32 * 2), ("parent", C.c_uint32), ("pad", C.c_uint32)]


WIDE_DTYPE = np.dtype([("origin", "<f4", 3), ("meta", "<u4"), ("child", "<u4", 4),
                       ("qlo", "<u4", 3), ("qhi", "<u4", 3), ("pad", "<u4", 2)])


class Bvh(C.Structure):
    _fields_ = [("prim_count", C.c_int32), ("node_count", C.c_int32),
                ("nodes", C.POINTER(BvhNode)), ("sorted_prims", C.POINTER(C.c_uint32)),
                ("morton", C.POINTER(C.c_uint32)), ("max_depth", C.c_int32),
                ("wide_count", C.c_int32), ("wide", C.c_void_p), ("wide_levels", C.c_int32),
                ("pad", C.c_int32)]


class Generated(C.Structure):
    _fields_ = [("scene", Scene), ("camera_count", C.c_int32), ("pad", C.c_int32),
                ("cameras", Camera * 4), ("profile", Profile)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays_closest", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64)]


class Reservoir(C.Structure):
    """RestirReservoir field order (src/common/restir_di.h:36-48), 84 bytes."""
    _fields_ = [("light_index", C.c_int32), ("sample_u0", C.c_float), ("sample_u1", C.c_float),
                ("position", f3), ("normal", f3), ("material_type", C.c_int32),
                ("primitive_type", C.c_int32), ("primitive_index", C.c_int32),
                ("weight_sum", C.c_float), ("selected_target", C.c_float),
                ("estimator_weight", C.c_float), ("candidate_count", C.c_int32),
                ("temporal_candidate_count", C.c_int32), ("spatial_candidate_count", C.c_int32),
                ("bias_correction_source_count", C.c_int32), ("age", C.c_int32),
                ("valid", C.c_int32)]


class RestirIO(C.Structure):
    _fields_ = [("prev_cam", C.c_void_p), ("history", C.c_void_p), ("out", C.c_void_p),
                ("history_length", C.c_int32), ("pad", C.c_int32)]


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


class CacheGrid(C.Structure):
    _fields_ = [("levels", C.c_int32), ("log2_table", C.c_int32), ("res", C.c_float * 16),
                ("dense", C.c_int32 * 16), ("lo", C.c_float * 3), ("inv_extent", C.c_float * 3)]


def cache_grid(info: dict) -> CacheGrid:
    """Grid constants (levels, log2_table, res, dense, lo, inv_extent)."""
    g = CacheGrid()
    g.levels = int(info["levels"])
    g.log2_table = int(info["log2_table"])
    for l in range(16):
        g.res[l] = float(info["res"][l])
        g.dense[l] = int(info["dense"][l])
    for a in range(3):
        g.lo[a] = float(info["lo"][a])
        g.inv_extent[a] = float(info["inv_extent"][a])
    return g


def cache_corners(g: CacheGrid, level: int, p):
    p = np.ascontiguousarray(p, np.float32)
    idx = np.zeros(8, np.uint32)
    w = np.zeros(8, np.float32)
    lib().rto_cache_corners(C.byref(g), level, ptr(p), ptr(idx), ptr(w))
    return idx, w


def cache_features(g: CacheGrid, table, query, bf16=True) -> np.ndarray:
    t = np.ascontiguousarray(table, np.float32)
    q = np.ascontiguousarray(query, np.float32)
    x = np.zeros(64, np.float32)
    lib().rto_cache_features(C.byref(g), ptr(t), ptr(q), 1 if bf16 else 0, ptr(x))
    return x


def cache_predict(g: CacheGrid, params, inputs, bf16=True) -> np.ndarray:
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 16)
    out = np.zeros((x.shape[0], 3), np.float64)
    lib().rto_cache_predict(C.byref(g), ptr(p), x.shape[0], ptr(x), 1 if bf16 else 0, ptr(out))
    return out


def cache_loss_grad(g: CacheGrid, params, inputs, targets, bf16=True, want_grad=True):
    """Oracle radiance cache: (loss, gradient (float64) or None)."""
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 16)
    t = np.ascontiguousarray(targets, np.float32).reshape(-1, 4)
    grad = np.zeros(p.size, np.float64) if want_grad else None
    loss = lib().rto_cache_loss_grad(C.byref(g), ptr(p), x.shape[0], ptr(x), ptr(t),
                                     1 if bf16 else 0, ptr(grad) if grad is not None else None)
    return loss, grad


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc) and, when the
    reference tree is present, oracle/_ref/libref.so from its headers."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.rto_generate_scene.argtypes = [C.c_int, C.c_uint64, P(Generated)]
        L.rto_free_generated.argtypes = [P(Generated)]
        L.rto_build_lbvh.argtypes = [P(Scene), P(Bvh)]
        L.rto_free_lbvh.argtypes = [P(Bvh)]
        L.rto_refit_lbvh.argtypes = [P(Scene), P(Bvh)]
        L.rto_bvh_validate.argtypes = [P(Scene), P(Bvh)]
        L.rto_trace_closest.argtypes = [P(Scene), P(Bvh), C.c_void_p, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int]
        L.rto_trace_any.argtypes = [P(Scene), P(Bvh), C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_int]
        L.rto_brute_closest.argtypes = [P(Scene), C.c_void_p, C.c_int, C.c_void_p]
        L.rto_rng4.argtypes = [C.c_uint32] * 6 + [C.c_void_p]
        L.rto_pcg4d.argtypes = [C.c_void_p]
        L.rto_sincos_turn.argtypes = [C.c_float, P(C.c_float), P(C.c_float)]
        L.rto_cbrt01.argtypes = [C.c_float]
        L.rto_cbrt01.restype = C.c_float
        L.rto_expand_bits10.argtypes = [C.c_uint32]
        L.rto_expand_bits10.restype = C.c_uint32
        L.rto_unproject.argtypes = [P(Camera), C.c_double, C.c_double, C.c_void_p]
        L.rto_camera_ray.argtypes = [P(Camera), C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.rto_project.argtypes = [P(Camera), C.c_void_p, C.c_void_p]
        L.rto_make_equi_lut.argtypes = [P(Camera)]
        L.rto_resolve.argtypes = [P(Camera), P(Camera), C.c_int, C.c_int] + [C.c_void_p] * 8
        L.rto_resolve.restype = None
        L.rto_denoise_pass.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 5
        L.rto_denoise_pass.restype = None
        L.rto_denoise.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6
        L.rto_denoise.restype = None
        L.rto_noise_sin.argtypes = [C.c_float]
        L.rto_noise_sin.restype = C.c_float
        L.rto_turbulence.argtypes = [C.c_void_p]
        L.rto_turbulence.restype = C.c_float
        L.rto_eval_texture.argtypes = [P(Scene), C.c_int, C.c_float, C.c_float, C.c_void_p, C.c_void_p]
        L.rto_eval_texture.restype = None
        L.rto_surface_uv.argtypes = [P(Scene), C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                     P(C.c_float), P(C.c_float)]
        L.rto_surface_uv.restype = None
        L.rto_display_u8.argtypes = [C.c_float]
        L.rto_display_u8.restype = C.c_uint8
        L.rto_tone_map_rgba8.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.rto_tone_map_rgba8.restype = None
        L.rto_nn_features.argtypes = [C.c_int] + [C.c_void_p] * 5
        L.rto_nn_features.restype = None
        L.rto_nn_loss_grad.argtypes = [C.c_void_p] * 7 + [C.c_int]
        L.rto_nn_loss_grad.restype = C.c_double
        L.rto_bf16_round.argtypes = [C.c_float]
        L.rto_bf16_round.restype = C.c_float
        L.rto_primary_rays.argtypes = [P(Camera), C.c_int, C.c_int, C.c_uint32, C.c_int,
                                       C.c_void_p]
        L.rto_build_lights.argtypes = [P(Scene), C.c_void_p, C.c_int]
        L.rto_hit_material.argtypes = [P(Scene), C.c_void_p, C.c_void_p, C.c_float, C.c_int, C.c_void_p]
        L.rto_hit_material.restype = None
        L.rto_render_camera.argtypes = [P(Scene), P(Bvh), P(Camera), C.c_int, P(Profile),
                                        C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_int, C.c_int, C.c_int, P(RenderStats)]
        L.rto_trace_path.argtypes = [P(Scene), P(Bvh), P(Camera), C.c_int, P(Profile),
                                     C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.rto_openpbr_evaluate.argtypes = [P(OpenPbr), C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]
        L.rto_openpbr_sample_fn.argtypes = [P(OpenPbr), C.c_void_p, C.c_void_p, C.c_float,
                                            C.c_float, C.c_float, C.c_void_p]
        L.rto_openpbr_emission.argtypes = [P(OpenPbr), C.c_void_p]
        L.rto_openpbr_evaluate_ctx.argtypes = [P(OpenPbr)] + [C.c_void_p] * 5
        L.rto_openpbr_sample_ctx.argtypes = [P(OpenPbr), C.c_void_p, C.c_void_p, C.c_float,
                                             C.c_float, C.c_float, C.c_void_p, C.c_void_p]
        L.rto_openpbr_subsurface_medium.argtypes = [P(OpenPbr), C.c_void_p]
        L.rto_openpbr_subsurface_segment.argtypes = [P(OpenPbr), C.c_float, C.c_float,
                                                     C.c_void_p, C.c_void_p]
        L.rto_openpbr_hg.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_void_p]
        L.rto_generate_rig.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_int, P(Camera),
                                       P(C.c_int)]
        L.rto_render_camera_ex.argtypes = [P(Scene), P(Bvh), P(Camera), C.c_int, P(Profile),
                                           C.c_uint32, P(RestirIO), C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                           P(RenderStats)]
        L.rto_restir_update.argtypes = [P(Reservoir), C.c_int, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_int, C.c_float]
        L.rto_restir_finalize.argtypes = [P(Reservoir)]
        L.rto_restir_finalize_basic.argtypes = [P(Reservoir), C.c_float, C.c_float, C.c_int]
        L.rto_restir_merge.argtypes = [P(Reservoir), P(Reservoir), C.c_float, C.c_int, C.c_float,
                                       C.c_int, P(C.c_int)]
        L.rto_analytic_lights_compile.argtypes = [C.c_void_p, C.c_int]
        L.rto_direction_debug.argtypes = [P(Camera), C.c_void_p]
        L.rto_log_fixed.argtypes = [C.c_float]
        L.rto_log_fixed.restype = C.c_float
        L.rto_analytic_lights_compile.restype = C.c_int
        L.rto_cache_mlp_param_count.argtypes = []
        L.rto_cache_mlp_param_count.restype = C.c_int
        L.rto_cache_corners.argtypes = [P(CacheGrid), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.rto_cache_corners.restype = None
        L.rto_cache_features.argtypes = [P(CacheGrid), C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.rto_cache_features.restype = None
        L.rto_cache_predict.argtypes = [P(CacheGrid), C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                        C.c_void_p]
        L.rto_cache_loss_grad.argtypes = [P(CacheGrid), C.c_void_p, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_int, C.c_void_p]
        L.rto_cache_loss_grad.restype = C.c_double
        for name in ("rto_power_heuristic", "rto_area_pdf", "rto_cone_pdf",
                     "rto_uniform_sphere_pdf"):
            getattr(L, name).restype = C.c_float
        L.rto_power_heuristic.argtypes = [C.c_float, C.c_float]
        L.rto_cone_pdf.argtypes = [C.c_float]
        L.rto_area_pdf.argtypes = [C.c_float, C.c_float, C.c_float]
        L.rto_uniform_sphere_pdf.argtypes = []
        L.rto_pick_cdf.argtypes = [C.c_void_p, C.c_int, C.c_float]
        L.rto_triangle_barycentrics.argtypes = [C.c_float, C.c_float, C.c_void_p]
        L.rto_triangle_barycentrics.restype = None
        L.rto_sample_light_at.argtypes = [P(Scene), P(Bvh), C.c_int] + [C.c_void_p] * 5
        L.rto_light_pdf_for_hit.argtypes = [P(Scene), P(Bvh), C.c_int, C.c_int] + [C.c_void_p] * 3
        L.rto_light_pdf_for_hit.restype = C.c_float
        L.rto_scatter.argtypes = [P(Scene), C.c_int] + [C.c_void_p] * 5
        _lib = L
    return _lib


REF_PATH = os.path.join(HERE, "_ref", "libref.so")
_ref = None


def ref_lib():
    """The reference's own headers compiled in place (oracle/_ref), or None when absent."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_openpbr_evaluate.argtypes = [C.c_void_p] * 5
        R.ref_openpbr_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                         C.c_float, C.c_float, C.c_void_p]
        R.ref_openpbr_emission.argtypes = [C.c_void_p, C.c_void_p]
        R.ref_openpbr_evaluate_ctx.argtypes = [C.c_void_p] * 6
        R.ref_openpbr_sample_ctx.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                             C.c_float, C.c_float, C.c_void_p, C.c_void_p]
        R.ref_openpbr_subsurface_medium.argtypes = [C.c_void_p, C.c_void_p]
        R.ref_openpbr_subsurface_segment.argtypes = [C.c_void_p, C.c_float, C.c_float,
                                                     C.c_void_p, C.c_void_p]
        R.ref_openpbr_hg.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_void_p]
        R.ref_openpbr_material_floats.restype = C.c_int
        for name in ("ref_light_uniform_cone_pdf", "ref_light_power_heuristic"):
            getattr(R, name).restype = C.c_float
        R.ref_light_uniform_cone_pdf.argtypes = [C.c_float]
        R.ref_light_power_heuristic.argtypes = [C.c_float, C.c_float]
        R.ref_light_area_to_solid_angle_pdf.argtypes = [C.c_float, C.c_float, C.c_float]
        R.ref_light_area_to_solid_angle_pdf.restype = C.c_float
        R.ref_sample_packed_light.argtypes = [C.c_void_p, C.c_int, C.c_float]
        R.ref_sample_uniform_triangle.argtypes = [C.c_float, C.c_float, C.c_void_p]
        R.ref_restir_update.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_int,
                                        C.c_float, C.POINTER(C.c_int)]
        R.ref_restir_finalize.argtypes = [C.c_void_p]
        R.ref_linear_to_display.argtypes = [C.c_float]
        R.ref_linear_to_display.restype = C.c_float
        R.ref_linear_to_display_u8.argtypes = [C.c_float]
        R.ref_linear_to_display_u8.restype = C.c_uint8
        _ref = R
    return _ref


def openpbr_material(**kw) -> "OpenPbr":
    """OpenPBR parameters with the reference defaults (openpbr_core.h:33-77)."""
    m = OpenPbr()
    defaults = dict(base_weight=1.0, base_color=(0.8, 0.8, 0.8), specular_weight=1.0,
                    specular_color=(1, 1, 1), specular_roughness=0.3, specular_ior=1.5,
                    transmission_color=(1, 1, 1), transmission_dispersion_abbe_number=20.0,
                    subsurface_color=(0.8, 0.8, 0.8), subsurface_radius=1.0,
                    subsurface_radius_scale=(1.0, 0.5, 0.25), fuzz_color=(1, 1, 1),
                    fuzz_roughness=0.5, coat_color=(1, 1, 1), coat_ior=1.6, coat_darkening=1.0,
                    thin_film_thickness=0.5, thin_film_ior=1.4, emission_color=(1, 1, 1),
                    geometry_opacity=1.0)
    defaults.update(kw)
    for k, v in defaults.items():
        if isinstance(v, (tuple, list)):
            setattr(m, k, f3(*v))
        else:
            setattr(m, k, v)
    return m


def openpbr_evaluate(m, n, wo, wi, thr=None):
    """evaluate_openpbr_core; thr = the transport context's path throughput."""
    out = np.zeros(4, np.float32)
    a = [np.asarray(x, np.float32) for x in (n, wo, wi)]
    if thr is None:
        lib().rto_openpbr_evaluate(C.byref(m), ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(out))
    else:
        t = np.asarray(thr, np.float32)
        lib().rto_openpbr_evaluate_ctx(C.byref(m), ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(t),
                                       ptr(out))
    return out[:3].copy(), float(out[3])


def openpbr_sample(m, n, wo, u_lobe, u1, u2, thr=None):
    class _S(C.Structure):
        _fields_ = [("wi", f3), ("weight", f3), ("pdf", C.c_float), ("delta", C.c_int32),
                    ("valid", C.c_int32), ("event", C.c_int32)]
    s = _S()
    a = [np.asarray(x, np.float32) for x in (n, wo)]
    if thr is None:
        lib().rto_openpbr_sample_fn(C.byref(m), ptr(a[0]), ptr(a[1]), u_lobe, u1, u2, C.byref(s))
    else:
        t = np.asarray(thr, np.float32)
        lib().rto_openpbr_sample_ctx(C.byref(m), ptr(a[0]), ptr(a[1]), u_lobe, u1, u2, ptr(t),
                                     C.byref(s))
    return {"wi": np.array(s.wi[:], np.float32), "weight": np.array(s.weight[:], np.float32),
            "pdf": s.pdf, "delta": s.delta, "valid": s.valid, "event": s.event}


def openpbr_emission(m):
    out = np.zeros(3, np.float32)
    lib().rto_openpbr_emission(C.byref(m), ptr(out))
    return out


def openpbr_subsurface_medium(m):
    """(extinction[3], albedo[3], anisotropy, active) of openpbr_subsurface_medium."""
    out = np.zeros(8, np.float32)
    lib().rto_openpbr_subsurface_medium(C.byref(m), ptr(out))
    return out


def openpbr_subsurface_segment(m, boundary, u, thr):
    """(weight[3], distance, scattered) of one random-walk segment."""
    out = np.zeros(5, np.float32)
    t = np.asarray(thr, np.float32)
    lib().rto_openpbr_subsurface_segment(C.byref(m), boundary, u, ptr(t), ptr(out))
    return out


def openpbr_hg(incident, g, u1, u2):
    out = np.zeros(3, np.float32)
    a = np.asarray(incident, np.float32)
    lib().rto_openpbr_hg(ptr(a), g, u1, u2, ptr(out))
    return out


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class GeneratedScene:
    """A synthetic scene generated by the oracle's own generator."""

    def __init__(self, config: int, seed: int = 0):
        self.g = Generated()
        rc = lib().rto_generate_scene(config, seed, C.byref(self.g))
        if rc != 0:
            raise ValueError(f"unknown scene config {config}")

    def __del__(self):
        try:
            lib().rto_free_generated(C.byref(self.g))
        except Exception:
            pass

    @property
    def scene(self) -> Scene:
        return self.g.scene

    def cameras(self):
        return [self.g.cameras[i] for i in range(self.g.camera_count)]


def scene_from_arrays(spheres=None, quads=None, triangles=None, materials=None, openpbr=None,
                      analytic=None, background=(0.0, 0.0, 0.0), background_top=None, media=None):
    """Builds an oracle Scene over numpy structured/raw byte arrays; returns (scene, keepalive)."""
    keep = []
    s = Scene()

    def attach(name, cnt_name, arr, ctype):
        if arr is None or len(arr) == 0:
            setattr(s, cnt_name, 0)
            setattr(s, name, C.POINTER(ctype)())
            return
        a = np.ascontiguousarray(arr)
        keep.append(a)
        setattr(s, cnt_name, len(a))
        setattr(s, name, a.ctypes.data_as(C.POINTER(ctype)))

    attach("spheres", "sphere_count", spheres, Sphere)
    attach("quads", "quad_count", quads, Quad)
    attach("triangles", "triangle_count", triangles, Triangle)
    attach("materials", "material_count", materials, Material)
    attach("openpbr", "openpbr_count", openpbr, OpenPbr)
    attach("analytic_lights", "analytic_light_count", analytic, AnalyticLight)
    if media is not None and len(media):
        a = np.ascontiguousarray(media)
        keep.append(a)
        s.medium_count, s.media = len(a), a.ctypes.data
    s.background_mode = 1 if background_top is not None else 0
    s.background = f3(*background)
    s.background_top = f3(*(background_top if background_top is not None else (0, 0, 0)))
    return s, keep


class OracleBvh:
    def __init__(self, scene: Scene):
        self.scene = scene
        self.bvh = Bvh()
        lib().rto_build_lbvh(C.byref(scene), C.byref(self.bvh))

    def __del__(self):
        try:
            lib().rto_free_lbvh(C.byref(self.bvh))
        except Exception:
            pass

    def refit(self, scene: "Scene"):
        """Refit to `scene` (same primitive counts), keeping the topology."""
        if lib().rto_refit_lbvh(C.byref(scene), C.byref(self.bvh)) != 0:
            raise ValueError("refit needs a scene with the built primitive counts (> 1)")
        self.scene = scene

    def nodes(self) -> np.ndarray:
        n = self.bvh.node_count
        if n == 0:
            return np.zeros(0, NODE_DTYPE)
        buf = C.string_at(self.bvh.nodes, n * NODE_DTYPE.itemsize)
        return np.frombuffer(buf, NODE_DTYPE).copy()

    def sorted_prims(self) -> np.ndarray:
        n = self.bvh.prim_count
        return np.ctypeslib.as_array(self.bvh.sorted_prims, (n,)).copy() if n else np.zeros(0, np.uint32)

    def wide(self) -> np.ndarray:
        n = self.bvh.wide_count
        if n == 0:
            return np.zeros(0, WIDE_DTYPE)
        return np.frombuffer(C.string_at(self.bvh.wide, n * 64), WIDE_DTYPE).copy()

    def morton(self) -> np.ndarray:
        n = self.bvh.prim_count
        return np.ctypeslib.as_array(self.bvh.morton, (n,)).copy() if n else np.zeros(0, np.uint32)

    def validate(self) -> int:
        return lib().rto_bvh_validate(C.byref(self.scene), C.byref(self.bvh))

    def trace_closest(self, rays: np.ndarray, threads: int = 8, counts: bool = False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = len(rays)
        hits = np.zeros(n, HIT_DTYPE)
        nv = np.zeros(n, np.uint32)
        nt = np.zeros(n, np.uint32)
        lib().rto_trace_closest(C.byref(self.scene), C.byref(self.bvh), ptr(rays), n, ptr(hits),
                                ptr(nv), ptr(nt), threads)
        return (hits, nv, nt) if counts else hits

    def trace_any(self, rays: np.ndarray, threads: int = 8, counts: bool = False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = len(rays)
        occ = np.zeros(n, np.int32)
        nv = np.zeros(n, np.uint32)
        nt = np.zeros(n, np.uint32)
        lib().rto_trace_any(C.byref(self.scene), C.byref(self.bvh), ptr(rays), n, ptr(occ),
                            ptr(nv), ptr(nt), threads)
        return (occ, nv, nt) if counts else occ


def brute_closest(scene: Scene, rays: np.ndarray) -> np.ndarray:
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(rays), HIT_DTYPE)
    lib().rto_brute_closest(C.byref(scene), ptr(rays), len(rays), ptr(hits))
    return hits


def primary_rays(cam: Camera, cam_index: int, sample: int, stream: int, jitter: int) -> np.ndarray:
    rays = np.zeros(cam.width * cam.height, RAY_DTYPE)
    lib().rto_primary_rays(C.byref(cam), cam_index, sample, stream, jitter, ptr(rays))
    return rays


def render_camera(scene: Scene, bvh: OracleBvh, cam: Camera, cam_index: int, profile: Profile,
                  stream: int, threads: int = 8, rows=None):
    w, h = cam.width, cam.height
    beauty = np.zeros((h, w, 4), np.float32)
    normal = np.zeros((h, w, 4), np.float32)
    albedo = np.zeros((h, w, 4), np.float32)
    depth = np.zeros((h, w), np.float32)
    st = RenderStats()
    r0, r1 = (0, h) if rows is None else rows
    rc = lib().rto_render_camera(C.byref(scene), C.byref(bvh.bvh), C.byref(cam), cam_index,
                                 C.byref(profile), stream, ptr(beauty), ptr(normal), ptr(albedo),
                                 ptr(depth), threads, r0, r1, C.byref(st))
    if rc != 0:
        raise RuntimeError("oracle render failed")
    return {"beauty": beauty, "normal": normal, "albedo": albedo, "depth": depth,
            "paths": st.paths, "rays_closest": st.rays_closest, "rays_shadow": st.rays_shadow,
            "node_visits": st.node_visits, "prim_tests": st.prim_tests}


def generate_rig(config: int, seed: int, frame: int, frames: int):
    cams = (Camera * 4)()
    n = C.c_int()
    if lib().rto_generate_rig(config, seed, frame, frames, cams, C.byref(n)) != 0:
        raise ValueError(f"config {config} has no orbit rig")
    return [cams[i] for i in range(n.value)]


def resolve(cam: Camera, prev_cam, history_length: int, restir_enabled: bool, frame: dict,
            history: dict):
    """Temporal reprojection of one camera image in place; returns (flow, trust)."""
    w, h = cam.width, cam.height
    flow = np.zeros((h, w, 2), np.float32)
    trust = np.zeros((h, w), np.float32)
    lib().rto_resolve(C.byref(cam), C.byref(prev_cam) if prev_cam is not None else None,
                      history_length, 1 if restir_enabled else 0, ptr(frame["beauty"]),
                      ptr(frame["normal"]), ptr(frame["depth"]), ptr(history["beauty"]),
                      ptr(history["normal"]), ptr(history["depth"]), ptr(flow), ptr(trust))
    return flow, trust


def denoise(frame: dict, iterations: int = 4) -> np.ndarray:
    """The edge-aware a-trous filter over one camera image."""
    h, w = frame["depth"].shape
    out = np.zeros((h, w, 4), np.float32)
    scratch = np.zeros((h, w, 4), np.float32)
    lib().rto_denoise(w, h, iterations, ptr(np.ascontiguousarray(frame["beauty"])),
                      ptr(frame["normal"]), ptr(frame["albedo"]), ptr(frame["depth"]), ptr(out),
                      ptr(scratch))
    return out


class Images(C.Structure):
    _fields_ = [("count", C.c_int32), ("offset", C.c_int32 * 17), ("width", C.c_int32 * 16)]


def eval_texture(scene: Scene, index: int, u: float, v: float, p) -> np.ndarray:
    pp = np.asarray(p, np.float32)
    out = np.zeros(3, np.float32)
    lib().rto_eval_texture(C.byref(scene), index, u, v, ptr(pp), ptr(out))
    return out


def surface_uv(scene: Scene, o, d, prim_type: int, prim_index: int, pos):
    o, d, pos = (np.asarray(x, np.float32) for x in (o, d, pos))
    u, v = C.c_float(), C.c_float()
    lib().rto_surface_uv(C.byref(scene), ptr(o), ptr(d), prim_type, prim_index, ptr(pos),
                         C.byref(u), C.byref(v))
    return u.value, v.value


def tone_map_rgba8(rgba: np.ndarray) -> np.ndarray:
    """float4 pixels (..., 4) -> uint8 RGBA (..., 4), alpha 255."""
    src = np.ascontiguousarray(rgba, dtype=np.float32)
    out = np.empty(src.shape, np.uint8)
    lib().rto_tone_map_rgba8(ptr(src), src.size // 4, ptr(out))
    return out


def images(sizes):
    """Image table for [(width, height), ...] laid back to back."""
    im = Images()
    im.count = len(sizes)
    o = 0
    for i, (w, h) in enumerate(sizes):
        im.offset[i] = o
        im.width[i] = w
        o += w * h
    im.offset[len(sizes)] = o
    return im


def nn_loss_grad(im, params, frame, target=None, bf16=True, want_grad=True):
    """Oracle denoiser: returns (loss, grad or None, out rgba)."""
    P = im.offset[im.count]
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    beauty, normal, albedo = f32(frame["beauty"]), f32(frame["normal"]), f32(frame["albedo"])
    depth = f32(frame["depth"])
    feat = np.zeros((P, 16), np.float32)
    lib().rto_nn_features(P, ptr(beauty), ptr(normal), ptr(albedo), ptr(depth), ptr(feat))
    if not bf16:
        # the same features without bf16 rounding
        b = beauty.reshape(P, 4)
        feat[:, 0:3] = np.log1p(np.maximum(b[:, :3], 0))
        feat[:, 3:6] = np.clip(albedo.reshape(P, 4)[:, :3], 0, 1)
        feat[:, 6:9] = normal.reshape(P, 4)[:, :3] * 2 - 1
        feat[:, 9] = np.log1p(np.maximum(depth.reshape(P), 0)) * 0.125
    params = f32(params)
    out = np.zeros((P, 4), np.float32)
    grad = np.zeros(len(params), np.float32) if (want_grad and target is not None) else None
    tgt = f32(target) if target is not None else None
    loss = lib().rto_nn_loss_grad(C.byref(im), ptr(params), ptr(feat), ptr(beauty),
                                  ptr(tgt) if tgt is not None else None, ptr(out),
                                  ptr(grad) if grad is not None else None, 1 if bf16 else 0)
    return loss, grad, out


class RestirSequence:
    """Multi-frame oracle render keeping per-camera history as the renderer
    does between frames: ReSTIR reservoirs (capture_launch_history) and the
    resolved beauty / normal / depth for temporal reprojection
    (launch_resolve_kernel + copy_frame_to_history)."""

    def __init__(self, scene: Scene, bvh: "OracleBvh", profile: Profile, threads: int = 8):
        self.scene, self.bvh, self.profile, self.threads = scene, bvh, profile, threads
        self.prev = {}
        self.frame = 0
        min_l = profile.restir_min_analytic_lights if profile.restir_min_analytic_lights > 0 else 1
        self.restir = bool(profile.enable_restir_di) and scene.analytic_light_count >= min_l

    def render(self, cams, stream: int, rows=None):
        outs = []
        new_prev = {}
        for ci, cam in enumerate(cams):
            w, h = cam.width, cam.height
            res = np.zeros(w * h, RESERVOIR_DTYPE)
            io = RestirIO()
            keep = []
            prev = self.prev.get(ci) if self.frame > 0 else None
            if prev is not None:
                pc, hist, _ = prev
                keep += [pc, hist]
                io.prev_cam = C.cast(C.pointer(pc), C.c_void_p)
                io.history = hist.ctypes.data
                io.history_length = self.frame
            io.out = res.ctypes.data
            beauty = np.zeros((h, w, 4), np.float32)
            normal = np.zeros((h, w, 4), np.float32)
            albedo = np.zeros((h, w, 4), np.float32)
            depth = np.zeros((h, w), np.float32)
            st = RenderStats()
            r0, r1 = (0, h) if rows is None else rows
            rc = lib().rto_render_camera_ex(C.byref(self.scene), C.byref(self.bvh.bvh),
                                            C.byref(cam), ci, C.byref(self.profile), stream,
                                            C.byref(io), ptr(beauty), ptr(normal), ptr(albedo),
                                            ptr(depth), self.threads, r0, r1, C.byref(st))
            if rc != 0:
                raise RuntimeError("oracle render failed")
            out = {"beauty": beauty, "normal": normal, "albedo": albedo, "depth": depth,
                   "reservoirs": res}
            if prev is not None:
                out["flow"], out["trust"] = resolve(cam, prev[0], self.frame, self.restir, out,
                                                    prev[2])
            else:
                out["flow"] = np.zeros((h, w, 2), np.float32)
                out["trust"] = np.zeros((h, w), np.float32)
            cam_copy = Camera()
            C.memmove(C.byref(cam_copy), C.byref(cam), C.sizeof(Camera))
            new_prev[ci] = (cam_copy, res, {k: out[k].copy() for k in ("beauty", "normal", "depth")})
            outs.append(out)
        self.prev = new_prev
        self.frame += 1
        return outs


def rng4(pixel, sample, camera, bounce, group, stream) -> np.ndarray:
    out = np.zeros(4, np.float32)
    lib().rto_rng4(pixel, sample, camera, bounce, group, stream, ptr(out))
    return out


def sincos_turn(u: float):
    s = C.c_float()
    c = C.c_float()
    lib().rto_sincos_turn(u, C.byref(s), C.byref(c))
    return s.value, c.value


def pinhole_camera(fx, fy, cx, cy, k1=0.0, k2=0.0, k3=0.0, p1=0.0, p2=0.0, width=640,
                   height=480) -> Camera:
    """A pinhole32 camera at the origin looking down +z (identity basis)."""
    c = Camera()
    c.enabled, c.width, c.height, c.model = 1, width, height, 0
    c.basis_x[:] = (1.0, 0.0, 0.0)
    c.basis_y[:] = (0.0, 1.0, 0.0)
    c.basis_z[:] = (0.0, 0.0, 1.0)
    c.fx, c.fy, c.cx, c.cy = fx, fy, cx, cy
    c.k1, c.k2, c.k3, c.p1, c.p2 = k1, k2, k3, p1, p2
    return c


def equi_camera(width, height, fx, fy, cx, cy, radial=(0.0,) * 6, tangential=(0.0, 0.0),
                build_lut=True) -> Camera:
    """An equi62_lut1d camera (identity basis); the LUT is built by the
    oracle's restatement of make_equi62_lut1d_params unless build_lut=False."""
    c = Camera()
    c.enabled, c.width, c.height, c.model = 1, width, height, 1
    c.basis_x[:] = (1.0, 0.0, 0.0)
    c.basis_y[:] = (0.0, 1.0, 0.0)
    c.basis_z[:] = (0.0, 0.0, 1.0)
    c.fx, c.fy, c.cx, c.cy = fx, fy, cx, cy
    c.radial[:] = tuple(radial)
    c.tangential[:] = tuple(tangential)
    if build_lut:
        lib().rto_make_equi_lut(C.byref(c))
    return c


def unproject(cam: Camera, px: float, py: float) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().rto_unproject(C.byref(cam), float(px), float(py), ptr(out))
    return out


def project(cam: Camera, d) -> np.ndarray:
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros(2, np.float32)
    lib().rto_project(C.byref(cam), ptr(d), ptr(out))
    return out


def sample_light_at(scene: Scene, bvh: "OracleBvh", prim: int, pos, gn, d, r):
    """NEE sample from a surface point: (light index or -1, dir, emission, dist, pdf)."""
    a = [np.ascontiguousarray(v, np.float32) for v in (pos, gn, d, r)]
    out = np.zeros(8, np.float32)
    li = lib().rto_sample_light_at(C.byref(scene), C.byref(bvh.bvh), prim, *[ptr(v) for v in a],
                                   ptr(out))
    return li, out[0:3], out[3:6], float(out[6]), float(out[7])


def hit_material(scene: Scene, o, d, t: float, prim: int):
    """make_hit at o + t d on global primitive prim: (base colour, emission, (u, v))."""
    a = [np.ascontiguousarray(v, np.float32) for v in (o, d)]
    out = np.zeros(8, np.float32)
    lib().rto_hit_material(C.byref(scene), ptr(a[0]), ptr(a[1]), float(t), int(prim), ptr(out))
    return out[0:3], out[3:6], out[6:8]


class LightRec(C.Structure):
    _fields_ = [("type", C.c_int32), ("prim", C.c_int32), ("selection_pdf", C.c_float),
                ("cdf", C.c_float)]


def build_lights(scene: Scene):
    """The emissive light table: [(type, primitive, selection pdf, cdf)], type 0 sphere,
    1 quad, 2 triangle, 3 environment."""
    n = lib().rto_build_lights(C.byref(scene), None, 0)
    arr = (LightRec * max(n, 1))()
    lib().rto_build_lights(C.byref(scene), C.cast(arr, C.c_void_p), n)
    return [(r.type, r.prim, r.selection_pdf, r.cdf) for r in arr[:n]]


def light_pdf_for_hit(scene: Scene, bvh: "OracleBvh", prim: int, prev_mat: int, pos, sn, d) -> float:
    a = [np.ascontiguousarray(v, np.float32) for v in (pos, sn, d)]
    return float(lib().rto_light_pdf_for_hit(C.byref(scene), C.byref(bvh.bvh), prim, prev_mat,
                                             *[ptr(v) for v in a]))


def scatter(scene: Scene, prim: int, pos, gn, d, r):
    """One BSDF sample: (alive, new direction, throughput factor, pdf, delta)."""
    a = [np.ascontiguousarray(v, np.float32) for v in (pos, gn, d, r)]
    out = np.zeros(8, np.float32)
    alive = lib().rto_scatter(C.byref(scene), prim, *[ptr(v) for v in a], ptr(out))
    return bool(alive), out[0:3], out[3:6], float(out[6]), bool(out[7])
