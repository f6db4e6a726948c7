"""Bridges between the product C ABI structs and the oracle's identical structs."""
from __future__ import annotations

import ctypes as C

import numpy as np


def oracle_scene(oracle, desc):
    """Oracle Scene viewing the same host arrays as a product SceneDesc (same layout)."""
    s = oracle.Scene()
    for name in ("sphere_count", "quad_count", "triangle_count", "material_count",
                 "openpbr_count", "analytic_light_count", "background_mode"):
        setattr(s, name, getattr(desc, name))
    s.background = oracle.f3(*desc.background)
    s.background_top = oracle.f3(*desc.background_top)
    s.spheres = C.cast(desc.spheres, C.POINTER(oracle.Sphere))
    s.quads = C.cast(desc.quads, C.POINTER(oracle.Quad))
    s.triangles = C.cast(desc.triangles, C.POINTER(oracle.Triangle))
    s.materials = C.cast(desc.materials, C.POINTER(oracle.Material))
    s.openpbr = C.cast(desc.openpbr, C.POINTER(oracle.OpenPbr))
    s.analytic_lights = C.cast(desc.analytic_lights, C.POINTER(oracle.AnalyticLight))
    s.texture_count, s.image_texel_count = desc.texture_count, desc.image_texel_count
    s.textures, s.image_texels = desc.textures, desc.image_texels
    s.material_albedo_texture = desc.material_albedo_texture
    s.triangle_uvs = desc.triangle_uvs
    s.triangle_normals = desc.triangle_normals
    s.medium_count, s.media = desc.medium_count, desc.media
    s.material_emission_texture = desc.material_emission_texture
    s.openpbr_textures = desc.openpbr_textures
    return s


def oracle_camera(oracle, cam):
    c = oracle.Camera()
    C.memmove(C.byref(c), C.byref(cam), C.sizeof(c))
    return c


def oracle_profile(oracle, prof):
    p = oracle.Profile()
    C.memmove(C.byref(p), C.byref(prof), C.sizeof(p))
    return p


def product_desc(rt, arrays: dict, background=(0, 0, 0), background_top=None):
    """SceneDesc over numpy structured arrays (kept alive by the returned list)."""
    d = rt.SceneDesc()
    keep = []

    def attach(field, count_field, arr):
        if arr is None or len(arr) == 0:
            setattr(d, count_field, 0)
            setattr(d, field, None)
            return
        a = np.ascontiguousarray(arr)
        keep.append(a)
        setattr(d, count_field, len(a))
        setattr(d, field, a.ctypes.data)

    attach("spheres", "sphere_count", arrays.get("spheres"))
    attach("quads", "quad_count", arrays.get("quads"))
    attach("triangles", "triangle_count", arrays.get("triangles"))
    attach("materials", "material_count", arrays.get("materials"))
    attach("openpbr", "openpbr_count", arrays.get("openpbr"))
    attach("analytic_lights", "analytic_light_count", arrays.get("analytic"))
    attach("media", "medium_count", arrays.get("media"))
    d.background_mode = 1 if background_top is not None else 0
    d.background = (C.c_float * 3)(*background)
    d.background_top = (C.c_float * 3)(*(background_top or (0, 0, 0)))
    return d, keep


SPHERE_DT = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("material", "<i4")])
QUAD_DT = np.dtype([("origin", "<f4", 3), ("edge_u", "<f4", 3), ("edge_v", "<f4", 3),
                    ("material", "<i4")])
TRI_DT = np.dtype([("p0", "<f4", 3), ("p1", "<f4", 3), ("p2", "<f4", 3), ("material", "<i4")])
MAT_DT = np.dtype([("type", "<i4"), ("albedo", "<f4", 3), ("emission", "<f4", 3),
                   ("fuzz", "<f4"), ("ior", "<f4"), ("openpbr_index", "<i4")])
ALIGHT_DT = np.dtype([("type", "<i4"), ("position", "<f4", 3), ("basis_x", "<f4", 3),
                      ("basis_y", "<f4", 3), ("basis_z", "<f4", 3), ("radiance", "<f4", 3),
                      ("radius", "<f4"), ("width", "<f4"), ("height", "<f4"), ("length", "<f4"),
                      ("world_area", "<f4"), ("cos_theta_max", "<f4"), ("selection_pdf", "<f4"),
                      ("cdf", "<f4"), ("delta", "<i4"), ("treat_as_point", "<i4"),
                      ("treat_as_line", "<i4")])
ALIGHT_TYPES = {"sphere": 0, "disk": 1, "rect": 2, "cylinder": 3, "distant": 4, "dome": 5}
MEDIUM_DT = np.dtype([("local_center_or_min", "<f4", 3), ("radius", "<f4"), ("local_max", "<f4", 3),
                      ("density", "<f4"), ("rotation_row0", "<f4", 3), ("material_index", "<i4"),
                      ("rotation_row1", "<f4", 3), ("boundary_type", "<i4"),
                      ("rotation_row2", "<f4", 3), ("pad0", "<f4"), ("translation", "<f4", 3),
                      ("pad1", "<f4")])


def medium(kind, material, density, centre=(0, 0, 0), radius=1.0, lo=(-1, -1, -1), hi=(1, 1, 1),
           rotation=np.eye(3), translation=(0, 0, 0)):
    """One PackedMedium record: kind "sphere" (local centre, radius) or "box"
    (local corners); world -> local = rotation (p - translation)."""
    m = np.zeros(1, MEDIUM_DT)
    m["boundary_type"] = 0 if kind == "sphere" else 1
    m["local_center_or_min"] = centre if kind == "sphere" else lo
    m["radius"] = radius
    m["local_max"] = hi
    m["density"], m["material_index"] = density, material
    R = np.asarray(rotation, np.float64)
    m["rotation_row0"], m["rotation_row1"], m["rotation_row2"] = R[0], R[1], R[2]
    m["translation"] = translation
    return m


def analytic_lights(oracle, specs):
    """Analytic lights from dicts (type name, position, radiance and any of
    radius / width / height / length / basis (three columns) /
    cos_theta_max / treat_as_point / treat_as_line), compiled by the
    oracle's restatement of the light compiler."""
    a = np.zeros(len(specs), ALIGHT_DT)
    for i, sp in enumerate(specs):
        r = a[i]
        r["type"] = ALIGHT_TYPES[sp["type"]]
        r["position"] = sp.get("position", (0.0, 0.0, 0.0))
        bx, by, bz = sp.get("basis", ((1, 0, 0), (0, 1, 0), (0, 0, 1)))
        r["basis_x"], r["basis_y"], r["basis_z"] = bx, by, bz
        r["radiance"] = sp["radiance"]
        for k, d in (("radius", 0.5), ("width", 1.0), ("height", 1.0), ("length", 1.0),
                     ("cos_theta_max", 1.0), ("treat_as_point", 0), ("treat_as_line", 0)):
            r[k] = sp.get(k, d)
    assert oracle.lib().rto_analytic_lights_compile(a.ctypes.data, len(a)) == 0
    return a


def lambert_materials(n, albedo=(0.5, 0.5, 0.5)):
    m = np.zeros(n, MAT_DT)
    m["type"] = 0
    m["albedo"] = albedo
    m["ior"] = 1.0
    m["openpbr_index"] = -1
    return m


def random_triangles(rng, n, scale=10.0, size=1.0, material=0):
    t = np.zeros(n, TRI_DT)
    c = rng.uniform(-scale, scale, (n, 3)).astype(np.float32)
    t["p0"] = c + rng.uniform(-size, size, (n, 3)).astype(np.float32)
    t["p1"] = c + rng.uniform(-size, size, (n, 3)).astype(np.float32)
    t["p2"] = c + rng.uniform(-size, size, (n, 3)).astype(np.float32)
    t["material"] = material
    return t


def random_rays(rng, n, origin_scale=15.0, target_scale=10.0):
    from oracle_ctypes import RAY_DTYPE
    r = np.zeros(n, RAY_DTYPE)
    o = rng.uniform(-origin_scale, origin_scale, (n, 3)).astype(np.float32)
    t = rng.uniform(-target_scale, target_scale, (n, 3)).astype(np.float32)
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["o"] = o
    r["d"] = d.astype(np.float32)
    r["tmin"] = 1e-3
    r["tmax"] = 1e30
    return r


def rel_mse(x, ref, eps=1e-2):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))


def threshold_neighbourhood(ulps: int = 64) -> np.ndarray:
    """Every float within `ulps` of each input where the output steps."""
    k = np.arange(255, dtype=np.float64)
    t = np.float32(((k + 0.5) / 255.0) ** 2)
    bits = t.view(np.int32)[:, None] + np.arange(-ulps, ulps + 1, dtype=np.int32)[None, :]
    return bits.ravel().view(np.float32)


def sweep() -> np.ndarray:
    rng = np.random.default_rng(5)
    dense = np.linspace(-2.0, 8.0, 2_000_001, dtype=np.float32)
    rand = rng.uniform(0.0, 1.0, 1_000_000).astype(np.float32)
    special = np.array([0.0, -0.0, 1e-45, 1e-38, 0.998001, 0.999, 1.0, 3.4e38], np.float32)
    return np.concatenate([threshold_neighbourhood(), dense, rand, special])


TEX_DT = np.dtype([("type", "<i4"), ("even_texture", "<i4"), ("odd_texture", "<i4"),
                   ("image_offset", "<i4"), ("image_width", "<i4"), ("image_height", "<i4"),
                   ("u_address_mode", "<i4"), ("v_address_mode", "<i4"), ("filter_type", "<i4"),
                   ("scale", "<f4"), ("color", "<f4", 3)])
PBR_TEX_DT = np.dtype([(n, "<i4") for n in (
    "base_color", "base_color_space", "specular_color", "specular_color_space",
    "transmission_color", "transmission_color_space", "emission_color", "emission_color_space",
    "base_metalness", "specular_roughness")])


def texture(kind, color=(0, 0, 0), even=-1, odd=-1, scale=1.0, offset=0, w=0, h=0, um=1, vm=1,
            filt=0):
    """One PackedTexture record: kind "constant", "checker", "image" or "noise"."""
    t = np.zeros(1, TEX_DT)
    t["type"] = {"constant": 0, "checker": 1, "image": 2, "noise": 3}[kind]
    t["color"], t["even_texture"], t["odd_texture"], t["scale"] = color, even, odd, scale
    t["image_offset"], t["image_width"], t["image_height"] = offset, w, h
    t["u_address_mode"], t["v_address_mode"], t["filter_type"] = um, vm, filt
    return t


def pbr_bindings(n, **kw):
    """n OpenPBR texture-binding records, all unbound (-1, linear sRGB) but for
    the per-material overrides kw[name] = [(material, texture[, space]), ...]."""
    b = np.zeros(n, PBR_TEX_DT)
    for name in ("base_color", "specular_color", "transmission_color", "emission_color",
                 "base_metalness", "specular_roughness"):
        b[name] = -1
    for name in ("base_color_space", "specular_color_space", "transmission_color_space",
                 "emission_color_space"):
        b[name] = 1
    for name, items in kw.items():
        for it in items:
            b[name][it[0]] = it[1]
            if len(it) > 2:
                b[name + "_space"][it[0]] = it[2]
    return b


def attach_textures(scene, keep, textures=None, texels=None, albedo=None, emission=None,
                    pbr=None, triangle_uvs=None):
    """Sets the texture fields of an oracle Scene or a product SceneDesc (same
    layout) over numpy arrays, which `keep` keeps alive."""
    def ref(a, dtype=None):
        a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
        keep.append(a)
        return a.ctypes.data
    if textures is not None:
        t = np.concatenate(textures) if isinstance(textures, list) else textures
        scene.texture_count, scene.textures = len(t), ref(t)
    if texels is not None:
        px = np.asarray(texels, np.float32).reshape(-1, 3)
        scene.image_texel_count, scene.image_texels = len(px), ref(px)
    if albedo is not None:
        scene.material_albedo_texture = ref(albedo, np.int32)
    if emission is not None:
        scene.material_emission_texture = ref(emission, np.int32)
    if pbr is not None:
        scene.openpbr_textures = ref(pbr)
    if triangle_uvs is not None:
        scene.triangle_uvs = ref(triangle_uvs, np.float32)
