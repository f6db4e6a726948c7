"""Scene ingest: the reference's YAML scene files (format_version 1, the
`assets/scenes/*/scene.yaml` family) to this library's C ABI scene
description and camera rig.

Parsing follows rt::scene::load_scene_definition
(src/scene/yaml_scene_loader.cpp:575-709): `includes` are loaded first and
share the id tables of the including file (duplicate ids are errors, an
include cycle is an error); `scene.textures`, `scene.materials`,
`scene.shapes`, `scene.instances` and `scene.media` are parsed in that order
(:215-395); `imports` of type obj_mtl append an OBJ/MTL file's triangles
with one diffuse material per MTL material (:583-619,
src/scene/obj_mtl_importer.cpp:86-230); `cpu_presets` and
`realtime.default_view` give the cameras (:396-528).  Every error message
is prefixed by the scene file's path, as the reference's loader does.

Flattening into primitives follows the realtime scene adapter
(src/scene/realtime_scene_adapter.cpp:24-109): an instance's transform maps
points by rotation * p + translation and vectors by the rotation; a box
becomes six quads in the adapter's face order.  Cameras follow
compile_realtime_camera_transform / compile_cpu_camera_transform
(src/scene/scene_definition.cpp:86-139) and the body pose of
src/realtime/viewer/body_pose.cpp:36-85; the camera axes are x right,
y down, z forward, as the renderer's.

Albedo and emission textures of every type (constant, checker, image,
noise) are rendered, and so are participating media (`scene.media` over
sphere or box shapes, isotropic materials); media bounded by quads or meshes
raise SceneFileError when a scene is flattened.

    python -m ray_tracing_amd.scene_yaml path/to/scene.yaml --spp 16 --out frame.npy
"""
from __future__ import annotations

import ctypes as C
import math
import os
import sys
from dataclasses import dataclass, field

import numpy as np
import yaml

from . import abi

MAT_LAMBERTIAN, MAT_METAL, MAT_DIELECTRIC, MAT_DIFFUSE_LIGHT, MAT_ISOTROPIC = 0, 1, 2, 3, 4
CAM_PINHOLE32, CAM_EQUI62_LUT1D = 0, 1


class SceneFileError(RuntimeError):
    """A scene file could not be loaded; the message starts with its path."""


# --------------------------------------------------------------- definition

@dataclass
class Transform:
    rotation: np.ndarray = field(default_factory=lambda: np.eye(3))
    translation: np.ndarray = field(default_factory=lambda: np.zeros(3))

    def point(self, p) -> np.ndarray:
        return self.rotation @ np.asarray(p, np.float64) + self.translation

    def vector(self, v) -> np.ndarray:
        return self.rotation @ np.asarray(v, np.float64)

    def normal(self, n) -> np.ndarray:
        """transform_normal (src/scene/realtime_scene_adapter.cpp:112-120): the
        inverse transpose of the linear part applied to `n`, normalized; a
        singular transform or a zero-length or non-finite normal is refused."""
        with np.errstate(all="ignore"):
            try:
                m = np.linalg.inv(self.rotation).T
            except np.linalg.LinAlgError:
                m = np.full((3, 3), np.nan)
            t = m @ np.asarray(n, np.float64)
            length = float(np.sqrt(t @ t))
        if not np.isfinite(length) or length <= 1e-12:
            raise ValueError("SceneIR v2 contains a singular or zero-length mesh normal")
        return t / length


@dataclass
class CameraSpec:
    model: int
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    T_bc: Transform
    pinhole32: dict = field(default_factory=lambda: dict(k1=0.0, k2=0.0, k3=0.0, p1=0.0, p2=0.0))
    radial: list = field(default_factory=lambda: [0.0] * 6)
    tangential: list = field(default_factory=lambda: [0.0, 0.0])


@dataclass
class CpuPreset:
    preset_id: str
    samples_per_pixel: int
    camera: CameraSpec
    lookfrom: np.ndarray
    lookat: np.ndarray
    vup: np.ndarray
    max_depth: int


@dataclass
class RealtimePreset:
    position: np.ndarray
    yaw_deg: float
    pitch_deg: float
    frame_convention: str
    camera: CameraSpec
    base_move_speed: float


@dataclass
class SceneDefinition:
    id: str = ""
    label: str = ""
    background: np.ndarray = field(default_factory=lambda: np.zeros(3))
    textures: list = field(default_factory=list)     # dicts with "type"
    materials: list = field(default_factory=list)    # dicts with "type"
    shapes: list = field(default_factory=list)       # dicts with "type"
    instances: list = field(default_factory=list)    # (shape, material, Transform)
    media: list = field(default_factory=list)
    cpu_presets: list = field(default_factory=list)
    realtime: RealtimePreset | None = None
    dependencies: list = field(default_factory=list)

    @property
    def supports_cpu_render(self) -> bool:
        return bool(self.cpu_presets)

    @property
    def supports_realtime(self) -> bool:
        return self.realtime is not None


# ------------------------------------------------------------------ parsing

def _map(node, name):
    if node is not None and not isinstance(node, dict):
        raise ValueError(f"{name} must be a map")
    return node


def _seq(node, name):
    if node is not None and not isinstance(node, list):
        raise ValueError(f"{name} must be a sequence")
    return node


def _vec(node, n=3):
    if not isinstance(node, list) or len(node) != n:
        raise ValueError(f"expected a {n}-element vector")
    return np.array([float(v) for v in node], np.float64)


def _mat3(node):
    if not isinstance(node, list) or len(node) != 3:
        raise ValueError("expected a 3x3 matrix")
    m = np.eye(3)
    for r, row in enumerate(node):
        if not isinstance(row, list) or len(row) != 3:
            raise ValueError("expected a 3x3 matrix")
        m[r] = [float(v) for v in row]
    return m


def _transform(node) -> Transform:
    t = Transform()
    if node is None:
        return t
    _map(node, "transform")
    if node.get("translation") is not None:
        t.translation = _vec(node["translation"])
    if node.get("rotation") is not None:
        t.rotation = _mat3(node["rotation"])
    return t


def _isometry(node, name) -> Transform:
    if node is None:
        raise ValueError(f"{name} must be a map")
    return _transform(_map(node, name))


def _unique(ids, key, kind):
    if key in ids:
        raise ValueError(f"duplicate {kind} id: {key}")


def _require(ids, key, kind):
    if key not in ids:
        raise ValueError(f"unknown {kind} id: {key}")
    return ids[key]


def _camera_spec(node) -> CameraSpec:
    """parse_camera_spec (yaml_scene_loader.cpp:147-207)"""
    if node is None:
        raise ValueError("camera is required")
    _map(node, "camera")
    if any(k in node for k in ("vfov", "vfov_deg", "use_default_viewer_intrinsics")):
        raise ValueError("legacy camera fields are not supported; use explicit camera model and intrinsics")
    model = node.get("model")
    if model is None:
        raise ValueError("camera.model is required")
    if model == "pinhole32":
        m = CAM_PINHOLE32
    elif model == "equi62_lut1d":
        m = CAM_EQUI62_LUT1D
    else:
        raise ValueError(f"unsupported camera model: {model}")
    if any(node.get(k) is None for k in ("width", "height", "fx", "fy", "cx", "cy")):
        raise ValueError("camera requires width, height, fx, fy, cx, and cy")
    spec = CameraSpec(m, int(node["width"]), int(node["height"]), float(node["fx"]), float(node["fy"]),
                      float(node["cx"]), float(node["cy"]), _isometry(node.get("T_bc"), "camera.T_bc"))
    ph = node.get("pinhole32")
    if ph is not None:
        _map(ph, "camera.pinhole32")
        for k in ("k1", "k2", "k3", "p1", "p2"):
            if ph.get(k) is not None:
                spec.pinhole32[k] = float(ph[k])
    eq = node.get("equi62_lut1d")
    if eq is not None:
        _map(eq, "camera.equi62_lut1d")
        if eq.get("radial") is not None:
            _seq(eq["radial"], "camera.equi62_lut1d.radial")
            if len(eq["radial"]) != 6:
                raise ValueError("camera.equi62_lut1d.radial must have 6 entries")
            spec.radial = [float(v) for v in eq["radial"]]
        if eq.get("tangential") is not None:
            spec.tangential = list(_vec(eq["tangential"], 2))
    return spec


def _textures(node, ids, out: SceneDefinition, scene_dir):
    """parse_textures (yaml_scene_loader.cpp:209-262)"""
    if node is None:
        return
    _map(node, "scene.textures")
    for key, t in node.items():
        key = str(key)
        _unique(ids, key, "texture")
        _map(t, "texture")
        kind = t.get("type")
        if kind == "constant":
            desc = {"type": "constant", "color": _vec(t["color"])}
        elif kind == "checker":
            desc = {"type": "checker", "scale": float(t.get("scale", 1.0)),
                    "even": _require(ids, str(t["even"]), "texture"),
                    "odd": _require(ids, str(t["odd"]), "texture")}
        elif kind == "image":
            p = str(t["path"])
            desc = {"type": "image", "path": os.path.normpath(os.path.join(scene_dir, p))}
        elif kind == "noise":
            desc = {"type": "noise", "scale": float(t.get("scale", 1.0))}
        else:
            raise ValueError(f"unsupported texture type: {kind}")
        ids[key] = len(out.textures)
        out.textures.append(desc)


def _materials(node, tex_ids, ids, out: SceneDefinition):
    """parse_materials (yaml_scene_loader.cpp:264-310)"""
    if node is None:
        return
    _map(node, "scene.materials")
    for key, m in node.items():
        key = str(key)
        _unique(ids, key, "material")
        _map(m, "material")
        kind = m.get("type")
        if kind == "diffuse":
            desc = {"type": "diffuse", "albedo": _require(tex_ids, str(m["albedo"]), "texture")}
        elif kind == "metal":
            desc = {"type": "metal", "albedo": _require(tex_ids, str(m["albedo"]), "texture"),
                    "fuzz": float(m.get("fuzz", 0.0))}
        elif kind == "dielectric":
            desc = {"type": "dielectric", "ior": float(m.get("ior", 1.0))}
        elif kind == "emissive":
            desc = {"type": "emissive", "emission": _require(tex_ids, str(m["emission"]), "texture")}
        elif kind == "isotropic":
            desc = {"type": "isotropic", "albedo": _require(tex_ids, str(m["albedo"]), "texture")}
        else:
            raise ValueError(f"unsupported material type: {kind}")
        ids[key] = len(out.materials)
        out.materials.append(desc)


def _shapes(node, ids, out: SceneDefinition):
    """parse_shapes (yaml_scene_loader.cpp:312-350)"""
    if node is None:
        return
    _map(node, "scene.shapes")
    for key, s in node.items():
        key = str(key)
        _unique(ids, key, "shape")
        _map(s, "shape")
        kind = s.get("type")
        if kind == "sphere":
            desc = {"type": "sphere", "center": _vec(s["center"]), "radius": float(s["radius"])}
        elif kind == "quad":
            desc = {"type": "quad", "origin": _vec(s["origin"]), "edge_u": _vec(s["edge_u"]),
                    "edge_v": _vec(s["edge_v"])}
        elif kind == "box":
            lo = s.get("min_corner", s.get("min"))
            hi = s.get("max_corner", s.get("max"))
            desc = {"type": "box", "min": _vec(lo), "max": _vec(hi)}
        else:
            raise ValueError(f"unsupported shape type: {kind}")
        ids[key] = len(out.shapes)
        out.shapes.append(desc)


def _instances(node, shape_ids, mat_ids, out: SceneDefinition):
    """parse_instances (yaml_scene_loader.cpp:352-366)"""
    if node is None:
        return
    _seq(node, "scene.instances")
    for inst in node:
        _map(inst, "instance")
        out.instances.append((_require(shape_ids, str(inst["shape"]), "shape"),
                              _require(mat_ids, str(inst["material"]), "material"),
                              _transform(inst.get("transform"))))


def _media(node, shape_ids, mat_ids, ids, out: SceneDefinition):
    """parse_media (yaml_scene_loader.cpp:368-388)"""
    if node is None:
        return
    _map(node, "scene.media")
    for key, m in node.items():
        key = str(key)
        _unique(ids, key, "medium")
        ids[key] = 0
        _map(m, "medium")
        out.media.append({"shape": _require(shape_ids, str(m["shape"]), "shape"),
                          "material": _require(mat_ids, str(m["material"]), "material"),
                          "density": float(m["density"]),
                          "transform": _transform(m.get("transform"))})


def _cpu_presets(node, ids, out: SceneDefinition):
    """parse_cpu_presets / parse_camera_preset (yaml_scene_loader.cpp:390-448)"""
    if node is None:
        return
    _map(node, "cpu_presets")
    for key, p in node.items():
        key = str(key)
        _unique(ids, key, "cpu preset")
        ids[key] = 0
        _map(p, "cpu preset")
        if p.get("camera") is None:
            raise ValueError("cpu preset camera is required")
        cam = p["camera"]
        spec = _camera_spec(cam)
        out.cpu_presets.append(CpuPreset(
            key, int(p.get("samples_per_pixel", 0)), spec,
            _vec(cam["lookfrom"]) if cam.get("lookfrom") is not None else np.zeros(3),
            _vec(cam["lookat"]) if cam.get("lookat") is not None else np.array([0.0, 0.0, -1.0]),
            _vec(cam["vup"]) if cam.get("vup") is not None else np.array([0.0, 1.0, 0.0]),
            int(cam.get("max_depth", 50))))


def _realtime(node, out: SceneDefinition):
    """parse_realtime_section / parse_realtime_preset (yaml_scene_loader.cpp:450-528)"""
    if node is None:
        return
    if not isinstance(node, dict):
        raise ValueError("realtime must be a map")
    view = node.get("default_view")
    if view is None:
        return
    if out.realtime is not None:
        raise ValueError("duplicate realtime preset")
    _map(view, "default_view")
    pos, yaw, pitch = np.array([0.0, -0.8, 0.35]), 0.0, 0.0   # default_spawn_pose
    pose = view.get("initial_body_pose")
    if pose is not None:
        _map(pose, "initial_body_pose")
        if pose.get("position") is not None:
            pos = _vec(pose["position"])
        yaw = float(pose.get("yaw_deg", 0.0))
        pitch = float(pose.get("pitch_deg", 0.0))
    conv = view.get("frame_convention", "world_z_up")
    if conv == "world_y_up":
        conv = "legacy_y_up"
    if conv not in ("world_z_up", "legacy_y_up"):
        raise ValueError(f"unsupported frame convention: {conv}")
    if view.get("camera") is None:
        raise ValueError("realtime default_view.camera is required")
    out.realtime = RealtimePreset(pos, yaw, pitch, conv, _camera_spec(view["camera"]),
                                  float(view.get("base_move_speed", 0.0)))


def _parse_mtl(path):
    """newmtl / Kd / map_Kd of one MTL file.  A material without Kd is black
    (tinyobj's InitMaterial zeroes the diffuse colour), and map_Kd is
    resolved against the directory of the MTL file that declares it
    (obj_mtl_importer.cpp:94-99)."""
    mats, cur = {}, None
    mtl_dir = os.path.dirname(path)
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if tok[0] == "newmtl":
                cur = " ".join(tok[1:])
                mats[cur] = {"Kd": np.zeros(3), "map_Kd": None}
            elif cur is not None and tok[0] == "Kd":
                mats[cur]["Kd"] = np.array([float(v) for v in tok[1:4]])
            elif cur is not None and tok[0] == "map_Kd":
                mats[cur]["map_Kd"] = os.path.normpath(os.path.join(mtl_dir, tok[-1]))
    return mats


def _obj_index(text, count, what):
    """An OBJ index: 1-based from the start, or negative from the end of the
    elements read so far (-1 is the last), as tinyobj resolves them; 0 or
    beyond the elements read is an error."""
    i = int(text)
    if i > 0:
        return i - 1
    if i < 0 and count + i >= 0:
        return count + i
    raise ValueError(f"obj {what} index {i} out of range")


def import_obj_mtl(path, out: SceneDefinition):
    """import_obj_mtl (src/scene/obj_mtl_importer.cpp:118-230): positions and
    polygon faces (fan-triangulated) grouped by material; one diffuse
    material per group (Kd, or an image texture for map_Kd, or white
    without a material); appended with an identity transform.  A group keeps
    its `vt` texture coordinates (and `vn` normals) only when every corner
    of every triangle in it has one (obj_mtl_importer.cpp:202-224).  Normals
    are normalized when flattened (Transform.normal)."""
    path = os.path.normpath(path)
    root = os.path.dirname(path)
    verts, uvs, norms, groups, order, mats, cur = [], [], [], {}, [], {}, -1
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if tok[0] == "v":
                verts.append([float(v) for v in tok[1:4]])
            elif tok[0] == "vt":
                uvs.append([float(v) for v in tok[1:3]])
            elif tok[0] == "vn":
                norms.append([float(v) for v in tok[1:4]])
            elif tok[0] == "mtllib":
                mp = os.path.normpath(os.path.join(root, " ".join(tok[1:])))
                mats.update(_parse_mtl(mp))
                out.dependencies.append(mp)
            elif tok[0] == "usemtl":
                # material index = declaration order in the MTL files (-1: unknown)
                name = " ".join(tok[1:])
                cur = list(mats).index(name) if name in mats else -1
            elif tok[0] == "f":
                idx, tidx, nidx = [], [], []
                for c in tok[1:]:
                    parts = c.split("/")
                    idx.append(_obj_index(parts[0], len(verts), "vertex"))
                    tidx.append(_obj_index(parts[1], len(uvs), "texture coordinate")
                                if len(parts) > 1 and parts[1] else -1)
                    nidx.append(_obj_index(parts[2], len(norms), "normal")
                                if len(parts) > 2 and parts[2] else -1)
                if len(idx) < 3:
                    raise ValueError("only triangulated faces are supported")
                if cur not in groups:
                    groups[cur] = []
                    order.append(cur)
                for k in range(1, len(idx) - 1):
                    groups[cur].append((idx[0], idx[k], idx[k + 1], tidx[0], tidx[k], tidx[k + 1],
                                        nidx[0], nidx[k], nidx[k + 1]))
    V = np.array(verts, np.float64).reshape(-1, 3)
    UV = np.array(uvs, np.float64).reshape(-1, 2)
    N = np.array(norms, np.float64).reshape(-1, 3)
    for gi in sorted(order):
        rows = np.array(groups[gi], np.int64).reshape(-1, 9)
        tris, tuv, tn = rows[:, :3], rows[:, 3:6], rows[:, 6:]
        if tris.size and (tris.max() >= len(V)):
            raise ValueError("obj face index out of range")
        if tuv.size and tuv.max() >= len(UV):
            raise ValueError("obj texture coordinate index out of range")
        if tn.size and tn.max() >= len(N):
            raise ValueError("obj normal index out of range")
        texcoords = UV[tuv] if len(UV) and tuv.size and (tuv >= 0).all() else None
        normals = N[tn] if len(N) and tn.size and (tn >= 0).all() else None
        m = mats[list(mats)[gi]] if gi >= 0 else None
        if m is not None and m["map_Kd"] is not None:
            out.textures.append({"type": "image", "path": m["map_Kd"]})
        else:
            out.textures.append({"type": "constant", "color": m["Kd"] if m is not None else np.ones(3)})
        out.materials.append({"type": "diffuse", "albedo": len(out.textures) - 1})
        out.shapes.append({"type": "mesh", "positions": V, "triangles": tris, "texcoords": texcoords,
                           "normals": normals})
        out.instances.append((len(out.shapes) - 1, len(out.materials) - 1, Transform()))


def _imports(node, scene_dir, out: SceneDefinition):
    """parse_imports (yaml_scene_loader.cpp:583-619)"""
    if node is None:
        return
    _map(node, "imports")
    for _, imp in node.items():
        _map(imp, "import")
        kind = imp.get("type")
        if kind != "obj_mtl":
            raise ValueError(f"unsupported import type: {kind}")
        if imp.get("mtl") is not None:
            raise ValueError("explicit mtl overrides are not supported")
        p = str(imp["obj"])
        if not os.path.isabs(p):
            p = os.path.join(scene_dir, p)
        try:
            import_obj_mtl(p, out)
        except (ValueError, OSError) as e:
            raise ValueError(f"{os.path.normpath(p)}: {e}") from None
        out.dependencies.append(os.path.normpath(p))


def _load_file(path, require_version, parse_meta, out, ids, active):
    path = os.path.normpath(path)
    try:
        if path in active:
            raise ValueError("include cycle detected")
        active.add(path)
        with open(path) as f:
            root = yaml.safe_load(f)
        if root is None:
            root = {}
        _map(root, "root")
        if root.get("format_version") is not None:
            if int(root["format_version"]) != 1:
                raise ValueError("unsupported format_version")
        elif require_version:
            raise ValueError("unsupported format_version")
        if path not in out.dependencies:
            out.dependencies.append(path)
        scene = root.get("scene")
        _map(scene, "scene")
        if parse_meta:
            if scene is None:
                raise ValueError("scene must be a map")
            out.id = str(scene["id"])
            out.label = str(scene["label"])
            if scene.get("background") is not None:
                out.background = _vec(scene["background"])
        includes = root.get("includes")
        _seq(includes, "includes")
        for inc in includes or []:
            if not isinstance(inc, str):
                raise ValueError("include must be a scalar path")
            p = inc if os.path.isabs(inc) else os.path.join(os.path.dirname(path), inc)
            _load_file(p, False, False, out, ids, active)
        d = os.path.dirname(path)
        if scene is not None:
            _textures(scene.get("textures"), ids["texture"], out, d)
            _materials(scene.get("materials"), ids["texture"], ids["material"], out)
            _shapes(scene.get("shapes"), ids["shape"], out)
            _instances(scene.get("instances"), ids["shape"], ids["material"], out)
            _media(scene.get("media"), ids["shape"], ids["material"], ids["medium"], out)
        _imports(root.get("imports"), d, out)
        _cpu_presets(root.get("cpu_presets"), ids["preset"], out)
        _realtime(root.get("realtime"), out)
        active.discard(path)
    except SceneFileError:
        active.discard(path)
        raise
    except (ValueError, KeyError, TypeError, OSError, yaml.YAMLError) as e:
        active.discard(path)
        msg = f"missing field {e}" if isinstance(e, KeyError) else str(e)
        raise SceneFileError(f"{path}: {msg}") from None


def load_scene_definition(path) -> SceneDefinition:
    """Loads a scene file and its includes and imports."""
    out = SceneDefinition()
    ids = {k: {} for k in ("texture", "material", "shape", "medium", "preset")}
    _load_file(str(path), True, True, out, ids, set())
    return out


# --------------------------------------------------------------- flattening

TEX_TYPES = {"constant": 0, "checker": 1, "image": 2, "noise": 3}
NO_UV = np.full(6, np.nan)      # a triangle without texture coordinates
NO_N = np.full(9, np.nan)       # a triangle without vertex normals


def _albedo(defn: SceneDefinition, tex: int, out: abi.Material) -> int:
    """A constant albedo is stored in the material (binding -1); any other
    texture is bound by its index in the scene's texture table."""
    t = defn.textures[tex]
    if t["type"] == "constant":
        out.albedo = abi.f3(*t["color"])
        return -1
    return tex


def _emission(defn: SceneDefinition, tex: int, out: abi.Material) -> int:
    """A constant emission is stored in the material (binding -1); any other
    texture is bound by its index (MaterialSample::emission_texture)."""
    t = defn.textures[tex]
    if t["type"] == "constant":
        out.emission = abi.f3(*t["color"])
        return -1
    return tex


def _material(defn: SceneDefinition, m: dict, path: str):
    """(C ABI material, albedo texture binding, emission texture binding)"""
    out = abi.Material()
    out.openpbr_index = -1
    out.ior = 1.0
    binding = -1
    emission = -1
    if m["type"] == "diffuse":
        out.type = MAT_LAMBERTIAN
        binding = _albedo(defn, m["albedo"], out)
    elif m["type"] == "metal":
        out.type = MAT_METAL
        binding = _albedo(defn, m["albedo"], out)
        out.fuzz = m["fuzz"]
    elif m["type"] == "dielectric":
        out.type = MAT_DIELECTRIC
        out.ior = m["ior"]
    elif m["type"] == "emissive":
        out.type = MAT_DIFFUSE_LIGHT
        emission = _emission(defn, m["emission"], out)
    elif m["type"] == "isotropic":
        # IsotropicVolumeMaterial (realtime_scene_adapter.cpp:67-68): the
        # phase function's albedo, constant or textured
        out.type = MAT_ISOTROPIC
        binding = _albedo(defn, m["albedo"], out)
    else:
        raise SceneFileError(f"{path}: material type {m['type']} is not supported by this build")
    return out, binding, emission


def load_image_rgb(path: str) -> np.ndarray:
    """load_texture_image_rgb32f (packed_scene_preparation.cpp:126-147):
    8-bit RGB scaled by 1/255, rows top to bottom."""
    try:
        from PIL import Image
        with Image.open(path) as im:
            return np.asarray(im.convert("RGB"), np.float32) / np.float32(255.0)
    except (OSError, ValueError) as e:
        raise SceneFileError(f"{path}: failed to load texture image ({e})") from None


def _texture_table(defn: SceneDefinition):
    """pack_texture (packed_scene_preparation.cpp:149-190): one rt_texture
    per scene texture, in definition order; image texels back to back."""
    table, texels, offset = [], [], 0
    for t in defn.textures:
        r = abi.Texture()
        r.type = TEX_TYPES[t["type"]]
        r.even_texture = r.odd_texture = -1
        r.u_address_mode = r.v_address_mode = 1      # clamp
        r.filter_type = 0                            # closest (ImageTextureDesc defaults)
        r.scale = 1.0
        if t["type"] == "constant":
            r.color = abi.f3(*t["color"])
        elif t["type"] == "checker":
            r.scale, r.even_texture, r.odd_texture = t["scale"], t["even"], t["odd"]
        elif t["type"] == "noise":
            r.scale = t["scale"]
        else:
            img = load_image_rgb(t["path"])
            r.image_offset = offset
            r.image_height, r.image_width = img.shape[:2]
            texels.append(img.reshape(-1, 3))
            offset += img.shape[0] * img.shape[1]
        table.append(r)
    flat = np.ascontiguousarray(np.concatenate(texels) if texels else np.zeros((0, 3), np.float32))
    return table, flat


def box_quads(lo, hi, tf: Transform):
    """The six faces of a box (add_box_quads, realtime_scene_adapter.cpp:76-109):
    (origin, edge_u, edge_v) after the instance transform."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)

    def c(x, y, z):
        return np.array([(lo, hi)[x][0], (lo, hi)[y][1], (lo, hi)[z][2]])

    p000, p001, p010, p011 = c(0, 0, 0), c(0, 0, 1), c(0, 1, 0), c(0, 1, 1)
    p100, p101, p110, p111 = c(1, 0, 0), c(1, 0, 1), c(1, 1, 0), c(1, 1, 1)
    faces = [(p001, p101 - p001, p011 - p001), (p100, p000 - p100, p110 - p100),
             (p000, p001 - p000, p010 - p000), (p101, p100 - p101, p111 - p101),
             (p010, p110 - p010, p011 - p010), (p000, p100 - p000, p001 - p000)]
    return [(tf.point(o), tf.vector(u), tf.vector(v)) for o, u, v in faces]


@dataclass
class FlatScene:
    """Host arrays of a flattened scene; `desc` points into them."""
    desc: abi.SceneDesc
    spheres: object
    quads: object
    triangles: object
    materials: object
    textures: object = None
    texels: object = None
    albedo_bindings: object = None
    emission_bindings: object = None
    triangle_uvs: object = None
    triangle_normals: object = None
    media: object = None


def flatten(defn: SceneDefinition) -> FlatScene:
    """Primitives of every instance in instance order (spheres, quads and
    triangles each in their own array), one C ABI material per scene
    material, the scene background as a constant background."""
    path = defn.dependencies[0] if defn.dependencies else defn.id
    media = _media_records(defn, path)
    mats, bindings, emissions = [], [], []
    for m in defn.materials:
        mat, binding, emission = _material(defn, m, path)
        mats.append(mat)
        bindings.append(binding)
        emissions.append(emission)
    spheres, quads, tris, tri_uv, tri_n = [], [], [], [], []
    for shape_i, mat_i, tf in defn.instances:
        s = defn.shapes[shape_i]
        if s["type"] == "sphere":
            spheres.append(abi.Sphere(abi.f3(*tf.point(s["center"])), s["radius"], mat_i))
        elif s["type"] == "quad":
            quads.append(abi.Quad(abi.f3(*tf.point(s["origin"])), abi.f3(*tf.vector(s["edge_u"])),
                                  abi.f3(*tf.vector(s["edge_v"])), mat_i))
        elif s["type"] == "box":
            for o, u, v in box_quads(s["min"], s["max"], tf):
                quads.append(abi.Quad(abi.f3(*o), abi.f3(*u), abi.f3(*v), mat_i))
        else:
            P = s["positions"]
            tc = s.get("texcoords")
            for k, (a, b, c) in enumerate(s["triangles"]):
                tris.append(abi.Triangle(abi.f3(*tf.point(P[a])), abi.f3(*tf.point(P[b])),
                                         abi.f3(*tf.point(P[c])), mat_i))
                tri_uv.append(tc[k].reshape(6) if tc is not None else NO_UV)
                vn = s.get("normals")
                tri_n.append(np.concatenate([tf.normal(x) for x in vn[k]]) if vn is not None else NO_N)

    def arr(T, items):
        return (T * len(items))(*items) if items else None

    fs = FlatScene(abi.SceneDesc(), arr(abi.Sphere, spheres), arr(abi.Quad, quads),
                   arr(abi.Triangle, tris), arr(abi.Material, mats))
    d = fs.desc
    d.sphere_count, d.quad_count, d.triangle_count = len(spheres), len(quads), len(tris)
    d.material_count = len(mats)
    d.spheres = C.cast(fs.spheres, C.c_void_p) if fs.spheres is not None else None
    d.quads = C.cast(fs.quads, C.c_void_p) if fs.quads is not None else None
    d.triangles = C.cast(fs.triangles, C.c_void_p) if fs.triangles is not None else None
    d.materials = C.cast(fs.materials, C.c_void_p) if fs.materials is not None else None
    d.background_mode = 0
    d.background = abi.f3(*defn.background)
    if any(b >= 0 for b in bindings + emissions):
        table, texels = _texture_table(defn)
        fs.textures = (abi.Texture * len(table))(*table)
        fs.texels = texels
        d.texture_count, d.image_texel_count = len(table), texels.shape[0]
        d.textures = C.cast(fs.textures, C.c_void_p)
        d.image_texels = texels.ctypes.data_as(C.c_void_p) if texels.size else None
        if any(b >= 0 for b in bindings):
            fs.albedo_bindings = np.asarray(bindings, np.int32)
            d.material_albedo_texture = fs.albedo_bindings.ctypes.data_as(C.c_void_p)
        if any(b >= 0 for b in emissions):
            fs.emission_bindings = np.asarray(emissions, np.int32)
            d.material_emission_texture = fs.emission_bindings.ctypes.data_as(C.c_void_p)
    if any(not np.isnan(u[0]) for u in tri_uv):
        fs.triangle_uvs = np.ascontiguousarray(np.array(tri_uv, np.float32).reshape(-1, 6))
        d.triangle_uvs = fs.triangle_uvs.ctypes.data_as(C.c_void_p)
    if any(not np.isnan(n[0]) for n in tri_n):
        fs.triangle_normals = np.ascontiguousarray(np.array(tri_n, np.float32).reshape(-1, 9))
        d.triangle_normals = fs.triangle_normals.ctypes.data_as(C.c_void_p)
    if media:
        fs.media = (abi.Medium * len(media))(*media)
        d.medium_count = len(media)
        d.media = C.cast(fs.media, C.c_void_p)
    return fs


def _media_records(defn: SceneDefinition, path: str) -> list:
    """HomogeneousMediumPrimitive / pack_medium (realtime_scene_adapter.cpp:
    450-481, packed_scene_preparation.cpp:79-92): a sphere or box shape in
    the medium's own frame, world -> local = the transform's rotation
    transposed applied to p - translation; quad and mesh boundaries are
    refused with the reference's messages."""
    out = []
    for m in defn.media:
        shape = defn.shapes[m["shape"]]
        tf = m["transform"]
        rec = abi.Medium()
        rec.material_index = m["material"]
        rec.density = m["density"]
        R = np.asarray(tf.rotation, np.float64).T
        rec.rotation_row0, rec.rotation_row1, rec.rotation_row2 = (abi.f3(*R[k]) for k in range(3))
        rec.translation = abi.f3(*tf.translation)
        if shape["type"] == "sphere":
            rec.boundary_type = 0
            rec.local_center_or_min = abi.f3(*shape["center"])
            rec.radius = shape["radius"]
        elif shape["type"] == "box":
            rec.boundary_type = 1
            rec.local_center_or_min = abi.f3(*shape["min"])
            rec.local_max = abi.f3(*shape["max"])
        elif shape["type"] == "quad":
            raise SceneFileError(f"{path}: quad boundaries are unsupported for homogeneous media")
        else:
            raise SceneFileError(f"{path}: triangle mesh boundaries are unsupported for homogeneous media")
        out.append(rec)
    return out


# ------------------------------------------------------------------ cameras

def _rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], np.float64)


def _ry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float64)


# camera_to_world_matrix / body_to_world_matrix (src/realtime/frame_convention.h:7-21)
CAMERA_TO_RENDERER = np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], np.float64)
BODY_TO_RENDERER = np.array([[0, -1, 0], [0, 0, -1], [1, 0, 0]], np.float64)


def _world_up(conv):
    return np.array([0.0, 1.0, 0.0]) if conv == "legacy_y_up" else np.array([0.0, 0.0, 1.0])


def _body_to_world(p: RealtimePreset) -> np.ndarray:
    """body_to_world_rotation (body_pose.cpp:36-53)"""
    nf = np.array([0.0, 0.0, -1.0]) if p.frame_convention == "legacy_y_up" else np.array([0.0, 1.0, 0.0])
    basis = np.column_stack([_world_up(p.frame_convention), -np.array([1.0, 0.0, 0.0]), -nf])
    return basis @ (_rx(math.radians(p.yaw_deg)) @ _ry(math.radians(-p.pitch_deg)))


def realtime_pose(p: RealtimePreset):
    """(camera-to-world rotation with columns right, down, forward; position)
    of the realtime default view (compile_realtime_camera_transform,
    scene_definition.cpp:114-139)."""
    rot = _body_to_world(p)
    forward = rot @ np.array([0.0, 0.0, -1.0])
    forward /= np.linalg.norm(forward)
    right = np.cross(forward, _world_up(p.frame_convention))
    if right @ right < 1e-12:
        right = np.array([1.0, 0.0, 0.0]) if p.frame_convention == "legacy_y_up" else rot @ np.array([0.0, -1.0, 0.0])
    right /= np.linalg.norm(right)
    if not (np.isfinite(forward).all() and np.isfinite(right).all()):
        raise SceneFileError("realtime camera pose must define finite camera axes")
    up = np.cross(right, forward)
    up /= np.linalg.norm(up)
    r_cw = np.column_stack([right, -up, forward])
    # pose_camera_to_body * T_bc, mapped back to renderer axes
    rp = CAMERA_TO_RENDERER.T @ r_cw
    tp = BODY_TO_RENDERER.T @ p.position
    r = rp @ p.camera.T_bc.rotation
    t = rp @ p.camera.T_bc.translation + tp
    return CAMERA_TO_RENDERER @ r, BODY_TO_RENDERER @ t


def cpu_pose(p: CpuPreset):
    """compile_cpu_camera_transform (scene_definition.cpp:86-112)"""
    back = p.lookfrom - p.lookat
    if not (np.isfinite(back).all() and np.isfinite(p.vup).all()):
        raise SceneFileError("CPU camera look-at vectors must be finite")
    if back @ back <= 1e-20:
        return (CAMERA_TO_RENDERER @ p.camera.T_bc.rotation,
                BODY_TO_RENDERER @ p.camera.T_bc.translation)
    w = back / np.linalg.norm(back)
    right = np.cross(p.vup, w)
    if right @ right <= 1e-20:
        raise SceneFileError("CPU camera up vector must not be parallel to its view direction")
    u = right / np.linalg.norm(right)
    v = np.cross(w, u)
    v /= np.linalg.norm(v)
    return np.column_stack([u, -v, -w]), p.lookfrom


def make_camera(spec: CameraSpec, rotation, position) -> abi.Camera:
    c = abi.Camera()
    c.enabled, c.width, c.height, c.model = 1, spec.width, spec.height, spec.model
    c.origin = abi.d3(*position)
    c.basis_x = abi.d3(*rotation[:, 0])
    c.basis_y = abi.d3(*rotation[:, 1])
    c.basis_z = abi.d3(*rotation[:, 2])
    c.fx, c.fy, c.cx, c.cy = spec.fx, spec.fy, spec.cx, spec.cy
    c.k1, c.k2, c.k3 = spec.pinhole32["k1"], spec.pinhole32["k2"], spec.pinhole32["k3"]
    c.p1, c.p2 = spec.pinhole32["p1"], spec.pinhole32["p2"]
    c.radial = (C.c_double * 6)(*spec.radial)
    c.tangential = (C.c_double * 2)(*spec.tangential)
    if spec.model == CAM_EQUI62_LUT1D:
        from . import check, lib
        check(lib().rt_camera_build_lut(C.byref(c)))
    return c


def camera_rig(defn: SceneDefinition, preset: str = "realtime") -> abi.CameraRig:
    """One-camera rig: the realtime default view, or the cpu preset `preset`."""
    if preset == "realtime":
        if defn.realtime is None:
            raise SceneFileError(f"{defn.id}: the scene has no realtime default view")
        rot, pos = realtime_pose(defn.realtime)
        cam = make_camera(defn.realtime.camera, rot, pos)
    else:
        match = [p for p in defn.cpu_presets if p.preset_id == preset]
        if not match:
            raise SceneFileError(f"{defn.id}: no cpu preset {preset}")
        rot, pos = cpu_pose(match[0])
        cam = make_camera(match[0].camera, rot, pos)
    rig = abi.CameraRig()
    rig.active_count = 1
    rig.cameras[0] = cam
    return rig


def main():
    import argparse
    import time

    import ray_tracing_amd as rt

    ap = argparse.ArgumentParser(description="render one frame of a YAML scene file")
    ap.add_argument("scene")
    ap.add_argument("--preset", default="realtime", help="'realtime' or a cpu preset id")
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--out", default="")
    ap.add_argument("--bundle", default="",
                    help="write the flattened scene, rig and profile to this scene bundle and exit "
                         "(no device is used)")
    a = ap.parse_args()
    defn = load_scene_definition(a.scene)
    flat = flatten(defn)
    rig = camera_rig(defn, a.preset)
    prof = rt.Generated("cornell_mesh").profile   # the plain tracer's settings
    prof.samples_per_pixel, prof.max_bounces = a.spp, a.depth
    if a.bundle:
        from . import scene_bundle
        scene_bundle.write(a.bundle, flat.desc, rig, prof, defn.id)
        print(f"[scene] {defn.id}: bundle {a.bundle}", file=sys.stderr)
        return
    dev = rt.Device(0)
    sc = rt.Scene(dev, flat.desc, keepalive=flat)
    sc.build_bvh()
    ren = rt.Renderer(dev, sc)
    t0 = time.time()
    tm = ren.render(rig, prof, 0)
    img = ren.download(0)["beauty"]
    print(f"[scene] {defn.id}: {flat.desc.sphere_count} spheres, {flat.desc.quad_count} quads, "
          f"{flat.desc.triangle_count} triangles; {rig.cameras[0].width}x{rig.cameras[0].height} "
          f"{a.spp} spp in {tm.render_ms:.2f} ms (wall {1e3 * (time.time() - t0):.1f} ms)")
    if a.out:
        np.save(a.out, img)


if __name__ == "__main__":
    main()
