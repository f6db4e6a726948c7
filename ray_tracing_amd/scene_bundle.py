"""Scene bundles: a flattened scene (every array of an `rt_scene_desc`), its
camera rig and a render profile in one binary file, so that programs without
the YAML ingest (the C++ `render_realtime`, through
`rt_amd::load_scene_bundle`) render the scenes this package ingests.

Layout (little-endian):
    magic b"RTSB", u32 version (1), u32 section count, then per section
    u32 tag, u32 element size, u64 element count, element bytes.
Element sizes are the C ABI struct sizes; a reader refuses a section whose
element size differs from its own struct's.  Absent arrays are absent
sections.  Tags are the `TAG_*` constants below.
"""
from __future__ import annotations

import ctypes as C
import struct

from . import abi

MAGIC = b"RTSB"
VERSION = 1

TAG_BACKGROUND = 1      # i32 mode, 3 floats background, 3 floats background_top (28 B, one element)
TAG_SPHERES = 2
TAG_QUADS = 3
TAG_TRIANGLES = 4
TAG_MATERIALS = 5
TAG_OPENPBR = 6
TAG_ANALYTIC = 7
TAG_TEXTURES = 8
TAG_TEXELS = 9          # 3 floats per texel
TAG_ALBEDO_TEX = 10     # i32 per material
TAG_TRI_UVS = 11        # 6 floats per triangle
TAG_TRI_NORMALS = 12    # 9 floats per triangle
TAG_MEDIA = 13
TAG_EMISSION_TEX = 14   # i32 per material
TAG_PBR_TEX = 15        # rt_openpbr_textures per OpenPBR record
TAG_CAMERAS = 16        # rt_camera per active camera
TAG_PROFILE = 17        # one rt_profile
TAG_NAME = 18           # scene id, UTF-8 bytes (element size 1)


def _array(ptr, count, size):
    if not ptr or count <= 0:
        return None
    return C.string_at(ptr, count * size)


def sections(desc: abi.SceneDesc, rig: abi.CameraRig, profile: abi.Profile, name: str):
    d = desc
    bg = struct.pack("<i3f3f", d.background_mode, *d.background, *d.background_top)
    out = [(TAG_BACKGROUND, len(bg), 1, bg)]
    arrays = [
        (TAG_SPHERES, d.spheres, d.sphere_count, C.sizeof(abi.Sphere)),
        (TAG_QUADS, d.quads, d.quad_count, C.sizeof(abi.Quad)),
        (TAG_TRIANGLES, d.triangles, d.triangle_count, C.sizeof(abi.Triangle)),
        (TAG_MATERIALS, d.materials, d.material_count, C.sizeof(abi.Material)),
        (TAG_OPENPBR, d.openpbr, d.openpbr_count, C.sizeof(abi.OpenPbr)),
        (TAG_ANALYTIC, d.analytic_lights, d.analytic_light_count, C.sizeof(abi.AnalyticLight)),
        (TAG_TEXTURES, d.textures, d.texture_count, C.sizeof(abi.Texture)),
        (TAG_TEXELS, d.image_texels, d.image_texel_count, 12),
        (TAG_ALBEDO_TEX, d.material_albedo_texture, d.material_count, 4),
        (TAG_TRI_UVS, d.triangle_uvs, d.triangle_count, 24),
        (TAG_TRI_NORMALS, d.triangle_normals, d.triangle_count, 36),
        (TAG_MEDIA, d.media, d.medium_count, C.sizeof(abi.Medium)),
        (TAG_EMISSION_TEX, d.material_emission_texture, d.material_count, 4),
        (TAG_PBR_TEX, d.openpbr_textures, d.openpbr_count, C.sizeof(abi.OpenPbrTextures)),
    ]
    for tag, ptr, count, size in arrays:
        raw = _array(ptr, count, size)
        if raw is not None:
            out.append((tag, size, count, raw))
    n = rig.active_count
    cams = b"".join(bytes(rig.cameras[i]) for i in range(n))
    out.append((TAG_CAMERAS, C.sizeof(abi.Camera), n, cams))
    out.append((TAG_PROFILE, C.sizeof(abi.Profile), 1, bytes(profile)))
    nb = name.encode()
    out.append((TAG_NAME, 1, len(nb), nb))
    return out


def write(path: str, desc: abi.SceneDesc, rig: abi.CameraRig, profile: abi.Profile, name: str):
    secs = sections(desc, rig, profile, name)
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<II", VERSION, len(secs)))
        for tag, size, count, raw in secs:
            assert len(raw) == size * count
            f.write(struct.pack("<IIQ", tag, size, count))
            f.write(raw)


def read(path: str) -> dict:
    """{tag: (element size, count, bytes)}; raises ValueError on a malformed file."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != MAGIC:
        raise ValueError(f"{path}: not a scene bundle")
    version, n = struct.unpack_from("<II", data, 4)
    if version != VERSION:
        raise ValueError(f"{path}: bundle version {version}, expected {VERSION}")
    off, out = 12, {}
    for _ in range(n):
        if off + 16 > len(data):
            raise ValueError(f"{path}: truncated section header")
        tag, size, count = struct.unpack_from("<IIQ", data, off)
        off += 16
        end = off + size * count
        if end > len(data):
            raise ValueError(f"{path}: truncated section {tag}")
        out[tag] = (size, count, data[off:end])
        off = end
    if off != len(data):
        raise ValueError(f"{path}: trailing bytes")
    return out
