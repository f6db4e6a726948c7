"""Scene bundles (ray_tracing_amd.scene_bundle): the YAML ingest's flattened
scene written to a bundle reads back section for section equal to the scene
description's arrays and rig; the C++ loader (render_realtime --print-scene,
which exits before any device use) reports the same counts for every golden
scene, resolves scene ids in a scene directory and YAML paths through the
ingest run as a child process, and refuses malformed bundles and element
sizes that differ from its own."""
import ctypes as C
import os
import struct
import subprocess

import pytest

from ray_tracing_amd import abi, scene_bundle, scene_yaml

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(HERE, "golden", "scenes")
IDS = ["cornell_box", "final_room", "imported_obj_smoke", "simple_light"]


def flat_and_rig(sid):
    defn = scene_yaml.load_scene_definition(os.path.join(SCENES, sid, "scene.yaml"))
    return defn, scene_yaml.flatten(defn), scene_yaml.camera_rig(defn)


def profile():
    p = abi.Profile()
    p.samples_per_pixel, p.max_bounces = 4, 3
    return p


def cli():
    exe = os.path.join(os.path.dirname(scene_yaml.__file__), "render_realtime")
    if not os.path.exists(exe):
        pytest.skip("render_realtime is not built")
    return exe


@pytest.mark.parametrize("sid", IDS)
def test_bundle_round_trip(sid, tmp_path):
    defn, flat, rig = flat_and_rig(sid)
    path = str(tmp_path / f"{sid}.rtsb")
    scene_bundle.write(path, flat.desc, rig, profile(), defn.id)
    secs = scene_bundle.read(path)
    d = flat.desc
    checks = [(scene_bundle.TAG_SPHERES, d.spheres, d.sphere_count, C.sizeof(abi.Sphere)),
              (scene_bundle.TAG_QUADS, d.quads, d.quad_count, C.sizeof(abi.Quad)),
              (scene_bundle.TAG_TRIANGLES, d.triangles, d.triangle_count, C.sizeof(abi.Triangle)),
              (scene_bundle.TAG_MATERIALS, d.materials, d.material_count, C.sizeof(abi.Material)),
              (scene_bundle.TAG_TEXTURES, d.textures, d.texture_count, C.sizeof(abi.Texture)),
              (scene_bundle.TAG_TEXELS, d.image_texels, d.image_texel_count, 12),
              (scene_bundle.TAG_ALBEDO_TEX, d.material_albedo_texture, d.material_count, 4),
              (scene_bundle.TAG_TRI_UVS, d.triangle_uvs, d.triangle_count, 24),
              (scene_bundle.TAG_TRI_NORMALS, d.triangle_normals, d.triangle_count, 36),
              (scene_bundle.TAG_MEDIA, d.media, d.medium_count, C.sizeof(abi.Medium)),
              (scene_bundle.TAG_EMISSION_TEX, d.material_emission_texture, d.material_count, 4)]
    for tag, ptr, count, size in checks:
        if not ptr or count == 0:
            assert tag not in secs
            continue
        esize, n, raw = secs[tag]
        assert (esize, n) == (size, count)
        assert raw == C.string_at(ptr, size * count)
    _, n, raw = secs[scene_bundle.TAG_CAMERAS]
    assert n == rig.active_count
    assert raw == b"".join(bytes(rig.cameras[i]) for i in range(n))
    assert secs[scene_bundle.TAG_NAME][2].decode() == defn.id
    assert secs[scene_bundle.TAG_PROFILE][2] == bytes(profile())


@pytest.mark.parametrize("sid", IDS)
def test_cli_loads_yaml_scene(sid):
    exe = cli()
    defn, flat, rig = flat_and_rig(sid)
    d = flat.desc
    forms = [["--scene", os.path.join(SCENES, sid, "scene.yaml")]]
    if sid != "final_room":   # a built-in scene id resolves to the built-in generator first
        forms.append(["--scene", sid, "--scene-dir", SCENES])
    for args in forms:
        r = subprocess.run([exe, *args, "--print-scene"], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        want = (f"scene {defn.id}: {d.sphere_count} spheres, {d.quad_count} quads, "
                f"{d.triangle_count} triangles, {d.material_count} materials, {d.openpbr_count} openpbr, "
                f"{d.analytic_light_count} analytic lights, {d.texture_count} textures, "
                f"{d.image_texel_count} texels, {d.medium_count} media, {rig.active_count} cameras")
        assert r.stdout.startswith(want), (r.stdout, want)


def test_cli_refuses_bad_bundles(tmp_path):
    exe = cli()
    defn, flat, rig = flat_and_rig("cornell_box")
    good = str(tmp_path / "good.rtsb")
    scene_bundle.write(good, flat.desc, rig, profile(), defn.id)
    r = subprocess.run([exe, "--scene-bundle", good, "--print-scene"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("scene cornell_box: 0 spheres, 18 quads"), r.stdout
    data = open(good, "rb").read()
    cases = {"not a scene bundle": b"XXXX" + data[4:], "trailing bytes": data + b"\0",
             "truncated": data[:-10]}
    # a quad section whose element size is not the C struct's
    off = 12
    while True:
        tag, size, count = struct.unpack_from("<IIQ", data, off)
        if tag == scene_bundle.TAG_QUADS:
            break
        off += 16 + size * count
    bad = bytearray(data)
    struct.pack_into("<IIQ", bad, off, tag, size // 2, count * 2)
    cases["quad elements are"] = bytes(bad)
    for msg, blob in cases.items():
        p = tmp_path / "bad.rtsb"
        p.write_bytes(blob)
        r = subprocess.run([exe, "--scene-bundle", str(p), "--print-scene"], capture_output=True, text=True)
        assert r.returncode != 0 and msg in r.stderr, (msg, r.stderr)
    r = subprocess.run([exe, "--scene", "no_such_scene", "--scene-dir", SCENES, "--print-scene"],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "registered realtime scene" in r.stderr
