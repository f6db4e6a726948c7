"""CPU-side checks of the C ABI library: it loads, exports every declared entry point,
and its host-side scene generators equal the oracle's generators byte for byte."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "rt", "rt_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:rt_status|int|int32_t|int64_t|void|const char\*)\s+(rt_\w+)\(", src, re.M)))


def test_exports_every_declared_symbol(rt):
    names = declared_entry_points()
    assert len(names) >= 20
    assert sorted(rt.EXPORTED) == names
    for n in names:
        assert hasattr(rt.lib(), n), f"{n} not exported"
    assert rt.lib().rt_abi_version() == 3


def test_struct_layouts_match_oracle(rt, oracle):
    pairs = [(rt.abi.NnImages, oracle.Images, 4 + 17 * 4 + 16 * 4), (rt.abi.Sphere, oracle.Sphere, 20), (rt.abi.Quad, oracle.Quad, 40),
             (rt.abi.Triangle, oracle.Triangle, 40), (rt.abi.Material, oracle.Material, 40),
             (rt.abi.OpenPbr, oracle.OpenPbr, 204), (rt.abi.Camera, oracle.Camera, None),
             (rt.abi.Profile, oracle.Profile, None), (rt.abi.SceneDesc, oracle.Scene, None)]
    for a, b, size in pairs:
        assert C.sizeof(a) == C.sizeof(b)
        if size is not None:
            assert C.sizeof(a) == size
    assert rt.NODE_DTYPE.itemsize == 64


def test_invalid_arguments_report_errors(rt):
    h = C.c_void_p()
    st = rt.lib().rt_scene_upload(None, None, C.byref(h))
    assert st == 1
    assert b"requires" in rt.lib().rt_last_error()
    p = C.POINTER(rt.abi.GeneratedScene)()
    assert rt.lib().rt_scene_generate(99, 0, C.byref(p)) == 1
    assert b"unknown scene config" in rt.lib().rt_last_error()


@pytest.mark.parametrize("config", [0, 1, 2, 3, 4, 5, 6])
def test_generators_match_oracle(rt, oracle, config):
    g = rt.Generated(config, 0)
    og = oracle.GeneratedScene(config, 0)
    arr = g.arrays()
    d, od = g.desc, og.scene
    for field, ctype, cnt in (("spheres", oracle.Sphere, "sphere_count"),
                              ("quads", oracle.Quad, "quad_count"),
                              ("triangles", oracle.Triangle, "triangle_count"),
                              ("materials", oracle.Material, "material_count"),
                              ("openpbr", oracle.OpenPbr, "openpbr_count"),
                              ("analytic", oracle.AnalyticLight, "analytic_light_count")):
        n = getattr(od, cnt)
        assert getattr(d, cnt) == n, field
        if n:
            src = "analytic_lights" if field == "analytic" else field
            ref = C.string_at(getattr(od, src), n * C.sizeof(ctype))
            assert arr[field].tobytes() == ref, field
    assert g.g.camera_count == og.g.camera_count
    for i in range(g.g.camera_count):
        assert bytes(g.g.cameras[i]) == bytes(og.g.cameras[i])
    assert bytes(g.g.profile) == bytes(og.g.profile)
    assert tuple(d.background) == tuple(od.background)


@pytest.mark.parametrize("frame", [0, 7, 59])
def test_orbit_rigs_match_oracle(rt, oracle, frame):
    g = rt.Generated("instanced_mesh_restir", 0)
    rig = g.rig_at(frame, 60)
    cams = oracle.generate_rig(2, 0, frame, 60)
    assert rig.active_count == 4
    for i in range(4):
        assert bytes(rig.cameras[i]) == bytes(cams[i])
    assert rig.cameras[2].model == 1 and rig.cameras[2].lut[1023] < 1.7


def test_cornell_mesh_size(rt):
    g = rt.Generated("cornell_mesh", 0)
    assert g.desc.triangle_count == 81920
    assert g.desc.quad_count == 6
    assert g.g.camera_count == 4
    g1 = rt.Generated("random_spheres", 0)
    assert g1.desc.sphere_count == 488
