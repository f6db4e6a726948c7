"""The OpenPBR lobes beyond the opaque core on the GPU against the oracle: a
lit room with spheres of rough glass, smooth dispersive glass, a subsurface
random walk, thin-walled subsurface and transmission, fuzz over an
anisotropic coat, and a thin-film metal.

- ReSTIR off, depth 6, 4 spp: at least 99.9 % of the beauty values within
  1e-4 relative of the oracle and at least 95 % bit-exact (the host libm and
  the device library may round exp / log / cos / pow differently in the last
  place), normals / albedo / depth equal, frame means within 1e-3;
- the same spheres reduced to the opaque core render through the core-only
  kernels bit for bit like the oracle;
- ReSTIR DI over two frames with analytic lights on these materials:
  reservoirs equal and frames within tolerance;
- the random walk is exercised: turning subsurface off changes the
  subsurface sphere's pixels, and the walk keeps the frame finite.
The oracle's lobes are pinned against the reference header in
test_oracle_openpbr_full.py."""
import sys

import numpy as np
import pytest

from helpers import MAT_DT, QUAD_DT, SPHERE_DT, analytic_lights, oracle_camera, oracle_profile, oracle_scene, product_desc
from test_gpu_analytic_lights import profile, rig
from test_gpu_restir import compare_frame, compare_reservoirs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pbr_materials(oracle, core=False, subsurface=True):
    m = oracle.openpbr_material
    mats = [
        m(base_color=(0.9, 0.9, 0.9), specular_roughness=0.3, specular_ior=1.5, transmission_weight=1.0,
          transmission_color=(0.9, 0.8, 0.7)),
        m(base_color=(0.9, 0.9, 0.9), specular_roughness=0.0, specular_ior=1.6, transmission_weight=1.0,
          transmission_dispersion_scale=1.0, transmission_dispersion_abbe_number=20.0),
        m(base_color=(0.8, 0.6, 0.5), specular_roughness=0.4, subsurface_weight=1.0 if subsurface else 0.0,
          subsurface_color=(0.9, 0.5, 0.3), subsurface_radius=0.3,
          subsurface_radius_scale=(1.0, 0.5, 0.25), subsurface_scatter_anisotropy=0.3),
        m(base_color=(0.6, 0.8, 0.6), specular_roughness=0.5, subsurface_weight=0.6,
          subsurface_color=(0.7, 0.9, 0.6), transmission_weight=0.3, geometry_thin_walled=1),
        m(base_color=(0.3, 0.3, 0.8), specular_roughness=0.4, specular_roughness_anisotropy=0.6,
          fuzz_weight=0.8, fuzz_color=(0.9, 0.7, 0.9), fuzz_roughness=0.4, coat_weight=0.7,
          coat_roughness=0.3, coat_roughness_anisotropy=0.5),
        m(base_color=(0.9, 0.7, 0.4), base_metalness=1.0, specular_roughness=0.25, thin_film_weight=1.0,
          thin_film_thickness=0.45, thin_film_ior=1.4),
    ]
    if core:
        core_mats = []
        for x in mats:
            for f in ("transmission_weight", "subsurface_weight", "fuzz_weight", "thin_film_weight",
                      "transmission_dispersion_scale", "specular_roughness_anisotropy",
                      "coat_roughness_anisotropy"):
                setattr(x, f, 0.0)
            x.geometry_thin_walled = 0
            core_mats.append(x)
        mats = core_mats
    raw = b"".join(bytes(x) for x in mats)
    return np.frombuffer(raw, dtype=np.dtype((np.void, len(bytes(mats[0]))))).copy()


def arrays(oracle, core=False, subsurface=True, lights=False):
    pbr = pbr_materials(oracle, core, subsurface)
    mats = np.zeros(2 + len(pbr), MAT_DT)
    mats["type"][0:2] = (0, 3)
    mats["albedo"][0] = (0.7, 0.7, 0.7)
    mats["emission"][1] = (6.0, 5.5, 5.0)
    mats["ior"] = 1.0
    mats["openpbr_index"] = -1
    for i in range(len(pbr)):
        mats["type"][2 + i] = 5
        mats["openpbr_index"][2 + i] = i
    floor = np.zeros(1, QUAD_DT)
    floor[0] = ((-6, 0, -6), (0, 0, 12), (12, 0, 0), 0)
    lamp = np.zeros(1, QUAD_DT)
    lamp[0] = ((-1.0, 3.2, -1.0), (2.0, 0, 0), (0, 0, 2.0), 1)     # faces down
    sph = np.zeros(len(pbr), SPHERE_DT)
    for i in range(len(pbr)):
        sph[i] = ((-1.9 + 0.75 * i, 0.36, 0.1 * (i % 2)), 0.34, 2 + i)
    a = {"spheres": sph, "quads": np.concatenate([floor, lamp]), "materials": mats, "openpbr": pbr}
    if lights:
        a["analytic"] = analytic_lights(oracle, [
            {"type": "rect", "position": (1.5, 2.4, -1.0), "width": 0.8, "height": 0.5,
             "radiance": (3, 3, 2.5), "basis": ((0, 0, 1), (1, 0, 0), (0, 1, 0))},
            {"type": "sphere", "position": (-1.0, 1.8, 1.5), "radius": 0.15, "radiance": (6, 5, 4)},
            {"type": "dome", "radiance": (0.05, 0.06, 0.08)},
        ])
    return a


def scene(rt, oracle, device, **kw):
    desc, keep = product_desc(rt, arrays(oracle, **kw), background=(0.15, 0.17, 0.2))
    sc = rt.Scene(device, desc, keep)
    sc.build_bvh()
    osc = oracle_scene(oracle, desc)
    return desc, keep, sc, osc, oracle.OracleBvh(osc)


def compare(gpu, ref, label, frac=0.999, exact_frac=0.95):
    a = gpu["beauty"][..., :3].astype(np.float64)
    b = ref["beauty"][..., :3].astype(np.float64)
    assert np.isfinite(a).all(), f"{label}: non-finite beauty"
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    close = float(np.mean(rel < 1e-4))
    exact = float(np.mean(a == b))
    say(f"  {label}: beauty exact {exact:.5f}, within 1e-4 {close:.5f}, mean {a.mean():.5f} vs {b.mean():.5f}")
    assert close >= frac, label
    assert exact >= exact_frac, label
    assert abs(a.mean() - b.mean()) <= 1e-3 * max(abs(b.mean()), 1e-3), label
    for key in ("normal", "albedo", "depth"):
        assert np.mean(gpu[key] == ref[key]) >= 0.999, f"{label}: {key}"
    return exact


def test_full_openpbr_frames_match_oracle(rt, oracle, device):
    say("[gpu] OpenPBR transmission / dispersion / subsurface / thin walls / fuzz / film, depth 6, 4 spp")
    desc, keep, sc, osc, ob = scene(rt, oracle, device)
    r = rig(rt, oracle)
    prof = profile(rt, 4, 6, False)
    ren = rt.Renderer(device, sc)
    ren.render(r, prof, 5)
    for ci in range(2):
        ref = oracle.render_camera(osc, ob, oracle_camera(oracle, r.cameras[ci]), ci,
                                   oracle_profile(oracle, prof), 5, threads=16)
        compare(ren.download(ci), ref, f"camera {ci}")


def test_core_materials_take_the_core_kernels(rt, oracle, device):
    say("[gpu] the same spheres reduced to the opaque core, depth 6, 4 spp")
    desc, keep, sc, osc, ob = scene(rt, oracle, device, core=True)
    r = rig(rt, oracle)
    prof = profile(rt, 4, 6, False)
    ren = rt.Renderer(device, sc)
    ren.render(r, prof, 5)
    for ci in range(2):
        ref = oracle.render_camera(osc, ob, oracle_camera(oracle, r.cameras[ci]), ci,
                                   oracle_profile(oracle, prof), 5, threads=16)
        compare(ren.download(ci), ref, f"core camera {ci}", exact_frac=1.0)


def test_full_openpbr_with_restir_match_oracle(rt, oracle, device):
    say("[gpu] full OpenPBR under ReSTIR DI, two frames")
    desc, keep, sc, osc, ob = scene(rt, oracle, device, lights=True)
    r = rig(rt, oracle)
    prof = profile(rt, 1, 3, True)
    ren = rt.Renderer(device, sc)
    seq = oracle.RestirSequence(osc, ob, oracle_profile(oracle, prof))
    for frame in range(2):
        ren.render(r, prof, 20 + frame)
        refs = seq.render([oracle_camera(oracle, r.cameras[i]) for i in range(2)], 20 + frame)
        for ci in range(2):
            compare_frame(ren.download(ci), refs[ci], f"frame{frame} cam{ci}", tol_frac=0.999)
            compare_reservoirs(ren.reservoirs(ci), refs[ci]["reservoirs"], f"frame{frame} cam{ci}")


def test_random_walk_changes_the_subsurface_sphere(rt, oracle, device):
    say("[gpu] subsurface random walk on / off")
    r = rig(rt, oracle)
    prof = profile(rt, 8, 8, False)
    frames = []
    for on in (True, False):
        desc, keep, sc, osc, ob = scene(rt, oracle, device, subsurface=on)
        ren = rt.Renderer(device, sc)
        ren.render(r, prof, 9)
        frames.append(ren.download(0)["beauty"][..., :3].astype(np.float64))
    assert all(np.isfinite(f).all() for f in frames)
    diff = np.abs(frames[0] - frames[1]).max(axis=-1)
    changed = float(np.mean(diff > 1e-3))
    say(f"  pixels changed by the random walk: {changed:.4f}")
    assert changed > 0.01
