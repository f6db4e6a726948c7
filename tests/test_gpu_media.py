"""Homogeneous media on the GPU against the oracle: a floor, a box, an
emissive quad, analytic lights and a sky, with an isotropic sphere medium
and a rotated, translated isotropic box medium; closest-hit rays scatter
inside them at free-flight distances, shadow rays pass through:
- ReSTIR off, depth 5, 4 spp: beauty / normal / albedo / depth bit-exact on
  at least 99.9 % of the values, and the media visibly change the frame;
- ReSTIR DI over two frames: frames within 1e-4 on 99 % of the pixels and
  reservoirs equal;
- the upload refuses a medium with an unknown boundary or a material index
  outside the table.
The oracle's media are pinned by closed forms in test_oracle_media.py."""
import math
import sys

import numpy as np
import pytest

from helpers import MAT_DT, QUAD_DT, SPHERE_DT, analytic_lights, medium, oracle_camera, oracle_profile, oracle_scene, product_desc
from test_gpu_analytic_lights import box_quads, profile, rig
from test_gpu_restir import compare_frame, compare_reservoirs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def arrays(oracle, with_media=True):
    mats = np.zeros(5, MAT_DT)
    mats["type"] = (0, 0, 3, 4, 4)
    mats["albedo"] = ((0.7, 0.7, 0.7), (0.8, 0.4, 0.3), (0, 0, 0), (0.9, 0.85, 0.8), (0.6, 0.7, 0.9))
    mats["emission"][2] = (8.0, 7.0, 6.0)
    mats["ior"] = 1.0
    mats["openpbr_index"] = -1
    floor = np.zeros(1, QUAD_DT)
    floor[0] = ((-6, 0, -6), (0, 0, 12), (12, 0, 0), 0)
    lamp = np.zeros(1, QUAD_DT)
    lamp[0] = ((-0.5, 3.0, -0.5), (1.0, 0, 0), (0, 0, 1.0), 2)     # faces down
    quads = np.concatenate([floor, box_quads((-2.2, 0, -1.2), (-1.4, 1.2, -0.4), 1), lamp])
    sph = np.zeros(1, SPHERE_DT)
    sph[0] = ((1.6, 0.4, 0.8), 0.4, 1)
    lights = analytic_lights(oracle, [
        {"type": "rect", "position": (1.5, 2.4, -1.0), "width": 0.8, "height": 0.5, "radiance": (3, 3, 2.5),
         "basis": ((0, 0, 1), (1, 0, 0), (0, 1, 0))},
        {"type": "sphere", "position": (-1.0, 1.8, 1.5), "radius": 0.15, "radiance": (6, 5, 4)},
        {"type": "dome", "radiance": (0.05, 0.06, 0.08)},
    ])
    a = {"spheres": sph, "quads": quads, "materials": mats, "analytic": lights}
    if with_media:
        c, s_ = math.cos(0.5), math.sin(0.5)
        rot = np.array([[c, 0, -s_], [0, 1, 0], [s_, 0, c]])
        a["media"] = np.concatenate([
            medium("sphere", 3, 1.4, centre=(0.0, 0.9, 0.2), radius=0.8),
            medium("box", 4, 0.8, lo=(-0.6, -0.5, -0.4), hi=(0.6, 0.5, 0.4), rotation=rot,
                   translation=(1.2, 0.5, -0.8)),
        ])
    return a


def scene(rt, oracle, device, with_media=True):
    desc, keep = product_desc(rt, arrays(oracle, with_media), background=(0.08, 0.09, 0.12))
    sc = rt.Scene(device, desc, keep)
    sc.build_bvh()
    osc = oracle_scene(oracle, desc)
    return desc, keep, sc, osc, oracle.OracleBvh(osc)


def test_media_frames_match_oracle(rt, oracle, device):
    say("[gpu] homogeneous media, ReSTIR off, depth 5, 4 spp")
    desc, keep, sc, osc, ob = scene(rt, oracle, device)
    r = rig(rt, oracle)
    prof = profile(rt, 4, 5, False)
    ren = rt.Renderer(device, sc)
    ren.render(r, prof, 2)
    for ci in range(2):
        gpu = ren.download(ci)
        ref = oracle.render_camera(osc, ob, oracle_camera(oracle, r.cameras[ci]), ci,
                                   oracle_profile(oracle, prof), 2, threads=16)
        for key in ("beauty", "normal", "albedo", "depth"):
            same = float(np.mean(gpu[key] == ref[key]))
            say(f"  camera {ci} {key}: bit-exact fraction {same:.6f}")
            assert same > 0.999, f"camera {ci} {key}"
    # without the media the same frame differs where the media stand
    _, keep2, sc2, _, _ = scene(rt, oracle, device, with_media=False)
    ren2 = rt.Renderer(device, sc2)
    ren2.render(r, prof, 2)
    diff = np.mean(ren2.download(0)["albedo"] != ren.download(0)["albedo"])
    say(f"  albedo values changed by the media: {diff:.3f}")
    assert diff > 0.02


def test_media_with_restir_match_oracle(rt, oracle, device):
    say("[gpu] homogeneous media, ReSTIR DI, two frames")
    desc, keep, sc, osc, ob = scene(rt, oracle, device)
    r = rig(rt, oracle)
    prof = profile(rt, 1, 3, True)
    ren = rt.Renderer(device, sc)
    seq = oracle.RestirSequence(osc, ob, oracle_profile(oracle, prof))
    for frame in range(2):
        ren.render(r, prof, 10 + frame)
        refs = seq.render([oracle_camera(oracle, r.cameras[i]) for i in range(2)], 10 + frame)
        for ci in range(2):
            compare_frame(ren.download(ci), refs[ci], f"frame{frame} cam{ci}", tol_frac=0.99)
            compare_reservoirs(ren.reservoirs(ci), refs[ci]["reservoirs"], f"frame{frame} cam{ci}")


def test_upload_refuses_bad_media(rt, oracle, device):
    a = arrays(oracle)
    for field, value, msg in (("boundary_type", 7, "unknown boundary"), ("material_index", 9, "material")):
        bad = dict(a)
        bad["media"] = a["media"].copy()
        bad["media"][field][0] = value
        desc, keep = product_desc(rt, bad)
        with pytest.raises(rt.RtError, match=msg):
            rt.Scene(device, desc, keep)
