"""GPU frames lit by analytic lights of every type (rect, disk, cylinder,
scaled and point spheres, cone and delta distant lights, a dome) against the
oracle, on a floor with a box and spheres that cast shadows:
- next-event estimation and BSDF hits under MIS, ReSTIR off, depth 4,
  4 spp: beauty / normal / albedo / depth bit-exact on at least 99.9 % of
  the values;
- ReSTIR DI over two frames (temporal and spatial reuse, bias correction):
  frames within 1e-4 on 99 % of the pixels and reservoirs equal;
- the upload refuses lights whose cdf decreases or whose type is unknown.
The oracle's light math is pinned separately (test_oracle_analytic_lights).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import MAT_DT, QUAD_DT, SPHERE_DT, analytic_lights, oracle_camera, oracle_profile, oracle_scene, product_desc
from test_gpu_restir import compare_frame, compare_reservoirs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def box_quads(lo, hi, mat):
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    q = np.zeros(6, QUAD_DT)
    q[0] = ((x0, y0, z1), (x1 - x0, 0, 0), (0, y1 - y0, 0), mat)
    q[1] = ((x1, y0, z0), (x0 - x1, 0, 0), (0, y1 - y0, 0), mat)
    q[2] = ((x0, y0, z0), (0, 0, z1 - z0), (0, y1 - y0, 0), mat)
    q[3] = ((x1, y0, z1), (0, 0, z0 - z1), (0, y1 - y0, 0), mat)
    q[4] = ((x0, y1, z0), (0, 0, z1 - z0), (x1 - x0, 0, 0), mat)
    q[5] = ((x0, y0, z0), (x1 - x0, 0, 0), (0, 0, z1 - z0), mat)
    return q


def light_specs():
    rng = np.random.default_rng(12)
    specs = []
    for k in range(4):
        specs.append({"type": "rect", "position": (-2.0 + 1.3 * k, 2.6, -1.0 + 0.4 * k), "width": 0.6,
                      "height": 0.4, "radiance": tuple(rng.uniform(1, 4, 3)),
                      "basis": ((0, 0, 1), (1, 0, 0), (0, 1, 0))})
        specs.append({"type": "disk", "position": (-1.8 + 1.2 * k, 2.2, 0.8 - 0.5 * k), "radius": 0.25,
                      "radiance": tuple(rng.uniform(1, 4, 3)),
                      "basis": ((0, 0.2, 1), (1, 0, 0), (0, 1, -0.2))})
    for k in range(3):
        specs.append({"type": "cylinder", "position": (-1.5 + 1.5 * k, 1.9, -1.8), "radius": 0.07,
                      "length": 0.9, "radiance": tuple(rng.uniform(1, 3, 3)),
                      "basis": ((0, 1, 0), (0, 0, 1), (1, 0, 0))})
    s = 1.5
    specs += [
        {"type": "sphere", "position": (1.6, 1.2, 0.9), "radius": 0.1, "radiance": (6, 5, 4),
         "basis": ((s, 0, 0), (0, s, 0), (0, 0, s))},
        {"type": "sphere", "position": (-1.4, 1.0, 1.2), "radius": 0.12, "radiance": (3, 4, 6)},
        {"type": "sphere", "position": (0.0, 2.0, 0.0), "radius": 0.05, "radiance": (8, 8, 8),
         "treat_as_point": 1},
        {"type": "sphere", "position": (2.2, 0.6, -0.6), "radius": 0.2, "radiance": (2, 2, 1)},
        {"type": "distant", "radiance": (0.8, 0.7, 0.6), "cos_theta_max": float(np.cos(np.radians(8.0))),
         "basis": ((1, 0, 0), (0, 0, 1), (0.4, 0.9, 0.3))},
        {"type": "distant", "radiance": (0.5, 0.5, 0.6), "basis": ((1, 0, 0), (0, 0, 1), (-0.5, 0.8, 0.2))},
        {"type": "dome", "radiance": (0.05, 0.07, 0.1)},
    ]
    return specs


@pytest.fixture(scope="module")
def lit_scene(rt, oracle, device):
    mats = np.zeros(3, MAT_DT)
    mats["type"] = (0, 0, 1)
    mats["albedo"] = ((0.7, 0.7, 0.7), (0.8, 0.4, 0.3), (0.9, 0.9, 0.9))
    mats["fuzz"] = (0, 0, 0.2)
    mats["ior"] = 1.0
    mats["openpbr_index"] = -1
    floor = np.zeros(1, QUAD_DT)
    floor[0] = ((-6, 0, -6), (0, 0, 12), (12, 0, 0), 0)
    quads = np.concatenate([floor, box_quads((-0.6, 0, -0.4), (0.2, 0.9, 0.3), 1)])
    sph = np.zeros(3, SPHERE_DT)
    sph[0] = ((1.0, 0.5, 0.2), 0.5, 2)
    sph[1] = ((-1.3, 0.35, 0.4), 0.35, 1)
    sph[2] = ((0.5, 0.25, 1.2), 0.25, 0)
    lights = analytic_lights(oracle, light_specs())
    desc, keep = product_desc(rt, {"spheres": sph, "quads": quads, "materials": mats, "analytic": lights})
    sc = rt.Scene(device, desc, keep)
    sc.build_bvh()
    osc = oracle_scene(oracle, desc)
    return desc, keep, sc, osc, oracle.OracleBvh(osc)


def rig(rt, oracle, w=160, h=120):
    r = rt.CameraRig()
    r.active_count = 2
    for i, (pos, z) in enumerate((((0.3, 1.6, 4.5), (0.0, -0.3, -1.0)), ((-3.5, 2.2, 2.5), (0.7, -0.4, -0.6)))):
        c = oracle.pinhole_camera(0.9 * w, 0.9 * w, w / 2, h / 2, width=w, height=h)
        zz = np.array(z, np.float64)
        zz /= np.linalg.norm(zz)
        xx = np.cross((0.0, 1.0, 0.0), zz) * -1.0
        xx /= np.linalg.norm(xx)
        yy = np.cross(zz, xx)
        c.origin[:] = pos
        c.basis_x[:], c.basis_y[:], c.basis_z[:] = tuple(xx), tuple(yy), tuple(zz)
        C.memmove(C.byref(r.cameras[i]), C.byref(c), C.sizeof(c))
    return r


def profile(rt, spp, depth, restir):
    p = rt.Profile()
    p.samples_per_pixel, p.max_bounces, p.rr_start_bounce = spp, depth, depth
    p.enable_denoise = 0
    p.enable_restir_di = 1 if restir else 0
    p.restir_initial_candidates, p.restir_temporal_reuse, p.restir_max_history_age = 4, 1, 20
    p.restir_max_temporal_candidates, p.restir_spatial_neighbors = 64, 1
    p.restir_max_spatial_candidates, p.restir_bias_correction, p.restir_min_analytic_lights = 4, 1, 1
    return p


def test_every_light_type_matches_oracle(rt, oracle, device, lit_scene):
    say("[gpu] analytic lights of every type, ReSTIR off, depth 4, 4 spp")
    desc, keep, sc, osc, ob = lit_scene
    r = rig(rt, oracle)
    prof = profile(rt, 4, 4, False)
    ren = rt.Renderer(device, sc)
    ren.render(r, prof, 3)
    for ci in range(2):
        gpu = ren.download(ci)
        ref = oracle.render_camera(osc, ob, oracle_camera(oracle, r.cameras[ci]), ci,
                                   oracle_profile(oracle, prof), 3, threads=16)
        assert float(ref["beauty"][..., :3].mean()) > 1e-3
        for key in ("beauty", "normal", "albedo", "depth"):
            same = float(np.mean(gpu[key] == ref[key]))
            say(f"  camera {ci} {key}: bit-exact fraction {same:.6f}")
            assert same > 0.999, f"camera {ci} {key}"


def test_restir_over_every_light_type_matches_oracle(rt, oracle, device, lit_scene):
    say("[gpu] analytic lights of every type, ReSTIR DI, two frames")
    desc, keep, sc, osc, ob = lit_scene
    r = rig(rt, oracle)
    prof = profile(rt, 1, 2, True)
    ren = rt.Renderer(device, sc)
    seq = oracle.RestirSequence(osc, ob, oracle_profile(oracle, prof))
    for frame in range(2):
        ren.render(r, prof, frame)
        refs = seq.render([oracle_camera(oracle, r.cameras[i]) for i in range(2)], frame)
        for ci in range(2):
            compare_frame(ren.download(ci), refs[ci], f"frame{frame} cam{ci}", tol_frac=0.99)
            compare_reservoirs(ren.reservoirs(ci), refs[ci]["reservoirs"], f"frame{frame} cam{ci}")


def test_upload_refuses_bad_light_tables(rt, oracle, device):
    lights = analytic_lights(oracle, light_specs()[:3])
    floor = np.zeros(1, QUAD_DT)
    floor[0] = ((-1, 0, -1), (0, 0, 2), (2, 0, 0), 0)
    mats = np.zeros(1, MAT_DT)
    mats["openpbr_index"] = -1
    for field, value, msg in (("cdf", 0.1, "cdf outside"), ("type", 9, "unknown type")):
        bad = lights.copy()
        bad[field][-1] = value
        desc, keep = product_desc(rt, {"quads": floor, "materials": mats, "analytic": bad})
        with pytest.raises(rt.RtError, match=msg):
            rt.Scene(device, desc, keep)
