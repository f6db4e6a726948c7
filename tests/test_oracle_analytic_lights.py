"""The oracle's analytic lights of every type (sphere, disk, rect, cylinder,
distant, dome; oracle_render.c after src/realtime/gpu/programs.cu:651-890,
1234-1346 and the light compiler src/scene/analytic_light_compiler.cpp:15-39,
114-240) pinned by what geometry fixes:

- compilation: world areas (4 pi (r s)^2, pi r^2 |bx x by|, w h |bx x by|,
  2 pi (r s)(l s)), delta distant lights, selection pdf = weight / total
  with weight = area, cone solid angle (1 for delta) or 4 pi times
  luminance, cdf closed at 1; non-similarity spheres refused; the
  library's rt_analytic_lights_compile returns the same bytes;
- radiance off a diffuse floor of albedo rho under one light, at depth 1
  (next-event only) and depth 3 (next-event and BSDF hits under MIS, whose
  weights need the right pdf of a hit for an unbiased sum):
  rect - Lambert's polygon formula; disk, cylinder - quadrature of
  L cos(x) cos(light) / d^2 over the emitting side; scaled sphere -
  pi (r s / d)^2 cos; point sphere - L A cos / d^2 (deterministic); delta
  distant - L cos (deterministic); cone distant - L pi sin^2(a) (a . n);
  dome - rho L;
- all six types together under next-event estimation: the sum of the
  single-light values (selection by the compiled cdf), also through
  ReSTIR's RIS without reuse.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as O
from helpers import ALIGHT_DT, MAT_DT, QUAD_DT, analytic_lights
from test_oracle_integrator import assert_z, down_camera, floor_points, floor_quad, material, mc, profile

RHO = (0.6, 0.5, 0.4)
LE = (3.0, 2.5, 2.0)
N = np.array([0.0, 1.0, 0.0])


def floor_scene(lights):
    mats = np.concatenate([material(0, albedo=RHO)])
    return O.scene_from_arrays(quads=floor_quad(0), materials=mats, analytic=lights)


def rho_pi():
    return np.array(RHO) / math.pi


# --------------------------------------------------------------- compile
def mixed_specs():
    s2 = 1.5 / math.sqrt(2.0)
    rot = ((s2, s2, 0.0), (-s2, s2, 0.0), (0.0, 0.0, 1.5))       # a rotation scaled by 1.5
    return [
        {"type": "sphere", "position": (0.4, 2.2, -0.3), "radius": 0.3, "radiance": LE, "basis": rot},
        {"type": "disk", "position": (-0.5, 1.8, 0.2), "radius": 0.35, "radiance": (2.0, 2.0, 2.0),
         "basis": ((0, 0, 1), (1, 0, 0), (0, 1, 0))},
        {"type": "rect", "position": (0.6, 2.0, 0.5), "width": 0.5, "height": 0.3, "radiance": (4.0, 3.0, 1.0),
         "basis": ((0, 0, 1.2), (0.8, 0, 0), (0, 1, 0))},
        {"type": "cylinder", "position": (-0.2, 2.4, -0.6), "radius": 0.1, "length": 0.8,
         "radiance": (1.0, 2.0, 3.0), "basis": ((0, 1, 0), (0, 0, 1), (1, 0, 0))},
        {"type": "distant", "radiance": (0.8, 0.8, 0.8), "cos_theta_max": math.cos(math.radians(10.0)),
         "basis": ((1, 0, 0), (0, 0, 1), (0.3, 0.9, 0.2))},
        {"type": "dome", "radiance": (0.2, 0.25, 0.3)},
    ]


def test_compile_closed_forms_and_library_agreement(rt):
    a = analytic_lights(O, mixed_specs())
    r = [0.3 * 1.5, 0.35, None, 0.1]
    assert a["world_area"][0] == pytest.approx(4 * math.pi * r[0] ** 2, rel=1e-6)
    assert a["world_area"][1] == pytest.approx(math.pi * 0.35 ** 2, rel=1e-6)
    assert a["world_area"][2] == pytest.approx(0.5 * 0.3 * 1.2 * 0.8, rel=1e-6)
    assert a["world_area"][3] == pytest.approx(2 * math.pi * 0.1 * 0.8, rel=1e-6)
    assert a["delta"].tolist() == [0, 0, 0, 0, 0, 0]
    lum = np.array([0.2126, 0.7152, 0.0722]) @ a["radiance"].astype(np.float64).T
    ca = math.cos(math.radians(10.0))
    w = np.array([a["world_area"][0] * lum[0], a["world_area"][1] * lum[1], a["world_area"][2] * lum[2],
                  a["world_area"][3] * lum[3], 2 * math.pi * (1 - np.float32(ca)) * lum[4], 4 * math.pi * lum[5]])
    assert a["selection_pdf"] == pytest.approx(w / w.sum(), rel=1e-5)
    assert np.all(np.diff(a["cdf"]) >= 0) and a["cdf"][-1] == 1.0
    # the library's compiler returns the same bytes
    b = np.zeros(len(a), ALIGHT_DT)
    for k in ("type", "position", "basis_x", "basis_y", "basis_z", "radiance", "radius", "width",
              "height", "length", "cos_theta_max", "treat_as_point", "treat_as_line"):
        b[k] = a[k]
    rt.compile_analytic_lights(b)
    assert b.tobytes() == a.tobytes()
    # a delta distant light, and a non-similarity sphere refused by both
    d = analytic_lights(O, [{"type": "distant", "radiance": (1, 1, 1)}, {"type": "dome", "radiance": (0, 0, 0)}])
    assert d["delta"].tolist() == [1, 0] and d["selection_pdf"].tolist() == [1.0, 0.0]
    bad = np.zeros(1, ALIGHT_DT)
    bad["type"], bad["basis_x"], bad["basis_y"], bad["basis_z"] = 0, (2, 0, 0), (0, 1, 0), (0, 0, 1)
    assert O.lib().rto_analytic_lights_compile(bad.ctypes.data, 1) == -1
    with pytest.raises(rt.RtError, match="similarity transform"):
        rt.compile_analytic_lights(bad)


# ----------------------------------------------------- single-light radiance
def quad_points(n_u, n_v, u0, u1, v0, v1):
    u = u0 + (np.arange(n_u) + 0.5) * (u1 - u0) / n_u
    v = v0 + (np.arange(n_v) + 0.5) * (v1 - v0) / n_v
    return np.meshgrid(u, v, indexing="ij")


def irradiance_quadrature(x, pts, normals, dA):
    """sum over emitting points of cos(x) cos(light) / d^2 dA, front faces only."""
    v = pts - x
    d2 = (v * v).sum(-1)
    d = np.sqrt(d2)
    cx = np.clip((v @ N) / d, 0.0, None)
    cl = np.clip(-(v * normals).sum(-1) / d, 0.0, None)
    return float((cx * cl / d2 * dA).sum())


def single(kind):
    """(lights, expect(x) -> outgoing radiance at floor point x, deterministic at depth 1)"""
    le = np.array(LE)
    if kind == "rect":
        c, bx, by, w, h = np.array([0.3, 2.0, -0.2]), np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]), 1.4, 0.9
        L = analytic_lights(O, [{"type": "rect", "position": tuple(c), "width": w, "height": h,
                                 "radiance": LE, "basis": (tuple(bx), tuple(by), (0, 1, 0))}])
        verts = [c - bx * w / 2 - by * h / 2, c + bx * w / 2 - by * h / 2, c + bx * w / 2 + by * h / 2,
                 c - bx * w / 2 + by * h / 2]
        from test_oracle_integrator import polygon_cos_integral
        return L, (lambda x: rho_pi() * le * polygon_cos_integral(x, N, verts)), False
    if kind == "disk":
        c, r = np.array([-0.2, 1.6, 0.3]), 0.45
        L = analytic_lights(O, [{"type": "disk", "position": tuple(c), "radius": r, "radiance": LE,
                                 "basis": ((0, 0, 1), (1, 0, 0), (0, 1, 0))}])
        rr, ph = quad_points(160, 360, 0.0, r, 0.0, 2 * math.pi)
        pts = np.stack([c[0] + rr * np.sin(ph), np.full_like(rr, c[1]), c[2] + rr * np.cos(ph)], -1).reshape(-1, 3)
        dA = (rr * (r / 160) * (2 * math.pi / 360)).reshape(-1)
        nrm = np.tile([0.0, -1.0, 0.0], (len(pts), 1))
        return L, (lambda x: rho_pi() * le * irradiance_quadrature(x, pts, nrm, dA)), False
    if kind == "cylinder":
        c, r, ln = np.array([0.1, 1.5, -0.1]), 0.2, 1.6
        L = analytic_lights(O, [{"type": "cylinder", "position": tuple(c), "radius": r, "length": ln,
                                 "radiance": LE, "basis": ((0, 1, 0), (0, 0, 1), (1, 0, 0))}])
        ph, z = quad_points(720, 200, 0.0, 2 * math.pi, -ln / 2, ln / 2)
        nrm = np.stack([np.zeros_like(ph), np.cos(ph), np.sin(ph)], -1).reshape(-1, 3)
        pts = (c + np.stack([z, r * np.cos(ph), r * np.sin(ph)], -1)).reshape(-1, 3)
        dA = np.full(len(pts), r * (2 * math.pi / 720) * (ln / 200))
        return L, (lambda x: rho_pi() * le * irradiance_quadrature(x, pts, nrm, dA)), False
    if kind == "sphere_scaled":
        c, r, s = np.array([0.4, 2.2, -0.3]), 0.25, 2.0
        L = analytic_lights(O, [{"type": "sphere", "position": tuple(c), "radius": r, "radiance": LE,
                                 "basis": ((s, 0, 0), (0, s, 0), (0, 0, s))}])

        def expect(x):
            v = c - x
            d2 = float(v @ v)
            return np.array(RHO) * le * (r * s) ** 2 * (v[1] / math.sqrt(d2)) / d2
        return L, expect, False
    if kind == "point":
        c, r = np.array([0.4, 2.2, -0.3]), 0.2
        L = analytic_lights(O, [{"type": "sphere", "position": tuple(c), "radius": r, "radiance": LE,
                                 "treat_as_point": 1}])
        area = 4 * math.pi * r * r

        def expect(x):
            v = c - x
            d2 = float(v @ v)
            return rho_pi() * le * area * (v[1] / math.sqrt(d2)) / d2
        return L, expect, True
    if kind == "distant_delta":
        a = np.array([0.3, 0.9, 0.2])
        L = analytic_lights(O, [{"type": "distant", "radiance": LE, "basis": ((1, 0, 0), (0, 0, 1), tuple(a))}])
        cos = a[1] / np.linalg.norm(a)
        return L, (lambda x: rho_pi() * le * cos), True
    if kind == "distant_cone":
        a, half = np.array([0.3, 0.9, 0.2]), math.radians(12.0)
        L = analytic_lights(O, [{"type": "distant", "radiance": LE, "cos_theta_max": math.cos(half),
                                 "basis": ((1, 0, 0), (0, 0, 1), tuple(a))}])
        cm = float(np.float32(math.cos(half)))
        e = math.pi * (1 - cm * cm) * a[1] / np.linalg.norm(a)
        return L, (lambda x: rho_pi() * le * e), False
    if kind == "dome":
        L = analytic_lights(O, [{"type": "dome", "radiance": LE}])
        return L, (lambda x: np.array(RHO) * le), False
    raise ValueError(kind)


KINDS = ["rect", "disk", "cylinder", "sphere_scaled", "point", "distant_delta", "distant_cone", "dome"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("depth", [1, 3])
def test_floor_radiance_under_one_analytic_light(kind, depth):
    L, expect, deterministic = single(kind)
    s, keep = floor_scene(L)
    cam = down_camera(1.2, w=20, h=14, f=12.0)
    x = floor_points(cam)
    ref = np.stack([[expect(x[j, i]) for i in range(cam.width)] for j in range(cam.height)])
    if deterministic and depth == 1:
        img = O.render_camera(s, O.OracleBvh(s), cam, 0, profile(4, 1), 0, threads=8)["beauty"][..., :3]
        assert np.allclose(img, ref, rtol=2e-5, atol=1e-7), kind
        return
    mean, se = mc(s, cam, profile(24, depth))
    assert_z(mean, se, ref, label=f"{kind} depth {depth}")


def test_all_types_together_add_up_under_next_event_estimation_and_ris():
    specs = []
    exps = []
    for kind in ("rect", "disk", "cylinder", "sphere_scaled", "distant_cone", "dome"):
        L, e, _ = single(kind)
        specs.append(L)
        exps.append(e)
    lights = np.concatenate(specs)
    assert O.lib().rto_analytic_lights_compile(lights.ctypes.data, len(lights)) == 0
    s, keep = floor_scene(lights)
    cam = down_camera(1.2, w=16, h=12, f=10.0)
    x = floor_points(cam)
    ref = np.stack([[sum(e(x[j, i]) for e in exps) for i in range(cam.width)] for j in range(cam.height)])
    # next-event estimation alone (depth 1): a shadow ray tests the scene's
    # surfaces only (is_occluded, programs.cu:2303-2306), so the lights add;
    # a BSDF ray ends at the first finite light it meets, which hides the
    # lights behind it, so with BSDF hits the lights do not add
    mean, se = mc(s, cam, profile(32, 1))
    assert_z(mean, se, ref, label="mixed depth 1")
    mean, se = mc(s, cam, profile(32, 1, restir=True, candidates=4))
    assert_z(mean, se, ref, label="mixed RIS")
