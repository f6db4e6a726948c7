"""The oracle's light sampling (oracle_render.c, after
src/realtime/gpu/programs.cu:559-647,1057-1232 and
src/common/light_sampling.h) pinned by

1. the reference's known answers (tests/test_light_sampling.cpp:48-201):
   CDF boundaries and clamps for both light tables, the area-to-solid-angle
   Jacobian (2, 8, 0.5) -> 8, the balanced and complementary power
   heuristic, the zero-density weight, the sphere pdf normalization, the
   triangle barycentrics;
2. the reference's light_sampling.h compiled in place (oracle/_ref) on
   random inputs, bit for bit;
3. the density of every next-event sample equals the density that the
   emission-MIS path computes for the surface point the sample's ray hits
   ("sampling and hit PDFs match", test_light_sampling.cpp:88-93) for
   sphere, quad, triangle and environment lights;
4. closed forms of the geometry: the mean of 1 / pdf over a stratified
   sample of one light is the solid angle it subtends (Van Oosterom and
   Strackee's formula for triangles and rectangles, 2 pi (1 - cos) for
   spheres); the triangle barycentrics are uniform in area.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as O
from helpers import MAT_DT, QUAD_DT, SPHERE_DT, TRI_DT

L = O.lib()
needs_ref = pytest.mark.skipif(O.ref_lib() is None,
                               reason="oracle/_ref not built (reference tree absent)")


def cdf_pick(cdf, u):
    c = np.ascontiguousarray(cdf, np.float32)
    return L.rto_pick_cdf(O.ptr(c) if len(c) else None, len(c), u)


def test_reference_known_answers():
    """test_light_sampling.cpp:48-66 and 166-185."""
    assert cdf_pick([0.2, 0.5, 1.0], 0.0) == 0          # CDF lower boundary
    assert cdf_pick([0.2, 0.5, 1.0], 0.2) == 1          # first boundary
    assert cdf_pick([0.2, 0.5, 1.0], 0.5) == 2          # second boundary
    assert cdf_pick([0.2, 0.5, 1.0], 1.0) == 2          # upper clamp
    assert cdf_pick([], 0.5) == -1                      # empty distribution
    assert cdf_pick([0.25, 0.75, 1.0], 0.25) == 1       # analytic table
    assert cdf_pick([0.25, 0.75, 1.0], 0.75) == 2
    assert cdf_pick([0.25, 0.75, 1.0], -3.0) == 0
    assert abs(L.rto_uniform_sphere_pdf() * 4.0 * math.pi - 1.0) < 1e-6
    assert abs(L.rto_area_pdf(2.0, 8.0, 0.5) - 8.0) < 1e-6
    assert abs(L.rto_power_heuristic(0.25, 0.25) - 0.5) < 1e-6
    assert abs(L.rto_power_heuristic(0.2, 0.7) + L.rto_power_heuristic(0.7, 0.2) - 1.0) < 1e-6
    assert abs(L.rto_power_heuristic(0.0, 1.0)) < 1e-6
    w = np.zeros(3, np.float32)
    L.rto_triangle_barycentrics(0.36, 0.25, O.ptr(w))
    assert abs(w.sum() - 1.0) < 1e-6 and (w >= 0).all()
    # the rect light of test_light_sampling.cpp:95-118: a 2 x 2 rect at
    # distance 4 seen head-on from its centre has pdf 16 / (4 * 1) = 4
    assert abs(L.rto_area_pdf(4.0, 16.0, 1.0) - 4.0) < 1e-6
    # cone pdf: uniform over the cap's solid angle 2 pi (1 - cos)
    for c in np.array([0.0, 0.5, 0.9, 0.999], np.float32):
        assert abs(L.rto_cone_pdf(c) * 2.0 * math.pi * (1.0 - float(c)) - 1.0) < 1e-5
    assert L.rto_cone_pdf(1.0) == 0.0


@needs_ref
def test_matches_reference_header_on_random_inputs():
    R = O.ref_lib()
    rng = np.random.default_rng(7)
    a = rng.uniform(-0.5, 20.0, 20000).astype(np.float32)
    b = rng.uniform(0.0, 20.0, 20000).astype(np.float32)
    for x, y in zip(a[:5000], b[:5000]):
        assert L.rto_power_heuristic(x, y) == R.ref_light_power_heuristic(x, y)
    cs = np.concatenate([rng.uniform(-1, 1, 3000), [1.0, 1 - 1e-7, 0.0]]).astype(np.float32)
    for c in cs:
        assert L.rto_cone_pdf(c) == R.ref_light_uniform_cone_pdf(c)
    for area, d2, cos in zip(a[:3000], b[:3000], rng.uniform(-0.1, 1, 3000).astype(np.float32)):
        assert L.rto_area_pdf(area, d2, cos) == R.ref_light_area_to_solid_angle_pdf(area, d2, cos)
    for _ in range(300):
        n = int(rng.integers(1, 40))
        w = rng.uniform(0, 1, n) * (rng.uniform(0, 1, n) > 0.2)
        cdf = (np.cumsum(w) / max(w.sum(), 1e-9)).astype(np.float32)
        for u in np.concatenate([rng.uniform(-0.1, 1.1, 20), cdf[:5]]).astype(np.float32):
            assert cdf_pick(cdf, u) == R.ref_sample_packed_light(O.ptr(cdf), n, float(u))
    w1 = np.zeros(3, np.float32)
    w2 = np.zeros(3, np.float32)
    for u0, u1 in rng.uniform(-0.1, 1.1, (2000, 2)).astype(np.float32):
        L.rto_triangle_barycentrics(u0, u1, O.ptr(w1))
        R.ref_sample_uniform_triangle(u0, u1, O.ptr(w2))
        assert w1.tobytes() == w2.tobytes()


def test_triangle_barycentrics_are_uniform_in_area():
    """Over a stratified grid of (u0, u1) each barycentric averages 1/3 and
    each of the four midpoint sub-triangles receives a quarter of the samples."""
    n = 256
    u = (np.arange(n) + 0.5) / n
    ws = np.zeros((n * n, 3))
    w = np.zeros(3, np.float32)
    k = 0
    for u0 in u:
        for u1 in u:
            L.rto_triangle_barycentrics(u0, u1, O.ptr(w))
            ws[k] = w
            k += 1
    assert np.abs(ws.mean(0) - 1.0 / 3.0).max() < 2e-3
    corner = [(ws[:, i] > 0.5).mean() for i in range(3)]
    assert np.abs(np.array(corner) - 0.25).max() < 5e-3


def material(kind, albedo=(0.5, 0.5, 0.5), emission=(0, 0, 0)):
    m = np.zeros(1, MAT_DT)
    m["type"], m["albedo"], m["emission"], m["ior"], m["openpbr_index"] = kind, albedo, emission, 1.0, -1
    return m


def light_scene(background=(0.0, 0.0, 0.0)):
    """A floor (prim 1, material 0) under an emissive sphere (prim 0), a
    quad light (prim 2) and a triangle light (prim 3) of different radiance."""
    mats = np.concatenate([material(0), material(3, emission=(4, 4, 4)),
                           material(3, emission=(2, 3, 1)), material(3, emission=(6, 1, 1))])
    sph = np.zeros(1, SPHERE_DT)
    sph["center"], sph["radius"], sph["material"] = (1.5, 3.0, 0.5), 0.6, 1
    quads = np.zeros(2, QUAD_DT)
    quads[0] = ((-5, 0, -5), (0, 0, 10), (10, 0, 0), 0)           # floor, normal +y
    quads[1] = ((-2.5, 4.0, -1.0), (1.6, 0, 0), (0, 0, 1.2), 2)   # normal -y (faces the floor)
    tri = np.zeros(1, TRI_DT)
    tri[0] = ((0.5, 3.5, -2.0), (2.5, 3.5, -2.0), (1.0, 3.0, -0.5), 3)
    s, keep = O.scene_from_arrays(spheres=sph, quads=quads, triangles=tri, materials=mats,
                                  background=background)
    return s, keep


def test_sample_pdf_equals_hit_pdf():
    """For every valid NEE sample from random floor points, the emission-MIS
    density of the surface point its ray hits equals the sample's density."""
    s, keep = light_scene(background=(0.3, 0.4, 0.5))
    bvh = O.OracleBvh(s)
    rng = np.random.default_rng(1)
    seen = {}
    errs = []
    for _ in range(4000):
        pos = np.array([rng.uniform(-4, 4), 0.0, rng.uniform(-4, 4)], np.float32)
        d = np.array([0.3, -1.0, 0.2], np.float32)
        r = rng.uniform(0, 1, 4).astype(np.float32)
        li, wi, em, dist, pdf = O.sample_light_at(s, bvh, 1, pos, (0, 1, 0), d, r)
        if li < 0:      # invalid, or the ray reaches another light first
            continue
        hp = O.light_pdf_for_hit(s, bvh, 1, 0, pos, (0, 1, 0), wi)
        seen[li] = seen.get(li, 0) + 1
        errs.append(abs(hp - pdf) / pdf)
    errs = np.array(errs)
    assert len(seen) == 4, f"lights sampled: {seen}"     # sphere, quad, triangle, environment
    assert np.quantile(errs, 0.99) < 1e-4 and errs.max() < 2e-3, errs.max()


def van_oosterom_strackee(p, a, b, c):
    """Solid angle of triangle abc seen from p."""
    r = [np.asarray(v, np.float64) - np.asarray(p, np.float64) for v in (a, b, c)]
    la, lb, lc = (np.linalg.norm(v) for v in r)
    num = abs(np.dot(r[0], np.cross(r[1], r[2])))
    den = la * lb * lc + np.dot(r[0], r[1]) * lc + np.dot(r[0], r[2]) * lb + np.dot(r[1], r[2]) * la
    return 2.0 * math.atan2(num, den)


def one_light(kind):
    mats = np.concatenate([material(0), material(3, emission=(1, 1, 1))])
    quads = np.zeros(1, QUAD_DT)
    quads[0] = ((-5, 0, -5), (0, 0, 10), (10, 0, 0), 0)
    sph = tri = None
    if kind == "quad":
        q = np.zeros(1, QUAD_DT)
        q[0] = ((-1.0, 2.0, -0.5), (1.7, 0, 0), (0, 0.3, 1.4), 1)   # tilted, normal toward -y
        quads = np.concatenate([quads, q])
        verts = [(-1.0, 2.0, -0.5), (0.7, 2.0, -0.5), (0.7, 2.3, 0.9), (-1.0, 2.3, 0.9)]
        geom = [verts[0], verts[1], verts[2]], [verts[0], verts[2], verts[3]]
    elif kind == "triangle":
        tri = np.zeros(1, TRI_DT)
        tri[0] = ((0.5, 2.5, -1.0), (1.0, 2.0, 1.0), (-1.0, 3.0, 0.2), 1)   # normal toward -y
        geom = ([tuple(tri[0]["p0"]), tuple(tri[0]["p1"]), tuple(tri[0]["p2"])],)
    else:
        sph = np.zeros(1, SPHERE_DT)
        sph["center"], sph["radius"], sph["material"] = (0.3, 2.5, -0.4), 0.7, 1
        geom = None
    s, keep = O.scene_from_arrays(spheres=sph, quads=quads, triangles=tri, materials=mats)
    floor = 1 if kind == "sphere" else 0
    return s, keep, geom, floor


@pytest.mark.parametrize("kind", ["quad", "triangle", "sphere"])
def test_mean_inverse_pdf_is_the_solid_angle(kind):
    s, keep, geom, floor = one_light(kind)
    bvh = O.OracleBvh(s)
    n = 96
    grid = (np.arange(n) + 0.5) / n
    for pos in ((0.2, 0.0, 0.1), (-2.0, 0.0, 1.5), (3.0, 0.0, -2.5)):
        pos = np.array(pos, np.float32)
        total = 0.0
        for u1 in grid:
            for u2 in grid:
                li, wi, em, dist, pdf = O.sample_light_at(s, bvh, floor, pos, (0, 1, 0),
                                                          (0, -1, 0), (0.5, u1, u2, 0.5))
                if li >= 0:
                    total += 1.0 / pdf
        est = total / (n * n)
        if kind == "sphere":
            c = np.array((0.3, 2.5, -0.4)) - pos
            omega = 2 * math.pi * (1 - math.sqrt(1 - 0.49 / np.dot(c, c)))
        else:
            omega = sum(van_oosterom_strackee(pos, *t) for t in geom)
        assert abs(est / omega - 1.0) < 2e-3, (kind, pos, est, omega)
