"""The oracle's BSDF sampling and integrator (oracle_render.c trace_one /
sample_bsdf / direct_light, after src/realtime/gpu/programs.cu:1764-1953,
2249-2313, 2451-2565) pinned by physical laws and closed-form radiance:

BSDFs (one scattering event at a time, rto_scatter):
- Lambert: directions in the upper hemisphere with the moments of the
  cosine density (E[cos] = 2/3, E[cos^2] = 1/2), pdf = cos / pi, weight = albedo;
- metal: with fuzz 0 the law of reflection exactly; with fuzz f every
  direction within asin(f) of the mirror direction; delta events;
- dielectric: Snell's law on every refracted direction, reflection with
  Schlick's probability ((1 - n) / (1 + n))^2 at normal incidence, total
  internal reflection beyond the critical angle.

Radiance (rto_render_camera, pixel centres, many samples, compared through a
z-test on K independent streams):
- a Lambertian sphere under a uniform sky reflects albedo x sky exactly in
  every sample (NEE and BSDF sampling both use cos / pi, so the MIS weights
  are 1/2 and 1/2), at depth 1 and 3;
- a mirror floor shows albedo x the sky gradient in the reflected direction;
- a diffuse floor lit by one quad, triangle, geometric sphere or analytic
  sphere light: albedo / pi x Le x the cosine-weighted solid angle of the
  light (Lambert's polygon formula; pi (r / d)^2 cos for a sphere), at depth
  1 (next-event only) and depth 3 (next-event and BSDF hits under MIS), and
  through ReSTIR's RIS without reuse;
- a white furnace: inside a closed box of emissive walls of albedo a the
  radiance is Le * sum_{k=0..D} a^k for depth D, with Russian roulette off
  and on (RR is unbiased).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ctypes as O
from helpers import ALIGHT_DT, MAT_DT, QUAD_DT, SPHERE_DT, TRI_DT, analytic_lights


def material(kind, albedo=(0.5, 0.5, 0.5), emission=(0, 0, 0), fuzz=0.0, ior=1.0, openpbr=-1):
    m = np.zeros(1, MAT_DT)
    m["type"], m["albedo"], m["emission"] = kind, albedo, emission
    m["fuzz"], m["ior"], m["openpbr_index"] = fuzz, ior, openpbr
    return m


def floor_quad(material_index=0, size=20.0):
    q = np.zeros(1, QUAD_DT)
    q[0] = ((-size / 2, 0, -size / 2), (0, 0, size), (size, 0, 0), material_index)   # normal +y
    return q


def profile(spp, depth, rr=None, restir=False, candidates=1):
    p = O.Profile()
    p.samples_per_pixel = spp
    p.max_bounces = depth
    p.rr_start_bounce = depth if rr is None else rr
    p.enable_restir_di = 1 if restir else 0
    p.restir_initial_candidates = candidates
    p.restir_min_analytic_lights = 1
    p.restir_max_history_age = 20
    p.pixel_jitter = 0
    return p


def down_camera(height, w=24, h=18, f=30.0, x0=0.0, z0=0.0):
    """A pinhole camera at (x0, height, z0) looking straight down."""
    c = O.pinhole_camera(f, f, w / 2, h / 2, width=w, height=h)
    c.origin[:] = (x0, height, z0)
    c.basis_x[:] = (1.0, 0.0, 0.0)
    c.basis_y[:] = (0.0, 0.0, 1.0)
    c.basis_z[:] = (0.0, -1.0, 0.0)
    return c


def render(scene, cam, prof, stream=0):
    bvh = O.OracleBvh(scene)
    return O.render_camera(scene, bvh, cam, 0, prof, stream, threads=8)


def mc(scene, cam, prof, k=16):
    """Mean image and its standard error from k independent streams."""
    bvh = O.OracleBvh(scene)
    imgs = np.stack([O.render_camera(scene, bvh, cam, 0, prof, s, threads=8)["beauty"][..., :3]
                     for s in range(k)]).astype(np.float64)
    return imgs.mean(0), imgs.std(0, ddof=1) / math.sqrt(k)


def floor_points(cam):
    """World-space hit points of the pixel-centre rays on the plane y = 0."""
    rays = O.primary_rays(cam, 0, 0, 0, 0)
    o = rays["o"].astype(np.float64)
    d = rays["d"].astype(np.float64)
    t = -o[:, 1] / d[:, 1]
    return (o + d * t[:, None]).reshape(cam.height, cam.width, 3)


def assert_z(mean, stderr, expect, zmax=6.0, label=""):
    """All but 2 % of the values within zmax standard errors (plus a 2e-5
    relative floor: the errors are estimated from k streams and next-event
    estimates of nearby area lights are heavy-tailed, so a few values fall
    outside by chance) and the image mean within 4 standard errors of its
    own (a bias of a fraction of a percent fails this)."""
    tol = zmax * stderr + 2e-5 * np.abs(expect) + 1e-7
    bad = np.abs(mean - expect) > tol
    assert bad.mean() < 2e-2, f"{label}: {bad.sum()} values outside {zmax} sigma"
    se_mean = math.sqrt(float((stderr ** 2).sum())) / stderr.size
    assert abs(float((mean - expect).mean())) < 4 * se_mean + 1e-6, label


# ------------------------------------------------------------------ BSDFs
def scatter_scene(mat):
    q = floor_quad(0)
    s, keep = O.scene_from_arrays(quads=q, materials=mat)
    return s, keep


def test_lambert_sampling_has_cosine_density():
    s, keep = scatter_scene(material(0, albedo=(0.2, 0.5, 0.9)))
    n = 128
    cos = []
    for u0 in (np.arange(n) + 0.5) / n:
        for u1 in (np.arange(n) + 0.5) / n:
            alive, wi, thr, pdf, delta = O.scatter(s, 0, (0, 0, 0), (0, 1, 0), (0.3, -1, 0.1),
                                                   (u0, u1, 0.5, 0.5))
            assert alive and not delta
            assert np.allclose(thr, (0.2, 0.5, 0.9))
            assert abs(np.linalg.norm(wi) - 1) < 1e-6 and wi[1] >= 0
            assert abs(pdf - wi[1] / math.pi) < 1e-6
            cos.append(wi[1])
    cos = np.array(cos)
    assert abs(cos.mean() - 2 / 3) < 2e-3
    assert abs((cos ** 2).mean() - 1 / 2) < 2e-3


def test_metal_reflection_law_and_fuzz_cone():
    rng = np.random.default_rng(2)
    s, keep = scatter_scene(material(1, albedo=(0.9, 0.8, 0.7), fuzz=0.0))
    for _ in range(500):
        d = np.array([rng.uniform(-1, 1), -rng.uniform(0.1, 1), rng.uniform(-1, 1)])
        d /= np.linalg.norm(d)
        alive, wi, thr, pdf, delta = O.scatter(s, 0, (0, 0, 0), (0, 1, 0), d,
                                               rng.uniform(0, 1, 4))
        assert alive and delta and np.allclose(thr, (0.9, 0.8, 0.7))
        assert np.abs(wi - (d[0], -d[1], d[2])).max() < 1e-6
    f = 0.3
    s, keep = scatter_scene(material(1, albedo=(1, 1, 1), fuzz=f))
    worst = 0.0
    for _ in range(3000):
        d = np.array([rng.uniform(-1, 1), -rng.uniform(0.3, 1), rng.uniform(-1, 1)])
        d /= np.linalg.norm(d)
        alive, wi, thr, pdf, delta = O.scatter(s, 0, (0, 0, 0), (0, 1, 0), d,
                                               rng.uniform(0, 1, 4))
        if not alive:
            continue
        mirror = np.array([d[0], -d[1], d[2]])
        worst = max(worst, math.acos(min(1.0, float(np.dot(wi, mirror)))))
        assert wi[1] > 0
    assert worst <= math.asin(f) + 1e-5 and worst > 0.8 * math.asin(f)


def test_dielectric_snell_and_schlick():
    ior = 1.5
    s, keep = scatter_scene(material(2, ior=ior))
    rng = np.random.default_rng(4)
    n = 400
    for theta in (0.0, 0.3, 0.7, 1.2, 1.5):
        d = np.array([math.sin(theta), -math.cos(theta), 0.0])
        refl = 0
        for u in (np.arange(n) + 0.5) / n:
            alive, wi, thr, pdf, delta = O.scatter(s, 0, (0, 0, 0), (0, 1, 0), d,
                                                   (u, rng.uniform(), rng.uniform(), 0.5))
            assert alive and delta
            assert np.allclose(thr, (0.92, 0.95, 1.0))
            if wi[1] > 0:
                refl += 1
                assert np.abs(wi - (d[0], -d[1], d[2])).max() < 1e-6
            else:
                # entering glass from air: sin(theta_i) = ior * sin(theta_t)
                assert abs(math.sin(theta) - ior * math.hypot(wi[0], wi[2])) < 2e-6
                assert abs(wi[2]) < 1e-7
        r0 = ((1 - ior) / (1 + ior)) ** 2
        schlick = r0 + (1 - r0) * (1 - math.cos(theta)) ** 5
        assert abs(refl / n - schlick) <= 1.0 / n + 1e-9, (theta, refl / n, schlick)
    assert abs(((1 - ior) / (1 + ior)) ** 2 - 0.04) < 1e-12
    # leaving the glass: beyond the critical angle asin(1 / ior) all light reflects
    crit = math.asin(1 / ior)
    for theta in (crit + 0.01, 1.2):
        d = np.array([math.sin(theta), math.cos(theta), 0.0])   # travels along +y: a back-face hit
        for u in (0.0, 0.5, 0.999):
            alive, wi, thr, pdf, delta = O.scatter(s, 0, (0, 0, 0), (0, 1, 0), d, (u, 0, 0, 0))
            assert wi[1] < 0, "total internal reflection expected"
    theta = crit - 0.2
    d = np.array([math.sin(theta), math.cos(theta), 0.0])
    alive, wi, thr, pdf, delta = O.scatter(s, 0, (0, 0, 0), (0, 1, 0), d, (0.999, 0, 0, 0))
    assert wi[1] > 0 and abs(ior * math.sin(theta) - math.hypot(wi[0], wi[2])) < 2e-6


# --------------------------------------------------------------- radiance
@pytest.mark.parametrize("depth", [1, 3])
def test_lambert_sphere_under_uniform_sky(depth):
    le, a = 0.7, (0.25, 0.5, 0.8)
    sph = np.zeros(1, SPHERE_DT)
    sph["center"], sph["radius"], sph["material"] = (0, 0, 0), 1.0, 0
    s, keep = O.scene_from_arrays(spheres=sph, materials=material(0, albedo=a),
                                  background=(le, le, le))
    cam = down_camera(4.0, w=32, h=32, f=40.0)
    out = render(s, cam, profile(4, depth))
    img = out["beauty"][..., :3].astype(np.float64)
    on = out["depth"] > 0
    assert on.sum() > 300 and (~on).sum() > 100
    assert np.abs(img[on] - np.array(a) * le).max() < 2e-6
    assert np.abs(img[~on] - le).max() < 1e-7


def test_mirror_floor_reflects_the_sky_gradient():
    bottom, top, c = np.array([1.0, 1.0, 1.0]), np.array([0.5, 0.7, 1.0]), np.array([0.9, 0.6, 0.3])
    s, keep = O.scene_from_arrays(quads=floor_quad(0, 200.0), materials=material(1, albedo=c),
                                  background=tuple(bottom), background_top=tuple(top))
    cam = down_camera(3.0, w=32, h=24, f=20.0)
    cam.basis_z[:] = tuple(np.array([0.0, -0.8, 0.6]))
    cam.basis_y[:] = tuple(np.array([0.0, 0.6, 0.8]))   # basis_x x basis_y = basis_z
    out = render(s, cam, profile(1, 2))
    img = out["beauty"][..., :3].astype(np.float64)
    d = O.primary_rays(cam, 0, 0, 0, 0)["d"].astype(np.float64).reshape(24, 32, 3)
    r = d * np.array([1, -1, 1])
    t = 0.5 * (r[..., 1:2] + 1)
    expect = c * ((1 - t) * bottom + t * top)
    hit = d[..., 1] < 0
    assert hit.all()
    assert np.abs(img - expect).max() < 3e-6


def polygon_cos_integral(x, n, verts):
    """Lambert's formula: the cosine-weighted solid angle of a polygon."""
    total = 0.0
    for i in range(len(verts)):
        a = np.asarray(verts[i], np.float64) - x
        b = np.asarray(verts[(i + 1) % len(verts)], np.float64) - x
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        g = np.cross(a, b)
        total += math.acos(np.clip(np.dot(a, b), -1, 1)) * np.dot(n, g / np.linalg.norm(g))
    return abs(total) / 2


def lit_floor(kind, le=(6.0, 5.0, 4.0), rho=(0.6, 0.5, 0.4)):
    """Floor of albedo rho under one light; returns scene, keepalive and
    the expected outgoing radiance at floor point x."""
    mats = np.concatenate([material(0, albedo=rho), material(3, emission=le)])
    quads, tris, sph, ana = [floor_quad(0)], None, None, None
    rho_pi = np.array(rho) / math.pi
    if kind == "quad":
        q = np.zeros(1, QUAD_DT)
        q[0] = ((-0.8, 2.0, -0.6), (1.7, 0, 0), (0, 0.25, 1.3), 1)    # normal toward -y
        quads.append(q)
        verts = [(-0.8, 2.0, -0.6), (0.9, 2.0, -0.6), (0.9, 2.25, 0.7), (-0.8, 2.25, 0.7)]
        expect = lambda x: rho_pi * np.array(le) * polygon_cos_integral(x, (0, 1, 0), verts)
    elif kind == "triangle":
        tris = np.zeros(1, TRI_DT)
        tris[0] = ((0.5, 2.5, -1.0), (1.0, 2.0, 1.0), (-1.0, 3.0, 0.2), 1)
        verts = [tuple(tris[0][k]) for k in ("p0", "p1", "p2")]
        expect = lambda x: rho_pi * np.array(le) * polygon_cos_integral(x, (0, 1, 0), verts)
    else:
        centre, r = np.array([0.4, 2.2, -0.3]), 0.5

        def expect(x):
            v = centre - x
            d2 = float(np.dot(v, v))
            return np.array(rho) * np.array(le) * r * r * (v[1] / math.sqrt(d2)) / d2
        if kind == "sphere":
            sph = np.zeros(1, SPHERE_DT)
            sph["center"], sph["radius"], sph["material"] = tuple(centre), r, 1
        else:
            ana = analytic_lights(O, [{"type": "sphere", "position": tuple(centre), "radius": r,
                                       "radiance": le}])
    s, keep = O.scene_from_arrays(spheres=sph, quads=np.concatenate(quads), triangles=tris,
                                  materials=mats, analytic=ana)
    return s, keep, expect


@pytest.mark.parametrize("kind", ["quad", "triangle", "sphere", "analytic"])
@pytest.mark.parametrize("depth", [1, 3])
def test_floor_irradiance_under_one_light(kind, depth):
    s, keep, expect = lit_floor(kind)
    cam = down_camera(1.2, w=24, h=18, f=14.0)
    x = floor_points(cam)
    ref = np.stack([[expect(x[j, i]) for i in range(cam.width)] for j in range(cam.height)])
    mean, se = mc(s, cam, profile(24, depth))
    assert_z(mean, se, ref, label=f"{kind} depth {depth}")


@pytest.mark.parametrize("candidates", [1, 8])
def test_restir_ris_without_reuse_is_unbiased(candidates):
    s, keep, expect = lit_floor("analytic")
    cam = down_camera(1.2, w=24, h=18, f=14.0)
    x = floor_points(cam)
    ref = np.stack([[expect(x[j, i]) for i in range(cam.width)] for j in range(cam.height)])
    mean, se = mc(s, cam, profile(24, 1, restir=True, candidates=candidates))
    assert_z(mean, se, ref, label=f"RIS {candidates}")


def furnace_scene(albedo, le):
    """A closed 2 x 2 x 2 box of six inward-facing emissive OpenPBR walls."""
    m = O.openpbr_material(base_color=(albedo,) * 3, specular_weight=0.0, emission_luminance=le,
                           emission_color=(1, 1, 1))
    pbr = np.frombuffer(bytes(m), dtype=np.dtype((np.void, C.sizeof(O.OpenPbr)))).copy()
    mats = material(5, openpbr=0)
    q = np.zeros(6, QUAD_DT)
    lo, e = -1.0, 2.0
    ex, ey, ez = np.array([e, 0, 0]), np.array([0, e, 0]), np.array([0, 0, e])
    corner, far = np.array([lo, lo, lo]), np.array([-lo, -lo, -lo])
    walls = [(corner, ex, ez), (far, -ex, -ez), (corner, ex, ey), (far, -ex, -ey),
             (corner, ey, ez), (far, -ey, -ez)]
    for i, (o, u, v) in enumerate(walls):
        centre = o + 0.5 * (u + v)
        if np.dot(np.cross(u, v), -centre) < 0:       # make every wall face the inside
            u, v = v, u
        q[i] = (tuple(o), tuple(u), tuple(v), 0)
        assert np.dot(np.cross(q[i]["edge_u"], q[i]["edge_v"]), -centre) > 0
    s, keep = O.scene_from_arrays(quads=q, materials=mats, openpbr=pbr)
    return s, keep + [pbr], float(O.openpbr_emission(m)[0])


def test_openpbr_furnace_wall_is_lambertian():
    """The furnace's wall material (base only, no specular or coat) reflects
    albedo / pi in every direction pair."""
    m = O.openpbr_material(base_color=(0.6,) * 3, specular_weight=0.0)
    rng = np.random.default_rng(9)
    for _ in range(200):
        wo = rng.normal(size=3)
        wi = rng.normal(size=3)
        wo[1], wi[1] = abs(wo[1]), abs(wi[1])
        wo /= np.linalg.norm(wo)
        wi /= np.linalg.norm(wi)
        value, pdf = O.openpbr_evaluate(m, (0, 1, 0), wo, wi)
        assert np.abs(np.asarray(value) - 0.6 / math.pi).max() < 2e-6
        assert abs(pdf - wi[1] / math.pi) < 2e-6


@pytest.mark.parametrize("depth, rr", [(1, None), (2, None), (4, None), (4, 1), (6, 2)])
def test_white_furnace(depth, rr):
    a, le_in = 0.6, 1.5
    s, keep, le = furnace_scene(a, le_in)
    cam = O.pinhole_camera(10.0, 10.0, 8.0, 6.0, width=16, height=12)
    cam.origin[:] = (0.1, -0.2, 0.05)
    mean, se = mc(s, cam, profile(16, depth, rr=rr))
    expect = le * sum(a ** k for k in range(depth + 1))
    assert_z(mean, se, np.full(mean.shape, expect), label=f"furnace D={depth} rr={rr}")
