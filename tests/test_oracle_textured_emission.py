"""The oracle's textured emission and OpenPBR texture bindings (oracle_render.c
make_hit / sample_light / material_emission_lum / scattering_material, after
src/realtime/gpu/programs.cu:944-1024,1057-1232 and
src/realtime/gpu/packed_scene_preparation.cpp:195-248), pinned by what the
definitions fix:

- the sRGB decode of an OpenPBR colour binding: the reference's known answers
  (0.04045, 0.5, 1) -> (0.0031308, 0.21404114, 1) and the raw / linear
  bypass (tests/test_openpbr_core.cpp:243-251);
- light selection by texture average: a constant (negative channels
  clamped), a checker (the mean of its two children, nothing past 8 levels),
  an image (its texel mean), noise (0.5), a bad index (no light), an OpenPBR
  emission binding in sRGB; selection pdfs equal area x luminance of the
  average, normalised together with the environment;
- the hit shows the texture: camera rays on an image-textured quad light see
  texel (floor(u W), floor((1 - v) H)) of their edge coordinates, black from
  behind; a material other than a light emits nothing whatever its texture;
- a light sample's emission equals what a ray to the sampled point sees, for
  sphere (sphere_uv), quad (edge coordinates) and triangle (per-vertex
  coordinates) lights;
- closed-form radiance: a diffuse floor under a quad light with a 3 x 2
  image emission, and under a triangle light whose per-vertex coordinates
  select 2 x 2 texels, reflects albedo / pi x the sum over texel regions of
  the texel x Lambert's polygon integral, at depth 1 (next-event only) and
  depth 3 (next-event and BSDF hits under MIS);
- each OpenPBR binding drives its own input: a constant texture bound to an
  input renders bit for bit like the input set to that value (the record
  holding another), an sRGB binding like the decoded value, and a scalar
  binding reads the first channel;
- a subsurface walk begun under a metalness texture of 0 on a record of
  metalness 1 renders like a record of metalness 0 (the walk keeps the
  material of its entry point).
A doubled texel, a v axis not flipped, swapped quad edges or a binding
applied to the wrong input each fail one of these."""
import math

import numpy as np
import pytest

import oracle_ctypes as O
from helpers import MAT_DT, QUAD_DT, SPHERE_DT, TRI_DT, attach_textures, pbr_bindings, texture
from test_oracle_integrator import (assert_z, down_camera, floor_points, floor_quad, material, mc,
                                    polygon_cos_integral, profile)


def lum(c):
    c = np.maximum(np.asarray(c, np.float64), 0.0)
    return max(0.0, 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2])


def srgb_decode(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def pbr_record(**kw):
    raw = bytes(O.openpbr_material(**kw))
    return np.frombuffer(raw, dtype=np.dtype((np.void, len(raw)))).copy()


def quad(origin, eu, ev, mat):
    q = np.zeros(1, QUAD_DT)
    q[0] = (origin, eu, ev, mat)
    return q


# ------------------------------------------------------------ sRGB decode
@pytest.mark.parametrize("space,expect", [(0, (0.04045, 0.5, 1.0)), (1, (0.04045, 0.5, 1.0)),
                                          (2, (0.0031308, 0.21404114, 1.0))])
def test_openpbr_emission_binding_decodes_srgb(space, expect):
    mats = material(5, openpbr=0)
    pbr = pbr_record(emission_luminance=1.0, emission_color=(0.7, 0.7, 0.7))
    s, keep = O.scene_from_arrays(quads=quad((-1, -1, 0), (2, 0, 0), (0, 2, 0), 0), materials=mats,
                                  openpbr=pbr)
    attach_textures(s, keep, textures=[texture("constant", (0.04045, 0.5, 1.0))],
                    pbr=pbr_bindings(1, emission_color=[(0, 0, space)]))
    _, emission, _ = O.hit_material(s, (0.2, 0.1, 3.0), (0, 0, -1), 3.0, 0)
    np.testing.assert_allclose(emission, expect, atol=1e-6)


# ------------------------------------------------------ light selection
def test_light_selection_uses_texture_averages():
    rng = np.random.default_rng(3)
    img = rng.uniform(0.0, 2.0, (2, 3, 3)).astype(np.float32)
    img2 = rng.uniform(0.0, 1.0, (3, 2, 3)).astype(np.float32)
    texels = np.concatenate([img.reshape(-1, 3), img2.reshape(-1, 3)])
    T = [texture("constant", (2.0, -1.0, 0.5)),                       # 0
         texture("constant", (1.0, 1.0, 1.0)),                        # 1
         texture("constant", (0.0, 0.5, 0.0)),                        # 2
         texture("checker", even=1, odd=2, scale=0.3),                # 3
         texture("image", offset=0, w=3, h=2),                        # 4
         texture("noise", scale=4.0),                                 # 5
         texture("image", offset=6, w=2, h=3)]                        # 6
    # checker chains: 7 levels then a constant (depth 7: counted), 8 levels (nothing)
    def chain(levels, leaf):
        first = len(T)
        for k in range(levels):
            nxt = first + k + 1 if k + 1 < levels else leaf
            T.append(texture("checker", even=nxt, odd=nxt, scale=1.0))
        return first
    c7 = chain(7, 1)
    c8 = chain(8, 1)
    # materials: lights 0..7 with emission textures, 8 a constant light, 9 OpenPBR
    tex_of = [0, 3, 4, 5, 99, c7, c8, -1]
    mats = np.concatenate([material(3, emission=(9.0, 9.0, 9.0)) for _ in tex_of]
                          + [material(3, emission=(1.0, 2.0, 3.0)), material(5, openpbr=0)])
    emission_tex = tex_of + [-1, -1]
    pbr = pbr_record(emission_luminance=2.0, emission_color=(0.3, 0.3, 0.3))
    quads = np.concatenate([quad((3.0 * i, 0, 0), (1.0 + 0.25 * i, 0, 0), (0, 0, 1.0), i)
                            for i in range(len(mats))])
    bg = (0.2, 0.3, 0.4)
    s, keep = O.scene_from_arrays(quads=quads, materials=mats, openpbr=pbr, background=bg)
    attach_textures(s, keep, textures=T, texels=texels, emission=emission_tex,
                    pbr=pbr_bindings(1, emission_color=[(0, 6, 2)]))
    avg = {0: np.maximum([2.0, -1.0, 0.5], 0), 3: 0.5 * (np.ones(3) + np.array([0, 0.5, 0])),
           4: img.reshape(-1, 3).astype(np.float64).mean(0), 5: np.full(3, 0.5), 99: np.zeros(3),
           c7: np.ones(3), c8: np.zeros(3)}
    weights = []
    for i in range(len(mats)):
        area = 1.0 + 0.25 * i
        if i < len(tex_of) and tex_of[i] >= 0:
            e = avg[tex_of[i]]
        elif i == len(tex_of) - 1:
            e = np.array([9.0, 9.0, 9.0])      # emission texture -1: the constant
        elif i == len(tex_of):
            e = np.array([1.0, 2.0, 3.0])
        else:
            e = 2.0 * srgb_decode(img2.reshape(-1, 3).astype(np.float64).mean(0))
        w = area * lum(e)
        if w > 1e-12:
            weights.append((1, i, w))
    weights.append((3, -1, 4 * math.pi * lum(bg)))
    total = sum(w for _, _, w in weights)
    got = O.build_lights(s)
    assert [(t, p) for t, p, _, _ in got] == [(t, p) for t, p, _ in weights]
    for (t, p, w), (_, _, pdf, _) in zip(weights, got):
        assert pdf == pytest.approx(w / total, rel=1e-5), (t, p)
    assert got[-1][3] == 1.0
    # the bad index and the 8-deep chain dropped their lights
    assert 4 not in [p for _, p, _, _ in got] and 6 not in [p for _, p, _, _ in got]


# --------------------------------------------------------- hit emission
def test_hit_shows_the_emission_texture():
    rng = np.random.default_rng(5)
    img = rng.uniform(0.1, 3.0, (2, 3, 3)).astype(np.float32)
    mats = np.concatenate([material(3, emission=(5, 5, 5)), material(0, albedo=(0.5, 0.5, 0.5))])
    quads = np.concatenate([quad((-1, -1, 0), (2, 0, 0), (0, 2, 0), 0),       # normal +z
                            quad((-1, -1, -2), (2, 0, 0), (0, 2, 0), 1)])
    s, keep = O.scene_from_arrays(quads=quads, materials=mats)
    attach_textures(s, keep, textures=[texture("image", w=3, h=2)], texels=img, emission=[0, 0])
    for _ in range(200):
        ix, iy = rng.integers(0, 3), rng.integers(0, 2)
        u = (ix + rng.uniform(0.02, 0.98)) / 3.0
        v = 1.0 - (iy + rng.uniform(0.02, 0.98)) / 2.0
        x, y = 2.0 * u - 1.0, 2.0 * v - 1.0
        _, e, uv = O.hit_material(s, (x, y, 4.0), (0, 0, -1), 4.0, 0)
        np.testing.assert_allclose(uv, (u, v), atol=1e-6)
        assert np.array_equal(e, img[iy, ix]), (ix, iy)
        _, back, _ = O.hit_material(s, (x, y, -1.0), (0, 0, 1), 1.0, 0)
        assert not back.any()
        _, dull, _ = O.hit_material(s, (x, y, 1.0), (0, 0, -1), 3.0, 1)
        assert not dull.any()


# ----------------------------------------------- light samples vs hits
def one_light_scene(kind, img):
    mats = np.concatenate([material(0, albedo=(0.5, 0.5, 0.5)), material(3, emission=(1, 1, 1))])
    sph = tri = None
    quads = [floor_quad(0)]
    uvs = None
    if kind == "sphere":
        sph = np.zeros(1, SPHERE_DT)
        sph[0] = ((0.3, 2.0, -0.2), 0.6, 1)
        prim = 0
    elif kind == "quad":
        quads.append(quad((-0.7, 1.8, -0.5), (1.4, 0.0, 0.1), (0.1, 0.3, 1.1), 1))
        prim = 1
    else:
        tri = np.zeros(1, TRI_DT)
        tri[0] = ((0.5, 2.5, -1.0), (1.0, 2.0, 1.0), (-1.0, 3.0, 0.2), 1)
        uvs = np.array([[0.1, 0.2, 0.9, 0.3, 0.4, 0.95]], np.float32)
        prim = 1
    s, keep = O.scene_from_arrays(spheres=sph, quads=np.concatenate(quads), triangles=tri,
                                  materials=mats)
    attach_textures(s, keep, textures=[texture("image", w=img.shape[1], h=img.shape[0], filt=1)],
                    texels=img, emission=[-1, 0], triangle_uvs=uvs)
    return s, keep, prim


@pytest.mark.parametrize("kind", ["sphere", "quad", "triangle"])
def test_light_sample_emission_equals_hit_emission(kind):
    rng = np.random.default_rng(11)
    img = rng.uniform(0.2, 4.0, (3, 4, 3)).astype(np.float32)
    s, keep, prim = one_light_scene(kind, img)
    bvh = O.OracleBvh(s)
    pos, gn = np.array([0.1, 0.0, 0.05], np.float32), np.array([0, 1, 0], np.float32)
    floor = 1 if kind == "sphere" else 0      # global primitive id of the receiving floor
    seen = 0
    for _ in range(300):
        r = rng.uniform(0, 1, 4).astype(np.float32)
        li, d, e, dist, pdf = O.sample_light_at(s, bvh, floor, pos, gn, (0, -1, 0), r)
        if li < 0:
            continue
        seen += 1
        _, he, _ = O.hit_material(s, pos, d, dist, prim)
        np.testing.assert_allclose(e, he, rtol=2e-4, atol=1e-6)
    assert seen > 200
    # the texture varies over the light: the samples see more than one colour
    assert len({tuple(np.round(O.sample_light_at(s, bvh, floor, pos, gn, (0, -1, 0),
                                                 rng.uniform(0, 1, 4).astype(np.float32))[2], 3))
                for _ in range(20)}) > 5


# ------------------------------------------------ closed-form radiance
def clip(poly, a, b, c):
    """Sutherland-Hodgman: the part of a convex 2-D polygon where a x + b y <= c."""
    out = []
    for i in range(len(poly)):
        p, q = poly[i], poly[(i + 1) % len(poly)]
        fp, fq = a * p[0] + b * p[1] - c, a * q[0] + b * q[1] - c
        if fp <= 0:
            out.append(p)
        if fp * fq < 0:
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def textured_floor(kind, rho=(0.6, 0.5, 0.4)):
    """A floor of albedo rho under one light with an image emission; returns
    scene, keepalive and the expected outgoing radiance at a floor point."""
    rng = np.random.default_rng(17)
    mats = np.concatenate([material(0, albedo=rho), material(3, emission=(50, 50, 50))])
    rho_pi = np.array(rho) / math.pi
    n = (0.0, 1.0, 0.0)
    if kind == "quad":
        img = rng.uniform(1.0, 9.0, (2, 3, 3)).astype(np.float32)
        o, eu, ev = np.array([-0.8, 2.0, -0.6]), np.array([1.7, 0, 0]), np.array([0, 0.25, 1.3])
        quads = np.concatenate([floor_quad(0), quad(tuple(o), tuple(eu), tuple(ev), 1)])
        s, keep = O.scene_from_arrays(quads=quads, materials=mats)
        regions = []
        for iy in range(2):
            for ix in range(3):
                u0, u1 = ix / 3, (ix + 1) / 3
                v0, v1 = 1 - (iy + 1) / 2, 1 - iy / 2      # image rows run down the v axis
                verts = [o + u * eu + v * ev for u, v in ((u0, v0), (u1, v0), (u1, v1), (u0, v1))]
                regions.append((img[iy, ix].astype(np.float64), verts))
        uvs = None
    else:
        img = rng.uniform(1.0, 9.0, (2, 2, 3)).astype(np.float32)
        P = np.array([(0.5, 2.5, -1.0), (1.0, 2.0, 1.0), (-1.0, 3.0, 0.2)])
        tri = np.zeros(1, TRI_DT)
        tri[0] = (tuple(P[0]), tuple(P[1]), tuple(P[2]), 1)
        # uv0 = (1, 0), uv1 = (0, 1), uv2 = (0, 0): u = w0, v = w1
        uvs = np.array([[1.0, 0.0, 0.0, 1.0, 0.0, 0.0]], np.float32)
        s, keep = O.scene_from_arrays(quads=floor_quad(0), triangles=tri, materials=mats)
        regions = []
        for iy in range(2):
            for ix in range(2):
                poly = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)]           # (w0, w1)
                poly = clip(poly, -1, 0, -0.5) if ix else clip(poly, 1, 0, 0.5)
                poly = clip(poly, 0, 1, 0.5) if iy else clip(poly, 0, -1, -0.5)
                if len(poly) < 3:
                    continue
                verts = [w0 * P[0] + w1 * P[1] + (1 - w0 - w1) * P[2] for w0, w1 in poly]
                regions.append((img[iy, ix].astype(np.float64), verts))
    attach_textures(s, keep, textures=[texture("image", w=img.shape[1], h=img.shape[0])],
                    texels=img, emission=[-1, 0], triangle_uvs=uvs)

    def expect(x):
        return rho_pi * sum(c * polygon_cos_integral(x, n, v) for c, v in regions)
    return s, keep, expect


@pytest.mark.parametrize("kind", ["quad", "triangle"])
@pytest.mark.parametrize("depth", [1, 3])
def test_floor_under_image_textured_light(kind, depth):
    s, keep, expect = textured_floor(kind)
    cam = down_camera(1.2, w=24, h=18, f=14.0)
    x = floor_points(cam)
    ref = np.stack([[expect(x[j, i]) for i in range(cam.width)] for j in range(cam.height)])
    mean, se = mc(s, cam, profile(24, depth))
    assert_z(mean, se, ref, label=f"{kind} depth {depth}")


# -------------------------------------------------------- OpenPBR bindings
def pbr_scene(record, bindings=None, textures=None):
    mats = np.concatenate([material(0, albedo=(0.6, 0.6, 0.6)), material(3, emission=(8, 7, 6)),
                           material(5, openpbr=0)])
    sph = np.zeros(1, SPHERE_DT)
    sph[0] = ((0.0, 0.6, 0.0), 0.6, 2)
    quads = np.concatenate([floor_quad(0), quad((-1.0, 3.0, -1.0), (2.0, 0, 0), (0, 0, 2.0), 1)])
    s, keep = O.scene_from_arrays(spheres=sph, quads=quads, materials=mats, openpbr=record,
                                  background=(0.1, 0.12, 0.15))
    if bindings is not None:
        attach_textures(s, keep, textures=textures, pbr=bindings)
    return s, keep


def side_camera():
    c = O.pinhole_camera(20.0, 20.0, 12.0, 9.0, width=24, height=18)
    c.origin[:] = (0.0, 0.8, -3.0)
    c.basis_x[:] = (1.0, 0.0, 0.0)
    c.basis_y[:] = (0.0, -1.0, 0.0)
    c.basis_z[:] = (0.0, 0.0, 1.0)
    return c


def pbr_render(s):
    bvh = O.OracleBvh(s)
    return O.render_camera(s, bvh, side_camera(), 0, profile(4, 4), 2, threads=8)["beauty"][..., :3]


BASE = dict(base_color=(0.7, 0.5, 0.3), specular_roughness=0.35, transmission_weight=0.4,
            transmission_color=(0.8, 0.9, 0.7), specular_color=(0.9, 0.95, 1.0), base_metalness=0.2,
            emission_luminance=1.5, emission_color=(0.3, 0.2, 0.1))
COLOR_INPUTS = {"base_color": (0.2, 0.8, 0.4), "specular_color": (0.5, 0.3, 0.9),
                "transmission_color": (0.3, 0.6, 0.95), "emission_color": (0.9, 0.1, 0.5)}
SCALAR_INPUTS = {"base_metalness": 0.75, "specular_roughness": 0.08}


@pytest.mark.parametrize("name", list(COLOR_INPUTS) + list(SCALAR_INPUTS))
def test_constant_texture_binding_equals_the_input(name):
    value = COLOR_INPUTS.get(name, SCALAR_INPUTS.get(name))
    direct = pbr_render(pbr_scene(pbr_record(**{**BASE, name: value}))[0])
    # the record holds the base value; a constant texture supplies `value`
    texel = value if name in COLOR_INPUTS else (value, 0.9, 0.9)
    bound = pbr_render(pbr_scene(pbr_record(**BASE), pbr_bindings(1, **{name: [(0, 0)]}),
                                 [texture("constant", texel)])[0])
    unbound = pbr_render(pbr_scene(pbr_record(**BASE))[0])
    assert np.array_equal(direct, bound), name
    assert not np.array_equal(direct, unbound), name


@pytest.mark.parametrize("name", list(COLOR_INPUTS))
def test_srgb_texture_binding_equals_the_decoded_input(name):
    t = np.array(COLOR_INPUTS[name], np.float32)
    direct = pbr_render(pbr_scene(pbr_record(**{**BASE, name: tuple(srgb_decode(t))}))[0])
    bound = pbr_render(pbr_scene(pbr_record(**BASE), pbr_bindings(1, **{name: [(0, 0, 2)]}),
                                 [texture("constant", tuple(t))])[0])
    np.testing.assert_allclose(bound, direct, rtol=2e-5, atol=1e-7)


def test_subsurface_walk_keeps_its_entry_material():
    sss = dict(base_color=(0.8, 0.6, 0.5), specular_roughness=0.4, subsurface_weight=1.0,
               subsurface_color=(0.9, 0.5, 0.3), subsurface_radius=0.3,
               subsurface_radius_scale=(1.0, 0.5, 0.25))
    direct = pbr_render(pbr_scene(pbr_record(**sss, base_metalness=0.0))[0])
    bound = pbr_render(pbr_scene(pbr_record(**sss, base_metalness=1.0),
                                 pbr_bindings(1, base_metalness=[(0, 0)]),
                                 [texture("constant", (0.0, 0.0, 0.0))])[0])
    metal = pbr_render(pbr_scene(pbr_record(**sss, base_metalness=1.0))[0])
    assert np.array_equal(direct, bound)
    assert not np.array_equal(direct, metal)
