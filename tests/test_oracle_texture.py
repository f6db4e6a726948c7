"""The oracle's texture evaluation (oracle/oracle_texture.c, after
src/realtime/gpu/programs.cu:183-302) pinned by what the definitions fix:
the noise sine against numpy's sin, checker parity on hand-counted cells
(negative coordinates, scale 0, the 8-deep chain limit, bad indices),
nearest and bilinear texel picks under each address mode, and the texture
coordinates of known points on a sphere, a quad and a triangle."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ctypes as O

MAGENTA = [1.0, 0.0, 1.0]


def scene(textures, texels=None, spheres=None, quads=None, triangles=None):
    s = O.Scene()
    keep = [(O.Texture * len(textures))(*textures)]
    s.texture_count = len(textures)
    s.textures = C.cast(keep[0], C.c_void_p)
    if texels is not None:
        t = np.ascontiguousarray(texels, np.float32).reshape(-1, 3)
        keep.append(t)
        s.image_texel_count = t.shape[0]
        s.image_texels = t.ctypes.data_as(C.c_void_p)
    for name, T, items in (("spheres", O.Sphere, spheres), ("quads", O.Quad, quads),
                           ("triangles", O.Triangle, triangles)):
        if items:
            arr = (T * len(items))(*items)
            keep.append(arr)
            setattr(s, name, C.cast(arr, C.POINTER(T)))
            setattr(s, name[:-1] + "_count" if name != "spheres" else "sphere_count", len(items))
    s._keep = keep
    return s


def tex(type=0, color=(0, 0, 0), even=-1, odd=-1, scale=1.0, offset=0, w=0, h=0, um=1, vm=1, f=0):
    t = O.Texture()
    t.type, t.color, t.even_texture, t.odd_texture, t.scale = type, O.f3(*color), even, odd, scale
    t.image_offset, t.image_width, t.image_height = offset, w, h
    t.u_address_mode, t.v_address_mode, t.filter_type = um, vm, f
    return t


def test_noise_sine_matches_sin():
    L = O.lib()
    x = np.concatenate([np.linspace(-50, 50, 20001), np.linspace(-3e4, 3e4, 20001)]).astype(np.float32)
    got = np.array([L.rto_noise_sin(float(v)) for v in x], np.float64)
    err = np.abs(got - np.sin(x.astype(np.float64)))
    tol = 3e-7 + 1.5e-8 * np.abs(x)          # argument rounding grows with |x|
    assert (err <= tol).all(), float(err.max())
    assert L.rto_noise_sin(0.0) == 0.0 and L.rto_noise_sin(-0.0) == 0.0
    assert math.isnan(L.rto_noise_sin(float("inf")))


def test_checker_cells():
    s = scene([tex(0, (1, 1, 1)), tex(0, (0, 0, 0)), tex(1, even=0, odd=1, scale=2.0),
               tex(1, even=0, odd=1, scale=0.0)])
    # cells floor(p / 2): (0,0,0) even, (1,0,0) odd, (-1,0,0) odd, (-1,-1,0) even
    for p, want in (((0.5, 0.5, 0.5), 1), ((2.5, 0.5, 0.5), 0), ((-0.5, 0.5, 0.5), 0),
                    ((-0.5, -0.5, 1.0), 1), ((-3.0, 5.0, -7.5), 1)):
        assert O.eval_texture(s, 2, 0, 0, p).tolist() == [want] * 3, p
    assert O.eval_texture(s, 3, 0, 0, (2.5, 0, 0)).tolist() == [1, 1, 1]   # scale 0 -> even
    assert O.eval_texture(s, 7, 0, 0, (0, 0, 0)).tolist() == MAGENTA
    assert O.eval_texture(s, -1, 0, 0, (0, 0, 0)).tolist() == MAGENTA


def test_checker_chain_depth():
    # texture i (1..9) is a scale-0 checker whose even child is i - 1; 0 is white
    chain = [tex(0, (1, 1, 1))] + [tex(1, even=i - 1, odd=i - 1, scale=0.0) for i in range(1, 10)]
    s = scene(chain)
    assert O.eval_texture(s, 7, 0, 0, (0, 0, 0)).tolist() == [1, 1, 1]    # 7 links + the leaf
    assert O.eval_texture(s, 8, 0, 0, (0, 0, 0)).tolist() == MAGENTA      # 9 resolutions


def test_image_lookup_and_addressing():
    # 3 x 2 image, texel (x, y) = (x, y, 10) with row 0 at the top
    texels = [(x, y, 10) for y in range(2) for x in range(3)]
    nearest = [tex(2, offset=0, w=3, h=2, um=m, vm=1, f=0) for m in range(4)]
    s = scene(nearest + [tex(2, offset=0, w=3, h=2, f=1)], texels)
    ev = lambda i, u, v: O.eval_texture(s, i, u, v, (0, 0, 0)).tolist()
    assert ev(1, 0.1, 0.9) == [0, 0, 10]      # v = 0.9 is the top row
    assert ev(1, 0.5, 0.1) == [1, 1, 10]
    assert ev(1, 1.0, 0.0) == [2, 1, 10]      # u = 1 clamps into the last column
    assert ev(0, 1.25, 0.5) == [0, 0, 0]      # border: black outside [0, 1]
    assert ev(2, 1.1, 0.9) == [0, 0, 10]      # repeat: 1.1 -> 0.1
    assert ev(3, 1.9, 0.9) == [0, 0, 10]      # mirror: 1.9 -> 0.1
    assert ev(3, 1.1, 0.9) == [2, 0, 10]      # mirror: 1.1 -> 0.9
    assert ev(1, -0.5, 0.9) == [0, 0, 10]     # clamp
    # bilinear: texel centres at u = x / 2, v = 1 - y; halfway averages
    assert ev(4, 0.5, 1.0) == [1, 0, 10]
    assert ev(4, 0.25, 0.5) == pytest.approx([0.5, 0.5, 10])
    # no texels: cyan; a texel index past the array: cyan
    s2 = scene([tex(2, offset=5, w=3, h=2)], texels)
    assert O.eval_texture(s2, 0, 0.9, 0.1, (0, 0, 0)).tolist() == [0, 1, 1]


def test_noise_texture_range_and_formula():
    s = scene([tex(3, scale=4.0)])
    rng = np.random.default_rng(3)
    L = O.lib()
    for p in rng.uniform(-20, 20, (200, 3)).astype(np.float32):
        c = O.eval_texture(s, 0, 0, 0, p)
        assert c[0] == c[1] == c[2] and 0.0 <= c[0] <= 1.0
        turb = L.rto_turbulence(p.ctypes.data_as(C.c_void_p))
        assert 0.0 <= turb <= 2.0                       # 7 octaves of weight 1, 1/2, ...
        want = 0.5 * (1.0 + math.sin(np.float32(4.0) * p[2] + np.float32(10.0) * turb))
        assert abs(c[0] - want) < 2e-6


def test_surface_uv():
    sp = O.Sphere(); sp.center, sp.radius = O.f3(1, 2, 3), 2.0
    q = O.Quad(); q.origin, q.edge_u, q.edge_v = O.f3(0, 0, 0), O.f3(4, 0, 0), O.f3(0, 0, 2)
    tr = O.Triangle(); tr.p0, tr.p1, tr.p2 = O.f3(0, 0, 0), O.f3(2, 0, 0), O.f3(0, 2, 0)
    s = scene([tex()], spheres=[sp], quads=[q], triangles=[tr])
    # sphere: +x is the equator at u = 1/2, -y the pole v = 0, +y the pole v = 1
    assert O.surface_uv(s, (0, 0, 0), (1, 0, 0), 0, 0, (3, 2, 3)) == pytest.approx((0.5, 0.5))
    assert O.surface_uv(s, (0, 0, 0), (1, 0, 0), 0, 0, (1, 0, 3))[1] == pytest.approx(0.0, abs=1e-6)
    assert O.surface_uv(s, (0, 0, 0), (1, 0, 0), 0, 0, (1, 4, 3))[1] == pytest.approx(1.0)
    # -z: phi = atan2(1, 0) + pi = 3 pi / 2 -> u = 3/4
    assert O.surface_uv(s, (0, 0, 0), (1, 0, 0), 0, 0, (1, 2, 1))[0] == pytest.approx(0.75)
    # quad: origin + 0.25 edge_u + 0.5 edge_v
    assert O.surface_uv(s, (0, 1, 0), (0, -1, 0), 1, 0, (1, 0, 1)) == pytest.approx((0.25, 0.5))
    # triangle: a ray hitting p0 + 0.3 e1 + 0.2 e2 = (0.6, 0.4, 0)
    uv = O.surface_uv(s, (0.6, 0.4, 5), (0, 0, -1), 2, 0, (0.6, 0.4, 0))
    assert uv == pytest.approx((0.3, 0.2), abs=1e-6)
    # per-vertex coordinates: uv0 w + uv1 u + uv2 v with (w, u, v) = (0.5, 0.3, 0.2)
    tuv = np.array([[0.1, 0.9, 0.7, 0.5, 0.2, 0.3]], np.float32)
    s.triangle_uvs = tuv.ctypes.data_as(C.c_void_p)
    uv = O.surface_uv(s, (0.6, 0.4, 5), (0, 0, -1), 2, 0, (0.6, 0.4, 0))
    assert uv == pytest.approx((0.1 * 0.5 + 0.7 * 0.3 + 0.2 * 0.2, 0.9 * 0.5 + 0.5 * 0.3 + 0.3 * 0.2), abs=1e-6)
    tuv[0, 0] = np.nan                     # no coordinates: barycentrics
    uv = O.surface_uv(s, (0.6, 0.4, 5), (0, 0, -1), 2, 0, (0.6, 0.4, 0))
    assert uv == pytest.approx((0.3, 0.2), abs=1e-6)


def test_vertex_normals_set_the_shading_normal():
    """programs.cu:1535-1545: with one tilted normal n at every vertex, each
    primary hit on that triangle gets the shading normal +-normalize(n)
    (the sign facing the viewer); the normal AOV (0.5 s + 0.5) changes
    exactly on those pixels."""
    import ray_tracing_amd as rt
    from helpers import oracle_camera, oracle_profile, oracle_scene
    from test_gpu_textures import textured_cornell
    desc, keep, rig = textured_cornell(rt)
    prof = rt.Generated("cornell_mesh").profile
    prof.samples_per_pixel, prof.max_bounces, prof.rr_start_bounce = 1, 1, 1
    cam = oracle_camera(O, rig.cameras[0])
    nv = np.array([0.3, 0.2, -1.0], np.float32)
    out = []
    for normals in (None, np.array([[np.nan] * 9, list(nv) * 3], np.float32)):
        desc.triangle_normals = normals.ctypes.data_as(C.c_void_p) if normals is not None else None
        osc = oracle_scene(O, desc)
        out.append(O.render_camera(osc, O.OracleBvh(osc), cam, 0, oracle_profile(O, prof), 3,
                                   threads=8)["normal"][..., :3])
    changed = np.any(out[0] != out[1], axis=-1)
    assert changed.sum() > 500
    s = nv / np.linalg.norm(nv)
    enc = out[1][changed]
    ok = np.minimum(np.abs(enc - (0.5 * s + 0.5)).max(-1), np.abs(enc - (-0.5 * s + 0.5)).max(-1))
    assert ok.max() < 2e-6
