"""GPU parity of albedo textures (evaluate_texture / sample_image_texture,
src/realtime/gpu/programs.cu:183-302) with the oracle: the reference's
simple_light scene file (turbulence-noise spheres) and the Cornell box
re-bound to a checker of an image, a repeat / bilinear image and a mirror /
nearest image, with an image-textured sphere and two triangles added (one
with per-vertex texture coordinates, one without).  Noise,
checker, quad and triangle coordinates use only +, -, *, / and floor, so
they are bit-exact; the sphere's coordinates go through acosf / atan2f,
whose device and libm results may differ in the last place, so a few
sphere pixels may pick a neighbouring texel."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from helpers import oracle_camera, oracle_profile, oracle_scene, rel_mse

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def render_both(rt, oracle, device, desc, keep, rig, spp, depth):
    sc = rt.Scene(device, desc, keepalive=keep)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    prof = rt.Generated("cornell_mesh").profile
    prof.samples_per_pixel, prof.max_bounces, prof.rr_start_bounce = spp, depth, depth
    ren.render(rig, prof, 3)
    gpu = ren.download(0)
    osc = oracle_scene(oracle, desc)
    ref = oracle.render_camera(osc, oracle.OracleBvh(osc), oracle_camera(oracle, rig.cameras[0]), 0,
                               oracle_profile(oracle, prof), 3, threads=8)
    return gpu, ref


def test_noise_textured_scene_matches_oracle(rt, oracle, device):
    from ray_tracing_amd import scene_yaml as S
    say("[gpu] simple_light.yaml (noise albedo) 2 spp depth 8")
    d = S.load_scene_definition(os.path.join(SCENES, "simple_light", "scene.yaml"))
    flat = S.flatten(d)
    gpu, ref = render_both(rt, oracle, device, flat.desc, flat, S.camera_rig(d, "default"), 2, 8)
    a, b = gpu["beauty"][..., :3], ref["beauty"][..., :3]
    same = np.mean(a == b)
    alb = np.mean(gpu["albedo"] == ref["albedo"])
    say(f"  beauty bit-exact {same:.6f}, albedo bit-exact {alb:.6f}, relMSE {rel_mse(a, b):.2e}")
    assert b.mean() > 0.01
    # the noise varies across the ground: many distinct albedo values
    assert len(np.unique(ref["albedo"][..., 0])) > 1000
    assert same > 0.999 and alb > 0.999


def textured_cornell(rt):
    from PIL import Image  # noqa: F401  (images are built in memory)
    from ray_tracing_amd import abi
    from ray_tracing_amd import scene_yaml as S
    d = S.load_scene_definition(os.path.join(SCENES, "cornell_box", "scene.yaml"))
    flat = S.flatten(d)
    rng = np.random.default_rng(4)
    img_a = rng.uniform(0.05, 0.95, (16, 24, 3)).astype(np.float32)
    img_b = rng.uniform(0.05, 0.95, (7, 5, 3)).astype(np.float32)
    texels = np.ascontiguousarray(np.concatenate([img_a.reshape(-1, 3), img_b.reshape(-1, 3)]))
    T = []

    def tex(**kw):
        t = abi.Texture()
        t.even_texture = t.odd_texture = -1
        t.u_address_mode = t.v_address_mode = 1
        t.scale = 1.0
        for k, v in kw.items():
            setattr(t, k, v)
        T.append(t)
        return len(T) - 1

    white = tex(type=0, color=abi.f3(0.73, 0.73, 0.73))
    a_near = tex(type=2, image_offset=0, image_width=24, image_height=16, filter_type=0)
    board = tex(type=1, scale=60.0, even_texture=white, odd_texture=a_near)
    a_lin = tex(type=2, image_offset=0, image_width=24, image_height=16, filter_type=1,
                u_address_mode=2, v_address_mode=2)
    b_mirror = tex(type=2, image_offset=24 * 16, image_width=5, image_height=7, filter_type=0,
                   u_address_mode=3, v_address_mode=3)
    mats = [flat.materials[i] for i in range(flat.desc.material_count)]
    bind = []
    for m in mats:
        r, g, b = tuple(m.albedo)
        if m.type != S.MAT_LAMBERTIAN:
            bind.append(-1)
        elif abs(r - g) < 1e-6:
            bind.append(board)       # white walls and boxes
        elif r > g:
            bind.append(a_lin)       # red wall
        else:
            bind.append(b_mirror)    # green wall
    extra = abi.Material()
    extra.type, extra.ior, extra.openpbr_index = S.MAT_LAMBERTIAN, 1.0, -1
    mats += [extra, extra]
    bind += [a_near, b_mirror]
    k = len(mats)
    spheres = (abi.Sphere * 1)(abi.Sphere(abi.f3(400.0, 90.0, 120.0), 80.0, k - 2))
    tris = (abi.Triangle * 2)(abi.Triangle(abi.f3(60.0, 1.0, 60.0), abi.f3(260.0, 1.0, 60.0),
                                           abi.f3(120.0, 260.0, 200.0), k - 1),
                              abi.Triangle(abi.f3(300.0, 300.0, 500.0), abi.f3(500.0, 300.0, 500.0),
                                           abi.f3(400.0, 500.0, 500.0), k - 1))
    # the first triangle has per-vertex coordinates (repeating the image 3 x 2
    # times), the second none (NaN: barycentrics)
    tri_uv = np.array([[0.0, 0.0, 3.0, 0.0, 1.5, 2.0], [np.nan] * 6], np.float32)
    # vertex normals on the second triangle (tilted, differing per vertex)
    tri_n = np.array([[np.nan] * 9, [0.3, 0.2, -1.0, -0.2, 0.1, -1.0, 0.0, -0.4, -1.0]], np.float32)
    mat_arr = (abi.Material * k)(*mats)
    tex_arr = (abi.Texture * len(T))(*T)
    bind_arr = np.asarray(bind, np.int32)
    desc = abi.SceneDesc()
    C.memmove(C.byref(desc), C.byref(flat.desc), C.sizeof(desc))
    desc.sphere_count, desc.triangle_count, desc.material_count = 1, 2, k
    desc.triangle_uvs = tri_uv.ctypes.data_as(C.c_void_p)
    desc.triangle_normals = tri_n.ctypes.data_as(C.c_void_p)
    desc.spheres = C.cast(spheres, C.c_void_p)
    desc.triangles = C.cast(tris, C.c_void_p)
    desc.materials = C.cast(mat_arr, C.c_void_p)
    desc.texture_count, desc.image_texel_count = len(T), texels.shape[0]
    desc.textures = C.cast(tex_arr, C.c_void_p)
    desc.image_texels = texels.ctypes.data_as(C.c_void_p)
    desc.material_albedo_texture = bind_arr.ctypes.data_as(C.c_void_p)
    keep = (flat, spheres, tris, tri_uv, tri_n, mat_arr, tex_arr, bind_arr, texels)
    return desc, keep, S.camera_rig(d, "default")


def test_checker_and_image_textures_match_oracle(rt, oracle, device):
    say("[gpu] textured Cornell box (checker of image, repeat/bilinear, mirror/nearest, "
        "sphere + triangle) 2 spp depth 1 and depth 6")
    desc, keep, rig = textured_cornell(rt)
    for depth in (1, 6):
        gpu, ref = render_both(rt, oracle, device, desc, keep, rig, 2, depth)
        a, b = gpu["beauty"][..., :3], ref["beauty"][..., :3]
        same = np.mean(a == b)
        alb = np.mean(gpu["albedo"] == ref["albedo"])
        e = rel_mse(a, b)
        say(f"  depth {depth}: beauty bit-exact {same:.6f}, albedo bit-exact {alb:.6f}, relMSE {e:.2e}")
        assert b.mean() > 0.01
        nrm = np.mean(gpu["normal"] == ref["normal"])
        say(f"  depth {depth}: normal bit-exact {nrm:.6f}")
        assert same > 0.99 and alb > 0.995 and e < 1e-3 and nrm > 0.995
    # the textures show: the primary albedo takes many values
    assert len(np.unique(ref["albedo"][..., 0])) > 200
