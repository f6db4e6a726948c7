"""Pins the oracle's per-frame post passes: temporal reprojection
(resolve_reprojection_kernel, src/realtime/gpu/programs.cu:1955-2067) and
the edge-aware denoise filter, by closed forms and invariants."""
import ctypes as C

import numpy as np

import oracle_ctypes as O


def pinhole(w=64, h=48, fx=60.0, origin=(0.0, 0.0, 0.0)):
    c = O.Camera()
    c.enabled, c.width, c.height, c.model = 1, w, h, 0
    c.origin[:] = origin
    c.basis_x[:] = (1.0, 0.0, 0.0)
    c.basis_y[:] = (0.0, 1.0, 0.0)
    c.basis_z[:] = (0.0, 0.0, 1.0)
    c.fx, c.fy, c.cx, c.cy = fx, fx, w / 2.0, h / 2.0
    return c


def plane_frame(cam, z_plane, rng):
    """A fronto-parallel plane at z = z_plane seen by `cam`: per-pixel depth is
    the distance along the pixel-centre ray, normals face the camera."""
    w, h = cam.width, cam.height
    depth = np.zeros((h, w), np.float32)
    for y in range(h):
        for x in range(w):
            o = np.zeros(3, np.float32)
            d = np.zeros(3, np.float32)
            O.lib().rto_camera_ray(C.byref(cam), x + 0.5, y + 0.5, o.ctypes.data, d.ctypes.data)
            depth[y, x] = (z_plane - o[2]) / d[2]
    normal = np.zeros((h, w, 4), np.float32)
    normal[..., :3] = (0.5, 0.5, 0.0)  # encoded (0, 0, -1)
    normal[..., 3] = 1.0
    beauty = rng.uniform(0.0, 2.0, (h, w, 4)).astype(np.float32)
    beauty[..., 3] = 1.0
    return {"beauty": beauty, "normal": normal, "depth": depth}


def test_resolve_static_camera_blends_exactly():
    rng = np.random.default_rng(1)
    cam = pinhole()
    cur = plane_frame(cam, 5.0, rng)
    cur["beauty"][0, 0, :3] = (20.0, -1.0, 3.0)  # clamped to [0, 10] before blending
    hist = {k: v.copy() for k, v in plane_frame(cam, 5.0, rng).items()}
    hist["beauty"] = rng.uniform(0.0, 1.0, cur["beauty"].shape).astype(np.float32)
    n = 3
    expect = cur["beauty"].copy()
    blend = np.float32(1.0) / np.float32(n + 1)
    cl = np.clip(cur["beauty"][..., :3], np.float32(0), np.float32(10))
    expect[..., :3] = hist["beauty"][..., :3] + (cl - hist["beauty"][..., :3]) * blend
    flow, trust = O.resolve(cam, cam, n, False, cur, hist)
    assert trust.min() == 1.0
    assert np.array_equal(cur["beauty"][..., :3], expect[..., :3])
    assert np.abs(flow).max() < 1e-3


def test_resolve_translation_flow_and_rejection():
    rng = np.random.default_rng(2)
    prev = pinhole()
    cam = pinhole(origin=(0.1, 0.0, 0.0))  # moved 0.1 along +x, plane at z = 5
    cur = plane_frame(cam, 5.0, rng)
    hist = plane_frame(prev, 5.0, rng)
    flow, trust = O.resolve(cam, prev, 1, False, cur, hist)
    # a point at depth 5 shifts by fx * 0.1 / 5 = 1.2 px between the cameras
    inner = trust[:, 4:-4] == 1.0
    assert inner.all()
    assert np.allclose(flow[:, 4:-4, 0], -1.2, atol=2e-3)
    assert np.allclose(flow[:, 4:-4, 1], 0.0, atol=2e-3)
    # texels that left the previous image are rejected
    assert trust[:, -1].max() == 0.0 or trust[:, 0].max() == 0.0
    # a history surface at another depth is rejected everywhere
    far = plane_frame(prev, 9.0, rng)
    cur2 = plane_frame(cam, 5.0, rng)
    _, trust2 = O.resolve(cam, prev, 1, False, cur2, far)
    assert trust2.max() == 0.0


def test_resolve_without_history_is_identity():
    rng = np.random.default_rng(3)
    cam = pinhole()
    cur = plane_frame(cam, 5.0, rng)
    before = {k: v.copy() for k, v in cur.items()}
    flow, trust = O.resolve(cam, None, 0, False, cur, plane_frame(cam, 5.0, rng))
    for k in before:
        assert np.array_equal(before[k], cur[k])
    assert trust.max() == 0.0 and np.abs(flow).max() == 0.0


def frame_with(beauty, normal_xyz, depth, albedo=(0.5, 0.5, 0.5)):
    h, w = depth.shape
    n = np.zeros((h, w, 4), np.float32)
    n[..., :3] = np.asarray(normal_xyz, np.float32) * 0.5 + 0.5
    a = np.zeros((h, w, 4), np.float32)
    a[..., :3] = albedo
    return {"beauty": beauty.astype(np.float32), "normal": n, "albedo": a,
            "depth": depth.astype(np.float32)}


def test_denoise_keeps_constant_images():
    h, w = 32, 40
    b = np.zeros((h, w, 4), np.float32)
    b[..., :3] = (0.3, 0.6, 0.9)
    b[..., 3] = 1.0
    f = frame_with(b, (0, 0, 1), np.full((h, w), 4.0))
    out = O.denoise(f, 4)
    assert np.allclose(out[..., :3], b[..., :3], rtol=1e-6, atol=0)


def test_denoise_does_not_cross_normal_edges():
    h, w = 32, 40
    b = np.zeros((h, w, 4), np.float32)
    b[:, : w // 2, :3] = 1.0
    b[:, w // 2:, :3] = 0.1
    n = np.zeros((h, w, 3), np.float32)
    n[:, : w // 2] = (1, 0, 0)
    n[:, w // 2:] = (0, 1, 0)
    f = frame_with(b, (0, 0, 1), np.full((h, w), 4.0))
    f["normal"][..., :3] = n * 0.5 + 0.5
    out = O.denoise(f, 4)
    assert np.allclose(out[..., :3], b[..., :3], rtol=1e-6)


def test_denoise_reduces_noise_and_skips_background():
    rng = np.random.default_rng(4)
    h, w = 48, 64
    b = np.zeros((h, w, 4), np.float32)
    b[..., :3] = rng.exponential(0.5, (h, w, 1)).astype(np.float32)
    depth = np.full((h, w), 3.0, np.float32)
    depth[:4] = 0.0  # rows without a surface pass through
    f = frame_with(b, (0, 0, 1), depth)
    out = O.denoise(f, 4)
    assert np.array_equal(out[:4], b[:4])
    assert out[8:-8, 8:-8, 0].var() < 0.1 * b[8:-8, 8:-8, 0].var()
    assert abs(out[8:-8, 8:-8, 0].mean() - b[8:-8, 8:-8, 0].mean()) < 0.05
