"""GPU parity of the viewer tone map (rt_presenter_* / rt_tone_map_rgba8,
after src/realtime/viewer/cuda_gl_presenter.cu:23-130) with the oracle:

- the reference's presenter test (tests/test_cuda_gl_presenter.cpp:44-96):
  a 2x2 source, a refused null view, a refused resolution mismatch, then the
  per-channel conversion and opaque alpha;
- every float around each of the 255 rounding thresholds, a dense sweep,
  NaN / inf / signed zero, bit-exact with the oracle;
- a rendered four-camera frame presented from its device view;
- a 272M-pixel launch (266,000 workgroups of four-pixel threads, a 4.4 GB
  frame) checked through periodicity: the input repeats with a prime
  period, so must the output.
Device buffers come from the HIP runtime of librt_hip.so (tests/hipmem.py)."""
import sys

import numpy as np
import pytest

from hipmem import DeviceArray, sync
from helpers import sweep

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def view_of(rt, d, width, height):
    v = rt.abi.DeviceFrameView()
    v.width, v.height = width, height
    v.beauty = d.ptr if d is not None else None
    return v


def test_presenter_matches_reference_presenter_test(rt, oracle, device):
    say("[gpu] presenter 2x2 (test_cuda_gl_presenter.cpp)")
    src = np.array([[0.0, 0.25, 1.0, 0.0], [4.0, -1.0, 0.01, 0.0],
                    [0.04, 0.16, 0.36, 0.0], [0.64, 0.81, 0.99, 0.0]], np.float32)
    d = DeviceArray.of(src)
    p = rt.Presenter(device, 2, 2)
    try:
        assert p.present(view_of(rt, None, 2, 2)) == rt.PRESENT_INVALID_DEVICE_VIEW
        assert p.present(None) == rt.PRESENT_INVALID_DEVICE_VIEW
        assert p.present(view_of(rt, d, 3, 2)) == rt.PRESENT_RESOLUTION_MISMATCH
        assert p.present(view_of(rt, d, 2, 2)) == rt.PRESENT_OK
        px = p.download().reshape(4, 4)
        assert np.array_equal(px, oracle.tone_map_rgba8(src))
        assert (px[:, 3] == 255).all()
        assert p.device_pixels() != 0
    finally:
        p.close()
    with pytest.raises(rt.RtError, match="positive resolution"):
        rt.Presenter(device, 0, 2)


def test_tone_map_sweep_bit_exact(rt, oracle, device):
    v = sweep()
    special = np.array([np.nan, -np.nan, np.inf, -np.inf, -0.0, 0.0, -1e30, 1e-45], np.float32)
    v = np.concatenate([v, special])
    n = (v.size + 2) // 3
    src = np.zeros((n, 4), np.float32)
    src[:, :3].flat[: v.size] = v
    src[:, 3] = np.float32(np.nan)           # alpha input is ignored
    say(f"[gpu] tone map sweep, {n} pixels")
    d_src = DeviceArray.of(src)
    d_out = DeviceArray(n * 4)
    rt.tone_map_rgba8(device, d_src.ptr, d_out.ptr, n)
    device.synchronize()
    got = d_out.download((n, 4), np.uint8)
    want = oracle.tone_map_rgba8(src)
    bad = np.argwhere(got != want)
    say(f"  mismatches {len(bad)}")
    assert len(bad) == 0, (bad[:5], src[bad[:5, 0]])
    # argument checks of the entry point
    with pytest.raises(rt.RtError, match="negative pixel count"):
        rt.tone_map_rgba8(device, d_src.ptr, d_out.ptr, -1)
    with pytest.raises(rt.RtError, match="aligned"):
        rt.tone_map_rgba8(device, d_src.ptr + 4, d_out.ptr, 1)
    rt.tone_map_rgba8(device, 0, 0, 0)       # nothing to do


def test_present_rendered_frame(rt, oracle, device):
    say("[gpu] present the four smoke cameras from their device views")
    g = rt.Generated("smoke", 3)
    sc = rt.Scene(device, g.desc, keepalive=g)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    prof = g.profile
    ren.render(g.rig(), prof, 5)
    levels = []
    for cam in range(4):
        v = ren.device_view(cam)
        p = rt.Presenter(device, v.width, v.height)
        try:
            assert p.present(v) == rt.PRESENT_OK
            got = p.download()
        finally:
            p.close()
        beauty = ren.download(cam)["beauty"]
        want = oracle.tone_map_rgba8(beauty)
        assert got.shape == want.shape
        assert np.array_equal(got, want), cam
        levels.append(float(got[..., :3].mean()))
        say(f"  camera {cam}: {v.width}x{v.height}, mean display level {levels[-1]:.2f}")
    # the surround rig's side cameras may face the black background only
    assert max(levels) > 1.0


def test_tone_map_beyond_one_grid_periodic(rt, oracle, device):
    period = 1009
    n = (1 << 28) + 3 * (1 << 20) + 77          # not a multiple of the 1024-pixel tile
    rng = np.random.default_rng(9)
    base = rng.uniform(-0.5, 2.0, (period, 4)).astype(np.float32)
    say(f"[gpu] tone map {n} pixels, period {period}")
    d_src = DeviceArray(n * 16)
    d_src.upload(base)
    filled = period                              # doubling copies keep the period
    while filled < n:
        k = min(filled, n - filled)
        d_src.copy_within(filled * 16, 0, k * 16)
        filled += k
    d_out = DeviceArray(n * 4)
    sync()
    rt.tone_map_rgba8(device, d_src.ptr, d_out.ptr, n)
    device.synchronize()
    d_src.free()
    out = d_out.download((n, 4), np.uint8)
    d_out.free()
    want = oracle.tone_map_rgba8(base)
    k = n // period
    assert np.array_equal(out[: k * period].reshape(k, period, 4), np.broadcast_to(want, (k, period, 4)))
    assert np.array_equal(out[k * period:], want[: n - k * period])
