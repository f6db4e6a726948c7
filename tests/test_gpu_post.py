"""GPU parity of the per-frame post passes against the CPU oracle: temporal
reprojection with flow / trust guides (resolve_reprojection_kernel,
src/realtime/gpu/programs.cu:1955-2067) and the edge-aware denoise filter."""
import sys

import numpy as np
import pytest

from helpers import oracle_camera, oracle_profile, oracle_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def moved_rig(rig, dx):
    cam = rig.cameras[0]
    cam.origin[0] += dx
    return rig


def sequence(rt, oracle, device, denoise, frames=3):
    g = rt.Generated("random_spheres", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    prof = g.profile
    prof.samples_per_pixel = 2
    prof.max_bounces = 4
    prof.rr_start_bounce = 2
    prof.enable_denoise = 1 if denoise else 0
    ren = rt.Renderer(device, sc)
    osc = oracle_scene(oracle, g.desc)
    ob = oracle.OracleBvh(osc)
    seq = oracle.RestirSequence(osc, ob, oracle_profile(oracle, prof))
    out = []
    for f in range(frames):
        rig = moved_rig(g.rig(), 0.02 * f)
        ren.render(rig, prof, f)
        ref = seq.render([oracle_camera(oracle, rig.cameras[0])], f)[0]
        gpu = ren.download(0)
        gpu.update(ren.guides(0))
        out.append((gpu, ref))
    return out


def test_resolve_sequence_matches_oracle(rt, oracle, device):
    say("[gpu] resolve: random_spheres 3 frames, camera moving")
    for f, (gpu, ref) in enumerate(sequence(rt, oracle, device, False)):
        for key in ("beauty", "normal", "depth"):
            same = np.mean(gpu[key] == ref[key])
            say(f"  frame{f} {key}: bit-exact {same:.6f}")
            assert same > 0.999, f"frame {f} {key}: {same}"
        assert np.mean(gpu["trust"] == ref["trust"]) > 0.999
        assert np.allclose(gpu["flow"], ref["flow"], atol=1e-3)
        if f > 0:
            assert gpu["trust"].mean() > 0.5  # most texels reproject


def test_denoise_filter_matches_oracle(rt, oracle, device):
    say("[gpu] denoise filter: random_spheres 2 frames")
    for f, (gpu, ref) in enumerate(sequence(rt, oracle, device, True, frames=2)):
        want = oracle.denoise(ref, 4)
        a = gpu["beauty"][..., :3].astype(np.float64)
        b = want[..., :3].astype(np.float64)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
        close = np.mean(rel < 1e-5)
        say(f"  frame{f}: within 1e-5 {close:.6f}, bit-exact {np.mean(a == b):.6f}")
        assert close > 0.999
        # the resolved (pre-filter) beauty stays available as a guide
        assert np.mean(gpu["resolved"] == ref["beauty"]) > 0.999
