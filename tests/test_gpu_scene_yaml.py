"""GPU parity of scenes ingested from the reference's YAML scene files: the
Cornell box (`assets/scenes/cornell_box`, 18 quads after flattening) from
its realtime default view, and the imported OBJ triangle, rendered through
the C ABI and compared with the oracle's render of the same flattened scene
and camera."""
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


def render_yaml(rt, oracle, device, name, preset, spp, depth, count=False):
    from ray_tracing_amd import scene_yaml as S
    d = S.load_scene_definition(os.path.join(SCENES, name, "scene.yaml"))
    flat = S.flatten(d)
    rig = S.camera_rig(d, preset)
    sc = rt.Scene(device, flat.desc, keepalive=flat)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    if count:
        ren.set_tuning(count_traversal=1, use_graph=0)
    prof = rt.Generated("cornell_mesh").profile
    prof.samples_per_pixel, prof.max_bounces, prof.rr_start_bounce = spp, depth, depth
    tm = ren.render(rig, prof, 3)
    gpu = ren.download(0)
    gpu["timing"] = tm
    osc = oracle_scene(oracle, flat.desc)
    ob = oracle.OracleBvh(osc)
    ref = oracle.render_camera(osc, ob, oracle_camera(oracle, rig.cameras[0]), 0,
                               oracle_profile(oracle, prof), 3, threads=8)
    return gpu, ref


def test_cornell_box_yaml_single_bounce_matches_oracle(rt, oracle, device):
    say("[gpu] cornell_box.yaml realtime view 640x480, 2 spp, depth 1")
    gpu, ref = render_yaml(rt, oracle, device, "cornell_box", "realtime", 2, 1)
    a = gpu["beauty"][..., :3].astype(np.float64)
    b = ref["beauty"][..., :3].astype(np.float64)
    assert a.shape == (480, 640, 3) and np.isfinite(a).all()
    assert b.mean() > 0.01, "the view sees the lit room"
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    say(f"  max rel err {err.max():.2e}")
    assert err.max() < 1e-4
    for key in ("normal", "albedo", "depth"):
        same = np.mean(gpu[key] == ref[key])
        say(f"  {key}: bit-exact fraction {same:.6f}")
        assert same > 0.999


def test_cornell_box_yaml_full_paths_match_oracle(rt, oracle, device):
    say("[gpu] cornell_box.yaml cpu preset 1280x720, 2 spp, depth 8")
    gpu, ref = render_yaml(rt, oracle, device, "cornell_box", "default", 2, 8)
    a, b = gpu["beauty"][..., :3], ref["beauty"][..., :3]
    same = np.mean(a == b)
    e = rel_mse(a, b)
    say(f"  beauty bit-exact fraction {same:.6f}, relMSE {e:.2e}")
    assert same > 0.99 and e < 1e-4


def test_traversal_counts_equal_the_oracles(rt, oracle, device):
    """The counting frame's child-box and primitive-test totals (the inputs of
    bench.py's algorithmic bytes) equal the oracle's for the same rays: the
    render is bit-exact, so every path-extension ray is the same and the
    traversal order is shared."""
    say("[gpu] cornell_box.yaml 1280x720, 2 spp, depth 8, counting traversal")
    gpu, ref = render_yaml(rt, oracle, device, "cornell_box", "default", 2, 8, count=True)
    tm = gpu["timing"]
    say(f"  GPU visits {tm.node_visits} tests {tm.prim_tests}; oracle visits {ref['node_visits']} "
        f"tests {ref['prim_tests']}; extend rays {tm.extend_rays} vs {ref['rays_closest']}")
    assert np.mean(gpu["beauty"] == ref["beauty"]) == 1.0
    assert tm.extend_rays == ref["rays_closest"]
    assert tm.node_visits == ref["node_visits"] and tm.prim_tests == ref["prim_tests"]
    assert tm.node_visits > 10 * tm.extend_rays   # a real traversal, not a stub

