"""GPU parity: LBVH build, closest/any-hit traversal and full renders against the CPU oracle.

Every call goes through the C ABI of librt_hip.so.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import (ALIGHT_DT, MAT_DT, QUAD_DT, SPHERE_DT, TRI_DT, lambert_materials,
                     oracle_camera, oracle_profile, oracle_scene, product_desc, random_rays,
                     random_triangles, rel_mse)

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def tiny_scenes():
    rng = np.random.default_rng(7)
    out = {}
    s = np.zeros(1, SPHERE_DT)
    s["center"] = (0, 0, 0)
    s["radius"] = 1.0
    out["one_sphere"] = {"spheres": s, "materials": lambert_materials(1)}
    t = random_triangles(rng, 2)
    out["two_tris"] = {"triangles": t, "materials": lambert_materials(1)}
    # duplicated centroids: identical Morton codes everywhere
    d = np.zeros(300, SPHERE_DT)
    d["center"] = (1.5, -2.0, 3.0)
    d["radius"] = rng.uniform(0.1, 1.0, 300).astype(np.float32)
    out["duplicates"] = {"spheres": d, "materials": lambert_materials(1)}
    out["soup_5000"] = {"triangles": random_triangles(rng, 5000), "materials": lambert_materials(1)}
    q = np.zeros(3, QUAD_DT)
    q["origin"] = [(-5, -1, -5), (-5, -1, -5), (-5, -1, 5)]
    q["edge_u"] = [(10, 0, 0), (0, 10, 0), (10, 0, 0)]
    q["edge_v"] = [(0, 0, 10), (0, 0, 10), (0, 10, 0)]
    mixed_s = np.zeros(20, SPHERE_DT)
    mixed_s["center"] = rng.uniform(-4, 4, (20, 3)).astype(np.float32)
    mixed_s["radius"] = 0.5
    out["mixed"] = {"spheres": mixed_s, "quads": q, "triangles": random_triangles(rng, 200, 4.0, 0.5),
                    "materials": lambert_materials(1)}
    return out


def gpu_scene(rt, device, desc, keep=None):
    sc = rt.Scene(device, desc, keep)
    sc.build_bvh()
    return sc


def compare_bvh(oracle, desc, sc):
    osc = oracle_scene(oracle, desc)
    ob = oracle.OracleBvh(osc)
    assert ob.validate() == 0
    nodes, sp, mc = sc.download_bvh()
    on = ob.nodes()
    assert sc.stats.node_count == ob.bvh.node_count
    assert sc.stats.max_depth == ob.bvh.max_depth
    assert np.array_equal(mc, ob.morton()), "Morton codes differ"
    assert np.array_equal(sp, ob.sorted_prims()), "sorted primitive order differs"
    assert nodes.tobytes() == on.tobytes(), "node arrays differ"
    assert sc.stats.wide_node_count == ob.bvh.wide_count
    assert sc.stats.wide_levels == ob.bvh.wide_levels
    assert sc.download_wide().tobytes() == ob.wide().tobytes(), "4-wide node arrays differ"
    return osc, ob


@pytest.mark.parametrize("config", ["random_spheres", "cornell_mesh"])
def test_bvh_bitexact_configs(rt, oracle, device, config):
    say(f"[gpu] bvh {config}")
    g = rt.Generated(config, 0)
    sc = gpu_scene(rt, device, g.desc)
    compare_bvh(oracle, g.desc, sc)


@pytest.mark.parametrize("name", list(tiny_scenes().keys()))
def test_bvh_bitexact_edge_cases(rt, oracle, device, name):
    say(f"[gpu] bvh edge case {name}")
    arrays = tiny_scenes()[name]
    desc, keep = product_desc(rt, arrays)
    sc = gpu_scene(rt, device, desc, keep)
    compare_bvh(oracle, desc, sc)


def test_bvh_empty_scene(rt, device):
    say("[gpu] bvh empty")
    desc, keep = product_desc(rt, {"materials": lambert_materials(1)})
    sc = gpu_scene(rt, device, desc, keep)
    assert sc.stats.node_count == 0
    rays = np.zeros(4, rt.RAY_DTYPE)
    rays["d"] = (0, 0, 1)
    rays["tmax"] = 1e30
    hits = sc.trace_closest(rays)
    assert (hits["prim"] == -1).all()


def primary_rays_all(oracle, g, jitter=0):
    out = []
    for ci in range(g.g.camera_count):
        cam = oracle_camera(oracle, g.g.cameras[ci])
        r = oracle.primary_rays(cam, ci, 0, 0, jitter)
        r["tmin"] = 1e-3
        out.append(r)
    return np.concatenate(out)


@pytest.mark.parametrize("config", ["random_spheres", "cornell_mesh"])
def test_closest_hit_bitexact_primary(rt, oracle, device, config):
    say(f"[gpu] closest hit primary rays {config}")
    g = rt.Generated(config, 0)
    sc = gpu_scene(rt, device, g.desc)
    osc, ob = compare_bvh(oracle, g.desc, sc)
    rays = primary_rays_all(oracle, g, jitter=1)
    gh = sc.trace_closest(rays)
    oh = ob.trace_closest(rays)
    assert np.array_equal(gh["prim"], oh["prim"]), "closest-hit primitive ids differ"
    assert gh.tobytes() == oh.tobytes(), "closest-hit t/u/v differ"
    # and the oracle's traversal equals brute force on a subset
    sub = rays[:: max(1, len(rays) // 3000)]
    bh = oracle.brute_closest(osc, sub)
    assert np.array_equal(ob.trace_closest(sub)["prim"], bh["prim"])


@pytest.mark.parametrize("name", ["soup_5000", "mixed", "duplicates"])
def test_closest_and_any_hit_random_rays(rt, oracle, device, name):
    say(f"[gpu] closest/any random rays {name}")
    arrays = tiny_scenes()[name]
    desc, keep = product_desc(rt, arrays)
    sc = gpu_scene(rt, device, desc, keep)
    osc, ob = compare_bvh(oracle, desc, sc)
    rng = np.random.default_rng(11)
    rays = random_rays(rng, 20000)
    gh = sc.trace_closest(rays)
    oh = ob.trace_closest(rays)
    assert gh.tobytes() == oh.tobytes()
    bh = oracle.brute_closest(osc, rays[:2000])
    assert np.array_equal(oh["prim"][:2000], bh["prim"])
    short = rays.copy()
    short["tmax"] = rng.uniform(0.5, 30.0, len(rays)).astype(np.float32)
    assert np.array_equal(sc.trace_any(short), ob.trace_any(short))


def render_both(rt, oracle, device, config, spp, depth, cams=None, stream=0, threads=8, rr=None):
    g = rt.Generated(config, 0)
    sc = gpu_scene(rt, device, g.desc)
    ren = rt.Renderer(device, sc)
    prof = g.profile
    prof.samples_per_pixel = spp
    prof.max_bounces = depth
    prof.rr_start_bounce = depth if rr is None else rr
    rig = g.rig(cams)
    tm = ren.render(rig, prof, stream)
    osc = oracle_scene(oracle, g.desc)
    ob = oracle.OracleBvh(osc)
    out = []
    for ci in range(rig.active_count):
        gpu = ren.download(ci)
        ref = oracle.render_camera(osc, ob, oracle_camera(oracle, rig.cameras[ci]), ci,
                                   oracle_profile(oracle, prof), stream, threads=threads)
        out.append((gpu, ref))
    return out, tm


def test_render_random_spheres_matches_oracle(rt, oracle, device):
    say("[gpu] render random_spheres 4 spp depth 10")
    pairs, tm = render_both(rt, oracle, device, "random_spheres", 4, 10)
    gpu, ref = pairs[0]
    for key in ("beauty", "normal", "albedo", "depth"):
        a, b = gpu[key], ref[key]
        assert np.isfinite(a).all()
        same = np.mean(a == b)
        say(f"  {key}: bit-exact fraction {same:.6f}")
        assert same > 0.999, f"{key}: only {same:.6f} of values bit-exact"
    assert rel_mse(gpu["beauty"][..., :3], ref["beauty"][..., :3]) < 1e-6
    assert tm.paths == 400 * 225 * 4


def test_render_single_bounce_cornell(rt, oracle, device):
    """Deterministic single-bounce pass (NEE only): per-pixel radiance within 1e-4 relative."""
    say("[gpu] render cornell_mesh 2 spp depth 1")
    pairs, tm = render_both(rt, oracle, device, "cornell_mesh", 2, 1)
    for gpu, ref in pairs:
        a = gpu["beauty"][..., :3].astype(np.float64)
        b = ref["beauty"][..., :3].astype(np.float64)
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
        assert err.max() < 1e-4, f"max rel err {err.max()}"


def test_bench_traversal_counts_match_oracle(rt, oracle, device):
    """The bench workload's counted child boxes and primitive tests (the
    algorithmic bytes of its roofline) against the oracle's counts for the
    same frame (camera 0, 1 spp, depth 8): the OpenPBR sphere keeps the
    render within 1e-4 of the oracle rather than bit-exact, so a few paths
    may branch differently; the totals agree to 0.1 %."""
    say("[gpu] cornell_mesh counts vs oracle, camera 0, 1 spp, depth 8")
    g = rt.Generated("cornell_mesh", 0)
    sc = gpu_scene(rt, device, g.desc)
    ren = rt.Renderer(device, sc)
    ren.set_tuning(count_traversal=1, use_graph=0)
    prof = g.profile
    prof.samples_per_pixel = 1
    rig = g.rig(1)
    tm = ren.render(rig, prof, 0)
    osc = oracle_scene(oracle, g.desc)
    ob = oracle.OracleBvh(osc)
    ref = oracle.render_camera(osc, ob, oracle_camera(oracle, rig.cameras[0]), 0,
                               oracle_profile(oracle, prof), 0, threads=8)
    say(f"  GPU visits {tm.node_visits} tests {tm.prim_tests} rays {tm.extend_rays}; "
        f"oracle {ref['node_visits']} {ref['prim_tests']} {ref['rays_closest']}")
    for a, b in ((tm.extend_rays, ref["rays_closest"]), (tm.node_visits, ref["node_visits"]),
                 (tm.prim_tests, ref["prim_tests"])):
        assert abs(a - b) <= 1e-3 * b


def test_handles_destroyed_out_of_order(rt):
    """A scene keeps its device alive and a renderer its scene: destroying the
    device first, then the scene, then the renderer frees everything once."""
    say("[gpu] handle lifetimes")
    dev = rt.Device(0)
    desc, keep = product_desc(rt, tiny_scenes()["mixed"])
    sc = gpu_scene(rt, dev, desc, keep)
    ren = rt.Renderer(dev, sc)
    dev.close()
    sc.close()
    prof = rt.Generated("random_spheres", 0).profile
    prof.samples_per_pixel = 1
    prof.max_bounces = 2
    rig = rt.Generated("random_spheres", 0).rig()
    ren.render(rig, prof, 0)
    out = ren.download(0)
    assert np.isfinite(out["beauty"]).all()
    ren.close()
