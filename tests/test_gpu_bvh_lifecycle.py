"""GPU acceleration lifecycle (rt_scene_update, after GpuSceneAcceleration::update,
src/realtime/gpu/gpu_scene_acceleration.cpp:247-367) against the oracle:

- the classification sequence of the reference: an unchanged description is
  reused, a material change is an update (BVH kept), moved primitives are a
  refit, a changed primitive count is a rebuild; the generation counts every
  update, refit and rebuild;
- a refit's binary and 4-wide node arrays are byte-identical to the oracle's
  refit of the same topology, and its closest hits equal brute force over
  the moved primitives;
- a rebuild equals the oracle's fresh build;
- a frame rendered after a refit equals the oracle's render of the moved scene;
- timings of each kind beside the reference's CPU lifecycle numbers
  (.planning/v2.0-BASELINE.json:430-441: 1,024 spheres, 0.29 / 0.16 / 0.18 /
  0.16 ms rebuild / update / refit / reuse) at 1,024 spheres and 1,024,000
  triangles.
"""
import sys

import numpy as np
import pytest

from helpers import (lambert_materials, oracle_camera, oracle_profile, oracle_scene, product_desc,
                     random_rays, random_triangles)
from test_gpu_parity import compare_bvh

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def soup(seed, n):
    rng = np.random.default_rng(seed)
    return {"triangles": random_triangles(rng, n), "materials": lambert_materials(2)}


def test_classification_and_refit_parity(rt, oracle, device):
    say("[gpu] acceleration lifecycle: reuse / update / refit / rebuild")
    arrays = soup(41, 4000)
    desc, keep = product_desc(rt, arrays)
    sc = rt.Scene(device, desc, keep)
    sc.build_bvh()
    u = sc.last_update()
    assert u["kind"] == "rebuild" and u["generation"] == 1
    assert sc.update(desc, keep)["kind"] == "reuse"
    assert sc.last_update()["generation"] == 1
    nodes0 = sc.download_bvh()[0].tobytes()

    mats = arrays["materials"].copy()
    mats["albedo"][0] = (0.9, 0.1, 0.1)
    d2, k2 = product_desc(rt, dict(arrays, materials=mats))
    u = sc.update(d2, k2)
    assert u["kind"] == "update" and u["generation"] == 2 and u["bvh_ms"] == 0.0
    assert sc.download_bvh()[0].tobytes() == nodes0

    rng = np.random.default_rng(3)
    moved = arrays["triangles"].copy()
    for k in ("p0", "p1", "p2"):
        moved[k] += rng.normal(0, 0.4, moved[k].shape).astype(np.float32)
    d3, k3 = product_desc(rt, dict(arrays, triangles=moved, materials=mats))
    u = sc.update(d3, k3)
    say(f"  refit: {u}")
    assert u["kind"] == "refit" and u["generation"] == 3
    # the oracle refits its build of the original soup to the moved one
    ob = oracle.OracleBvh(oracle_scene(oracle, desc))
    osc3 = oracle_scene(oracle, d3)
    ob.refit(osc3)
    nodes, sp, mc = sc.download_bvh()
    assert np.array_equal(sp, ob.sorted_prims()) and np.array_equal(mc, ob.morton())
    assert nodes.tobytes() == ob.nodes().tobytes(), "refit binary nodes differ"
    assert sc.download_wide().tobytes() == ob.wide().tobytes(), "refit 4-wide nodes differ"
    rays = random_rays(rng, 20000)
    gh = sc.trace_closest(rays)
    assert gh.tobytes() == ob.trace_closest(rays).tobytes()
    assert np.array_equal(gh["prim"][:3000], oracle.brute_closest(osc3, rays[:3000])["prim"])

    fewer = dict(arrays, triangles=moved[:-7], materials=mats)
    d4, k4 = product_desc(rt, fewer)
    u = sc.update(d4, k4)
    assert u["kind"] == "rebuild" and u["generation"] == 4 and u["primitive_reference_count"] == 3993
    nodes, sp, mc = sc.download_bvh()
    ob4 = oracle.OracleBvh(oracle_scene(oracle, d4))
    assert nodes.tobytes() == ob4.nodes().tobytes()
    assert sc.download_wide().tobytes() == ob4.wide().tobytes()


def test_render_after_refit_matches_oracle(rt, oracle, device):
    say("[gpu] random_spheres: spheres moved, refit, render vs oracle")
    g = rt.Generated("random_spheres", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    prof = g.profile
    prof.samples_per_pixel, prof.max_bounces, prof.rr_start_bounce = 2, 5, 5
    rig = g.rig()
    ren.render(rig, prof, 3)
    arrays = g.arrays()
    from helpers import SPHERE_DT
    sph = np.frombuffer(arrays["spheres"].tobytes(), SPHERE_DT).copy()
    rng = np.random.default_rng(8)
    small = sph["radius"] < 0.5
    sph["center"][small] += rng.uniform(-0.15, 0.15, (int(small.sum()), 3)).astype(np.float32)
    sph["center"][small, 1] = sph["radius"][small]       # stay on the ground
    desc, keep = product_desc(rt, {"spheres": sph, "materials": arrays["materials"]},
                              background=tuple(g.desc.background),
                              background_top=tuple(g.desc.background_top))
    u = sc.update(desc, keep)
    say(f"  {u}")
    assert u["kind"] == "refit"
    ren.reset_accumulation()
    ren.render(rig, prof, 3)
    gpu = ren.download(0)
    ref = oracle.render_camera(oracle_scene(oracle, desc), oracle.OracleBvh(oracle_scene(oracle, desc)),
                               oracle_camera(oracle, rig.cameras[0]), 0,
                               oracle_profile(oracle, prof), 3, threads=16)
    for key in ("beauty", "normal", "albedo", "depth"):
        same = float(np.mean(gpu[key] == ref[key]))
        say(f"  {key}: bit-exact fraction {same:.6f}")
        assert same > 0.999


@pytest.mark.parametrize("config", ["random_spheres", "instanced_mesh_restir"])
def test_lifecycle_timings(rt, device, config):
    """Timings of each kind, reported (no bound: a report beside the
    reference's CPU lifecycle numbers)."""
    g = rt.Generated(config, 0)
    arrays = g.arrays()
    from helpers import SPHERE_DT, TRI_DT
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    res = {"rebuild": sc.last_update()}
    res["reuse"] = sc.update(g.desc)
    base = dict(arrays)          # raw records of every array, as generated
    if config == "random_spheres":
        sph = np.frombuffer(arrays["spheres"].tobytes(), SPHERE_DT).copy()
        sph["center"][:, 0] += 0.01
        base["spheres"] = sph
    else:
        tri = np.frombuffer(arrays["triangles"].tobytes(), TRI_DT).copy()
        for k in ("p0", "p1", "p2"):
            tri[k][:, 1] += 0.01
        base["triangles"] = tri
    top = tuple(g.desc.background_top) if g.desc.background_mode == 1 else None
    d1, k1 = product_desc(rt, base, background=tuple(g.desc.background), background_top=top)
    res["refit"] = sc.update(d1, k1)
    raw = bytearray(arrays["materials"].tobytes())
    raw[4:8] = np.float32(0.123).tobytes()      # material 0's albedo.x
    base["materials"] = np.frombuffer(bytes(raw), arrays["materials"].dtype).copy()
    d2, k2 = product_desc(rt, base, background=tuple(g.desc.background), background_top=top)
    res["update"] = sc.update(d2, k2)
    for k, v in res.items():
        say(f"  {config} {k}: host {v['elapsed_ms']:.3f} ms, BVH kernels {v['bvh_ms']:.3f} ms, "
            f"{v['primitive_reference_count']} primitives, generation {v['generation']}")
    assert res["refit"]["kind"] == "refit" and res["update"]["kind"] == "update"
    assert res["reuse"]["kind"] == "reuse"
    assert res["refit"]["bvh_ms"] < res["rebuild"]["bvh_ms"] * 1.5 + 1.0
