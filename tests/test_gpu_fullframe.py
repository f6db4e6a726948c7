"""GPU parity at the benchmark workloads' own shapes:

- the radiance cache's kernels at the full configuration (16 levels x 2^19
  hashed entries, resolutions 16..2048, room for 65,536 training records)
  against the oracle's prediction, loss and gradients, on queries that
  collide in the hashed levels;
- the four-camera Cornell mesh frame at depth 8 with Russian roulette on,
  against the oracle: the GPU-oracle difference of every camera must sit far
  inside the Monte Carlo band (the difference between two independent
  oracle renders of the same frame);
- a frame replayed from the captured hipGraph against the oracle itself;
- the four-camera radiance-cache scene on a 1/8-scale rig with the cache and
  the learned denoiser attached: the cache becomes usable and answers
  queries, the frame stays finite, and its relMSE against a high-spp plain
  render (itself checked against the oracle) beats the plain 1-spp frame.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from helpers import oracle_camera, oracle_profile, oracle_scene, rel_mse
from test_oracle_cache import MLP

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_cache_full_configuration_matches_oracle(rt, oracle, device):
    say("[gpu] radiance cache at 16 levels x 2^19, res 16..2048, 65536 records")
    rng = np.random.default_rng(51)
    c = rt.RadianceCache(0, levels=16, log2_table=19, base_resolution=16, max_resolution=2048,
                         train_samples=65536, train_steps=1, learning_rate=1e-2, seed=3)
    lo, hi = np.array([-50.0, -50.0, -50.0], np.float32), np.array([50.0, 50.0, 50.0], np.float32)
    c.set_bounds(lo, hi)
    info = c.grid_info()
    assert info["res"][0] == 16.0 and info["res"][15] == 2048.0
    hashed = [l for l in range(16) if not info["dense"][l]]
    assert len(hashed) >= 10, info["dense"]
    n = c.n
    assert n == MLP + 16 * 2 ** 19 * 2
    params = np.concatenate([rng.uniform(-0.25, 0.25, MLP),
                             rng.uniform(-0.5, 0.5, n - MLP)]).astype(np.float32)
    c.set_params(params)
    q = np.zeros((3000, 16), np.float32)
    # clustered positions: many queries share coarse cells, the finest levels hash
    centres = rng.uniform(lo, hi, (40, 3))
    q[:, 0:3] = (centres[rng.integers(0, 40, 3000)] + rng.normal(0, 1.5, (3000, 3))).clip(lo, hi)
    d = rng.normal(size=(3000, 3))
    q[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    q[:, 7] = rng.random(3000)
    nn = rng.normal(size=(3000, 3))
    q[:, 8:11] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q[:, 12:15] = rng.random((3000, 3))
    g = oracle.cache_grid(info)
    # the finest level maps different cells onto shared table entries
    cells, slots = set(), {}
    for p in q[:, 0:3]:
        pn = (p - lo) / (hi - lo)
        idx, w = oracle.cache_corners(g, 15, pn)
        cell = tuple(np.floor(pn * info["res"][15]).astype(int))
        for i in idx:
            slots.setdefault(int(i), set()).add(cell)
        cells.add(cell)
    collisions = sum(1 for v in slots.values() if len(v) > 1)
    say(f"  {len(cells)} finest cells, {collisions} shared table entries")
    assert collisions > 0
    out = c.infer(q)[:, :3].astype(np.float64)
    ref = oracle.cache_predict(g, params, q, bf16=True)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    say(f"  inference rel L2 {rel:.2e}")
    assert np.isfinite(out).all() and rel < 5e-3
    t = np.zeros((len(q), 4), np.float32)
    t[:, :3] = rng.uniform(0.0, 3.0, (len(q), 3))
    loss, grad = c.train_step(q, t, want_grad=True)
    ref_loss, rgrad = oracle.cache_loss_grad(g, params, q, t, bf16=True)
    gm, rm = grad[:MLP].astype(np.float64), rgrad[:MLP]
    gg, rg = grad[MLP:].astype(np.float64), rgrad[MLP:]
    em = np.linalg.norm(gm - rm) / np.linalg.norm(rm)
    eg = np.linalg.norm(gg - rg) / np.linalg.norm(rg)
    say(f"  loss {loss:.6f} oracle {ref_loss:.6f}; MLP grad {em:.2e}, grid grad {eg:.2e}")
    assert abs(loss - ref_loss) <= 2e-3 * ref_loss
    assert em < 1e-2 and eg < 1e-2
    touched_gpu, touched_ref = np.abs(gg) > 0, np.abs(rg) > 0
    assert np.mean(touched_gpu == touched_ref) > 0.999999


def render_cornell(rt, device, spp, depth, rr, stream, graph=False, frames=1):
    g = rt.Generated("cornell_mesh", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    ren.set_tuning(use_graph=1 if graph else 0)
    prof = g.profile
    prof.samples_per_pixel, prof.max_bounces, prof.rr_start_bounce = spp, depth, rr
    rig = g.rig()
    for _ in range(frames):
        ren.reset_accumulation()
        ren.render(rig, prof, stream)
    return g, rig, prof, [ren.download(c)["beauty"][..., :3] for c in range(rig.active_count)]


def test_cornell_depth8_rr_matches_oracle_within_noise(rt, oracle, device):
    say("[gpu] cornell_mesh 4 x 640x480, 2 spp, depth 8, RR from bounce 2, vs oracle")
    g, rig, prof, gpu = render_cornell(rt, device, 2, 8, 2, 7)
    osc = oracle_scene(oracle, g.desc)
    ob = oracle.OracleBvh(osc)
    op = oracle_profile(oracle, prof)
    for ci in range(rig.active_count):
        cam = oracle_camera(oracle, rig.cameras[ci])
        a = oracle.render_camera(osc, ob, cam, ci, op, 7, threads=THREADS)["beauty"][..., :3]
        b = oracle.render_camera(osc, ob, cam, ci, op, 8, threads=THREADS)["beauty"][..., :3]
        band = rel_mse(a, b)             # two independent estimates of the same frame
        err = rel_mse(gpu[ci], a)
        same = float(np.mean(gpu[ci] == a))
        say(f"  camera {ci}: GPU vs oracle relMSE {err:.3e}, MC band {band:.3e}, "
            f"bit-exact fraction {same:.4f}, mean {gpu[ci].mean():.5f} / {a.mean():.5f}")
        assert np.isfinite(gpu[ci]).all()
        assert err <= 0.02 * band + 1e-9, (ci, err, band)


def test_graph_replay_matches_oracle(rt, oracle, device):
    say("[gpu] graph replay (third replay) vs oracle: random_spheres 2 spp depth 6")
    g = rt.Generated("random_spheres", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    ren.set_tuning(use_graph=1)
    prof = g.profile
    prof.samples_per_pixel, prof.max_bounces, prof.rr_start_bounce = 2, 6, 3
    rig = g.rig()
    for f in range(3):
        ren.reset_accumulation()
        ren.render(rig, prof, 40)
        sc.trace_closest(np.zeros(64, rt.RAY_DTYPE))     # eager work between replays
    gpu = ren.download(0)
    osc = oracle_scene(oracle, g.desc)
    ob = oracle.OracleBvh(osc)
    ref = oracle.render_camera(osc, ob, oracle_camera(oracle, rig.cameras[0]), 0,
                               oracle_profile(oracle, prof), 40, threads=THREADS)
    for key in ("beauty", "normal", "albedo", "depth"):
        same = float(np.mean(gpu[key] == ref[key]))
        say(f"  {key}: bit-exact fraction {same:.6f}")
        assert same > 0.999


def scaled(oracle, rig, scale):
    for i in range(rig.active_count):
        cam = rig.cameras[i]
        cam.width //= scale
        cam.height //= scale
        cam.fx /= scale
        cam.fy /= scale
        cam.cx /= scale
        cam.cy /= scale
        if cam.model == 1:
            oc = oracle_camera(oracle, cam)
            oracle.lib().rto_make_equi_lut(C.byref(oc))
            C.memmove(C.byref(cam), C.byref(oc), C.sizeof(oc))
    return rig


def test_radiance_cache_scene_small_rig(rt, oracle, device):
    say("[gpu] radiance_cache_realtime, rig at 1/8 scale (4 x 160x90), cache + denoiser")
    frame, seed = 12, 1000
    g = rt.Generated("radiance_cache_realtime", seed)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()

    def rig():
        return scaled(oracle, g.rig_at(frame, 60), 8)

    # high-spp plain reference, and its parity with the oracle on camera 0
    ren = rt.Renderer(device, sc)
    prof = g.profile
    prof.enable_restir_di, prof.enable_denoise = 0, 0
    prof.samples_per_pixel = 4
    r = rig()
    ren.render(r, prof, 5)
    gpu0 = ren.download(0)["beauty"][..., :3]
    osc = oracle_scene(oracle, g.desc)
    ob = oracle.OracleBvh(osc)
    ref0 = oracle.render_camera(osc, ob, oracle_camera(oracle, r.cameras[0]), 0,
                                oracle_profile(oracle, prof), 5, threads=THREADS)["beauty"][..., :3]
    e0 = rel_mse(gpu0, ref0)
    say(f"  4-spp plain camera 0: GPU vs oracle relMSE {e0:.2e}, bit-exact "
        f"{np.mean(gpu0 == ref0):.4f}")
    assert e0 < 1e-4
    prof.samples_per_pixel = 512
    ren.reset_accumulation()
    ren.render(r, prof, 77)
    ref = [ren.download(c)["beauty"][..., :3] for c in range(r.active_count)]
    prof.samples_per_pixel = 1
    ren.reset_accumulation()
    ren.render(r, prof, 9)
    plain = [ren.download(c)["beauty"][..., :3] for c in range(r.active_count)]
    ren.close()

    w = np.load(os.path.join(os.path.dirname(rt.__file__), "weights", "denoiser.npy"))
    nn = rt.Denoiser(0, w)
    # a fixed training share: on this small rig the automatic share would grow
    # to every path (fewer pixels than training records), leaving no queries
    cache = rt.RadianceCache(0, train_fraction=0.25, train_samples=16384)
    ren = rt.Renderer(device, sc)
    ren.set_denoiser(nn)
    ren.set_radiance_cache(cache)
    prof = g.profile
    prof.samples_per_pixel, prof.enable_denoise = 1, 1
    for f in range(frame - 11, frame + 1):
        ren.render(scaled(oracle, g.rig_at(f, 60), 8), prof, 100 + f)
    st = cache.stats()
    out = [ren.download(c)["beauty"][..., :3] for c in range(r.active_count)]
    e_plain = np.mean([rel_mse(p, q) for p, q in zip(plain, ref)])
    e_learn = np.mean([rel_mse(p, q) for p, q in zip(out, ref)])
    say(f"  cache stats {st}; relMSE plain 1 spp {e_plain:.4f}, cache + denoiser {e_learn:.4f}")
    assert st["usable"] == 1 and st["queries"] > 0 and st["records"] > 0
    assert all(np.isfinite(o).all() for o in out)
    assert e_learn < e_plain
