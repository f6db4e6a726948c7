"""GPU parity of the radiance cache's BF16 MFMA kernels against the CPU
oracle (inference, loss, MLP and hash-grid gradients, the Adam step), a
training run that learns a smooth function, and the cache inside the
renderer: attached but not yet usable it leaves every frame bit-identical
to the plain tracer; once usable it ends paths at the query vertex."""
import sys

import numpy as np
import pytest

from test_oracle_cache import MLP

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


LO = np.array([-1.0, -2.0, 0.5], np.float32)
HI = np.array([3.0, 2.0, 4.5], np.float32)


def small_cache(rt, **kw):
    cfg = dict(levels=4, log2_table=12, base_resolution=4, max_resolution=64, train_samples=4096,
               train_steps=1, learning_rate=1e-2, seed=7)
    cfg.update(kw)
    c = rt.RadianceCache(0, **cfg)
    c.set_bounds(LO, HI)
    return c


def random_queries(rng, n):
    q = np.zeros((n, 16), np.float32)
    q[:, 0:3] = rng.uniform(LO, HI, (n, 3))
    d = rng.normal(size=(n, 3))
    q[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    q[:, 7] = rng.random(n)
    nn = rng.normal(size=(n, 3))
    q[:, 8:11] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q[:, 12:15] = rng.random((n, 3))
    return q


def random_params(rng, c):
    p = np.concatenate([rng.uniform(-0.25, 0.25, MLP), rng.uniform(-0.5, 0.5, c.n - MLP)])
    return p.astype(np.float32)


def test_cache_inference_matches_oracle(rt, oracle, device):
    say("[gpu] cache inference vs oracle")
    rng = np.random.default_rng(21)
    c = small_cache(rt)
    params = random_params(rng, c)
    c.set_params(params)
    q = random_queries(rng, 777)
    out = c.infer(q)[:, :3].astype(np.float64)
    g = oracle.cache_grid(c.grid_info())
    ref = oracle.cache_predict(g, params, q, bf16=True)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    say(f"  rel L2 {rel:.2e}, max abs {np.abs(out - ref).max():.2e}, |ref| {np.abs(ref).mean():.3f}")
    assert np.isfinite(out).all()
    assert rel < 5e-3
    # the encoding's constants are the library's documented ones
    info = c.grid_info()
    assert list(info["res"][:4]) == [4.0, 10.0, 25.0, 64.0]
    assert list(info["dense"][:4]) == [1, 1, 0, 0]


def test_cache_gradient_and_adam_step_match_oracle(rt, oracle, device):
    say("[gpu] cache gradient vs oracle")
    rng = np.random.default_rng(22)
    c = small_cache(rt)
    params = random_params(rng, c)
    c.set_params(params)
    q = random_queries(rng, 1500)
    t = np.zeros((len(q), 4), np.float32)
    t[:, :3] = rng.uniform(0.0, 3.0, (len(q), 3))
    loss, grad = c.train_step(q, t, want_grad=True)
    g = oracle.cache_grid(c.grid_info())
    ref_loss, ref = oracle.cache_loss_grad(g, params, q, t, bf16=True)
    say(f"  loss {loss:.6f} oracle {ref_loss:.6f}")
    assert abs(loss - ref_loss) <= 2e-3 * ref_loss
    gm, rm = grad[:MLP].astype(np.float64), ref[:MLP]
    gg, rg = grad[MLP:].astype(np.float64), ref[MLP:]
    em = np.linalg.norm(gm - rm) / np.linalg.norm(rm)
    eg = np.linalg.norm(gg - rg) / np.linalg.norm(rg)
    say(f"  MLP grad rel L2 {em:.2e}, grid grad rel L2 {eg:.2e}")
    assert em < 1e-2 and eg < 1e-2
    # the grid gradient lands exactly on the cells the samples touch
    assert np.array_equal(np.flatnonzero(np.abs(gg) > 0), np.flatnonzero(np.abs(rg) > 0)) or \
        np.mean((np.abs(gg) > 0) == (np.abs(rg) > 0)) > 0.999
    # first Adam step: m / sqrt(v) = sign(g), so each parameter moves by lr
    after = c.params()
    moved = params.astype(np.float64) - after
    big = np.abs(grad) > 1e-6
    expect = 1e-2 * np.sign(grad[big]) * (np.abs(grad[big]) / (np.abs(grad[big]) + 1e-8))
    np.testing.assert_allclose(moved[big], expect, rtol=1e-4, atol=2e-7)


def test_cache_learns_a_smooth_function(rt, device):
    say("[gpu] cache training")
    rng = np.random.default_rng(23)
    c = small_cache(rt, levels=8, log2_table=14, max_resolution=128, train_samples=8192)

    def target(q):
        p = (q[:, 0:3] - LO) / (HI - LO)
        v = 0.6 + 0.4 * np.sin(4.0 * p[:, 0]) * np.cos(3.0 * p[:, 1]) + 0.3 * p[:, 2]
        t = np.zeros((len(q), 4), np.float32)
        t[:, 0] = v
        t[:, 1] = 0.5 * v
        t[:, 2] = 1.0 - 0.5 * p[:, 2]
        return t

    held = random_queries(rng, 2048)
    th = target(held)

    def held_loss():
        y = c.infer(held)[:, :3].astype(np.float64)
        return float(np.mean((y - th[:, :3]) ** 2 / (y ** 2 + 0.01)))

    before = held_loss()
    losses = []
    for _ in range(300):
        q = random_queries(rng, 4096)
        loss, _ = c.train_step(q, target(q))
        losses.append(loss)
    after = held_loss()
    say(f"  held-out relative loss {before:.4f} -> {after:.5f}; train {losses[0]:.4f} -> {losses[-1]:.5f}")
    assert np.isfinite(losses).all()
    assert after < 0.05 * before and after < 0.01
    assert c.stats()["steps"] == 300


def spheres_scene(rt, device):
    g = rt.Generated("random_spheres", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    prof = g.profile
    prof.samples_per_pixel = 2
    prof.max_bounces = 6
    return g, sc, prof


def test_cache_attached_before_warmup_leaves_frames_bit_identical(rt, device):
    say("[gpu] cache attached, not yet usable: plain frames")
    g, sc, prof = spheres_scene(rt, device)
    plain = rt.Renderer(device, sc)
    cached = rt.Renderer(device, sc)
    cache = rt.RadianceCache(0, train_samples=4096, warmup_frames=1000)
    cached.set_radiance_cache(cache)
    for f in range(3):
        plain.render(g.rig(), prof, 100 + f)
        cached.render(g.rig(), prof, 100 + f)
        a = plain.download(0)["beauty"]
        b = cached.download(0)["beauty"]
        assert np.array_equal(a, b), f"frame {f}"
    st = cache.stats()
    say(f"  stats {st}")
    assert st["frames"] == 3 and st["usable"] == 0 and st["queries"] == 0
    assert 0 < st["records"] <= 4096
    assert st["steps"] == 3 * cache.config.train_steps
    assert np.isfinite(st["last_loss"])


def test_cache_ends_paths_once_usable(rt, device):
    say("[gpu] cache usable: queries end paths")
    g, sc, prof = spheres_scene(rt, device)
    ren = rt.Renderer(device, sc)
    cache = rt.RadianceCache(0, train_samples=8192, warmup_frames=2, max_rel_loss=1e9)
    ren.set_radiance_cache(cache)
    del cache  # the renderer keeps its own reference
    stats = []
    for f in range(5):
        ren.render(g.rig(), prof, 200 + f)
        stats.append(ren._cache.stats())
    say(f"  stats {stats[-1]}")
    assert [s["usable"] for s in stats] == [0, 1, 1, 1, 1]
    assert stats[0]["queries"] == 0 and stats[1]["queries"] == 0
    assert all(s["queries"] > 0 for s in stats[2:])
    img = ren.download(0)["beauty"]
    assert np.isfinite(img).all() and img[..., :3].min() >= 0.0
    # a detached cache restores the plain tracer
    ren.set_radiance_cache(None)
    ren.render(g.rig(), prof, 300)
    assert np.isfinite(ren.download(0)["beauty"]).all()


def test_cache_frames_replay_in_a_graph(rt, device):
    say("[gpu] cache + graph replay")
    g, sc, prof = spheres_scene(rt, device)
    eager = rt.Renderer(device, sc)
    graph = rt.Renderer(device, sc)
    graph.set_tuning(use_graph=1)
    # records are appended in atomic order: with one batch and a capacity no
    # frame fills, both caches train on the same set and differ only in
    # summation order
    kw = dict(train_samples=65536, train_steps=1, train_fraction=0.03, warmup_frames=1,
              max_rel_loss=1e9)
    ce = rt.RadianceCache(0, **kw)
    cg = rt.RadianceCache(0, **kw)
    eager.set_radiance_cache(ce)
    graph.set_radiance_cache(cg)
    for f in range(4):
        eager.render(g.rig(), prof, 400 + f)
        graph.render(g.rig(), prof, 400 + f)
        a = eager.download(0)["beauty"][..., :3].astype(np.float64)
        b = graph.download(0)["beauty"][..., :3].astype(np.float64)
        if f == 0:
            # before the first prediction the frames are plain: identical
            assert np.array_equal(a, b), f"frame {f}"
        else:
            # gradients sum in atomic order, so the two caches agree to rounding
            rel = np.mean((a - b) ** 2) / max(np.mean(a ** 2), 1e-12)
            say(f"  frame {f}: eager vs graph relative MSE {rel:.2e}")
            assert rel < 1e-4, (f, rel)
        se, sg = ce.stats(), cg.stats()
        assert se["usable"] == sg["usable"] and se["queries"] == sg["queries"]


def test_cache_guard_falls_back_to_the_plain_tracer(rt, device):
    """The learned part's fallback for the cache: a cache whose running
    relative loss stays above max_rel_loss never ends a path, so the frames
    equal the plain tracer's bit for bit; a usable cache given NaN
    parameters (set_params restarts its training state, so it must earn
    usability again) reports a NaN loss, never becomes usable under a
    threshold it would otherwise pass, and the frames are plain."""
    say("[gpu] cache guard: loss over budget, diverged parameters")
    g, sc, prof = spheres_scene(rt, device)
    plain = rt.Renderer(device, sc)
    strict = rt.Renderer(device, sc)
    strict.set_radiance_cache(rt.RadianceCache(0, train_samples=8192, warmup_frames=1,
                                               max_rel_loss=1e-9))
    for f in range(4):
        plain.render(g.rig(), prof, 500 + f)
        strict.render(g.rig(), prof, 500 + f)
        st = strict._cache.stats()
        assert st["usable"] == 0 and st["queries"] == 0 and st["frames"] == f + 1
        assert strict.download(0)["beauty"].tobytes() == plain.download(0)["beauty"].tobytes(), f

    ren = rt.Renderer(device, sc)
    cache = rt.RadianceCache(0, train_samples=8192, warmup_frames=1, max_rel_loss=1e9)
    ren.set_radiance_cache(cache)
    for f in range(3):
        ren.render(g.rig(), prof, 600 + f)
    assert cache.stats()["usable"] == 1 and cache.stats()["queries"] > 0
    p = cache.params()
    p[:] = np.nan
    cache.set_params(p)
    ren.render(g.rig(), prof, 603)            # trains on the NaN cache: a NaN loss
    st = cache.stats()
    say(f"  after NaN parameters: {st}")
    assert st["usable"] == 0 and not np.isfinite(st["last_loss"])
    # both start a fresh history (the NaN frame is in this renderer's)
    ren.reset_accumulation()
    plain.reset_accumulation()
    for f in range(4, 6):
        ren.render(g.rig(), prof, 600 + f)
        plain.render(g.rig(), prof, 600 + f)
        st = cache.stats()
        assert st["usable"] == 0 and st["queries"] == 0
        # the same paths as the plain tracer: training paths go on with the
        # plain integrator
        assert ren.download(0)["beauty"].tobytes() == plain.download(0)["beauty"].tobytes(), f
