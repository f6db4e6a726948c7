"""Asynchronous frames (rt_renderer_set_async): untimed frames return once
enqueued, with the cameras and per-frame values staged in two host slots.
Four frames of a moving rig enqueued back to back, with no wait between
them, end in the same last frame (ReSTIR reservoirs, reprojected history,
denoiser) as four synchronous frames, eager and graph-captured; the radiance
cache's record and query counts of an asynchronous frame appear after
rt_renderer_synchronize and equal a synchronous frame's."""
import sys

import pytest

from test_gpu_restir import scaled_rig

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def last_frame(rt, oracle, device, asynchronous, graph, frames=4):
    g = rt.Generated("instanced_mesh_restir", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    ren.set_tuning(use_graph=1 if graph else 0)
    ren.set_async(asynchronous)
    prof = g.profile
    prof.enable_denoise = 1
    rig = None
    for f in range(frames):
        rig = scaled_rig(oracle, g, f)
        ren.render(rig, prof, 200 + f, timed=False)
    if asynchronous:
        ren.synchronize()
    out = [ren.download(c) for c in range(rig.active_count)]
    for c in range(rig.active_count):
        out[c]["reservoirs"] = ren.reservoirs(c)
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_async_frames_equal_sync(rt, oracle, device, graph):
    say(f"[gpu] four asynchronous frames vs synchronous, graph={graph}")
    a = last_frame(rt, oracle, device, False, graph)
    b = last_frame(rt, oracle, device, True, graph)
    for c, (x, y) in enumerate(zip(a, b)):
        for k in x:
            assert x[k].tobytes() == y[k].tobytes(), f"camera {c}: {k} differs"
    say("  last frames identical")


def test_async_cache_stats_after_synchronize(rt, device):
    say("[gpu] radiance cache statistics of an asynchronous frame")
    stats = []
    for asynchronous in (False, True):
        g = rt.Generated("radiance_cache_realtime", 0)
        sc = rt.Scene(device, g.desc)
        sc.build_bvh()
        ren = rt.Renderer(device, sc)
        cache = rt.RadianceCache(0)
        ren.set_radiance_cache(cache)
        ren.set_async(asynchronous)
        prof = g.profile
        prof.samples_per_pixel = 1
        ren.render(g.rig(), prof, 5, timed=False)
        ren.synchronize()
        st = cache.stats()
        say(f"  async={asynchronous}: records {st['records']}, queries {st['queries']}")
        stats.append((st["records"], st["queries"]))
        ren.set_radiance_cache(None)
    assert stats[0] == stats[1]
    assert stats[0][0] > 0  # training records (the cache answers no queries before its first step)
