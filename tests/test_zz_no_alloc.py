"""The allocation-free closest-hit entry point (rt_trace_closest_device_noalloc)
against rt_trace_closest_device, through the C ABI on device buffers: two
calls in a row, the second with fewer rays (its fetch cursor and stack
spill live in the library's one per-device scratch array, which keeps the
previous call's contents), on a shallow and on a deep (1,024,000-triangle)
tree; every hit record bit-exact."""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import random_rays
from hipmem import DeviceArray, sync

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("config, scale", [("random_spheres", 12.0),
                                           ("instanced_mesh_restir", 60.0)])
def test_noalloc_trace_matches_allocating_entry_point(rt, device, config, scale):
    g = rt.Generated(config, 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    rng = np.random.default_rng(90)
    L = rt.lib()
    for n in (300_000, 4_099):
        say(f"[gpu] noalloc trace {config}: {n} rays")
        rays = random_rays(rng, n, origin_scale=scale, target_scale=scale * 0.7)
        d_rays = DeviceArray.of(rays)
        d_a = DeviceArray(n * rt.HIT_DTYPE.itemsize)
        d_b = DeviceArray(n * rt.HIT_DTYPE.itemsize)
        ms = C.c_double()
        rt.check(L.rt_trace_closest_device(sc.h, d_rays.p, n, d_a.p, C.byref(ms)))
        rt.check(L.rt_trace_closest_device_noalloc(sc.h, d_rays.p, n, d_b.p))
        sync()
        a = d_a.download(n, rt.HIT_DTYPE)
        b = d_b.download(n, rt.HIT_DTYPE)
        hits = int((a["prim"] >= 0).sum())
        say(f"  {hits} hits; identical {a.tobytes() == b.tobytes()}")
        assert hits > n // 20
        assert a.tobytes() == b.tobytes()
