"""Pins the oracle's 4-wide quantized BVH (collapsed from the binary LBVH):
leaf ranges partition the primitives, every decoded child box contains the
exact bounds of its subtree, and traversal equals brute force."""
import numpy as np
import pytest

import oracle_ctypes as O
from helpers import TRI_DT, random_rays, random_triangles


def soup(rng, n, scale=10.0, size=1.0):
    t = random_triangles(rng, n, scale, size)
    m = np.zeros(1, np.dtype([("type", "<i4"), ("albedo", "<f4", 3), ("emission", "<f4", 3),
                              ("fuzz", "<f4"), ("ior", "<f4"), ("openpbr_index", "<i4")]))
    m["ior"] = 1.0
    m["openpbr_index"] = -1
    scene, keep = O.scene_from_arrays(triangles=t, materials=m)
    return scene, keep, t


def tri_boxes(t):
    p = np.stack([t["p0"], t["p1"], t["p2"]], 1).astype(np.float32)
    lo, hi = p.min(1), p.max(1)
    thin = (hi - lo) < np.float32(1e-5)
    lo = np.where(thin, lo - np.float32(1e-5), lo)
    hi = np.where(thin, hi + np.float32(1e-5), hi)
    return lo, hi


def decode(node, k):
    lo = np.zeros(3)
    hi = np.zeros(3)
    for a in range(3):
        e = int((node["meta"] >> (8 * a)) & 255) - 127
        ql = (int(node["qlo"][a]) >> (8 * k)) & 255
        qh = (int(node["qhi"][a]) >> (8 * k)) & 255
        lo[a] = float(node["origin"][a]) + ql * 2.0 ** e
        hi[a] = float(node["origin"][a]) + qh * 2.0 ** e
    return lo, hi


def test_wide_tree_partitions_and_bounds():
    rng = np.random.default_rng(21)
    scene, keep, t = soup(rng, 3000)
    b = O.OracleBvh(scene)
    w = b.wide()
    sp = b.sorted_prims()
    n = len(t)
    assert b.bvh.wide_count == len(w) > 0
    lo, hi = tri_boxes(t)
    slo, shi = lo[sp].astype(np.float64), hi[sp].astype(np.float64)
    # subtree slot ranges, children before parents (children have larger BFS
    # levels; resolve by repeated sweeps)
    rng_of = {}
    cover = np.zeros(n, int)
    for node in w:
        for ch in node["child"]:
            if ch != 0xFFFFFFFF and ch >> 31:
                f, c = int(ch & 0x0FFFFFFF), int((ch >> 28) & 7) + 1
                cover[f:f + c] += 1
    assert (cover == 1).all(), "leaf ranges must partition the primitives"
    pending = list(range(len(w)))
    while pending:
        nxt = []
        for i in pending:
            parts = []
            ok = True
            for ch in w[i]["child"]:
                if ch == 0xFFFFFFFF:
                    continue
                if ch >> 31:
                    f, c = int(ch & 0x0FFFFFFF), int((ch >> 28) & 7) + 1
                    parts.append((f, f + c))
                elif int(ch) in rng_of:
                    parts.append(rng_of[int(ch)])
                else:
                    ok = False
            if ok:
                rng_of[i] = (min(p[0] for p in parts), max(p[1] for p in parts))
            else:
                nxt.append(i)
        assert len(nxt) < len(pending)
        pending = nxt
    assert rng_of[0] == (0, n)
    for i, node in enumerate(w):
        for k, ch in enumerate(node["child"]):
            if ch == 0xFFFFFFFF:
                continue
            if ch >> 31:
                f, c = int(ch & 0x0FFFFFFF), int((ch >> 28) & 7) + 1
                r0, r1 = f, f + c
            else:
                r0, r1 = rng_of[int(ch)]
            dlo, dhi = decode(node, k)
            assert (dlo <= slo[r0:r1].min(0)).all() and (dhi >= shi[r0:r1].max(0)).all()
            assert (dhi - dlo <= (shi[r0:r1].max(0) - slo[r0:r1].min(0)) + 2.0 ** (
                ((node["meta"] >> (8 * np.arange(3))) & 255).astype(int) - 126)).all()


def test_wide_traversal_equals_brute_force():
    rng = np.random.default_rng(22)
    for n, size in ((1, 2.0), (2, 1.0), (37, 1.0), (5000, 0.5)):
        scene, keep, t = soup(rng, n, 10.0, size)
        b = O.OracleBvh(scene)
        rays = random_rays(rng, 4000)
        h = b.trace_closest(rays)
        bh = O.brute_closest(scene, rays)
        assert np.array_equal(h["prim"], bh["prim"])
        assert h.tobytes() == bh.tobytes()
        short = rays.copy()
        short["tmax"] = rng.uniform(0.5, 30.0, len(rays)).astype(np.float32)
        occ = b.trace_any(short)
        bs = O.brute_closest(scene, short)
        assert np.array_equal(occ, (bs["prim"] >= 0).astype(np.int32))


def _soup(seed, n=3000):
    from helpers import lambert_materials, random_triangles
    rng = np.random.default_rng(seed)
    return random_triangles(rng, n), lambert_materials(1)


def test_refit_of_unchanged_geometry_is_the_build(oracle):
    """A refit to the primitives the tree was built from reproduces the
    build's binary and 4-wide node arrays byte for byte."""
    t, m = _soup(31)
    s, keep = oracle.scene_from_arrays(triangles=t, materials=m)
    a = oracle.OracleBvh(s)
    b = oracle.OracleBvh(s)
    b.refit(s)
    assert b.nodes().tobytes() == a.nodes().tobytes()
    assert b.wide().tobytes() == a.wide().tobytes()


def test_refit_of_moved_geometry_keeps_topology_and_traces_exactly(oracle):
    """After moving every vertex the refit tree keeps the Morton order and
    child references, its boxes contain their subtrees, and closest hits
    equal brute force over the moved primitives."""
    from helpers import random_rays
    t, m = _soup(32)
    s, keep = oracle.scene_from_arrays(triangles=t, materials=m)
    ob = oracle.OracleBvh(s)
    before = ob.nodes()
    rng = np.random.default_rng(5)
    t2 = t.copy()
    for k in ("p0", "p1", "p2"):
        t2[k] += rng.normal(0, 0.3, t2[k].shape).astype(np.float32)
    s2, keep2 = oracle.scene_from_arrays(triangles=t2, materials=m)
    ob.refit(s2)
    after = ob.nodes()
    assert np.array_equal(before["child"], after["child"]) and np.array_equal(before["parent"], after["parent"])
    assert ob.validate() == 0
    rays = random_rays(rng, 4000)
    assert np.array_equal(ob.trace_closest(rays)["prim"], oracle.brute_closest(s2, rays)["prim"])
    # a refit is not a rebuild: the fresh tree of the moved soup differs
    assert oracle.OracleBvh(s2).nodes().tobytes() != after.tobytes()
    with pytest.raises(ValueError):
        fewer, k3 = oracle.scene_from_arrays(triangles=t2[:-1], materials=m)
        ob.refit(fewer)
