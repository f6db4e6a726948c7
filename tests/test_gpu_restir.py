"""GPU parity of the analytic-light and ReSTIR DI paths against the CPU oracle.

The scene is the 1,024,000-triangle instanced torus field with 256 analytic
sphere lights; the orbiting 4-camera rig is scaled to 1/8 resolution so the
oracle renders each frame in seconds. Frames are rendered as a sequence so the
temporal and spatial reuse read the previous frame's reservoirs, exactly as
the renderer keeps them between frames (history captured after each frame,
ping-pong reservoir buffers).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import oracle_camera, oracle_profile, oracle_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SCALE = 8


def scaled_rig(oracle, g, frame, frames=60):
    """Orbit rig at frame `frame`, every camera scaled down by SCALE (the
    equidistant LUT is rebuilt for the new image extent)."""
    rig = g.rig_at(frame, frames)
    for i in range(rig.active_count):
        cam = rig.cameras[i]
        cam.width //= SCALE
        cam.height //= SCALE
        cam.fx /= SCALE
        cam.fy /= SCALE
        cam.cx /= SCALE
        cam.cy /= SCALE
        if cam.model == 1:
            oc = oracle_camera(oracle, cam)
            oracle.lib().rto_make_equi_lut(C.byref(oc))
            C.memmove(C.byref(cam), C.byref(oc), C.sizeof(oc))
    return rig


@pytest.fixture(scope="module")
def instanced(rt, oracle, device):
    g = rt.Generated("instanced_mesh_restir", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    osc = oracle_scene(oracle, g.desc)
    ob = oracle.OracleBvh(osc)
    return g, sc, osc, ob


def test_instanced_bvh_bitexact(rt, oracle, instanced):
    say("[gpu] bvh instanced_mesh_restir (1,024,000 triangles)")
    g, sc, osc, ob = instanced
    nodes, sp, mc = sc.download_bvh()
    assert sc.stats.node_count == ob.bvh.node_count == 1024000 - 1  # internal nodes
    assert sc.stats.max_depth == ob.bvh.max_depth
    assert np.array_equal(mc, ob.morton())
    assert np.array_equal(sp, ob.sorted_prims())
    assert nodes.tobytes() == ob.nodes().tobytes()
    assert sc.stats.wide_node_count == ob.bvh.wide_count
    assert sc.download_wide().tobytes() == ob.wide().tobytes()


def compare_frame(gpu, ref, label, tol_frac=0.995):
    a = gpu["beauty"][..., :3].astype(np.float64)
    b = ref["beauty"][..., :3].astype(np.float64)
    assert np.isfinite(a).all(), f"{label}: non-finite beauty"
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    close = np.mean(np.all(rel < 1e-4, axis=-1))
    exact = np.mean(np.all(a == b, axis=-1))
    say(f"  {label}: bit-exact pixels {exact:.5f}, within 1e-4 {close:.5f}, "
        f"mean {a.mean():.5f} vs {b.mean():.5f}")
    assert close >= tol_frac, f"{label}: only {close:.5f} of pixels within 1e-4"
    # frame means agree far tighter than the Monte-Carlo noise of a 1-spp frame
    assert abs(a.mean() - b.mean()) <= 2e-3 * max(abs(b.mean()), 1e-3)
    for key in ("normal", "albedo", "depth"):
        assert np.mean(gpu[key] == ref[key]) >= 0.999, f"{label}: {key} differs"


def test_analytic_nee_without_restir(rt, oracle, device, instanced):
    say("[gpu] instanced_mesh_restir analytic NEE, ReSTIR off, 1/8 rig, 2 spp")
    g, sc, osc, ob = instanced
    prof = g.profile
    prof.enable_restir_di = 0
    prof.samples_per_pixel = 2
    rig = scaled_rig(oracle, g, 3)
    ren = rt.Renderer(device, sc)
    ren.render(rig, prof, 5)
    for ci in range(rig.active_count):
        ref = oracle.render_camera(osc, ob, oracle_camera(oracle, rig.cameras[ci]), ci,
                                   oracle_profile(oracle, prof), 5)
        compare_frame(ren.download(ci), ref, f"cam{ci}")


def compare_reservoirs(gr, orr, label):
    same_light = np.mean(gr["light_index"] == orr["light_index"])
    same_valid = np.mean(gr["valid"] == orr["valid"])
    w_ok = np.isclose(gr["estimator_weight"], orr["estimator_weight"], rtol=1e-4, atol=1e-7)
    counts = [np.mean(gr[k] == orr[k]) for k in ("candidate_count", "temporal_candidate_count",
                                                  "spatial_candidate_count", "age")]
    say(f"  {label}: light {same_light:.5f} valid {same_valid:.5f} W {np.mean(w_ok):.5f} "
        f"counts {['%.5f' % c for c in counts]}")
    assert same_valid >= 0.995 and same_light >= 0.99 and np.mean(w_ok) >= 0.99
    assert min(counts) >= 0.99


@pytest.mark.parametrize("bias", [0, 1])
def test_restir_sequence_matches_oracle(rt, oracle, device, instanced, bias):
    say(f"[gpu] instanced_mesh_restir ReSTIR 3-frame sequence, bias_correction={bias}")
    g, sc, osc, ob = instanced
    prof = g.profile
    assert prof.enable_restir_di == 1 and prof.samples_per_pixel == 1
    prof.restir_bias_correction = bias
    prof.restir_spatial_radius = 4  # 30 px at full resolution
    ren = rt.Renderer(device, sc)
    seq = oracle.RestirSequence(osc, ob, oracle_profile(oracle, prof))
    for frame in range(3):
        rig = scaled_rig(oracle, g, frame)
        ren.render(rig, prof, frame)
        refs = seq.render([oracle_camera(oracle, rig.cameras[i]) for i in range(rig.active_count)],
                          frame)
        for ci in range(rig.active_count):
            compare_frame(ren.download(ci), refs[ci], f"frame{frame} cam{ci}", tol_frac=0.99)
            gr = ren.reservoirs(ci)
            compare_reservoirs(gr, refs[ci]["reservoirs"], f"frame{frame} cam{ci} reservoirs")
            # RestirDiagnostics counts (optix_renderer.cpp:183-215) over the same reservoirs
            st = ren.restir_stats(ci)
            valid = gr["valid"] != 0
            assert st["pixel_count"] == len(gr)
            assert st["active_reservoir_count"] == int(valid.sum())
            assert st["temporal_reuse_count"] == int((valid & (gr["temporal_candidate_count"] > 0)).sum())
            assert st["spatial_reuse_count"] == int((valid & (gr["spatial_candidate_count"] > 0)).sum())
            assert st["bias_corrected_count"] == int(
                (valid & (gr["bias_correction_source_count"] > 0)).sum())
            assert st["max_candidate_count"] == int(gr["candidate_count"][valid].max(initial=0))
            assert st["max_age"] == int(gr["age"][valid].max(initial=0))
            if frame > 0:
                assert st["temporal_reuse_count"] > 0 and st["spatial_reuse_count"] > 0
            if bias and frame > 0:
                assert st["bias_corrected_count"] > 0
