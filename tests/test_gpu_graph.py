"""The frame's launch sequence captured in a hipGraph and replayed gives the
same frames as eager launches, frame after frame (ping-pong history,
ReSTIR reservoirs, denoiser), with eager work between replays."""
import sys

import numpy as np
import pytest

from test_gpu_restir import scaled_rig

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_sequence(rt, oracle, device, config, graph, frames, denoise=False, nn=None):
    g = rt.Generated(config, 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    ren.set_tuning(use_graph=1 if graph else 0)
    if nn is not None:
        ren.set_denoiser(nn)
    prof = g.profile
    prof.enable_denoise = 1 if denoise else 0
    if config == "random_spheres":
        prof.samples_per_pixel = 2
        prof.max_bounces = 4
    out = []
    for f in range(frames):
        rig = scaled_rig(oracle, g, f) if config == "instanced_mesh_restir" else g.rig()
        ren.render(rig, prof, 100 + f)
        frame = [ren.download(c) for c in range(rig.active_count)]
        if prof.enable_restir_di:
            for c in range(rig.active_count):
                frame[c]["reservoirs"] = ren.reservoirs(c)
        out.append(frame)
        # eager work between replays: an unrelated trace on the same device
        sc.trace_closest(np.zeros(64, rt.RAY_DTYPE))
    return out


@pytest.mark.parametrize("config,denoise", [("instanced_mesh_restir", False),
                                            ("random_spheres", True)])
def test_graph_replay_equals_eager(rt, oracle, device, config, denoise):
    say(f"[gpu] graph replay vs eager: {config} denoise={denoise}")
    eager = run_sequence(rt, oracle, device, config, False, 4, denoise)
    graph = run_sequence(rt, oracle, device, config, True, 4, denoise)
    for f, (fe, fg) in enumerate(zip(eager, graph)):
        for c, (a, b) in enumerate(zip(fe, fg)):
            for k in a:
                assert a[k].tobytes() == b[k].tobytes(), f"frame {f} camera {c}: {k} differs"
    say("  all frames identical")


def test_graph_with_learned_denoiser(rt, oracle, device):
    say("[gpu] graph replay with the learned denoiser")
    import os
    w = np.load(os.path.join(os.path.dirname(rt.__file__), "weights", "denoiser.npy"))
    nn = rt.Denoiser(0, w)
    eager = run_sequence(rt, oracle, device, "random_spheres", False, 3, True, nn)
    graph = run_sequence(rt, oracle, device, "random_spheres", True, 3, True, nn)
    for fe, fg in zip(eager, graph):
        assert fe[0]["beauty"].tobytes() == fg[0]["beauty"].tobytes()
    assert np.isfinite(graph[-1][0]["beauty"]).all()


def test_render_realtime_cli(rt, tmp_path):
    """The C++ render_realtime CLI renders, reports and writes its artifacts."""
    import json
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(rt.__file__), "render_realtime")
    if not os.path.exists(exe):
        pytest.skip("render_realtime not built")
    say("[gpu] render_realtime CLI")
    out = tmp_path / "rr"
    r = subprocess.run([exe, "--scene", "random_spheres", "--profile", "scene", "--frames", "3",
                        "--warmup-frames", "1", "--camera-count", "1", "--skip-image-write",
                        "--output-dir", str(out)], capture_output=True, text=True, timeout=120)
    say(r.stdout[-600:])
    assert r.returncode == 0, r.stderr
    assert "summary profile=scene" in r.stdout and "frame=2 sample_stream=3" in r.stdout
    s = json.load(open(out / "benchmark_summary.json"))
    assert s["schema_version"] == 3 and len(s["frames"]) == 3 and len(s["per-camera"]) == 3
    assert s["aggregate"]["frame_ms"]["avg"] > 0
    rows = open(out / "benchmark_frames.csv").read().strip().splitlines()
    assert len(rows) == 4
    m = json.load(open(out / "benchmark_manifest.json"))
    assert m["measured_frame_count"] == 3 and len(m["artifacts"]) == 2
