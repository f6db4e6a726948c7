"""The learned denoiser's fallback to the reference filter, decided on the
GPU inside the frame: a network with a NaN weight (non-finite output) or
with all weights zero (the identity: it removes no noise) is replaced, per
camera, by the a-trous filter, and the frame equals the filter's frame bit
for bit, in eager launches and in graph replays; the shipped network on the
workload it was trained for passes the guard and its output is kept."""
import os
import sys

import numpy as np
import pytest

from test_gpu_restir import scaled_rig

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def weights(rt):
    return np.load(os.path.join(os.path.dirname(rt.__file__), "weights", "denoiser.npy"))


def frames(rt, oracle, device, nn, graph=False, n=1, budget=None):
    g = rt.Generated("instanced_mesh_restir", 0)
    sc = rt.Scene(device, g.desc)
    sc.build_bvh()
    ren = rt.Renderer(device, sc)
    ren.set_tuning(use_graph=1 if graph else 0)
    if nn is not None:
        ren.set_denoiser(nn)
    if budget is not None:
        ren.set_denoiser_budget(*budget)
    prof = g.profile
    prof.samples_per_pixel, prof.enable_denoise = 1, 1
    out = []
    for f in range(n):
        rig = scaled_rig(oracle, g, f)
        ren.render(rig, prof, 300 + f)
        out.append([ren.download(c)["beauty"] for c in range(rig.active_count)])
        sc.trace_closest(np.zeros(64, rt.RAY_DTYPE))     # eager work between replays
    guard = ren.denoiser_guard()
    return out, guard


@pytest.fixture(scope="module")
def filtered(rt, oracle, device):
    return frames(rt, oracle, device, None, n=3)[0]


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("kind", ["nan", "zero"])
def test_broken_network_falls_back_to_the_filter(rt, oracle, device, filtered, kind, graph):
    say(f"[gpu] denoiser guard: {kind} weights, graph={graph}")
    p = weights(rt).copy()
    if kind == "nan":
        p[-3:] = np.nan          # the FP32 head's biases (a NaN inside the ReLU layers is clipped)
    else:
        p[:] = 0.0
    out, guard = frames(rt, oracle, device, rt.Denoiser(0, p), graph=graph, n=3)
    say(f"  guard {guard}")
    assert guard["frames_checked"] == 3
    assert guard["fallback"] == [1, 1, 1, 1]
    assert guard["frames_fallback"] == 12
    if kind == "nan":
        # NaN outputs, or zeros where the output stage clamps them: either way caught
        assert all(n > 0 or m > 0.5 for n, m in zip(guard["nonfinite"], guard["mean_shift"]))
    else:
        assert all(v > 0.9 for v in guard["residual_noise"])
    for f in range(3):
        for c in range(4):
            assert out[f][c].tobytes() == filtered[f][c].tobytes(), f"frame {f} camera {c}"


def test_trained_network_passes_the_guard(rt, oracle, device, filtered):
    say("[gpu] denoiser guard: the shipped weights on their workload")
    out, guard = frames(rt, oracle, device, rt.Denoiser(0, weights(rt)), n=3)
    say(f"  guard {guard}")
    assert guard["fallback"] == [0, 0, 0, 0] and guard["frames_fallback"] == 0
    assert all(v < 0.5 for v in guard["residual_noise"])
    assert all(np.isfinite(o).all() for o in out[-1])
    assert any(out[-1][c].tobytes() != filtered[-1][c].tobytes() for c in range(4))


def test_budgets_are_honoured(rt, oracle, device):
    """With unbounded budgets the identity network's output is kept (it
    reproduces the input to rounding); with a zero noise budget even the
    trained network is replaced."""
    say("[gpu] denoiser guard budgets")
    out, guard = frames(rt, oracle, device, rt.Denoiser(0, np.zeros_like(weights(rt))),
                        budget=(1e30, 1e30))
    assert guard["fallback"] == [0, 0, 0, 0] and guard["max_residual_noise"] == np.float32(1e30)
    assert all(abs(v - 1.0) < 1e-2 for v in guard["residual_noise"])
    assert all(v < 1e-3 for v in guard["mean_shift"])
    out, guard = frames(rt, oracle, device, rt.Denoiser(0, weights(rt)), budget=(0.5, 0.0))
    assert guard["fallback"] == [1, 1, 1, 1]
    with pytest.raises(rt.RtError):
        frames(rt, oracle, device, None, budget=(-1.0, 0.5))
