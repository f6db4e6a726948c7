"""GPU parity of the learned denoiser's BF16 MFMA kernels (4-level
encoder-decoder) against the CPU oracle: forward output, loss and the full
parameter gradient (backward data and weight passes through the pooling,
upsampling and skip routes), plus a short training run that lowers the
loss."""
import sys

import numpy as np
import pytest

from test_oracle_nn import random_batch, scaled_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SIZES = [(40, 24), (17, 33), (64, 16)]  # odd sizes: partial pools


def test_mfma_forward_matches_oracle(rt, oracle, device):
    say("[gpu] nn forward vs oracle")
    rng = np.random.default_rng(11)
    im, frame, target = random_batch(rng, SIZES)
    params = scaled_params(rng)
    nn = rt.Denoiser(0, params)
    out = nn.infer(SIZES, frame)
    _, _, ref = oracle.nn_loss_grad(oracle.images(SIZES), params, frame, None, bf16=True,
                                    want_grad=False)
    a, b = out[:, :3].astype(np.float64), ref[:, :3].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-2)
    say(f"  max rel {rel.max():.2e}, mean {rel.mean():.2e}")
    assert np.isfinite(a).all()
    assert np.mean(rel < 2e-2) > 0.995 and rel.mean() < 2e-3


def test_mfma_gradient_matches_oracle(rt, oracle, device):
    say("[gpu] nn gradient vs oracle")
    rng = np.random.default_rng(12)
    im, frame, target = random_batch(rng, SIZES)
    params = scaled_params(rng)
    nn = rt.Denoiser(0, params)
    loss, grad = nn.train_step(SIZES, frame, target, 0.0, want_grad=True)
    rloss, rgrad, _ = oracle.nn_loss_grad(oracle.images(SIZES), params, frame, target, bf16=True)
    say(f"  loss {loss:.6f} vs {rloss:.6f}")
    assert abs(loss - rloss) <= 1e-3 * abs(rloss)
    assert np.isfinite(grad).all()
    err = np.linalg.norm(grad - rgrad) / np.linalg.norm(rgrad)
    say(f"  gradient relative L2 error {err:.2e}")
    assert err < 2e-2
    # every layer's block of the gradient agrees, not just the large ones
    from test_oracle_nn import layer_ranges
    offs = [a for a, _ in layer_ranges()] + [layer_ranges()[-1][1], len(params)]
    for i in range(len(offs) - 1):
        g, r = grad[offs[i]:offs[i + 1]], rgrad[offs[i]:offs[i + 1]]
        e = np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-12)
        say(f"  block {i}: rel L2 {e:.2e}")
        assert e < 5e-2
    # the parameters are unchanged by a zero learning rate
    assert np.array_equal(nn.params(), params)


def test_training_lowers_loss(rt, oracle, device):
    say("[gpu] nn training steps")
    rng = np.random.default_rng(13)
    im, frame, _ = random_batch(rng, SIZES)
    # a target the network can represent: a function of the albedo buffer
    target = np.zeros_like(frame["albedo"])
    target[:, :3] = 0.8 * frame["albedo"][:, :3] + 0.05
    nn = rt.Denoiser(0, scaled_params(rng, 0.5))
    first, _ = nn.train_step(SIZES, frame, target, 1e-3)
    last = first
    for _ in range(80):
        last, _ = nn.train_step(SIZES, frame, target, 1e-3)
    say(f"  loss {first:.5f} -> {last:.5f}")
    assert last < 0.5 * first


def test_tile_shapes_give_identical_outputs_and_autotune(rt, oracle, device):
    """The convolution tile rows change the launch shape, not the arithmetic:
    every output accumulates its taps and channel chunks in the same order,
    so tilings of 4 and 8 rows agree bit for bit (forward and gradient), and
    the autotuner applies one of them."""
    say("[gpu] nn tile shapes 4 / 8 rows, autotune")
    rng = np.random.default_rng(12)
    sizes = [(200, 70), (90, 41)]            # widths on both sides of the 160-pixel rule
    im, frame, target = random_batch(rng, sizes)
    params = scaled_params(rng)
    nn = rt.Denoiser(0, params)
    outs, grads = [], []
    for rows in ([4, 4, 4, 4], [8, 8, 8, 8], [8, 4, 8, 4], [0, 0, 0, 0]):
        nn.set_tiling(rows)
        assert nn.tiling() == rows
        outs.append(nn.infer(sizes, frame))
        loss, g = nn.train_step(sizes, frame, target, lr=0.0, want_grad=True)
        grads.append(g)
    for o, g in zip(outs[1:], grads[1:]):
        assert np.array_equal(o, outs[0])
        assert np.array_equal(g, grads[0])
    res = nn.autotune(sizes, reps=2)
    say(f"  autotune {res}")
    assert all(r in (4, 8) for r in res["conv_rows"]) and res["trials"] == 5
    assert nn.tiling() == res["conv_rows"]
    assert res["best_ms"] <= res["baseline_ms"]
    assert np.array_equal(nn.infer(sizes, frame), outs[0])
