"""Pins the oracle's learned-denoiser restatement (4-level encoder-decoder:
forward, loss, backward): the analytic gradient equals central finite
differences of the loss for parameters of every convolution (so the pooling,
upsampling and skip routes of the backward pass are checked), the parameter
layout matches the product's, and bf16 rounding is round-to-nearest-even."""
import numpy as np

import oracle_ctypes as O

CONVS = [(16, 0, 32), (32, 0, 32), (32, 0, 48), (48, 0, 48), (48, 0, 64), (64, 0, 64),
         (64, 0, 96), (96, 0, 96), (96, 0, 96), (64, 96, 64), (64, 0, 64), (48, 64, 48),
         (48, 0, 48), (32, 48, 32), (32, 0, 32)]


def random_batch(rng, sizes):
    im = O.images(sizes)
    P = im.offset[im.count]
    beauty = np.zeros((P, 4), np.float32)
    beauty[:, :3] = rng.exponential(0.5, (P, 3))
    beauty[:, 3] = 1
    n = rng.normal(size=(P, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    normal = np.zeros((P, 4), np.float32)
    normal[:, :3] = n * 0.5 + 0.5
    normal[:, 3] = 1
    albedo = np.zeros((P, 4), np.float32)
    albedo[:, :3] = rng.uniform(0, 1, (P, 3))
    depth = rng.uniform(1, 10, P).astype(np.float32)
    target = np.zeros((P, 4), np.float32)
    target[:, :3] = rng.exponential(0.4, (P, 3))
    return im, {"beauty": beauty, "normal": normal, "albedo": albedo, "depth": depth}, target


def random_params(rng, scale=0.15):
    n = O.lib().rto_nn_param_count()
    return (rng.normal(size=n) * scale).astype(np.float32)


def scaled_params(rng, gain=1.0):
    """Weights N(0, gain / fan_in), biases N(0, 0.02), head N(0, 0.05): keeps
    the activations of all 15 convolutions in range."""
    n = O.lib().rto_nn_param_count()
    p = np.zeros(n, np.float32)
    o = 0
    for c0, c1, co in CONVS:
        fan = 9 * (c0 + c1)
        nw = co * fan
        p[o:o + nw] = rng.normal(size=nw) * np.sqrt(gain / fan)
        p[o + nw:o + nw + co] = rng.normal(size=co) * 0.02
        o += nw + co
    p[o:] = rng.normal(size=n - o) * 0.05
    return p.astype(np.float32)


def test_param_count_matches_product(rt):
    assert O.lib().rto_nn_param_count() == rt.lib().rt_nn_param_count() == 565443


def test_bf16_rounding_is_nearest_even():
    r = O.lib().rto_bf16_round
    assert r(1.0) == 1.0
    assert r(1.0 + 2.0 ** -9) == 1.0          # halfway, even stays
    assert r(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7 * 1  # halfway between 1+2^-8 and 1+2^-7 -> even
    assert r(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7  # above half -> up
    assert r(-3.14159) == -3.140625


def layer_ranges():
    o, out = 0, []
    for c0, c1, co in CONVS:
        n = co * 9 * (c0 + c1) + co
        out.append((o, o + n))
        o += n
    return out


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    # odd sizes: partial 2x2 pools at every level
    im, frame, target = random_batch(rng, [(12, 9), (7, 10)])
    params = random_params(rng, 0.12)
    loss, grad, _ = O.nn_loss_grad(im, params, frame, target, bf16=False)
    assert loss > 0 and np.isfinite(grad).all()
    idx = []
    for a, b in layer_ranges():
        idx += list(rng.choice(np.arange(a, b - 1), 6, replace=False)) + [b - 1]  # weights + a bias
    idx += list(range(len(params) - 99, len(params), 7))
    h = 1e-3
    ok = 0
    for i in idx:
        p1 = params.copy()
        p1[i] += h
        p2 = params.copy()
        p2[i] -= h
        fd = (O.nn_loss_grad(im, p1, frame, target, bf16=False, want_grad=False)[0]
              - O.nn_loss_grad(im, p2, frame, target, bf16=False, want_grad=False)[0]) / (2 * h)
        if abs(fd - grad[i]) <= 2e-4 * np.abs(grad).max() + 3e-2 * abs(grad[i]):
            ok += 1
    assert ok >= 0.95 * len(idx), f"{ok} of {len(idx)} gradient entries match finite differences"
    # every convolution receives gradient
    for a, b in layer_ranges():
        assert np.abs(grad[a:b]).max() > 0


def test_zero_network_is_identity():
    rng = np.random.default_rng(6)
    im, frame, target = random_batch(rng, [(8, 8)])
    params = np.zeros(O.lib().rto_nn_param_count(), np.float32)
    _, _, out = O.nn_loss_grad(im, params, frame, target)
    b = frame["beauty"][:, :3]
    assert np.allclose(out[:, :3], b, rtol=1e-5, atol=1e-6)


def test_loss_is_relative_squared_error_to_first_order():
    """The zero network returns its input, so its loss on a target that
    differs by a small relative amount d equals the relMSE of the input,
    mean (b - t)^2 / (t^2 + 1e-2), up to O(d); a dropped (1 + t) factor or a
    wrong epsilon moves it by far more."""
    rng = np.random.default_rng(7)
    im, frame, target = random_batch(rng, [(9, 7)])
    t = target[:, :3].astype(np.float64)
    t[::5] = 0.0                                   # black pixels: the epsilon term alone
    d = rng.uniform(-1, 1, t.shape) * 1e-3
    b = t * (1 + d) + 1e-4 * (t == 0)
    frame["beauty"][:, :3] = b
    target[:, :3] = t
    params = np.zeros(O.lib().rto_nn_param_count(), np.float32)
    loss, _, _ = O.nn_loss_grad(im, params, frame, target, bf16=False, want_grad=False)
    b32 = frame["beauty"][:, :3].astype(np.float64)
    rel = np.mean((b32 - t) ** 2 / (t ** 2 + 1e-2))
    assert abs(loss - rel) <= 5e-3 * rel, (loss, rel)
