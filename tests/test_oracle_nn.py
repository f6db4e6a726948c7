"""Pins the oracle's learned-denoiser restatement (forward, loss, backward):
the analytic gradient equals central finite differences of the loss, the
parameter layout matches the product's, and bf16 rounding is round-to-nearest-
even."""
import numpy as np

import oracle_ctypes as O


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


def test_param_count_matches_product(rt):
    assert O.lib().rto_nn_param_count() == rt.lib().rt_nn_param_count() == 32483


def test_bf16_rounding_is_nearest_even():
    r = O.lib().rto_bf16_round
    assert r(1.0) == 1.0
    assert r(1.0 + 2.0 ** -9) == 1.0          # halfway, even stays
    assert r(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7 * 1  # halfway between 1+2^-8 and 1+2^-7 -> even
    assert r(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7  # above half -> up
    assert r(-3.14159) == -3.140625


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    im, frame, target = random_batch(rng, [(12, 9), (7, 10)])
    params = random_params(rng)
    loss, grad, _ = O.nn_loss_grad(im, params, frame, target, bf16=False)
    assert loss > 0 and np.isfinite(grad).all()
    idx = np.concatenate([rng.choice(len(params), 80, replace=False),
                          np.arange(len(params) - 99, len(params))])
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
    assert ok >= 0.97 * len(idx), f"{ok} of {len(idx)} gradient entries match finite differences"


def test_zero_network_is_identity():
    rng = np.random.default_rng(6)
    im, frame, target = random_batch(rng, [(8, 8)])
    params = np.zeros(O.lib().rto_nn_param_count(), np.float32)
    _, _, out = O.nn_loss_grad(im, params, frame, target)
    b = frame["beauty"][:, :3]
    assert np.allclose(out[:, :3], b, rtol=1e-5, atol=1e-6)
