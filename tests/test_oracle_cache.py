"""Pins of the radiance-cache oracle (oracle/oracle_cache.c): hash-grid
indexing and interpolation known answers, and the analytic gradient against
finite differences of the forward pass."""
import numpy as np
import pytest

from oracle import oracle_ctypes as O

MLP = 4 * 64 * 64 + 3 * 64


def grid(levels=4, log2_table=10, base=4, mx=32, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    b = np.exp((np.log(mx) - np.log(base)) / (levels - 1)) if levels > 1 else 1.0
    res = np.zeros(16, np.float32)
    dense = np.zeros(16, np.int32)
    for l in range(levels):
        r = np.floor(base * b ** l + 1e-6)
        res[l] = r
        dense[l] = 1 if (int(r) + 1) ** 3 <= (1 << log2_table) else 0
    inv = np.array([np.float32(1.0 / (np.float64(h) - np.float64(l))) for l, h in zip(lo, hi)],
                   np.float32)
    return O.cache_grid({"levels": levels, "log2_table": log2_table, "res": res,
                         "dense": dense, "lo": np.array(lo, np.float32), "inv_extent": inv})


def query(pos, d=(0.0, 0.0, 1.0), rough=0.5, n=(0.0, 1.0, 0.0), alb=(0.5, 0.5, 0.5)):
    return np.array([*pos, 0.0, *d, rough, *n, 0.0, *alb, 0.0], np.float32)


def test_mlp_param_count():
    assert O.lib().rto_cache_mlp_param_count() == MLP


def test_dense_vertex_lookup_is_exact():
    g = grid(levels=1, log2_table=12, base=8, mx=8)  # 9^3 = 729 <= 4096: dense
    assert g.dense[0] == 1 and g.res[0] == 8.0
    for (i, j, k) in [(0, 0, 0), (3, 5, 7), (7, 0, 2)]:
        idx, w = O.cache_corners(g, 0, [i / 8, j / 8, k / 8])
        assert idx[0] == i + 9 * j + 81 * k
        assert w[0] == 1.0 and np.all(w[1:] == 0.0)
        # the 8 corners are the unit cube's vertices in x-fastest order
        expect = [(i + (c & 1)) + 9 * (j + ((c >> 1) & 1)) + 81 * (k + (c >> 2)) for c in range(8)]
        assert list(idx) == expect


def test_upper_face_clamps_into_the_last_cell():
    g = grid(levels=1, log2_table=12, base=8, mx=8)
    idx, w = O.cache_corners(g, 0, [1.0, 1.0, 1.0])
    # cell 7, fraction 1: all weight on the far corner (8, 8, 8)
    assert w[7] == 1.0 and np.all(w[:7] == 0.0)
    assert idx[7] == 8 + 9 * 8 + 81 * 8


def test_spatial_hash_known_answers():
    # a level too fine for dense storage uses the published primes
    # (1, 2654435761, 805459861) xor-ed and masked
    g = grid(levels=1, log2_table=8, base=64, mx=64)
    assert g.dense[0] == 0
    mask = 255
    idx, w = O.cache_corners(g, 0, [1 / 64, 0.0, 0.0])   # cell (1, 0, 0)
    assert idx[0] == (1 ^ 0 ^ 0) & mask
    assert idx[1] == 2 & mask
    assert idx[2] == (1 ^ 2654435761) & mask
    assert idx[4] == (1 ^ 805459861) & mask
    assert idx[7] == (2 ^ 2654435761 ^ 805459861) & mask


def test_weights_partition_unity():
    g = grid()
    rng = np.random.default_rng(3)
    for _ in range(50):
        p = rng.random(3).astype(np.float32)
        for l in range(4):
            idx, w = O.cache_corners(g, l, p)
            assert np.all(w >= 0.0)
            assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6
            assert np.all(idx < (1 << 10))


def test_features_are_linear_in_the_table_and_exact_at_vertices():
    g = grid(levels=2, log2_table=12, base=8, mx=16)
    rng = np.random.default_rng(5)
    T = 2 * (1 << 12)
    t1 = rng.normal(size=2 * T).astype(np.float32)
    t2 = rng.normal(size=2 * T).astype(np.float32)
    q = query([0.31, 0.77, 0.52])
    f1 = O.cache_features(g, t1, q, bf16=False)[:4].astype(np.float64)
    f2 = O.cache_features(g, t2, q, bf16=False)[:4].astype(np.float64)
    f12 = O.cache_features(g, (t1.astype(np.float64) + t2).astype(np.float32), q, bf16=False)[:4]
    np.testing.assert_allclose(f12, f1 + f2, rtol=1e-5, atol=1e-5)
    # at a level-0 vertex the features are that vertex's table entry
    qv = query([2 / 8, 3 / 8, 5 / 8])
    fv = O.cache_features(g, t1, qv, bf16=False)
    v = 2 + 9 * 3 + 81 * 5
    assert fv[0] == t1[2 * v] and fv[1] == t1[2 * v + 1]
    # non-grid inputs land on their channels; channel 42 is the constant 1
    assert fv[32:35].tolist() == [0.0, 0.0, 1.0] and fv[38] == np.float32(0.5)
    assert fv[42] == 1.0 and np.all(fv[43:] == 0.0) and np.all(fv[2 * 2:32] == 0.0)


def random_problem(seed=0, n=24, levels=3, log2_table=9):
    g = grid(levels=levels, log2_table=log2_table, base=4, mx=32,
             lo=(-1.0, -2.0, 0.5), hi=(3.0, 2.0, 4.5))
    rng = np.random.default_rng(seed)
    ng = (levels << log2_table) * 2
    params = np.concatenate([rng.uniform(-0.3, 0.3, MLP), rng.uniform(-0.5, 0.5, ng)]).astype(np.float32)
    qs = []
    for _ in range(n):
        pos = rng.uniform([-1, -2, 0.5], [3, 2, 4.5])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        nn = rng.normal(size=3)
        nn /= np.linalg.norm(nn)
        qs.append(query(pos, d, rng.random(), nn, rng.random(3)))
    inputs = np.stack(qs)
    targets = np.zeros((n, 4), np.float32)
    targets[:, :3] = rng.uniform(0.0, 2.0, (n, 3))
    return g, params, inputs, targets


def fixed_den_loss(g, params, inputs, targets, den):
    y = O.cache_predict(g, params, inputs, bf16=False).astype(np.float64)
    return float(np.sum((y - targets[:, :3]) ** 2 / den) / (3 * len(inputs)))


def test_gradient_matches_finite_differences():
    g, params, inputs, targets = random_problem()
    loss, grad = O.cache_loss_grad(g, params, inputs, targets, bf16=False)
    y0 = O.cache_predict(g, params, inputs, bf16=False).astype(np.float64)
    den = y0 ** 2 + 0.01
    # the reported loss is the relative L2 at the unperturbed point
    ref = float(np.sum((y0 - targets[:, :3]) ** 2 / den) / (3 * len(inputs)))
    assert abs(loss - ref) <= 1e-6 * ref
    rng = np.random.default_rng(1)
    mlp_idx = rng.choice(MLP, 24, replace=False)
    grid_idx = MLP + np.argsort(-np.abs(grad[MLP:]))[:16]
    for i in list(mlp_idx) + list(grid_idx):
        h = 1e-4 * max(abs(float(params[i])), 1e-2)
        pp = params.astype(np.float64).copy()
        pm = params.astype(np.float64).copy()
        pp[i] += h
        pm[i] -= h
        hp = np.float32(pp[i]) - np.float32(params[i])
        hm = np.float32(params[i]) - np.float32(pm[i])
        fd = (fixed_den_loss(g, pp.astype(np.float32), inputs, targets, den)
              - fixed_den_loss(g, pm.astype(np.float32), inputs, targets, den)) / float(hp + hm)
        assert abs(fd - grad[i]) <= 2e-3 * abs(grad[i]) + 1e-7, (i, fd, grad[i])
    # every used grid entry with a gradient is one a sample's corner touches
    assert np.count_nonzero(grad[MLP:]) <= len(inputs) * 8 * 3 * 2


def test_bf16_mode_tracks_exact_mode():
    g, params, inputs, targets = random_problem(seed=2)
    l0, g0 = O.cache_loss_grad(g, params, inputs, targets, bf16=False)
    l1, g1 = O.cache_loss_grad(g, params, inputs, targets, bf16=True)
    assert abs(l1 - l0) <= 0.05 * l0
    rel = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    assert rel < 0.05, rel


def test_zero_output_layer_predicts_zero_and_gradient_only_reaches_it():
    g, params, inputs, targets = random_problem(seed=4)
    p = params.copy()
    p[4 * 64 * 64:MLP] = 0.0
    y = O.cache_predict(g, p, inputs, bf16=False)
    assert np.all(y == 0.0)
    _, grad = O.cache_loss_grad(g, p, inputs, targets, bf16=False)
    # with W4 = 0 no gradient flows below the output layer
    assert np.all(grad[:4 * 64 * 64] == 0.0) and np.all(grad[MLP:] == 0.0)
    assert np.any(grad[4 * 64 * 64:MLP] != 0.0)
