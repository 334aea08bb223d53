"""Pins the radiance-cache oracle (oracle/nrc.py): analytic gradients of the
relative L2 loss against central finite differences in float64, the hash
encoding's partition of unity and dense/hashed indexing, and Adam's update."""
import numpy as np

from oracle import nrc


def _setup(seed=0, n=24, levels=3, table=64, extra=5, widths=(12, 8, 3)):
    rng = np.random.default_rng(seed)
    res = nrc.level_resolutions(levels, 2, 9)
    grid = rng.uniform(-0.5, 0.5, (levels, table, 2))
    x = rng.uniform(0, 1, (n, 3))
    ex = rng.uniform(-1, 1, (n, extra))
    k0 = levels * 2 + extra
    dims = [k0, *widths]
    weights = [rng.normal(0, 0.5, (dims[i + 1], dims[i])) for i in range(len(dims) - 1)]
    targets = rng.uniform(0, 2, (n, 3))
    return x, ex, targets, grid, weights, res


def test_encoding_partition_of_unity():
    x, ex, t, grid, w, res = _setup()
    ones = np.ones_like(grid)
    feats, cache = nrc.encode(x, ones, res)
    assert np.allclose(feats, 1.0)
    for idx, wts in cache:
        assert np.allclose(wts.sum(1), 1.0)


def test_dense_and_hashed_levels():
    # (res + 1)^3 <= T: dense, distinct corners; beyond: hashed into [0, T)
    assert nrc.corner_index(np.array([[1, 2, 3]]), 3, 64)[0] == 1 + 2 * 4 + 3 * 16
    h = nrc.corner_index(np.array([[5, 7, 11]]), 100, 1 << 10)[0]
    expect = (5 * 1 ^ (7 * 2654435761) ^ (11 * 805459861)) % (1 << 32) & ((1 << 10) - 1)
    assert h == expect


def test_gradients_match_finite_differences():
    x, ex, t, grid, w, res = _setup(seed=3)
    loss, dW, dgrid, _ = nrc.loss_and_grads(x, ex, t, grid, w, res)
    h = 1e-6
    rng = np.random.default_rng(7)

    def num(fn):
        return (fn(+h) - fn(-h)) / (2 * h)

    # The loss's denominator is held constant in its gradient (stop-gradient);
    # finite differences of the same surrogate use the fixed denominator.
    feats, _ = nrc.encode(x, grid, res)
    pred0 = nrc.forward(np.concatenate([feats, ex], 1), w)[-1][:, :3]
    denom = pred0 ** 2 + 0.01

    def surrogate(grid_, w_):
        f, _ = nrc.encode(x, grid_, res)
        p = nrc.forward(np.concatenate([f, ex], 1), w_)[-1][:, :3]
        return np.sum((p - t) ** 2 / denom) / (x.shape[0] * 3)

    for li in range(len(w)):
        for _ in range(4):
            i, j = rng.integers(w[li].shape[0]), rng.integers(w[li].shape[1])

            def f(d):
                ww = [m.copy() for m in w]
                ww[li][i, j] += d
                return surrogate(grid, ww)

            assert abs(num(f) - dW[li][i, j]) < 1e-6 * max(1.0, abs(dW[li][i, j])), (li, i, j)
    for _ in range(12):
        l, e, c = rng.integers(grid.shape[0]), rng.integers(grid.shape[1]), rng.integers(2)

        def g(d):
            gg = grid.copy()
            gg[l, e, c] += d
            return surrogate(gg, w)

        assert abs(num(g) - dgrid[l, e, c]) < 1e-6 * max(1.0, abs(dgrid[l, e, c])), (l, e, c)


def test_adam_first_step_moves_by_lr():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    m = np.zeros(3)
    v = np.zeros(3)
    nrc.adam(p, g, m, v, 1, 0.01)
    # bias-corrected first step: p -= lr * sign(g)
    assert np.allclose(p, [0.99, -1.99, 2.99])


def test_adam_sparse_skips_untouched_entries():
    p = np.array([1.0, 2.0])
    g = np.array([0.0, 1.0])
    m = np.array([0.5, 0.0])
    v = np.array([0.25, 0.0])
    nrc.adam(p, g, m, v, 1, 0.1, sparse=True)
    assert p[0] == 1.0 and m[0] == 0.5 and v[0] == 0.25
    assert np.isclose(p[1], 1.9)


def _device_setup(seed=0, n=64, table=256):
    rng = np.random.default_rng(seed)
    res = nrc.level_resolutions(16, 2, 40)
    grid = rng.uniform(-1, 1, (16, table, 2))
    w = np.concatenate([rng.uniform(-1, 1, r * c) * np.sqrt(6.0 / (r + c)) for r, c in nrc.SHAPES])
    q = np.zeros((n, 20))
    q[:, 0:3] = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    q[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    q[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q[:, 9:14] = rng.uniform(0, 1, (n, 5))
    q[:, 16:19] = rng.uniform(0, 2, (n, 3))
    return q, w, grid, res


def test_device_rounding_model_tracks_exact_gradients():
    # The FP16 rounding points change the gradient by about FP16's precision:
    # the device model's gradients match the float64 ones to a few 1e-3.
    q, w, grid, res = _device_setup()
    loss_h, dw_h, dg_h, pred_h = nrc.device_loss_and_grads(q, w, grid, res)
    x, extra = nrc.query_inputs(q[:, :16])
    loss, dW, dgrid, pred = nrc.loss_and_grads(x, extra, q[:, 16:19], grid, nrc.unflatten(w), res)
    dw = np.concatenate([g.reshape(-1) for g in dW])
    assert abs(loss_h - loss) < 1e-2 * loss
    assert np.linalg.norm(pred_h - pred) < 1e-2 * np.linalg.norm(pred)
    assert np.linalg.norm(dw_h - dw) < 2e-2 * np.linalg.norm(dw)
    assert np.linalg.norm(dg_h - dgrid) < 2e-2 * np.linalg.norm(dgrid)
