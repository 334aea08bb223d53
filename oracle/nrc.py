"""CPU oracle of the radiance cache network — TEST INFRASTRUCTURE ONLY.

The reference has no radiance cache (SURVEY: a neural radiance cache is only a
roadmap candidate, docs/design/ROADMAP_POLICY.md:526-535), so this module
states the network the product trains from its published sources, in float64:
  * multiresolution hash encoding (Mueller, Evans, Schied, Keller, "Instant
    Neural Graphics Primitives with a Multiresolution Hash Encoding", SIGGRAPH
    2022): L levels of resolution N_l = floor(N_min * b^l), trilinear
    interpolation of F features at the 8 corners of the cell, corners indexed
    densely while (N_l + 1)^3 <= T, else by the spatial hash
    x * 1 ^ y * 2654435761 ^ z * 805459861 mod T;
  * the MLP of "Real-time Neural Radiance Caching for Path Tracing" (Mueller,
    Rousselle, Novak, Keller, SIGGRAPH 2021): no biases, ReLU hidden layers,
    linear output, relative L2 loss (pred - y)^2 / (sg(pred)^2 + 0.01).
Pinned by central finite differences of the loss (tests/test_oracle_nrc.py).
"""
from __future__ import annotations

import numpy as np

PRIMES = (1, 2654435761, 805459861)


def level_resolutions(levels: int, n_min: int, n_max: int):
    b = np.exp((np.log(n_max) - np.log(n_min)) / max(levels - 1, 1))
    return [int(np.floor(n_min * b ** l)) for l in range(levels)]


def corner_index(c, res, table):
    """Index of integer corner(s) c[..., 3] at a level of resolution res."""
    c = np.asarray(c, np.uint64)
    if (res + 1) ** 3 <= table:
        return (c[..., 0] + c[..., 1] * (res + 1) + c[..., 2] * (res + 1) ** 2).astype(np.int64)
    h = (c[..., 0] * np.uint64(PRIMES[0])) ^ (c[..., 1] * np.uint64(PRIMES[1])) ^ (c[..., 2] * np.uint64(PRIMES[2]))
    return (h & np.uint64(0xFFFFFFFF) & np.uint64(table - 1)).astype(np.int64)


def encode(x, grid, resolutions):
    """x: [n, 3] in [0, 1]; grid: [L, T, F] -> features [n, L*F] and the
    corner indices/weights for the backward pass."""
    n = x.shape[0]
    L, T, F = grid.shape
    feats = np.zeros((n, L * F))
    cache = []
    for l, res in enumerate(resolutions):
        p = x * res
        p0 = np.floor(p)
        p0 = np.minimum(p0, res - 1)
        w = p - p0
        idx, wts = [], []
        for corner in range(8):
            o = np.array([(corner >> 0) & 1, (corner >> 1) & 1, (corner >> 2) & 1])
            c = (p0 + o).astype(np.int64)
            ww = np.prod(np.where(o == 1, w, 1.0 - w), axis=1)
            i = corner_index(c, res, T)
            feats[:, l * F:(l + 1) * F] += ww[:, None] * grid[l, i, :]
            idx.append(i)
            wts.append(ww)
        cache.append((np.stack(idx, 1), np.stack(wts, 1)))
    return feats, cache


def forward(inputs, weights):
    """inputs [n, K0]; weights: list of [out, in] matrices. Returns the
    activations list (input, hidden..., output)."""
    acts = [inputs]
    h = inputs
    for li, W in enumerate(weights):
        z = h @ W.T
        h = np.maximum(z, 0.0) if li < len(weights) - 1 else z
        acts.append(h)
    return acts


def loss_and_grads(x, extra, targets, grid, weights, resolutions, outputs=3, eps=0.01):
    """Relative L2 loss over n samples and `outputs` channels, its gradient
    with respect to every weight matrix and every grid feature."""
    feats, cache = encode(x, grid, resolutions)
    inputs = np.concatenate([feats, extra], axis=1)
    acts = forward(inputs, weights)
    pred = acts[-1][:, :outputs]
    n = x.shape[0]
    denom = pred ** 2 + eps
    loss = np.sum((pred - targets) ** 2 / denom) / (n * outputs)
    dpred = np.zeros_like(acts[-1])
    dpred[:, :outputs] = 2.0 * (pred - targets) / denom / (n * outputs)  # denominator held constant
    dW = [None] * len(weights)
    delta = dpred
    for li in range(len(weights) - 1, -1, -1):
        dW[li] = delta.T @ acts[li]
        delta = delta @ weights[li]
        if li > 0:
            delta = delta * (acts[li] > 0)
    dgrid = np.zeros_like(grid)
    L, T, F = grid.shape
    for l in range(L):
        idx, wts = cache[l]
        for corner in range(8):
            np.add.at(dgrid[l], idx[:, corner], wts[:, corner, None] * delta[:, l * F:(l + 1) * F])
    return loss, dW, dgrid, pred


def adam(param, grad, m, v, step, lr, beta1=0.9, beta2=0.99, eps=1e-15, sparse=False):
    """One Adam step (Kingma and Ba 2015) with bias correction; `sparse`
    leaves entries with an exactly zero gradient (and their moments) alone."""
    sel = grad != 0 if sparse else np.ones(grad.shape, bool)
    m[sel] = beta1 * m[sel] + (1 - beta1) * grad[sel]
    v[sel] = beta2 * v[sel] + (1 - beta2) * grad[sel] * grad[sel]
    mh = m[sel] / (1 - beta1 ** step)
    vh = v[sel] / (1 - beta2 ** step)
    param[sel] -= lr * mh / (np.sqrt(vh) + eps)
    return param


# --- the device network's shapes and FP16 rounding points -----------------
# Queries are 16 floats (position in [0, 1]^3, direction, normal, roughness,
# base colour, metallic, 2 unused); the network's 16 extra inputs are query
# floats 3..13, then 1, then 4 zeros. Weights are one flat vector of
# [out][in] matrices. The device rounds to FP16: the grid features it reads,
# each level's interpolated features, the extra inputs, the weights, every
# hidden activation (after ReLU) and every back-propagated output gradient
# (after the ReLU mask), the last clamped per sample to +-DELTA_CLAMP first.
SHAPES = [(64, 48), (64, 64), (64, 64), (64, 64), (16, 64)]
DELTA_CLAMP = 1024.0


def f16(x):
    return np.clip(x, -65504.0, 65504.0).astype(np.float16).astype(np.float64)


def unflatten(w):
    out, o = [], 0
    for r, c in SHAPES:
        out.append(np.asarray(w[o:o + r * c], np.float64).reshape(r, c))
        o += r * c
    return out


def query_inputs(q):
    q = np.asarray(q, np.float64)
    x = np.clip(q[:, 0:3], 0.0, 1.0)
    extra = np.zeros((q.shape[0], 16))
    extra[:, 0:11] = q[:, 3:14]
    extra[:, 11] = 1.0
    return x, extra


def device_forward(queries, weights, grid, resolutions):
    """The device network's output [n, 3] and what backprop needs."""
    x, extra = query_inputs(queries)
    feats, cache = encode(x, f16(grid), resolutions)
    inputs = np.concatenate([f16(feats), f16(extra)], axis=1)
    Ws = [f16(W) for W in unflatten(weights)]
    acts = [inputs]
    h = inputs
    for li, W in enumerate(Ws):
        z = h @ W.T
        h = f16(np.maximum(z, 0.0)) if li < len(Ws) - 1 else z
        acts.append(h)
    return acts[-1][:, :3], dict(acts=acts, Ws=Ws, cache=cache, x=x)


def device_loss_and_grads(samples, weights, grid, resolutions):
    """Mean relative L2 loss over samples [n, 20] and its mean gradients with
    respect to the flat weights and the grid, at the device's rounding points
    (see above). The device sums per-sample gradients and divides by n."""
    samples = np.asarray(samples, np.float64)
    n = samples.shape[0]
    pred, st = device_forward(samples[:, :16], weights, grid, resolutions)
    y = samples[:, 16:19]
    den = pred * pred + 0.01
    loss = np.sum((pred - y) ** 2 / den) / (3 * n)
    d = np.clip(2.0 * (pred - y) / den / 3.0, -DELTA_CLAMP, DELTA_CLAMP)
    acts, Ws = st["acts"], st["Ws"]
    dz = np.zeros((n, 16))
    dz[:, :3] = f16(d)
    dW = [None] * 5
    for li in range(4, -1, -1):
        dW[li] = dz.T @ acts[li]
        if li == 0:
            break
        dz = f16((dz @ Ws[li]) * (acts[li] > 0))
    dfeat = dz @ Ws[0][:, :32]
    dgrid = np.zeros_like(np.asarray(grid, np.float64))
    L, T, F = dgrid.shape
    for l in range(L):
        idx, wts = st["cache"][l]
        for corner in range(8):
            np.add.at(dgrid[l], idx[:, corner], wts[:, corner, None] * dfeat[:, l * F:(l + 1) * F])
    flat = np.concatenate([g.reshape(-1) for g in dW]) / n
    return loss, flat, dgrid / n, pred
