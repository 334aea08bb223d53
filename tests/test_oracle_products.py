"""Pins the oracle's render-buffer writes (adapters/hydra2/src/adapter.cpp:
2544-2596, 2713-2731) by known answers and an independent numpy
restatement: rows flipped bottom-up, RGBA32F to UNorm8 as lround(clamp(v,
0, 1) * 255) (halves round up, NaN gives 0), UNorm8 to RGBA32F as v / 255,
and the pairs the reference refuses."""
import numpy as np

import oracle

F32 = np.float32


def _np_unorm(v):
    p = np.clip(v, F32(0), F32(1)).astype(F32) * F32(255)
    out = np.floor(p.astype(np.float64) + 0.5)
    out[np.isnan(v)] = 0
    return out.astype(np.uint8)


def test_known_answers():
    v = np.array([[[0.5 / 255 * 255 / 255, 127.5 / 255, 1.0, -3.0], [2.0, np.nan, 0.0, 254.5 / 255]]], F32)
    out = oracle.write_render_buffer(v, 1, 0).reshape(1, 2, 4)
    assert out[0, 0, 2] == 255 and out[0, 0, 3] == 0 and out[0, 1, 0] == 255 and out[0, 1, 1] == 0
    assert out[0, 1, 2] == 0
    assert np.array_equal(out, _np_unorm(v))
    b = np.array([[[0, 1, 128, 255]]], np.uint8)
    f = oracle.write_render_buffer(b, 0, 1).view(F32).reshape(1, 1, 4)
    assert f[0, 0, 0] == 0 and f[0, 0, 3] == 1.0 and f[0, 0, 1] == F32(1) / F32(255)


def test_flips_and_conversions_match_numpy():
    rng = np.random.default_rng(3)
    h, w = 37, 53
    v = rng.uniform(-0.5, 1.5, (h, w, 4)).astype(F32)
    v[rng.uniform(size=(h, w, 4)) < 0.05] = np.nan
    v[:3] = (np.arange(3 * w * 4).reshape(3, w, 4) % 256 + 0.5).astype(F32) / F32(255)  # exact halves
    assert np.array_equal(oracle.write_render_buffer(v, 1, 0).reshape(h, w, 4), np.flipud(_np_unorm(v)))
    assert np.array_equal(oracle.write_render_buffer(v, 1, 1).view(F32).reshape(h, w, 4), np.flipud(v),
                          equal_nan=True)
    b = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    assert np.array_equal(oracle.write_render_buffer(b, 0, 1).view(F32).reshape(h, w, 4),
                          np.flipud(b.astype(F32) / F32(255)))
    assert np.array_equal(oracle.write_render_buffer(b, 0, 0).reshape(h, w, 4), np.flipud(b))
    d = rng.uniform(0, 1, (h, w)).astype(F32)
    assert np.array_equal(oracle.write_render_buffer(d, 2, 2).view(F32).reshape(h, w), np.flipud(d))


def test_refused_pairs():
    d = np.zeros((4, 4), F32)
    assert oracle.write_render_buffer(d, 2, 3) is None
    assert oracle.write_render_buffer(np.zeros((4, 4, 4), F32), 1, 2) is None
