"""Pins the oracle's SVGF accumulation with history rectification
(oracle/temporal.py, restating temporal.slang:143-267, which has no CPU
counterpart in the reference) by closed forms: no rectification is the plain
blend of the history into the fresh batch with weight kept / (kept + fresh);
a neighbourhood of one fresh value clamps any history to that value; a
history equal to the fresh neighbourhood keeps its whole count under
anti-lag and one four standard errors or more away drops it; a reflection
weight of one half keeps half the count, rounded to the nearest sample."""
import numpy as np

from lotus_amd import hip
from oracle import temporal as ot

W, H = 9, 7


def _records(value=(0.3, 0.5, 0.7), history=(0.6, 0.2, 0.9), old_count=12, noise=0.0, seed=0):
    rng = np.random.default_rng(seed)
    surf = np.zeros((H, W), hip.TEMPORAL_PIXEL)
    surf["flags"] = 7
    surf["instance_id"] = 5
    surf["albedo"] = 255 | (255 << 8) | (255 << 16)  # white: demodulation is the identity
    batch = np.zeros((H, W, 4), np.float32)
    batch[..., :3] = np.asarray(value, np.float32) * (1 + noise * rng.standard_normal((H, W, 3))).astype(np.float32)
    batch[..., 3] = 1
    valid = np.zeros((H, W), hip.TEMPORAL_VALIDITY)
    valid["valid"] = 1
    valid["previous_index"] = np.arange(H * W).reshape(H, W)
    old = np.zeros((H, W), hip.TEMPORAL_RADIANCE)
    old["color"][..., :3] = history
    old["color"][..., 3] = 1
    old["sample_count"] = old_count
    lum = ot.luminance(np.asarray(history, np.float32))
    old["first_moment"] = lum
    old["second_moment"] = lum * lum
    return batch, surf, valid, old


def test_plain_blend():
    batch, surf, valid, old = _records()
    out = ot.accumulate(batch, surf, valid, old, 4, 32)
    blend = np.float32(12) / np.float32(16)
    expect = batch[..., :3] + (old["color"][..., :3] - batch[..., :3]) * blend
    assert np.all(out["sample_count"] == 16)
    assert np.allclose(out["color"][..., :3], expect, rtol=1e-6)
    # the cap bounds the kept history: 4 fresh + at most 6 kept
    assert np.all(ot.accumulate(batch, surf, valid, old, 4, 10)["sample_count"] == 10)


def test_clamp_to_a_uniform_neighbourhood():
    batch, surf, valid, old = _records()
    out = ot.accumulate(batch, surf, valid, old, 4, 32, ot.CLAMP)
    # mean = the fresh value, deviation 0 up to the float rounding of
    # E[v^2] - E[v]^2 (its square root is ~1e-4): the history clamps to the
    # fresh value
    assert np.allclose(out["color"][..., :3], batch[..., :3], rtol=0, atol=3e-4)
    plain = ot.accumulate(batch, surf, valid, old, 4, 32)
    assert np.abs(plain["color"][..., :3] - batch[..., :3]).max() > 0.1
    # a history inside the band stays
    batch, surf, valid, old = _records(history=(0.3, 0.5, 0.7), noise=0.1, seed=3)
    inside = ot.accumulate(batch, surf, valid, old, 4, 32, ot.CLAMP)
    plain = ot.accumulate(batch, surf, valid, old, 4, 32)
    assert np.mean(np.all(inside["color"] == plain["color"], -1)) > 0.3


def test_anti_lag_keeps_a_matching_history_and_drops_a_far_one():
    batch, surf, valid, old = _records(history=(0.3, 0.5, 0.7), noise=0.05, seed=1)
    out = ot.accumulate(batch, surf, valid, old, 4, 32, ot.ANTI_LAG)
    assert np.all(out["sample_count"] == 16)  # z well below 2
    batch, surf, valid, old = _records(history=(3.0, 5.0, 7.0), noise=0.05, seed=1)
    out = ot.accumulate(batch, surf, valid, old, 4, 32, ot.ANTI_LAG)
    assert np.all(out["sample_count"] == 4)  # z far above 4: the history is dropped
    assert np.array_equal(out["color"], batch)


def test_reflection_weight_scales_the_kept_count():
    batch, surf, valid, old = _records(old_count=13)
    surf["reserved"] = np.array([0.5], np.float32).view(np.uint32)[0]
    out = ot.accumulate(batch, surf, valid, old, 4, 32, ot.REFLECTION)
    assert np.all(out["sample_count"] == 4 + 6)  # 6.5 rounds to the even 6
    surf["reserved"] = np.array([1.0], np.float32).view(np.uint32)[0]
    assert np.array_equal(ot.accumulate(batch, surf, valid, old, 4, 32, ot.REFLECTION),
                          ot.accumulate(batch, surf, valid, old, 4, 32))


def test_settle_closed_forms_and_reference():
    # SettleTemporalRadiance (temporal.slang:481-493): half way through its
    # span the output is the midpoint of prior and still; from the span on,
    # on a miss (span 0) and where the prior is not finite it is `still` bit
    # for bit; max_samples caps the span. The restatement equals the
    # reference's CPU form (temporal.cpp:514-528) bit for bit.
    import oracle
    rng = np.random.default_rng(4)
    n = 4096
    prior = rng.lognormal(-1, 1, (n, 4)).astype(np.float32)
    still = rng.lognormal(-1, 1, (n, 4)).astype(np.float32)
    spans = np.where(rng.random(n) < 0.2, 0, 32).astype(np.uint32)
    prior[:50, 1] = np.inf
    prior[50:60, 2] = np.nan
    for samples, cap in ((16, 0), (16, 40), (7, 10), (32, 0), (40, 40), (1, 0)):
        got = ot.settle(prior, still, spans, samples, cap)
        span = np.minimum(spans, cap) if cap else spans
        ends = (samples >= span) | ~np.all(np.isfinite(prior), -1)
        assert np.array_equal(got[ends], still[ends])
        if samples == 16 and cap == 0:
            mid = ~ends
            assert mid.sum() > 1000
            assert np.allclose(got[mid], (prior[mid] + still[mid]) / 2, rtol=1e-6)
        if oracle.ref() is not None:
            ref = ot.ref_settle(prior, still, spans, samples, cap)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (samples, cap)
