"""SVGF temporal accumulation with history rectification — TEST
INFRASTRUCTURE ONLY.

A float32 numpy restatement of AccumulateTemporalRadiance with ClampHistory,
AntiLagDrop and ScaleKept (backend/vulkan/shaders/common/temporal.slang:
143-267), which the reference has no CPU counterpart for. Pinned by closed
forms in tests/test_oracle_temporal.py: a uniform neighbourhood clamps a
history to its value, a history equal to the fresh mean keeps its anti-lag
count, a far one drops it, a reflection weight of one half halves the kept
count, and no rectification is the plain blend.

SettleTemporalRadiance (temporal.slang:477-493) is restated as `settle`
and checked against the reference's CPU form (backend/vulkan/temporal.cpp:
514-528, built into oracle/_ref) and the closed forms of its ends.

Records are numpy structured arrays in the reference's layouts (the
TEMPORAL_PIXEL, TEMPORAL_VALIDITY and TEMPORAL_RADIANCE dtypes of
lotus_amd.hip), rows from the top.
"""
from __future__ import annotations

import numpy as np

F = np.float32
NONE, CLAMP, ANTI_LAG, REFLECTION = range(4)


def unpack_albedo(a):
    a = int(a)
    return np.array([a & 0xFF, (a >> 8) & 0xFF, (a >> 16) & 0xFF], F) / F(255.0)


def demodulate(c, albedo):
    a = unpack_albedo(albedo)
    c = np.asarray(c, F)
    return np.where(a > 0, c / np.where(a > 0, a, F(1)), c).astype(F)


def remodulate(c, albedo):
    a = unpack_albedo(albedo)
    c = np.asarray(c, F)
    return np.where(a > 0, c * a, c).astype(F)


def luminance(v):
    return F(F(F(0.2126) * v[0]) + F(F(0.7152) * v[1])) + F(F(0.0722) * v[2])


def _neighbour(surface, x, y, instance):
    h, w = surface.shape
    if x < 0 or y < 0 or x >= w or y >= h:
        return False
    s = surface[y, x]
    return (int(s["flags"]) & 1) != 0 and int(s["instance_id"]) == instance


def clamp_history(batch, surface, x, y, old_color):
    """temporal.slang:152-172."""
    s = surface[y, x]
    inst = int(s["instance_id"])
    total = np.zeros(3, F)
    squares = np.zeros(3, F)
    count = F(0)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if not _neighbour(surface, x + dx, y + dy, inst):
                continue
            v = demodulate(batch[y + dy, x + dx, :3], surface[y + dy, x + dx]["albedo"])
            total = (total + v).astype(F)
            squares = (squares + v * v).astype(F)
            count = F(count + F(1))
    mean = (total / count).astype(F)
    dev = np.sqrt(np.maximum((squares / count).astype(F) - (mean * mean).astype(F), F(0))).astype(F)
    o = demodulate(old_color[:3], s["albedo"])
    kept = np.minimum(np.maximum(o, (mean - dev).astype(F)), (mean + dev).astype(F)).astype(F)
    out = np.empty(4, F)
    out[:3] = remodulate(kept, s["albedo"])
    out[3] = old_color[3]
    return out


def anti_lag_drop(batch, surface, validity, old, x, y):
    """temporal.slang:177-212."""
    h, w = surface.shape
    inst = int(surface[y, x]["instance_id"])
    fresh = squares = count = history = histories = F(0)
    flat_old = old.reshape(-1)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            if not _neighbour(surface, x + dx, y + dy, inst):
                continue
            q = (y + dy, x + dx)
            albedo = surface[q]["albedo"]
            l = luminance(demodulate(batch[q][:3], albedo))
            fresh = F(fresh + l)
            squares = F(squares + F(l * l))
            count = F(count + F(1))
            v = validity[q]
            if int(v["valid"]) == 0:
                continue
            o = flat_old[int(v["previous_index"])]
            if int(o["sample_count"]) == 0 or not np.all(np.isfinite(o["color"])):
                continue
            history = F(history + luminance(demodulate(o["color"][:3], albedo)))
            histories = F(histories + F(1))
    if histories == 0 or count < 2:
        return F(0)
    mean = F(fresh / count)
    variance = F(F(max(F(squares - F(F(count * mean) * mean)), F(0))) / F(count - F(1)))
    z = F(F(abs(F(F(history / histories) - mean))) / F(np.sqrt(F(F(variance / count) + F(1e-12)))))
    return F(min(max(F(F(z - F(2)) * F(0.5)), F(0)), F(1)))


def scale_kept(kept, scale):
    """temporal.slang:214-217: to the nearest sample (halves to even)."""
    return int(np.rint(F(F(kept) * F(min(max(F(scale), F(0)), F(1))))))


def accumulate(batch, surface, validity, old, fresh_samples, cap, rectification=NONE):
    """temporal.slang:219-267: this frame's radiance records."""
    from lotus_amd import hip
    h, w = surface.shape
    out = np.zeros((h, w), hip.TEMPORAL_RADIANCE)
    flat_old = old.reshape(-1)
    for y in range(h):
        for x in range(w):
            o = out[y, x]
            color = np.array(batch[y, x], F)
            s = surface[y, x]
            count = 0
            if int(s["flags"]) & 1:
                count = min(fresh_samples, cap)
                l = luminance(demodulate(color[:3], s["albedo"]))
                m1, m2 = l, F(l * l)
                v = validity[y, x]
                if int(v["valid"]):
                    prev = flat_old[int(v["previous_index"])]
                    kept = min(int(prev["sample_count"]), cap - count)
                    old_color = np.array(prev["color"], F)
                    finite = np.all(np.isfinite(old_color)) and np.isfinite(prev["first_moment"]) and \
                        np.isfinite(prev["second_moment"])
                    if kept and finite and rectification:
                        if rectification == CLAMP:
                            old_color = clamp_history(batch, surface, x, y, old_color)
                        elif rectification == ANTI_LAG:
                            kept = scale_kept(kept, F(1) - anti_lag_drop(batch, surface, validity, old, x, y))
                        elif rectification == REFLECTION:
                            kept = scale_kept(kept, np.array([s["reserved"]], np.uint32).view(F)[0])
                    if kept and finite:
                        total = kept + count
                        blend = F(F(kept) / F(total))
                        color = (color + ((old_color - color).astype(F) * blend).astype(F)).astype(F)
                        m1 = F(m1 + F(F(F(prev["first_moment"]) - m1) * blend))
                        m2 = F(m2 + F(F(F(prev["second_moment"]) - m2) * blend))
                        count = total
                o["first_moment"], o["second_moment"] = m1, m2
            o["color"] = color
            o["sample_count"] = count
            out[y, x] = o
    return out


def settle(prior, still, prior_samples, still_samples, max_samples):
    """temporal.slang:481-493 over whole images: prior and still [..., 4]
    float32 colours, prior_samples the per-pixel span before the cap (the
    history cap where the last changing frame had history, else 0)."""
    prior = np.asarray(prior, F)
    still = np.asarray(still, F)
    span = np.asarray(prior_samples, np.uint64)
    if max_samples:
        span = np.minimum(span, max_samples)
    keep = (still_samples >= span) | ~np.all(np.isfinite(prior), -1)
    t = (F(still_samples) / np.maximum(span, 1).astype(F)).astype(F)
    mixed = (prior + ((still - prior).astype(F) * t[..., None]).astype(F)).astype(F)
    return np.where(keep[..., None], still, mixed).astype(F)


def ref_settle(prior, still, prior_samples, still_samples, max_samples):
    """The reference's SettleTemporalRadiance (temporal.cpp:514-528) through
    oracle/_ref, pixel by pixel."""
    import ctypes as C
    import oracle
    R = oracle.ref()
    p = np.ascontiguousarray(prior, F)
    s = np.ascontiguousarray(still, F)
    n = np.ascontiguousarray(prior_samples, np.uint32)
    out = np.empty_like(s)
    fp, up = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    R.ref_settle(p.ctypes.data_as(fp), s.ctypes.data_as(fp), n.ctypes.data_as(up), n.size, still_samples,
                 max_samples, out.ctypes.data_as(fp))
    return out
