"""Shared helpers of the ReSTIR DI reuse tests: the oracle's temporal and
spatial passes (restir.slang:302-491) over a whole target, and the
closed-form re-derivation of validation/restir-gpu-test.cpp:124-206 and
435-515 on the parallel-light plane, through the reference's own
reservoir.cpp (oracle/_ref)."""
import ctypes as C

import numpy as np

import oracle

NO = 0xFFFFFFFF
TEMPORAL_SEED = 0x94D049BB  # restir.slang:350
SPATIAL_SEED = 0x369DEA0F   # restir.slang:437


def f32(words):
    return np.ascontiguousarray(words, np.uint32).view(np.float32)


def seed_random(pixel, sample):
    return oracle.lib().oracle_seed_random(pixel, sample)


def _next(state):
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return float(np.float32(((word >> 22) ^ word) >> 8) * np.float32(1.0 / 16777216.0))


class Rng:
    """sampling.slang:9-25: seeded by SeedRandom, stepped by the LCG."""

    def __init__(self, pixel, sample):
        self.state = seed_random(pixel, sample)

    def next(self):
        self.state = (self.state * 747796405 + 2891336453) & 0xFFFFFFFF
        return _next(self.state)


def oracle_temporal_frame(o, fr, sample, view, records_prev, surfaces_prev, validity, centre, history, cap):
    """The oracle's temporal pass over every pixel: records [h, w, 16] and
    receiver surfaces [h, w, 32]. `validity` holds (valid, previous index)
    per pixel and `centre` the pixel-centre surface's (instance id, material
    id), both the renderer's products."""
    h, w = fr.height, fr.width
    rec = np.zeros((h, w, 16), np.uint32)
    surf = np.zeros((h, w, 40), np.float32)
    hits = np.zeros((h, w), bool)
    for y in range(h):
        for x in range(w):
            valid, pi = (int(validity[y, x, 0]), int(validity[y, x, 1])) if history else (0, NO)
            use = valid != 0 and pi < w * h and records_prev is not None
            prev_r = records_prev.reshape(-1, 16)[pi] if use else None
            prev_s = surfaces_prev.reshape(-1, 40)[pi] if use else None
            r, s, hit = o.restir_temporal(fr, x, y, sample, view, use, pi, centre[y, x], prev_r, prev_s, cap)
            rec[y, x], surf[y, x], hits[y, x] = r, s, hit
    return rec, surf, hits


def oracle_spatial_frame(o, w, h, sample, radius, cap, records, surfaces, disk=False):
    """The oracle's spatial pass over every pixel of the temporal pass's
    records and surfaces: records [h, w, 16] and sources [h, w, 5] (the four
    axis neighbours and an unused fifth, or the disk mode's five)."""
    records = np.ascontiguousarray(records, np.uint32).reshape(-1, 16)
    surfaces = np.ascontiguousarray(surfaces, np.float32).reshape(-1, 40)
    rec = np.zeros((h, w, 16), np.uint32)
    src = np.zeros((h, w, 5), np.uint32)
    for y in range(h):
        for x in range(w):
            rec[y, x], src[y, x] = o.restir_spatial(w, h, x, y, sample, radius, cap, records, surfaces, disk)
    return rec, src


def ref_merge(destination, source_words, normalization, target, cap, random):
    """reservoir.cpp's MergeReevaluatedReservoir (in double) on 4-word
    reservoirs; returns whether the source was selected (None: rejected)."""
    R = oracle.ref()
    d = np.ascontiguousarray(destination, np.uint32)
    s = np.ascontiguousarray(source_words[:4], np.uint32)
    r = R.ref_merge_reevaluated(d.ctypes.data_as(C.POINTER(C.c_uint32)), s.ctypes.data_as(C.POINTER(C.c_uint32)),
                                float(normalization), float(np.float32(target)), cap, float(random))
    destination[:] = d
    return None if r < 0 else bool(r)


def ref_reuse_normalization(reservoir, supporting):
    R = oracle.ref()
    d = np.ascontiguousarray(reservoir[:4], np.uint32)
    return R.ref_reuse_normalization(d.ctypes.data_as(C.POINTER(C.c_uint32)), supporting)


def ref_normalization(reservoir):
    R = oracle.ref()
    d = np.ascontiguousarray(reservoir[:4], np.uint32)
    return R.ref_reservoir_normalization(d.ctypes.data_as(C.POINTER(C.c_uint32)))


def compare_records(got, want, what, rtol=4e-6):
    """Discrete fields exact (sample, M, u, v, proposal bits, flags,
    supporting candidates, previous index, age); weight sum, target,
    normalization and contribution within rtol. Returns the share of records
    equal bit for bit."""
    got = np.ascontiguousarray(got, np.uint32).reshape(-1, 16)
    want = np.ascontiguousarray(want, np.uint32).reshape(-1, 16)
    for k, name in ((0, "sample"), (3, "M"), (4, "u"), (5, "v"), (6, "proposal"), (7, "flags"),
                    (13, "supporting"), (14, "previous"), (15, "age")):
        bad = np.flatnonzero(got[:, k] != want[:, k])
        assert bad.size == 0, f"{what}: {name} differs at {bad[:8]} ({got[bad[:4], k]} vs {want[bad[:4], k]})"
    for k, name in ((1, "weight sum"), (2, "target"), (8, "r"), (9, "g"), (10, "b"), (11, "visibility"),
                    (12, "normalization")):
        a, b = f32(got[:, k]).astype(np.float64), f32(want[:, k]).astype(np.float64)
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
        err[(a == b)] = 0
        assert err.max() <= rtol, f"{what}: {name} off by {err.max():.3g} at {np.argmax(err)}"
    return float(np.mean(np.all(got == want, axis=1)))
