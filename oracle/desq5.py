"""DES-Q5 statistical image comparison and PFM reading — TEST INFRASTRUCTURE.

Restates adapters/common/reference.cpp:286-379 (Compare) and :104-130
(ReadPfm); pinned against the reference's own compiled Compare in
tests/test_oracle_transport.py when oracle/_ref is built.
"""
from __future__ import annotations

import dataclasses

import numpy as np

MATCH_Z = 5.0  # adapters/common/reference.hpp:80


def read_pfm(path: str) -> np.ndarray:
    """A little-endian RGB portable float map as float32 [h, w, 3], rows from
    the top."""
    with open(path, "rb") as f:
        data = f.read()
    parts = data.split(maxsplit=4)
    if parts[0] != b"PF":
        raise ValueError(f"{path} is not an RGB portable float map")
    w, h, scale = int(parts[1]), int(parts[2]), float(parts[3])
    if scale > 0:
        raise ValueError(f"{path} is big-endian")
    header = data.index(parts[3]) + len(parts[3]) + 1
    img = np.frombuffer(data, "<f4", count=w * h * 3, offset=header).reshape(h, w, 3)
    return np.ascontiguousarray(img[::-1]).astype(np.float32)


@dataclasses.dataclass
class Comparison:
    samples: int
    tile: int
    image_z: np.ndarray
    max_tile_z: float
    mean_tile_z2: float
    failure: str


def _z(diff, var, mag, n):
    floor = 1e-6 + 1e-5 * mag / n
    return diff / n / np.sqrt(var / (n * n) + floor * floor)


def compare(image: np.ndarray, samples: int, mean: np.ndarray, variance: np.ndarray,
            reference_samples: int) -> Comparison:
    """image, mean, variance: float32 [h, w, 3]."""
    h, w, _ = mean.shape
    if not np.all(np.isfinite(image)):
        bad = int(np.argmax(~np.isfinite(image).reshape(-1))) // 3
        return Comparison(samples, 0, np.zeros(3), 0.0, 0.0, f"a non-finite value at pixel {bad}")
    weight = 1.0 / samples + 1.0 / reference_samples
    diff = image.astype(np.float64) - mean.astype(np.float64)
    var = variance.astype(np.float64) * weight
    mag = np.abs(mean.astype(np.float64))
    failure = ""
    image_z = np.zeros(3)
    n = float(h * w)
    for c in range(3):
        image_z[c] = _z(diff[..., c].sum(), var[..., c].sum(), mag[..., c].sum(), n)
        if not abs(image_z[c]) <= MATCH_Z and not failure:
            failure = f"{samples} spp: the image, channel {c}: {image_z[c]:.4f} standard errors"
    tile = 8
    while tile * tile * samples < 1024 and tile < max(w, h):
        tile *= 2
    z2, tests, max_z = 0.0, 0, 0.0
    for y0 in range(0, h, tile):
        for x0 in range(0, w, tile):
            sl = (slice(y0, min(h, y0 + tile)), slice(x0, min(w, x0 + tile)))
            npx = float(diff[sl][..., 0].size)
            for c in range(3):
                z = _z(diff[sl][..., c].sum(), var[sl][..., c].sum(), mag[sl][..., c].sum(), npx)
                z2 += z * z
                tests += 1
                max_z = max(max_z, abs(z))
                if not abs(z) <= MATCH_Z and not failure:
                    failure = f"{samples} spp: the tile at {x0},{y0}, channel {c}: {z:.4f} standard errors"
    return Comparison(samples, tile, image_z, max_z, z2 / tests, failure)
