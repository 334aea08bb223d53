"""TEST INFRASTRUCTURE ONLY: CPU restatement of the hdLotus render buffer's
named temporal and ReSTIR AOV writes, HdLotusRenderBuffer::WriteTemporal
(adapters/hydra2/src/adapter.cpp:2598-2630) and WriteRestir (:2632-2672),
each followed by WriteRowsFlippedLocked's bottom-up row order (:2713-2731).
Records are the reference's layouts (include/lotus/temporal.hpp:193-234,
include/lotus/vulkan_backend.hpp:152-172) as numpy structured dtypes."""
import numpy as np

F32 = np.float32

TEMPORAL_PIXEL = np.dtype([("normal_depth", "<f4", 4), ("motion_depth", "<f4", 4), ("instance_id", "<u8"),
                           ("material_id", "<u8"), ("flags", "<u4"), ("albedo", "<u4"), ("shading_normal", "<u4"),
                           ("reserved", "<u4")])
VALIDITY_PIXEL = np.dtype([("valid", "<u4"), ("rejection", "<u4"), ("previous_index", "<u4"), ("reserved", "<u4")])
RADIANCE_PIXEL = np.dtype([("color", "<f4", 4), ("sample_count", "<u4"), ("moments", "<f4", 2), ("variance", "<f4")])
RESTIR_PIXEL = np.dtype([("light", "<u4"), ("weight_sum", "<f4"), ("selected_target", "<f4"),
                         ("candidate_count", "<u4"), ("uv", "<f4", 2), ("proposal", "<f4"), ("flags", "<u4"),
                         ("contribution", "<f4", 4)])
RESTIR_REUSE_PIXEL = np.dtype([("reservoir", RESTIR_PIXEL), ("normalization", "<f4"),
                               ("supporting_candidates", "<u4"), ("previous_index", "<u4"), ("age", "<u4")])
assert (TEMPORAL_PIXEL.itemsize, VALIDITY_PIXEL.itemsize, RADIANCE_PIXEL.itemsize, RESTIR_PIXEL.itemsize,
        RESTIR_REUSE_PIXEL.itemsize) == (64, 16, 32, 48, 64)

TEMPORAL_AOVS = ("lotus:motion", "lotus:historyLength", "lotus:historyValidity", "lotus:historyRejection")
RESTIR_AOVS = ("lotus:restirLight", "lotus:restirCandidates", "lotus:restirSupport", "lotus:restirAge",
               "lotus:restirWeight", "lotus:restirDirect")


def temporal_aov(name, current, validity, radiance=None):
    """WriteTemporal of `name` over [height, width] records: the bottom-up
    buffer, float32 [h, w, 4] for lotus:motion, else int32 [h, w]."""
    if name == "lotus:motion":
        values = current["motion_depth"].astype(F32)
    else:
        if name == "lotus:historyLength":
            values = (radiance["sample_count"] if radiance is not None
                      else np.zeros(current.shape, np.uint32))
        elif name == "lotus:historyValidity":
            values = validity["valid"]
        else:
            values = validity["rejection"]
        values = values.astype(np.uint32).view(np.int32)
    return np.ascontiguousarray(np.flipud(values))


def restir_aov(name, initial=None, resolved=None, shape=None):
    """WriteRestir of `name`: per pixel the resolved reuse record (spatial,
    else temporal), else the initial record, else an empty RestirPixel.
    float32 [h, w, 4] for lotus:restirWeight and lotus:restirDirect, else
    int32 [h, w]; bottom-up."""
    if resolved is not None:
        r = resolved["reservoir"]
        shape = resolved.shape
    elif initial is not None:
        r = initial
        shape = initial.shape
    else:
        r = np.zeros(shape, RESTIR_PIXEL)
        r["light"] = 0xFFFFFFFF
    m = r["candidate_count"]
    if name == "lotus:restirDirect":
        values = r["contribution"].astype(F32)
    elif name == "lotus:restirWeight":
        if resolved is not None:
            norm = resolved["normalization"].astype(F32)
        else:
            target = r["selected_target"].astype(F32)
            ok = (m != 0) & (target > 0)
            with np.errstate(divide="ignore", invalid="ignore"):
                q = (r["weight_sum"].astype(F32) / m.astype(F32)).astype(F32) / target
            norm = np.where(ok, q, F32(0)).astype(F32)
        values = np.stack([r["weight_sum"], r["selected_target"], norm, r["proposal"]], -1).astype(F32)
    else:
        if name == "lotus:restirLight":
            values = r["light"]
        elif name == "lotus:restirCandidates":
            values = m
        elif name == "lotus:restirSupport":
            values = resolved["supporting_candidates"] if resolved is not None else m
        else:
            values = resolved["age"] if resolved is not None else np.zeros(shape, np.uint32)
        values = values.astype(np.uint32).view(np.int32)
    return np.ascontiguousarray(np.flipud(values))
