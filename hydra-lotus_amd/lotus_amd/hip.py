"""ctypes binding of liblotus_hip.so (include/lotus_hip.h).

The library is the product; this module only marshals arrays through its
C ABI. Loading fails loudly when the library has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "liblotus_hip.so")

PASS, FAIL, SKIP = 0, 1, 77
MISS = 0xFFFFFFFF
OUTPUT_RADIANCE, OUTPUT_BARYCENTRICS, OUTPUT_SHADING_NORMAL, OUTPUT_ALBEDO, OUTPUT_ROUGHNESS_METALLIC = range(5)
LIGHT_BSDF, LIGHT_MIS, LIGHT_RESTIR = 0, 1, 2
FRAME_SPLIT_KINDS = 1
FRAME_RADIANCE_CACHE = 2
CACHE_PARAMS = 16384
CACHE_LEVELS = 16
# weight matrices [out][in] of the cache MLP, in parameter order
CACHE_SHAPES = [(64, 48), (64, 64), (64, 64), (64, 64), (16, 64)]

SYMBOLS = ["lotus_hip_probe", "lotus_hip_create", "lotus_hip_destroy", "lotus_hip_scene_apply",
           "lotus_hip_render", "lotus_hip_trace", "lotus_hip_trace_device", "lotus_hip_light_table",
           "lotus_hip_read_accumulation", "lotus_hip_set_tuning", "lotus_hip_get_tuning",
           "lotus_hip_cache_configure", "lotus_hip_cache_infer", "lotus_hip_cache_gradients",
           "lotus_hip_cache_train", "lotus_hip_cache_parameters", "lotus_hip_render_temporal", "lotus_hip_svgf",
           "lotus_hip_svgf_static", "lotus_hip_accumulate",
           "lotus_hip_reproject", "lotus_hip_render_batch", "lotus_hip_autotune", "lotus_hip_readback_scene",
           "lotus_hip_write_render_buffer", "lotus_hip_write_ids", "lotus_hip_write_temporal_aov",
           "lotus_hip_write_restir_aov"]
# Named AOVs (lotus_hip_write_temporal_aov, lotus_hip_write_restir_aov).
AOV_MOTION, AOV_HISTORY_LENGTH, AOV_HISTORY_VALIDITY, AOV_HISTORY_REJECTION = range(4)
AOV_RESTIR_LIGHT, AOV_RESTIR_CANDIDATES, AOV_RESTIR_SUPPORT, AOV_RESTIR_AGE, AOV_RESTIR_WEIGHT, \
    AOV_RESTIR_DIRECT = range(6)
BATCH_GRAPH = 1
BATCH_FUSED = 2
# Spatial reuse neighbours (lotus_hip_temporal_settings.restir_spatial_neighbours).
SPATIAL_AXES, SPATIAL_DISK = 0, 1
# History rectification (lotus_hip_temporal_settings.history_rectification).
RECTIFY_NONE, RECTIFY_CLAMP, RECTIFY_ANTI_LAG, RECTIFY_REFLECTION = range(4)
# Render-buffer formats (HdFormat UNorm8Vec4, Float32Vec4, Float32, Int32).
BUFFER_UNORM8X4, BUFFER_FLOAT32X4, BUFFER_FLOAT32, BUFFER_INT32 = range(4)
# What a scene apply did (lotus_hip_scene_evidence.update).
SCENE_UNCHANGED, SCENE_TABLES, SCENE_REFIT, SCENE_BUILD = range(4)
# BVH topology builders (lotus_hip_scene_plan.bvh_builder).
BVH_LBVH, BVH_SAH = 0, 1

# The reference's temporal records (include/lotus/temporal.hpp:193-245).
TEMPORAL_PIXEL = np.dtype([("normal_depth", "<f4", 4), ("motion_depth", "<f4", 4), ("instance_id", "<u8"),
                           ("material_id", "<u8"), ("flags", "<u4"), ("albedo", "<u4"), ("shading_normal", "<u4"),
                           ("reserved", "<u4")])
TEMPORAL_VALIDITY = np.dtype([("valid", "<u4"), ("rejection", "<u4"), ("previous_index", "<u4"),
                              ("reserved", "<u4")])
TEMPORAL_RADIANCE = np.dtype([("color", "<f4", 4), ("sample_count", "<u4"), ("first_moment", "<f4"),
                              ("second_moment", "<f4"), ("variance", "<f4")])

_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)


class ScenePlan(C.Structure):
    _fields_ = [
        ("geometry_count", C.c_uint32),
        ("positions", _f32p),
        ("triangles", _u32p),
        ("geometry_vertex_offset", _u32p),
        ("geometry_vertex_count", _u32p),
        ("geometry_triangle_offset", _u32p),
        ("geometry_triangle_count", _u32p),
        ("instance_count", C.c_uint32),
        ("instance_geometry", _u32p),
        ("instance_transform", _f32p),
        ("instance_material", _u32p),
        ("material_count", C.c_uint32),
        ("materials", _u32p),
        ("environment", C.c_float * 3),
        ("bvh_leaf_size", C.c_uint32),
        ("light_count", C.c_uint32),
        ("lights", C.c_void_p),
        ("environment_map", C.c_void_p),
        ("texture_count", C.c_uint32),
        ("textures", C.c_void_p),
        ("binding_count", C.c_uint32),
        ("bindings", C.c_void_p),
        ("texcoords", C.c_void_p),
        ("geometry_texcoord_offset", C.c_void_p),
        ("instance_texcoord_offset", C.c_void_p),
        ("normals", C.c_void_p),
        ("geometry_normal_offset", C.c_void_p),
        ("colors", C.c_void_p),
        ("instance_color_offset", C.c_void_p),
        ("instance_color_constant", C.c_void_p),
        ("bvh_builder", C.c_uint32),
        ("unchanged", C.c_uint32),
        ("moved_geometry_count", C.c_uint32),
        ("moved_geometries", C.c_void_p),
    ]


# lotus_hip_scene_plan.unchanged bits: what a plan vouches has not changed
UNCHANGED_TOPOLOGY, UNCHANGED_POSITIONS, UNCHANGED_TRANSFORMS, UNCHANGED_MATERIALS, UNCHANGED_LIGHTS = 1, 2, 4, 8, 16


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32), ("texels", C.c_void_p)]


def texture_structs(textures):
    """A C array of texture records for packed textures (None when empty)."""
    if not textures:
        return None
    arr = (Texture * len(textures))()
    for i, t in enumerate(textures):
        arr[i].height, arr[i].width = t["texels"].shape[:2]
        arr[i].format = t["format"]
        arr[i].texels = t["texels"].ctypes.data
    return arr


def set_texture_fields(desc, p):
    """Fills a plan's (or the oracle's) texture lookup fields; returns what
    must stay alive while it is used."""
    tex = texture_structs(p.get("textures"))
    desc.texture_count = len(p.get("textures") or [])
    desc.textures = C.addressof(tex) if tex is not None else None
    b = p.get("bindings")
    desc.binding_count = 0 if b is None else len(b) // 24
    desc.bindings = b.ctypes.data if desc.binding_count else None
    off = p.get("geometry_texcoord_offset")
    ioff = p.get("instance_texcoord_offset")
    desc.texcoords = p["texcoords"].ctypes.data if off is not None or ioff is not None else None
    desc.geometry_texcoord_offset = off.ctypes.data if off is not None else None
    desc.instance_texcoord_offset = ioff.ctypes.data if ioff is not None else None
    noff = p.get("geometry_normal_offset")
    desc.normals = p["normals"].ctypes.data if noff is not None else None
    desc.geometry_normal_offset = noff.ctypes.data if noff is not None else None
    coff = p.get("instance_color_offset")
    desc.colors = p["colors"].ctypes.data if coff is not None else None
    desc.instance_color_offset = coff.ctypes.data if coff is not None else None
    desc.instance_color_constant = p["instance_color_constant"].ctypes.data if coff is not None else None
    return tex


class EnvironmentMap(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("texels", C.c_void_p), ("scale", C.c_float * 3),
                ("world_from_environment", C.c_float * 16), ("format", C.c_uint32)]


def environment_struct(env):
    """The C record of a packed environment image (None: NULL): float texels
    are RGBA32F, 8-bit ones take env["format"] (sRGB unless it says
    otherwise)."""
    if env is None:
        return None
    m = EnvironmentMap()
    m.height, m.width = env["texels"].shape[:2]
    m.texels = env["texels"].ctypes.data
    m.format = 2 if env["texels"].dtype == np.float32 else env.get("format", 1)  # RGBA32F; 8-bit: sRGB unless stated
    for i in range(3):
        m.scale[i] = float(env["scale"][i])
    for i in range(16):
        m.world_from_environment[i] = float(env["transform"][i])
    return m


class SceneEvidence(C.Structure):
    _fields_ = [
        ("triangle_count", C.c_uint32),
        ("instance_count", C.c_uint32),
        ("material_count", C.c_uint32),
        ("light_count", C.c_uint32),
        ("node_count", C.c_uint32),
        ("bvh_depth", C.c_uint32),
        ("leaf_count", C.c_uint32),
        ("bvh_width", C.c_uint32),
        ("upload_ms", C.c_double),
        ("build_gpu_ms", C.c_double),
        ("device_bytes", C.c_uint64),
        ("update", C.c_uint32),
        ("moved_geometries", C.c_uint32),
        ("bvh_builds", C.c_uint64),
        ("bvh_refits", C.c_uint64),
        ("material_writes", C.c_uint64),
        ("light_writes", C.c_uint64),
        ("uploaded_bytes", C.c_uint64),
    ]


class SceneContents(C.Structure):
    _fields_ = [
        ("geometry_count", C.c_uint32),
        ("instance_count", C.c_uint32),
        ("material_count", C.c_uint32),
        ("triangle_count", C.c_uint32),
        ("node_count", C.c_uint32),
        ("light_count", C.c_uint32),
        ("position_floats", C.c_uint64),
        ("index_words", C.c_uint64),
        ("environment", C.c_float * 3),
        ("positions", _f32p),
        ("triangles", _u32p),
        ("geometry_vertex_offset", _u32p),
        ("geometry_vertex_count", _u32p),
        ("geometry_triangle_offset", _u32p),
        ("geometry_triangle_count", _u32p),
        ("instance_geometry", _u32p),
        ("instance_transform", _f32p),
        ("instance_material", _u32p),
        ("materials", _u32p),
        ("bvh_nodes", _u32p),
        ("bvh_triangles", _f32p),
    ]


class Tuning(C.Structure):
    _fields_ = [("camera_dynamic", C.c_uint32), ("trace_dynamic", C.c_uint32), ("tail_divisor", C.c_uint32),
                ("resident_percent", C.c_uint32), ("megakernel", C.c_uint32), ("camera_tile", C.c_uint32)]


class FrameParams(C.Structure):
    _fields_ = [
        ("world_to_clip", C.c_float * 16),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("display_window", C.c_float * 4),
        ("data_window", C.c_int32 * 4),
        ("clear_color", C.c_float * 4),
        ("frame_count", C.c_uint32),
        ("output", C.c_uint32),
        ("light_sampling", C.c_uint32),
        ("sample_index", C.c_uint32),
        ("max_bounces", C.c_uint32),
        ("max_samples", C.c_uint32),
        ("flags", C.c_uint32),
        ("restir_candidates", C.c_uint32),
        ("restir_proposal", C.c_uint32),
    ]


class FrameResult(C.Structure):
    _fields_ = [
        ("color", _f32p),
        ("depth", _f32p),
        ("path_counts", _u32p),
        ("path_counts_capacity", C.c_uint32),
        ("path_counts_size", C.c_uint32),
        ("wavefront_rounds", C.c_uint32),
        ("samples_per_pixel", C.c_uint32),
        ("frames_rendered", C.c_uint32),
        ("camera_rays", C.c_uint64),
        ("extension_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("camera_gpu_ms", C.c_double),
        ("intersect_gpu_ms", C.c_double),
        ("shade_gpu_ms", C.c_double),
        ("shadow_gpu_ms", C.c_double),
        ("trace_gpu_ms", C.c_double),
        ("tail_gpu_ms", C.c_double),
        ("resolve_gpu_ms", C.c_double),
        ("frame_gpu_ms", C.c_double),
        ("traversal_nodes", C.c_uint64 * 4),
        ("traversal_tris", C.c_uint64 * 4),
        ("traversal_rays", C.c_uint64 * 4),
        ("trace_launches", C.c_uint64),
        ("cache_queries", C.c_uint64),
        ("cache_samples", C.c_uint64),
        ("cache_loss", C.c_double),
        ("cache_infer_gpu_ms", C.c_double),
        ("cache_train_gpu_ms", C.c_double),
        ("restir_records", _u32p),
        ("restir_gpu_ms", C.c_double),
        ("cache_training", _f32p),
        ("cache_training_capacity", C.c_uint32),
        ("cache_training_count", C.c_uint32),
        ("temporal_surface_gpu_ms", C.c_double),
        ("svgf_gpu_ms", C.c_double),
        ("batch_graph", C.c_uint32),
    ]


class TemporalSettings(C.Structure):
    _fields_ = [
        ("frame_index", C.c_uint64),
        ("history_key", C.c_uint64),
        ("camera_cut", C.c_uint32),
        ("accumulate_radiance", C.c_uint32),
        ("max_history_samples", C.c_uint32),
        ("filter_iterations", C.c_uint32),
        ("view", C.c_float * 16),
        ("validation", C.c_float * 4),
        ("restir_temporal", C.c_uint32),
        ("restir_spatial", C.c_uint32),
        ("restir_spatial_radius", C.c_uint32),
        ("restir_max_history_candidates", C.c_uint32),
        ("instance_ids", C.POINTER(C.c_uint64)),
        ("material_ids", C.POINTER(C.c_uint64)),
        ("instance_flags", _u32p),
        ("previous_from_current", _f32p),
        ("surface", C.c_void_p),
        ("validity", C.c_void_p),
        ("radiance", C.c_void_p),
        ("reset_reason", C.c_uint32),
        ("history_available", C.c_uint32),
        ("restir_temporal_records", C.c_void_p),
        ("restir_spatial_records", C.c_void_p),
        ("restir_spatial_sources", C.c_void_p),
        ("restir_surfaces", C.c_void_p),
        ("restir_spatial_neighbours", C.c_uint32),
        ("history_rectification", C.c_uint32),
        ("settle", C.c_uint32),
        ("settling", C.c_uint32),
        ("settled_samples", C.c_uint32),
        ("batch_sample_index", C.c_uint32),
        ("batch_samples", C.c_uint32),
        ("batch", _f32p),
    ]


class CacheConfig(C.Structure):
    _fields_ = [
        ("log2_table", C.c_uint32),
        ("min_resolution", C.c_uint32),
        ("max_resolution", C.c_uint32),
        ("batch_capacity", C.c_uint32),
        ("learning_rate", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("epsilon", C.c_float),
        ("seed", C.c_uint32),
        ("query_depth", C.c_uint32),
        ("train_paths", C.c_uint32),
        ("train_steps", C.c_uint32),
        ("resolution", C.c_uint32 * 16),
        ("dense_levels", C.c_uint32),
    ]


CACHE_DEFAULTS = dict(log2_table=19, min_resolution=16, max_resolution=2048, batch_capacity=65536,
                      learning_rate=5e-3, beta1=0.9, beta2=0.99, epsilon=1e-15, seed=1, query_depth=1,
                      train_paths=16384, train_steps=4)


_lib = None


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with hydra-lotus_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.lotus_hip_probe.restype = C.c_int
        L.lotus_hip_probe.argtypes = [C.c_char_p, C.c_size_t]
        L.lotus_hip_create.restype = C.c_void_p
        L.lotus_hip_create.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
        L.lotus_hip_destroy.argtypes = [C.c_void_p]
        L.lotus_hip_scene_apply.restype = C.c_int
        L.lotus_hip_scene_apply.argtypes = [C.c_void_p, C.POINTER(ScenePlan), C.POINTER(SceneEvidence),
                                            C.c_char_p, C.c_size_t]
        L.lotus_hip_render.restype = C.c_int
        L.lotus_hip_render.argtypes = [C.c_void_p, C.POINTER(FrameParams), C.POINTER(FrameResult),
                                       C.c_char_p, C.c_size_t]
        L.lotus_hip_trace.restype = C.c_int
        L.lotus_hip_trace.argtypes = [C.c_void_p, C.c_uint32, _f32p, _u32p, C.c_int, C.c_char_p, C.c_size_t]
        L.lotus_hip_trace_device.restype = C.c_int
        L.lotus_hip_trace_device.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int,
                                             C.c_void_p, C.c_char_p, C.c_size_t]
        L.lotus_hip_readback_scene.restype = C.c_int
        L.lotus_hip_readback_scene.argtypes = [C.c_void_p, C.POINTER(SceneContents), C.c_char_p, C.c_size_t]
        L.lotus_hip_light_table.restype = C.c_int
        L.lotus_hip_light_table.argtypes = [C.c_void_p, _u32p, _u32p, _u32p, C.c_char_p, C.c_size_t]
        L.lotus_hip_set_tuning.restype = C.c_int
        L.lotus_hip_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning), C.c_char_p, C.c_size_t]
        L.lotus_hip_get_tuning.restype = C.c_int
        L.lotus_hip_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.lotus_hip_read_accumulation.restype = C.c_int
        L.lotus_hip_read_accumulation.argtypes = [C.c_void_p, _f32p, _u32p, C.c_char_p, C.c_size_t]
        L.lotus_hip_render_temporal.restype = C.c_int
        L.lotus_hip_render_temporal.argtypes = [C.c_void_p, C.POINTER(FrameParams), C.POINTER(TemporalSettings),
                                                C.POINTER(FrameResult), C.c_char_p, C.c_size_t]
        L.lotus_hip_write_render_buffer.restype = C.c_int
        L.lotus_hip_write_render_buffer.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                                    C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_char_p,
                                                    C.c_size_t]
        L.lotus_hip_write_temporal_aov.restype = C.c_int
        L.lotus_hip_write_temporal_aov.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                   C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32,
                                                   C.c_uint32, C.c_char_p, C.c_size_t]
        L.lotus_hip_write_restir_aov.restype = C.c_int
        L.lotus_hip_write_restir_aov.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                                 C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                                 C.c_char_p, C.c_size_t]
        L.lotus_hip_write_ids.restype = C.c_int
        L.lotus_hip_write_ids.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32,
                                          C.c_char_p, C.c_size_t]
        for name in ("lotus_hip_svgf", "lotus_hip_svgf_static"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_void_p, C.c_char_p, C.c_size_t]
        L.lotus_hip_accumulate.restype = C.c_int
        L.lotus_hip_accumulate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                           C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_char_p,
                                           C.c_size_t]
        L.lotus_hip_reproject.restype = C.c_int
        L.lotus_hip_reproject.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, _f32p,
                                          C.c_void_p, C.c_char_p, C.c_size_t]
        L.lotus_hip_render_batch.restype = C.c_int
        L.lotus_hip_render_batch.argtypes = [C.POINTER(C.c_void_p), C.POINTER(FrameParams), C.POINTER(FrameResult),
                                             C.c_uint32, C.c_uint32, C.POINTER(C.c_double), C.c_char_p, C.c_size_t]
        L.lotus_hip_autotune.restype = C.c_int
        L.lotus_hip_autotune.argtypes = [C.c_void_p, C.POINTER(FrameParams), C.c_uint32, C.POINTER(Tuning),
                                         C.c_char_p, C.c_size_t]
        L.lotus_hip_cache_configure.restype = C.c_int
        L.lotus_hip_cache_configure.argtypes = [C.c_void_p, C.POINTER(CacheConfig), C.c_char_p, C.c_size_t]
        L.lotus_hip_cache_infer.restype = C.c_int
        L.lotus_hip_cache_infer.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_char_p,
                                            C.c_size_t]
        L.lotus_hip_cache_gradients.restype = C.c_int
        L.lotus_hip_cache_gradients.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, _f32p, _f32p, _f32p,
                                                C.c_char_p, C.c_size_t]
        L.lotus_hip_cache_train.restype = C.c_int
        L.lotus_hip_cache_train.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, _f32p, C.c_char_p, C.c_size_t]
        L.lotus_hip_cache_parameters.restype = C.c_int
        L.lotus_hip_cache_parameters.argtypes = [C.c_void_p, _f32p, _f32p, C.c_int, C.c_char_p, C.c_size_t]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class LotusError(RuntimeError):
    def __init__(self, status, detail):
        super().__init__(f"status {status}: {detail}")
        self.status = status
        self.detail = detail


def probe() -> tuple:
    buf = C.create_string_buffer(256)
    s = lib().lotus_hip_probe(buf, 256)
    return s, buf.value.decode()


def frame_params(world_to_clip, width, height, frame_count=1, sample_index=0, max_bounces=8,
                 light_sampling=LIGHT_MIS, output=OUTPUT_RADIANCE, clear_color=(0.0, 0.0, 0.0, 0.0),
                 data_window=(0, 0, 0, 0), display_window=(0.0, 0.0, 0.0, 0.0), max_samples=0, split_kinds=False,
                 cache=False, restir_candidates=8, restir_proposal=0) -> FrameParams:
    fp = FrameParams()
    w2c = np.ascontiguousarray(world_to_clip, np.float32).reshape(16)
    for i in range(16):
        fp.world_to_clip[i] = float(w2c[i])
    fp.width, fp.height = width, height
    for i in range(4):
        fp.display_window[i] = display_window[i]
        fp.data_window[i] = data_window[i]
        fp.clear_color[i] = clear_color[i]
    fp.frame_count = frame_count
    fp.output = output
    fp.light_sampling = light_sampling
    fp.sample_index = sample_index
    fp.max_bounces = max_bounces
    fp.max_samples = max_samples
    fp.flags = (FRAME_SPLIT_KINDS if split_kinds else 0) | (FRAME_RADIANCE_CACHE if cache else 0)
    fp.restir_candidates = restir_candidates
    fp.restir_proposal = restir_proposal
    return fp


class Batch:
    """A batched frame whose parameter and result records are built once:
    run() renders one frame of every view (optionally each at a new sample
    index) and returns the rays the views traced and the batch's GPU ms,
    with little host work, as a serving loop would call it."""

    def __init__(self, renderers, views):
        n = len(renderers)
        self.n = n
        self.handles = (C.c_void_p * n)(*[r.handle for r in renderers])
        self.ps = (FrameParams * n)()
        self.rs = (FrameResult * n)()
        self._counts = []
        for i, v in enumerate(views):
            self.ps[i] = frame_params(**v)
            cap = v.get("max_bounces", 8) + 2
            k = np.zeros(cap, np.uint32)
            self.rs[i].path_counts = _p(k, C.c_uint32)
            self.rs[i].path_counts_capacity = cap
            self._counts.append(k)
        self.ms = C.c_double(0.0)
        self.buf = C.create_string_buffer(512)

    def run(self, graph=False, sample_index=None, fused=False):
        if sample_index is not None:
            for i in range(self.n):
                self.ps[i].sample_index = sample_index
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_FUSED if fused else 0)
        s = lib().lotus_hip_render_batch(self.handles, self.ps, self.rs, self.n, flags, C.byref(self.ms), self.buf,
                                         512)
        if s != PASS:
            raise LotusError(s, self.buf.value.decode())
        rays = sum(r.camera_rays + r.extension_rays + r.shadow_rays for r in self.rs)
        return rays, self.ms.value, self.rs[0].batch_graph


def render_batch(renderers, views, graph=False, want_color=True, fused=False):
    """One batched frame: views[i] (frame_params() keyword dicts) through
    renderers[i]; `fused` runs the views' paths in shared path-megakernel
    launches fed by one device work queue. Returns (per-view results, GPU
    ms)."""
    n = len(renderers)
    handles = (C.c_void_p * n)(*[r.handle for r in renderers])
    ps = (FrameParams * n)()
    rs = (FrameResult * n)()
    colors, counts = [], []
    for i, v in enumerate(views):
        ps[i] = frame_params(**v)
        cap = v.get("max_bounces", 8) + 2
        c = np.zeros((v["height"], v["width"], 4), np.float32) if want_color else None
        k = np.zeros(cap, np.uint32)
        if c is not None:
            rs[i].color = _p(c, C.c_float)
        rs[i].path_counts = _p(k, C.c_uint32)
        rs[i].path_counts_capacity = cap
        colors.append(c)
        counts.append(k)
    ms = C.c_double(0.0)
    buf = C.create_string_buffer(512)
    flags = (BATCH_GRAPH if graph else 0) | (BATCH_FUSED if fused else 0)
    s = lib().lotus_hip_render_batch(handles, ps, rs, n, flags, C.byref(ms), buf, 512)
    if s != PASS:
        raise LotusError(s, buf.value.decode())
    out = []
    for i in range(n):
        out.append(dict(color=colors[i], counts=counts[i][:rs[i].path_counts_size].copy(), spp=rs[i].samples_per_pixel,
                        camera_rays=rs[i].camera_rays, extension_rays=rs[i].extension_rays,
                        shadow_rays=rs[i].shadow_rays, rounds=rs[i].wavefront_rounds,
                        batch_graph=rs[i].batch_graph))
    return out, ms.value


class Renderer:
    """One OffscreenRenderer-like handle over the C ABI."""

    def __init__(self, device: int = 0):
        st = C.c_int(0)
        buf = C.create_string_buffer(256)
        self.handle = lib().lotus_hip_create(device, C.byref(st), buf, 256)
        if not self.handle:
            raise LotusError(st.value, buf.value.decode())
        self.device_detail = buf.value.decode()
        self._scene_keep = None

    def close(self):
        if getattr(self, "handle", None):
            lib().lotus_hip_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def apply(self, scene, leaf_size: int = 0, builder: int = 0, unchanged: int = 0,
              moved_geometries=None) -> SceneEvidence:
        """UpdateScene with a whole plan; `builder` is BVH_LBVH or BVH_SAH;
        `unchanged` (UNCHANGED_* bits) vouches for parts of it, with
        UNCHANGED_POSITIONS `moved_geometries` the geometries whose
        positions may have changed."""
        p = scene.packed()
        plan = ScenePlan()
        plan.geometry_count = p["geometry_count"]
        plan.positions = _p(p["positions"], C.c_float)
        plan.triangles = _p(p["triangles"], C.c_uint32)
        plan.geometry_vertex_offset = _p(p["geometry_vertex_offset"], C.c_uint32)
        plan.geometry_vertex_count = _p(p["geometry_vertex_count"], C.c_uint32)
        plan.geometry_triangle_offset = _p(p["geometry_triangle_offset"], C.c_uint32)
        plan.geometry_triangle_count = _p(p["geometry_triangle_count"], C.c_uint32)
        plan.instance_count = p["instance_count"]
        plan.instance_geometry = _p(p["instance_geometry"], C.c_uint32)
        plan.instance_transform = _p(p["instance_transform"], C.c_float)
        plan.instance_material = _p(p["instance_material"], C.c_uint32)
        plan.material_count = p["material_count"]
        plan.materials = _p(p["materials"], C.c_uint32)
        for i in range(3):
            plan.environment[i] = float(p["environment"][i])
        plan.bvh_leaf_size = leaf_size
        plan.bvh_builder = builder
        plan.unchanged = unchanged
        moved = np.ascontiguousarray(moved_geometries if moved_geometries is not None else [], np.uint32)
        plan.moved_geometry_count = len(moved)
        plan.moved_geometries = moved.ctypes.data if len(moved) else None
        plan.light_count = len(p["lights"])
        plan.lights = p["lights"].ctypes.data if len(p["lights"]) else None
        env = environment_struct(p.get("environment_map"))
        plan.environment_map = C.addressof(env) if env is not None else None
        keep = set_texture_fields(plan, p)  # noqa: F841 (alive through the call)
        ev = SceneEvidence()
        buf = C.create_string_buffer(512)
        s = lib().lotus_hip_scene_apply(self.handle, C.byref(plan), C.byref(ev), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        self._scene_keep = p
        self.instance_count = p["instance_count"]
        return ev

    def render(self, world_to_clip, width, height, frame_count=1, sample_index=0, max_bounces=8,
               light_sampling=LIGHT_MIS, output=OUTPUT_RADIANCE, clear_color=(0.0, 0.0, 0.0, 0.0),
               data_window=(0, 0, 0, 0), display_window=(0.0, 0.0, 0.0, 0.0), max_samples=0,
               want_color=True, want_depth=False, count_capacity=None, split_kinds=False, cache=False,
               restir_candidates=8, restir_proposal=0, want_restir=False, temporal=None, want_training=0):
        """One RenderScene call; with `temporal` (a TemporalSettings), one
        RenderTemporalScene frame."""
        fp = frame_params(world_to_clip, width, height, frame_count, sample_index, max_bounces, light_sampling,
                          output, clear_color, data_window, display_window, max_samples, split_kinds, cache,
                          restir_candidates, restir_proposal)
        res = FrameResult()
        color = np.zeros((height, width, 4), np.float32) if want_color else None
        depth = np.zeros((height, width), np.float32) if want_depth else None
        cap = count_capacity or (max_bounces + 2)
        counts = np.zeros(cap, np.uint32)
        if color is not None:
            res.color = _p(color, C.c_float)
        if depth is not None:
            res.depth = _p(depth, C.c_float)
        res.path_counts = _p(counts, C.c_uint32)
        restir = np.zeros((height, width, 12), np.uint32) if want_restir else None
        if restir is not None:
            res.restir_records = _p(restir, C.c_uint32)
        res.path_counts_capacity = cap
        training = np.zeros((want_training, 20), np.float32) if want_training else None
        if training is not None:
            res.cache_training = _p(training, C.c_float)
            res.cache_training_capacity = want_training
        buf = C.create_string_buffer(512)
        if temporal is None:
            s = lib().lotus_hip_render(self.handle, C.byref(fp), C.byref(res), buf, 512)
        else:
            s = lib().lotus_hip_render_temporal(self.handle, C.byref(fp), C.byref(temporal), C.byref(res), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        return dict(color=color, depth=depth, counts=counts[:res.path_counts_size].copy(),
                    rounds=res.wavefront_rounds, spp=res.samples_per_pixel, frames=res.frames_rendered,
                    camera_rays=res.camera_rays, extension_rays=res.extension_rays,
                    shadow_rays=res.shadow_rays, camera_ms=res.camera_gpu_ms,
                    intersect_ms=res.intersect_gpu_ms, shade_ms=res.shade_gpu_ms,
                    shadow_ms=res.shadow_gpu_ms, trace_ms=res.trace_gpu_ms, tail_ms=res.tail_gpu_ms,
                    resolve_ms=res.resolve_gpu_ms, frame_ms=res.frame_gpu_ms,
                    nodes=list(res.traversal_nodes), tris=list(res.traversal_tris),
                    walked=list(res.traversal_rays), trace_launches=res.trace_launches,
                    cache_queries=res.cache_queries, cache_samples=res.cache_samples, cache_loss=res.cache_loss,
                    cache_infer_ms=res.cache_infer_gpu_ms, cache_train_ms=res.cache_train_gpu_ms,
                    restir=restir, restir_ms=res.restir_gpu_ms,
                    temporal_surface_ms=res.temporal_surface_gpu_ms, svgf_ms=res.svgf_gpu_ms,
                    training=None if training is None else training[:res.cache_training_count].copy())

    def trace(self, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros(n if any_hit else (n, 5), np.uint32)
        buf = C.create_string_buffer(512)
        s = lib().lotus_hip_trace(self.handle, n, _p(rays, C.c_float), _p(hits, C.c_uint32),
                                  int(any_hit), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        return hits

    def transmittance(self, rays: np.ndarray) -> np.ndarray:
        """ShadowTransmittance per ray (any_hit 2)."""
        return self.trace(rays, any_hit=2).view(np.float32)

    def trace_device(self, n: int, rays_ptr: int, hits_ptr: int, any_hit: bool, stream_ptr: int = 0):
        buf = C.create_string_buffer(512)
        s = lib().lotus_hip_trace_device(self.handle, n, C.c_void_p(rays_ptr), C.c_void_p(hits_ptr),
                                         int(any_hit), C.c_void_p(stream_ptr), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def readback_scene(self) -> dict:
        """ReadBackScene: the resident scene copied from the device."""
        c = SceneContents()
        buf = C.create_string_buffer(512)
        s = lib().lotus_hip_readback_scene(self.handle, C.byref(c), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        G, I, M = c.geometry_count, c.instance_count, c.material_count
        out = dict(
            positions=np.zeros(c.position_floats, np.float32), triangles=np.zeros(c.index_words, np.uint32),
            geometry_vertex_offset=np.zeros(G, np.uint32), geometry_vertex_count=np.zeros(G, np.uint32),
            geometry_triangle_offset=np.zeros(G, np.uint32), geometry_triangle_count=np.zeros(G, np.uint32),
            instance_geometry=np.zeros(I, np.uint32), instance_transform=np.zeros((I, 16), np.float32),
            instance_material=np.zeros(I, np.uint32), materials=np.zeros((M, 36), np.uint32),
            bvh_nodes=np.zeros((c.node_count, 32), np.uint32),
            bvh_triangles=np.zeros((c.triangle_count, 12), np.float32))
        for name, a in out.items():
            t = C.c_float if a.dtype == np.float32 else C.c_uint32
            setattr(c, name, _p(a, t) if a.size else None)
        s = lib().lotus_hip_readback_scene(self.handle, C.byref(c), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        out["environment"] = np.array(list(c.environment), np.float32)
        out["light_count"] = c.light_count
        return out

    def light_table(self):
        count = np.zeros(1, np.uint32)
        buf = C.create_string_buffer(256)
        lib().lotus_hip_light_table(self.handle, None, None, _p(count, C.c_uint32), buf, 256)
        offsets = np.zeros(self.instance_count, np.uint32)
        records = np.zeros((max(int(count[0]), 1), 4), np.uint32)
        s = lib().lotus_hip_light_table(self.handle, _p(offsets, C.c_uint32), _p(records, C.c_uint32),
                                        _p(count, C.c_uint32), buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        return offsets, records[:int(count[0])]

    def tuning(self) -> dict:
        t = Tuning()
        lib().lotus_hip_get_tuning(self.handle, C.byref(t))
        return {"camera_dynamic": t.camera_dynamic, "trace_dynamic": t.trace_dynamic,
                "tail_divisor": t.tail_divisor, "resident_percent": t.resident_percent,
                "megakernel": t.megakernel, "camera_tile": t.camera_tile}

    def set_tuning(self, **kw):
        t = Tuning()
        lib().lotus_hip_get_tuning(self.handle, C.byref(t))
        for k, v in kw.items():
            setattr(t, k, int(v))
        buf = C.create_string_buffer(256)
        s = lib().lotus_hip_set_tuning(self.handle, C.byref(t), buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def accumulation(self, width, height):
        sums = np.zeros((height, width, 4), np.float32)
        n = np.zeros(1, np.uint32)
        buf = C.create_string_buffer(256)
        s = lib().lotus_hip_read_accumulation(self.handle, _p(sums, C.c_float), _p(n, C.c_uint32), buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        return sums, int(n[0])

    # radiance cache ---------------------------------------------------------
    def cache_configure(self, **kw) -> dict:
        """Creates the cache; returns the configuration as applied (with each
        level's resolution and the dense-level mask)."""
        cfg = CacheConfig()
        opts = dict(CACHE_DEFAULTS)
        opts.update(kw)
        for k, v in opts.items():
            setattr(cfg, k, v)
        buf = C.create_string_buffer(512)
        s = lib().lotus_hip_cache_configure(self.handle, C.byref(cfg), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        out = dict(opts)
        out["resolution"] = [int(x) for x in cfg.resolution]
        out["dense_levels"] = int(cfg.dense_levels)
        self.cache_config = out
        return out

    def cache_infer(self, queries_ptr: int, n: int, out_ptr: int):
        """Device buffers: queries [n][16] float32 in, radiance [n][3] float32 out."""
        buf = C.create_string_buffer(512)
        s = lib().lotus_hip_cache_infer(self.handle, C.c_void_p(queries_ptr), n, C.c_void_p(out_ptr), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def cache_gradients(self, samples_ptr: int, n: int, want_grid: bool = True):
        """Mean loss gradients over device samples [n][20]: (weights, grid, loss)."""
        wg = np.zeros(CACHE_PARAMS, np.float32)
        gg = np.zeros(self._grid_shape(), np.float32) if want_grid else None
        loss = np.zeros(1, np.float32)
        buf = C.create_string_buffer(512)
        s = lib().lotus_hip_cache_gradients(self.handle, C.c_void_p(samples_ptr), n, _p(wg, C.c_float),
                                            _p(gg, C.c_float) if gg is not None else None,
                                            _p(loss, C.c_float), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        return wg, gg, float(loss[0])

    def cache_train(self, samples_ptr: int, n: int) -> float:
        loss = np.zeros(1, np.float32)
        buf = C.create_string_buffer(512)
        s = lib().lotus_hip_cache_train(self.handle, C.c_void_p(samples_ptr), n, _p(loss, C.c_float), buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        return float(loss[0])

    def _grid_shape(self):
        return (CACHE_LEVELS, 1 << self.cache_config["log2_table"], 2)

    def cache_parameters(self, weights=None, grid=None):
        """With arrays: writes them (float32, [16384] and [16][T][2]); without:
        returns the current (weights, grid)."""
        buf = C.create_string_buffer(512)
        if weights is None and grid is None:
            w = np.zeros(CACHE_PARAMS, np.float32)
            g = np.zeros(self._grid_shape(), np.float32)
            s = lib().lotus_hip_cache_parameters(self.handle, _p(w, C.c_float), _p(g, C.c_float), 0, buf, 512)
            if s != PASS:
                raise LotusError(s, buf.value.decode())
            return w, g
        w = None if weights is None else np.ascontiguousarray(weights, np.float32).reshape(CACHE_PARAMS)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32).reshape(self._grid_shape())
        s = lib().lotus_hip_cache_parameters(self.handle, _p(w, C.c_float) if w is not None else None,
                                             _p(g, C.c_float) if g is not None else None, 1, buf, 512)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    # temporal frames ------------------------------------------------------
    def render_temporal(self, world_to_clip, view, width, height, frame_index, sample_index=0,
                        accumulate=False, max_history=32, filter_iterations=0, restir_temporal=False,
                        restir_spatial=False, radius=30, max_history_candidates=20, history_key=0,
                        camera_cut=False, want_products=False, spatial_disk=False, rectification=0, settle=False,
                        instance_ids=None, material_ids=None, instance_flags=None, previous_from_current=None,
                        **kw):
        """One temporal frame; returns render()'s dict plus the reset reason,
        history availability, whether it settled and its still samples, the
        first index and count of the samples it added and, with
        want_products, the surface, validity and radiance records and the
        batch (the resolved colour before accumulation). `sample_index` is
        the stream's base: fresh batches continue from it by themselves."""
        t = TemporalSettings()
        t.settle = int(settle)
        t.frame_index, t.history_key, t.camera_cut = frame_index, history_key, int(camera_cut)
        t.accumulate_radiance, t.max_history_samples, t.filter_iterations = int(accumulate), max_history, \
            filter_iterations
        v = np.ascontiguousarray(view, np.float32).reshape(16)
        for i in range(16):
            t.view[i] = float(v[i])
        t.restir_temporal, t.restir_spatial = int(restir_temporal), int(restir_spatial)
        t.restir_spatial_radius, t.restir_max_history_candidates = radius, max_history_candidates
        t.restir_spatial_neighbours = SPATIAL_DISK if spatial_disk else SPATIAL_AXES
        t.history_rectification = rectification
        # per-instance identities: host arrays the call copies to the device
        ident = {}
        if instance_ids is not None:
            ident["i"] = np.ascontiguousarray(instance_ids, np.uint64)
            t.instance_ids = ident["i"].ctypes.data_as(C.POINTER(C.c_uint64))
        if material_ids is not None:
            ident["m"] = np.ascontiguousarray(material_ids, np.uint64)
            t.material_ids = ident["m"].ctypes.data_as(C.POINTER(C.c_uint64))
        if instance_flags is not None:
            ident["f"] = np.ascontiguousarray(instance_flags, np.uint32)
            t.instance_flags = _p(ident["f"], C.c_uint32)
        if previous_from_current is not None:
            ident["x"] = np.ascontiguousarray(previous_from_current, np.float32).reshape(-1)
            t.previous_from_current = _p(ident["x"], C.c_float)
        prod = {}
        if want_products:
            prod = dict(surface=np.zeros((height, width), TEMPORAL_PIXEL),
                        validity=np.zeros((height, width), TEMPORAL_VALIDITY),
                        radiance=np.zeros((height, width), TEMPORAL_RADIANCE),
                        restir_temporal=np.zeros((height, width, 16), np.uint32),
                        restir_spatial=np.zeros((height, width, 16), np.uint32),
                        restir_sources=np.zeros((height, width, 8 if spatial_disk else 4), np.uint32),
                        restir_surfaces=np.zeros((height, width, 40), np.float32))
            t.surface = prod["surface"].ctypes.data
            t.validity = prod["validity"].ctypes.data
            t.radiance = prod["radiance"].ctypes.data
            t.restir_temporal_records = prod["restir_temporal"].ctypes.data
            t.restir_spatial_records = prod["restir_spatial"].ctypes.data
            t.restir_spatial_sources = prod["restir_sources"].ctypes.data
            t.restir_surfaces = prod["restir_surfaces"].ctypes.data
            prod["batch"] = np.zeros((height, width, 4), np.float32)
            t.batch = _p(prod["batch"], C.c_float)
        out = self.render(world_to_clip, width, height, sample_index=sample_index, temporal=t, **kw)
        out.update(prod)
        out["reset_reason"] = t.reset_reason
        out["history_available"] = t.history_available
        out["settling"] = bool(t.settling)
        out["settled_samples"] = t.settled_samples
        out["batch_sample_index"] = t.batch_sample_index
        out["batch_samples"] = t.batch_samples
        return out

    def svgf(self, surface_ptr, radiance_ptr, width, height, batch_samples, max_history, iterations, color_ptr=0,
             static=False):
        """SVGF over device buffers; static=True takes lotus_hip_svgf_static,
        which allocates nothing."""
        buf = C.create_string_buffer(256)
        fn = lib().lotus_hip_svgf_static if static else lib().lotus_hip_svgf
        s = fn(self.handle, C.c_void_p(surface_ptr), C.c_void_p(radiance_ptr), width, height,
               batch_samples, max_history, iterations, C.c_void_p(color_ptr), buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def accumulate(self, batch_ptr, surface_ptr, validity_ptr, old_ptr, width, height, fresh_samples, max_history,
                   rectification, radiance_ptr):
        """AccumulateTemporalRadiance over device buffers (RECTIFY_* modes)."""
        buf = C.create_string_buffer(256)
        s = lib().lotus_hip_accumulate(self.handle, C.c_void_p(batch_ptr), C.c_void_p(surface_ptr),
                                       C.c_void_p(validity_ptr), C.c_void_p(old_ptr), width, height, fresh_samples,
                                       max_history, rectification, C.c_void_p(radiance_ptr), buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def write_render_buffer(self, product_ptr, product_format, width, height, buffer_ptr, buffer_format,
                            buffer_width=None, buffer_height=None):
        """A top-down product into a bottom-up render buffer, both device pointers
        (BUFFER_* formats; the colour formats convert into each other)."""
        buf = C.create_string_buffer(256)
        s = lib().lotus_hip_write_render_buffer(self.handle, C.c_void_p(product_ptr), product_format, width, height,
                                                C.c_void_p(buffer_ptr), buffer_format,
                                                width if buffer_width is None else buffer_width,
                                                height if buffer_height is None else buffer_height, buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def write_ids(self, buffer_ptr, buffer_format, width, height, value):
        buf = C.create_string_buffer(256)
        s = lib().lotus_hip_write_ids(self.handle, C.c_void_p(buffer_ptr), buffer_format, width, height, value, buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def write_temporal_aov(self, aov, surface_ptr, validity_ptr, radiance_ptr, width, height, buffer_ptr,
                           buffer_format, buffer_width=None, buffer_height=None):
        """A named temporal AOV from device records into a bottom-up render buffer."""
        buf = C.create_string_buffer(256)
        s = lib().lotus_hip_write_temporal_aov(self.handle, aov, C.c_void_p(surface_ptr or None),
                                               C.c_void_p(validity_ptr or None), C.c_void_p(radiance_ptr or None),
                                               width, height, C.c_void_p(buffer_ptr), buffer_format,
                                               width if buffer_width is None else buffer_width,
                                               height if buffer_height is None else buffer_height, buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def write_restir_aov(self, aov, initial_ptr, resolved_ptr, width, height, buffer_ptr, buffer_format,
                         buffer_width=None, buffer_height=None):
        """A named ReSTIR AOV from device records into a bottom-up render buffer."""
        buf = C.create_string_buffer(256)
        s = lib().lotus_hip_write_restir_aov(self.handle, aov, C.c_void_p(initial_ptr or None),
                                             C.c_void_p(resolved_ptr or None), width, height, C.c_void_p(buffer_ptr),
                                             buffer_format, width if buffer_width is None else buffer_width,
                                             height if buffer_height is None else buffer_height, buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def reproject(self, current_ptr, previous_ptr, width, height, validity_ptr, validation=None):
        buf = C.create_string_buffer(256)
        v = None if validation is None else np.ascontiguousarray(validation, np.float32)
        s = lib().lotus_hip_reproject(self.handle, C.c_void_p(current_ptr), C.c_void_p(previous_ptr or None),
                                      width, height, _p(v, C.c_float) if v is not None else None,
                                      C.c_void_p(validity_ptr), buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())

    def autotune(self, trials=2, **view) -> dict:
        """Times each traversal driver and tail rule choice on `view` and keeps the fastest."""
        fp = frame_params(**view)
        t = Tuning()
        buf = C.create_string_buffer(256)
        s = lib().lotus_hip_autotune(self.handle, C.byref(fp), trials, C.byref(t), buf, 256)
        if s != PASS:
            raise LotusError(s, buf.value.decode())
        return {"camera_dynamic": t.camera_dynamic, "trace_dynamic": t.trace_dynamic,
                "tail_divisor": t.tail_divisor, "resident_percent": t.resident_percent,
                "megakernel": t.megakernel, "camera_tile": t.camera_tile}
