"""CPU oracle for the wavefront path tracer — TEST INFRASTRUCTURE ONLY.

ctypes bindings over ``liblotus_oracle.so`` (a C++ restatement of
hydra-lotus's reference transport, see ``lotus_oracle.cpp``) and over the
reference's own std-only CPU functions built into ``_ref/libref.so`` when the
reference sources were present at build time. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import
this package, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblotus_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")
REFERENCE_ROOT = "/root/reference"

_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)


def build(reference: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if reference and os.path.isdir(REFERENCE_ROOT):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
        if os.path.exists(os.path.join(os.path.dirname(HERE), "hydra-lotus_amd", "liblotus_hip.so")):
            subprocess.run(["make", "-s", "-C", HERE, "facade"], check=True)


class SceneDesc(C.Structure):
    _fields_ = [
        ("geometry_count", C.c_uint32),
        ("positions", _f32p),
        ("triangles", _u32p),
        ("geometry_vertex_offset", _u32p),
        ("geometry_triangle_offset", _u32p),
        ("geometry_triangle_count", _u32p),
        ("instance_count", C.c_uint32),
        ("instance_geometry", _u32p),
        ("instance_transform", _f32p),
        ("instance_material", _u32p),
        ("material_count", C.c_uint32),
        ("materials", _u32p),
        ("environment", C.c_float * 3),
        ("geometry_vertex_count", _u32p),
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
    ]


class Frame(C.Structure):
    _fields_ = [
        ("clip_to_world", C.c_float * 16),
        ("ndc_from_pixel", C.c_float * 4),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("rect", C.c_uint32 * 4),
        ("sample_index", C.c_uint32),
        ("spp", C.c_uint32),
        ("max_bounces", C.c_uint32),
        ("light_sampling", C.c_uint32),
        ("restir_candidates", C.c_uint32),
        ("restir_proposal", C.c_uint32),
    ]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(reference=False)
        L = C.CDLL(LIB_PATH)
        L.oracle_scene_create.restype = C.c_void_p
        L.oracle_scene_create.argtypes = [C.POINTER(SceneDesc), C.c_char_p, C.c_size_t]
        L.oracle_scene_destroy.argtypes = [C.c_void_p]
        L.oracle_scene_triangle_count.restype = C.c_uint32
        L.oracle_scene_triangle_count.argtypes = [C.c_void_p]
        L.oracle_scene_world_triangles.argtypes = [C.c_void_p, _f32p]
        L.oracle_light_count.restype = C.c_uint32
        L.oracle_light_count.argtypes = [C.c_void_p]
        L.oracle_light_table.argtypes = [C.c_void_p, _u32p, _u32p]
        L.oracle_trace_closest.argtypes = [C.c_void_p, C.c_uint32, _f32p, _u32p, C.c_int, C.c_int]
        L.oracle_trace_occluded.argtypes = [C.c_void_p, C.c_uint32, _f32p, _u32p, C.c_int, C.c_int]
        L.oracle_trace_transmittance.argtypes = [C.c_void_p, C.c_uint32, _f32p, _f32p, C.c_int, C.c_int]
        L.oracle_render.argtypes = [C.c_void_p, C.POINTER(Frame), _f64p, _u32p, _u64p, C.c_int]
        L.oracle_trace_path.restype = C.c_uint32
        L.oracle_trace_path.argtypes = [C.c_void_p, C.POINTER(Frame), C.c_uint32, C.c_uint32,
                                        C.c_uint32, _f32p]
        L.oracle_camera_ray.argtypes = [_f32p, C.c_float, C.c_float, _f32p]
        L.oracle_reservoir_stream.restype = C.c_int
        L.oracle_reservoir_stream.argtypes = [_u32p, C.c_uint32, C.c_float, C.c_float, C.c_float]
        L.oracle_reservoir_merge.restype = C.c_int
        L.oracle_reservoir_merge.argtypes = [_u32p, _u32p, C.c_float, C.c_float, C.c_uint32, C.c_float]
        L.oracle_reservoir_normalization.restype = C.c_float
        L.oracle_reservoir_normalization.argtypes = [_u32p]
        L.oracle_write_render_buffer.restype = C.c_int
        L.oracle_write_render_buffer.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                                 C.c_uint32]
        L.oracle_material_inputs.restype = None
        L.oracle_material_inputs.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float, _f32p]
        L.oracle_reuse_normalization.restype = C.c_float
        L.oracle_reuse_normalization.argtypes = [_u32p, C.c_uint32]
        L.oracle_restir_initial.restype = C.c_int
        L.oracle_restir_initial.argtypes = [C.c_void_p, C.POINTER(Frame), C.c_uint32, C.c_uint32, C.c_uint32,
                                            _u32p]
        L.oracle_camera_constants.restype = C.c_int
        L.oracle_camera_constants.argtypes = [_f32p, _f32p, _f32p]
        L.oracle_pcg_hash.restype = C.c_uint32
        L.oracle_pcg_hash.argtypes = [C.c_uint32]
        L.oracle_seed_random.restype = C.c_uint32
        L.oracle_seed_random.argtypes = [C.c_uint32, C.c_uint32]
        L.oracle_random_sequence.argtypes = [C.c_uint32, C.c_uint32, _f32p]
        L.oracle_sincos_turns.argtypes = [C.c_uint32, _f32p, _f32p, _f32p]
        L.oracle_bsdf_eval.argtypes = [_u32p, C.c_int, _f32p, _f32p, _f32p, _f32p]
        L.oracle_bsdf_sample.argtypes = [_u32p, C.c_int, _f32p, _f32p, _f32p]
        L.oracle_offset_ray_origin.argtypes = [_f32p, _f32p, _f32p]
        L.oracle_restir_temporal.restype = C.c_int
        L.oracle_restir_temporal.argtypes = [C.c_void_p, C.POINTER(Frame), C.c_uint32, C.c_uint32, C.c_uint32,
                                             _f32p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64,
                                             C.c_uint64, _u32p, _f32p, C.c_uint32, _u32p, _f32p]
        L.oracle_restir_spatial.restype = None
        L.oracle_restir_spatial.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _f32p, _u32p, _u32p]
        _lib = L
    return _lib


def ref():
    """The reference's own CPU functions (oracle/_ref), or None when unbuilt."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = C.CDLL(REF_PATH)
        R.ref_compare.restype = C.c_int
        R.ref_compare.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _f32p, _f32p,
                                  C.c_uint32, _f64p, _f64p, _f64p, _u32p, C.c_char_p, C.c_size_t]
        R.ref_read_pfm.restype = C.c_int
        R.ref_read_pfm.argtypes = [C.c_char_p, _u32p, _u32p, _f32p]
        R.ref_cornell_scene.restype = C.c_size_t
        R.ref_cornell_scene.argtypes = [C.c_float, _u32p, C.c_size_t, _f32p]
        R.ref_stream_candidate.restype = C.c_int
        R.ref_stream_candidate.argtypes = [_u32p, C.c_uint32, C.c_float, C.c_float, C.c_float]
        R.ref_merge_reevaluated.restype = C.c_int
        R.ref_merge_reevaluated.argtypes = [_u32p, _u32p, C.c_double, C.c_float, C.c_uint32, C.c_float]
        R.ref_reservoir_normalization.restype = C.c_double
        R.ref_reservoir_normalization.argtypes = [_u32p]
        R.ref_estimate_variance.restype = C.c_float
        R.ref_estimate_variance.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_uint32]
        R.ref_filter_radiance.restype = C.c_int
        R.ref_filter_radiance.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_uint32, _f32p]
        R.ref_reproject.restype = None
        R.ref_reproject.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, _f32p,
                                    _u32p]
        R.ref_material_fixture.restype = C.c_size_t
        R.ref_material_fixture.argtypes = [C.c_char_p, C.c_int, _u32p, C.c_size_t]
        R.ref_temporal_advance.restype = None
        R.ref_temporal_advance.argtypes = [C.c_uint32, C.POINTER(C.c_uint64), _u32p, _u32p, _f32p,
                                           C.POINTER(C.c_int32), C.POINTER(C.c_uint64), _u32p]
        R.ref_settle.restype = None
        R.ref_settle.argtypes = [_f32p, _f32p, _u32p, C.c_size_t, C.c_uint32, C.c_uint32, _f32p]
        R.ref_reuse_normalization.restype = C.c_double
        R.ref_reuse_normalization.argtypes = [_u32p, C.c_uint32]
        _ref = R
    return _ref


def ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def decode_environment(texels: np.ndarray, fmt: int) -> np.ndarray:
    """TexelRgb (backend/vulkan/gpu_scene.cpp:612-629) over an 8-bit image:
    channel / 255, and for sRGB (format 1) the colour channels' transfer
    function, in float64, rounded once to float32; alpha stays linear."""
    v = texels.astype(np.float64) / 255.0
    if fmt == 1:
        rgb = v[..., :3]
        v[..., :3] = np.where(rgb <= 0.04045, rgb / 12.92, ((rgb + 0.055) / 1.055) ** 2.4)
    return np.ascontiguousarray(v.astype(np.float32))


class OracleScene:
    """A scene (see ``lotus_amd.scenes.Scene``) loaded into the oracle."""

    def __init__(self, scene):
        p = scene.packed()
        self._keep = p
        d = SceneDesc()
        d.geometry_count = p["geometry_count"]
        d.positions = ptr(p["positions"], C.c_float)
        d.triangles = ptr(p["triangles"], C.c_uint32)
        d.geometry_vertex_offset = ptr(p["geometry_vertex_offset"], C.c_uint32)
        d.geometry_triangle_offset = ptr(p["geometry_triangle_offset"], C.c_uint32)
        d.geometry_triangle_count = ptr(p["geometry_triangle_count"], C.c_uint32)
        d.instance_count = p["instance_count"]
        d.instance_geometry = ptr(p["instance_geometry"], C.c_uint32)
        d.instance_transform = ptr(p["instance_transform"], C.c_float)
        d.instance_material = ptr(p["instance_material"], C.c_uint32)
        d.material_count = p["material_count"]
        d.materials = ptr(p["materials"], C.c_uint32)
        for i in range(3):
            d.environment[i] = float(p["environment"][i])
        d.geometry_vertex_count = ptr(p["geometry_vertex_count"], C.c_uint32)
        lights = p.get("lights")
        d.light_count = 0 if lights is None else len(lights)
        d.lights = lights.ctypes.data if d.light_count else None
        from lotus_amd.hip import environment_struct
        env = p.get("environment_map")
        if env is not None and env["texels"].dtype == np.uint8:
            env = dict(env, texels=decode_environment(env["texels"], env.get("format", 1)))
        self._env = environment_struct(env)
        d.environment_map = C.addressof(self._env) if self._env is not None else None
        from lotus_amd.hip import set_texture_fields
        self._tex = set_texture_fields(d, p)
        err = C.create_string_buffer(256)
        self.handle = lib().oracle_scene_create(C.byref(d), err, 256)
        if not self.handle:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "handle", None):
            lib().oracle_scene_destroy(self.handle)
            self.handle = None

    @property
    def triangle_count(self) -> int:
        return lib().oracle_scene_triangle_count(self.handle)

    def world_triangles(self) -> np.ndarray:
        out = np.zeros((self.triangle_count, 9), np.float32)
        lib().oracle_scene_world_triangles(self.handle, ptr(out, C.c_float))
        return out

    def light_table(self):
        L = lib()
        n = L.oracle_light_count(self.handle)
        offsets = np.zeros(self._keep["instance_count"], np.uint32)
        records = np.zeros((max(n, 1), 4), np.uint32)
        L.oracle_light_table(self.handle, ptr(offsets, C.c_uint32), ptr(records, C.c_uint32))
        return offsets, records[:n]

    def trace(self, rays: np.ndarray, brute_force=False, threads=8) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 5), np.uint32)
        lib().oracle_trace_closest(self.handle, rays.shape[0], ptr(rays, C.c_float),
                                   ptr(hits, C.c_uint32), int(brute_force), threads)
        return hits

    def occluded(self, rays: np.ndarray, brute_force=False, threads=8) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint32)
        lib().oracle_trace_occluded(self.handle, rays.shape[0], ptr(rays, C.c_float),
                                    ptr(out, C.c_uint32), int(brute_force), threads)
        return out

    def transmittance(self, rays: np.ndarray, brute_force=False, threads=8) -> np.ndarray:
        """ShadowTransmittance per ray (scene.slang:445-460)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.float32)
        lib().oracle_trace_transmittance(self.handle, rays.shape[0], ptr(rays, C.c_float),
                                         ptr(out, C.c_float), int(brute_force), threads)
        return out

    def render(self, frame: Frame, threads=8):
        """FP64 accumulation: returns (accum[h, w, 8], counts[spp, bounces+2],
        shadow[spp])."""
        acc = np.zeros((frame.height, frame.width, 8), np.float64)
        stride = frame.max_bounces + 2
        counts = np.zeros((frame.spp, stride), np.uint32)
        shadow = np.zeros(frame.spp, np.uint64)
        lib().oracle_render(self.handle, C.byref(frame), ptr(acc, C.c_double),
                            ptr(counts, C.c_uint32), ptr(shadow, C.c_uint64), threads)
        return acc, counts, shadow

    def material_inputs(self, instance: int, primitive: int, u: float, v: float) -> np.ndarray:
        """Base colour, roughness, emission, metallic, specular colour at a hit."""
        out = np.zeros(12, np.float32)
        lib().oracle_material_inputs(self.handle, instance, primitive, u, v, ptr(out, C.c_float))
        return out

    def restir_initial(self, frame: Frame, x: int, y: int, sample_index: int):
        """The 12-word initial reservoir record of pixel (x, y), or None on a
        camera miss."""
        rec = np.zeros(12, np.uint32)
        hit = lib().oracle_restir_initial(self.handle, C.byref(frame), x, y, sample_index, ptr(rec, C.c_uint32))
        return rec if hit else None

    def restir_temporal(self, frame: Frame, x: int, y: int, sample_index: int, view, history_valid: bool,
                        previous_index: int, centre_ids, previous_record, previous_surface, max_history: int,
                        instance_ids=None, material_ids=None):
        """WriteTemporalReservoir at pixel (x, y): (record[16] words,
        receiver surface[40] floats, hit)."""
        view = np.ascontiguousarray(view, np.float32)
        prec = np.ascontiguousarray(previous_record if previous_record is not None else np.zeros(16), np.uint32)
        psurf = np.ascontiguousarray(previous_surface if previous_surface is not None else np.zeros(40), np.float32)
        rec = np.zeros(16, np.uint32)
        surf = np.zeros(40, np.float32)
        iid = None if instance_ids is None else np.ascontiguousarray(instance_ids, np.uint64)
        mid = None if material_ids is None else np.ascontiguousarray(material_ids, np.uint64)
        hit = lib().oracle_restir_temporal(
            self.handle, C.byref(frame), x, y, sample_index, ptr(view, C.c_float),
            None if iid is None else iid.ctypes.data, None if mid is None else mid.ctypes.data,
            int(bool(history_valid)), int(previous_index) & 0xFFFFFFFF, int(centre_ids[0]), int(centre_ids[1]),
            ptr(prec, C.c_uint32), ptr(psurf, C.c_float), max_history, ptr(rec, C.c_uint32), ptr(surf, C.c_float))
        return rec, surf, bool(hit)

    def restir_spatial(self, width: int, height: int, x: int, y: int, sample_index: int, radius: int,
                       max_history: int, records: np.ndarray, surfaces: np.ndarray, disk: bool = False):
        """WriteSpatialReservoir at pixel (x, y) over the temporal pass's
        records [h*w, 16] and surfaces [h*w, 32]: (record[16], sources[5]),
        the four axis neighbours or, with disk, five drawn within the radius."""
        rec = np.zeros(16, np.uint32)
        src = np.zeros(5, np.uint32)
        lib().oracle_restir_spatial(self.handle, width, height, x, y, sample_index, radius, max_history,
                                    int(disk), ptr(records, C.c_uint32), ptr(surfaces, C.c_float),
                                    ptr(rec, C.c_uint32), ptr(src, C.c_uint32))
        return rec, src


def camera_constants(world_to_clip_gl):
    """(world_to_clip_vk, clip_to_world) as float32[16] column-major."""
    gl = np.ascontiguousarray(world_to_clip_gl, np.float32)
    vk = np.zeros(16, np.float32)
    inv = np.zeros(16, np.float32)
    ok = lib().oracle_camera_constants(ptr(gl, C.c_float), ptr(vk, C.c_float), ptr(inv, C.c_float))
    if not ok:
        raise ValueError("singular camera")
    return vk, inv


def make_frame(world_to_clip_gl, width, height, spp=1, sample_index=0, max_bounces=8,
               light_sampling=1, rect=None, restir_candidates=8, restir_proposal=0, display=None) -> Frame:
    """`display`: the display window (x, y, w, h), the whole image when
    None; `rect`: the data window as DataScissor resolves it."""
    _, inv = camera_constants(world_to_clip_gl)
    f = Frame()
    for i in range(16):
        f.clip_to_world[i] = float(inv[i])
    # ndc = pixel * (2/w, 2/h) + (-2x/w - 1, -2y/h - 1) over the display
    # window (vulkan_backend.cpp:2408-2411), in float
    F = np.float32
    vx, vy, vw, vh = (F(0), F(0), F(width), F(height)) if display is None else tuple(F(v) for v in display)
    f.ndc_from_pixel[0] = F(2.0) / vw
    f.ndc_from_pixel[1] = F(2.0) / vh
    f.ndc_from_pixel[2] = F(F(F(-2.0) * vx) / vw) - F(1.0)
    f.ndc_from_pixel[3] = F(F(F(-2.0) * vy) / vh) - F(1.0)
    f.width, f.height = width, height
    r = rect or (0, 0, width, height)
    for i in range(4):
        f.rect[i] = r[i]
    f.sample_index = sample_index
    f.spp = spp
    f.max_bounces = max_bounces
    f.light_sampling = light_sampling
    f.restir_candidates = restir_candidates
    f.restir_proposal = restir_proposal
    return f


def pcg_hash(v: int) -> int:
    return lib().oracle_pcg_hash(v)


def random_sequence(state: int, count: int) -> np.ndarray:
    out = np.zeros(count, np.float32)
    lib().oracle_random_sequence(state, count, ptr(out, C.c_float))
    return out


def write_render_buffer(product: np.ndarray, product_format: int, buffer_format: int):
    """The reference's render-buffer write of a top-down [height, width, ...]
    product: the bottom-up buffer as bytes, or None where it refuses."""
    product = np.ascontiguousarray(product)
    h, w = product.shape[:2]
    out = np.zeros(h * w * (16 if buffer_format == 1 else 4), np.uint8)
    ok = lib().oracle_write_render_buffer(product.ctypes.data, product_format, w, h, out.ctypes.data, buffer_format)
    return out if ok else None
