"""Scene fixtures and workloads for the wavefront path tracer.

Every fixture is generated from a recorded master seed with the PCG hash the
renderer itself uses, so the CPU oracle and the GPU read identical bytes.
The reference's fixtures are restated from its sources:
  * ``cornell_v1`` — adapters/common/reference.cpp:221-280 (SetCornellBox);
  * ``grid`` indexing — adapters/common/workloads.cpp:10-37;
  * material records — backend/vulkan/gpu_scene.cpp:2041-2089
    (GpuMaterialRecord, gpu_scene.hpp:94-113);
  * mesh / material table order — the SceneExtraction plan order
    (include/lotus/extraction.hpp:107-117): instances in mesh key order,
    the default material at slot 0, then materials in key order.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

F32 = np.float32


def identity() -> np.ndarray:
    return np.eye(4, dtype=F32).reshape(-1, order="F")


def multiply(left: np.ndarray, right: np.ndarray) -> np.ndarray:
    """Column-major float32 product with the reference's summation order
    (core/render-world/render_world.cpp:288-300)."""
    out = np.zeros(16, F32)
    for c in range(4):
        for r in range(4):
            s = F32(0.0)
            for i in range(4):
                s = F32(s + F32(left[i * 4 + r] * right[c * 4 + i]))
            out[c * 4 + r] = s
    return out


# ------------------------------------------------------------- materials --
MATERIAL_WORDS = 36


def preview_material(base_color=(0.18, 0.18, 0.18), roughness=0.5, metallic=0.0, ior=1.5,
                     emission=(0.0, 0.0, 0.0), specular_workflow=False,
                     specular_color=(0.0, 0.0, 0.0), transmission=0.0,
                     transmission_color=(1.0, 1.0, 1.0), thin_walled=False,
                     opacity=1.0, opacity_threshold=0.0, normal=(0.0, 0.0, 1.0)) -> np.ndarray:
    """A UsdPreviewSurface GpuMaterialRecord (gpu_scene.hpp:94-113) as 36
    words, filled as gpu_scene.cpp:2041-2089 fills it. Defaults are
    PreviewSurface's (include/lotus/material.hpp:157-196)."""
    w = np.zeros(MATERIAL_WORDS, np.uint32)
    f = w.view(np.float32)
    flags = (1 if specular_workflow else 0) | (2 if thin_walled else 0)
    w[0] = 0 | (flags << 8)  # model kUsdPreviewSurface
    f[4:8] = (0.0, 0.0, opacity, opacity_threshold)
    f[8:12] = (*normal, 0.0)
    f[12:16] = (*base_color, roughness)
    f[16:20] = (*emission, metallic)
    f[20:24] = (*specular_color, ior)
    f[24:28] = (*transmission_color, transmission)
    return w


def gltf_material(base_color=(1.0, 1.0, 1.0), roughness=1.0, metallic=1.0, emissive=(0.0, 0.0, 0.0),
                  emissive_strength=1.0, ior=1.5, specular=1.0, specular_color=(1.0, 1.0, 1.0), transmission=0.0,
                  clearcoat=0.0, clearcoat_roughness=0.0, sheen_color=(0.0, 0.0, 0.0), sheen_roughness=0.0,
                  opacity=1.0, opacity_threshold=0.0, normal=(0.0, 0.0, 1.0)) -> np.ndarray:
    """A glTF PBR GpuMaterialRecord (gpu_scene.hpp:94-113) as 36 words, filled
    as gpu_scene.cpp:2050-2071 fills it: the dielectric's reflectance at
    normal incidence ((1 - ior) / (1 + ior))^2 computed in double. Defaults
    are GltfPbr's (include/lotus/material.hpp:180-219)."""
    w = np.zeros(MATERIAL_WORDS, np.uint32)
    f = w.view(np.float32)
    w[0] = 1  # model kGltfPbr
    f[4:8] = (0.0, 0.0, opacity, opacity_threshold)
    f[8:12] = (*normal, 0.0)
    root = (1.0 - float(np.float32(ior))) / (1.0 + float(np.float32(ior)))
    f[12:16] = (*base_color, roughness)
    f[16:20] = (*emissive, metallic)
    f[20:24] = (emissive_strength, np.float32(root * root), specular, transmission)
    f[24:28] = (*specular_color, clearcoat)
    f[28:32] = (*sheen_color, sheen_roughness)
    f[32:34] = (clearcoat_roughness, ior)
    return w


# ----------------------------------------------------------------- scene --
# lotus_hip_light (Light, include/lotus/render_world.hpp:89-114).
LIGHT_SPHERE, LIGHT_RECT, LIGHT_DISK, LIGHT_DISTANT = range(4)
LIGHT_DTYPE = np.dtype([("shape", "<u4"), ("normalize", "<u4"), ("world_from_light", "<f4", 16),
                        ("radiance", "<f4", 3), ("radius", "<f4"), ("width", "<f4"), ("height", "<f4"),
                        ("angle", "<f4")])


def analytic_light(shape, transform=None, radiance=(1.0, 1.0, 1.0), radius=0.5, width=1.0, height=1.0,
                   angle=0.53, normalize=False) -> np.ndarray:
    """One analytic light record: a sphere centred on its frame's origin, a
    rect or disk in its frame's XY plane emitting from its -Z side, or a
    distant light shining along its frame's -Z axis (angle 0: parallel)."""
    rec = np.zeros((), LIGHT_DTYPE)
    rec["shape"], rec["normalize"] = shape, int(normalize)
    rec["world_from_light"] = identity() if transform is None else np.asarray(transform, F32).reshape(16)
    rec["radiance"] = radiance
    rec["radius"], rec["width"], rec["height"], rec["angle"] = radius, width, height, angle
    return rec


# Material texture lookups (gpu_scene.hpp:56-84; material.hpp:15-50).
TEXTURE_RGBA8_UNORM, TEXTURE_RGBA8_SRGB, TEXTURE_RGBA32F = range(3)
WRAP_BLACK, WRAP_CLAMP, WRAP_REPEAT, WRAP_MIRROR = range(4)
SLOT_BASE_COLOR, SLOT_ROUGHNESS, SLOT_METALLIC, SLOT_EMISSION, SLOT_SPECULAR_COLOR = 2, 3, 4, 5, 6


def texture_binding(texture: int, wrap=(WRAP_REPEAT, WRAP_REPEAT), channel: int = 0, scale=(1, 1, 1, 1),
                    bias=(0, 0, 0, 0), fallback=(0, 0, 0, 1), st_s=(1.0, 0.0, 0.0), st_t=(0.0, 1.0, 0.0),
                    fallback_only: bool = False) -> np.ndarray:
    """One GpuTextureInputRecord (24 words): texture, sampler (wrap_s * 4 +
    wrap_t), channel, mode (1 lookup, 2 fallback), scale, bias, fallback and
    the coordinate rows s' = st_s . (s, t, 1), t' = st_t . (s, t, 1)."""
    w = np.zeros(24, np.uint32)
    f = w.view(F32)
    w[0], w[1], w[2], w[3] = texture, wrap[0] * 4 + wrap[1], channel, 2 if fallback_only else 1
    f[4:8], f[8:12], f[12:16] = scale, bias, fallback
    f[16:19], f[20:23] = st_s, st_t
    return w


def with_lookups(record: np.ndarray, first_binding: int, slots) -> np.ndarray:
    """A material record whose `slots` read bindings first_binding,
    first_binding + 1, ... in slot order (words 1 and 2)."""
    out = np.array(record, np.uint32)
    out[1] = first_binding
    out[2] = sum(1 << s for s in slots)
    return out


@dataclasses.dataclass
class Scene:
    """Object-space geometries placed by instances; materials as records;
    a camera as view and projection (include/lotus/render_world.hpp:31-37)."""

    geometries: List[Tuple[np.ndarray, np.ndarray]] = dataclasses.field(default_factory=list)
    instances: List[Tuple[int, np.ndarray, int]] = dataclasses.field(default_factory=list)
    materials: List[np.ndarray] = dataclasses.field(default_factory=lambda: [preview_material()])
    environment: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    lights: List[np.ndarray] = dataclasses.field(default_factory=list)
    # a latitude-longitude image: dict(texels=float32 [height, width, 4] rows
    # from the top, scale=(r, g, b), transform=16 floats), or None
    environment_map: Optional[dict] = None
    # texture lookups: textures as dict(texels, format) with texels
    # [height, width, 4] (uint8 or float32), binding records, and per
    # geometry its per-corner texture coordinates [triangles * 3, 2] or None
    textures: List[dict] = dataclasses.field(default_factory=list)
    bindings: List[np.ndarray] = dataclasses.field(default_factory=list)
    texcoords: Dict[int, np.ndarray] = dataclasses.field(default_factory=dict)
    # per instance the texture-coordinate set its material reads, when not
    # its geometry's ([triangles * 3, 2]); per geometry authored shading
    # normals ([triangles * 3, 3]); per instance its colour set, as
    # (colours [n, 3], constant) with n = 1 when constant, else one per corner
    instance_texcoords: Dict[int, np.ndarray] = dataclasses.field(default_factory=dict)
    normals: Dict[int, np.ndarray] = dataclasses.field(default_factory=dict)
    colors: Dict[int, Tuple[np.ndarray, bool]] = dataclasses.field(default_factory=dict)
    view: np.ndarray = dataclasses.field(default_factory=identity)
    projection: np.ndarray = dataclasses.field(default_factory=identity)
    name: str = "scene"
    _packed: Optional[dict] = None

    def add_mesh(self, positions, triangles, material: int, transform=None) -> int:
        positions = np.ascontiguousarray(positions, F32).reshape(-1, 3)
        triangles = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
        self.geometries.append((positions, triangles))
        t = identity() if transform is None else np.ascontiguousarray(transform, F32).reshape(16)
        self.instances.append((len(self.geometries) - 1, t, material))
        self._packed = None
        return len(self.instances) - 1

    def add_texture(self, texels: np.ndarray, format: int) -> int:
        self.textures.append(dict(texels=np.ascontiguousarray(texels), format=format))
        self._packed = None
        return len(self.textures) - 1

    def add_binding(self, record: np.ndarray) -> int:
        self.bindings.append(np.array(record, np.uint32))
        self._packed = None
        return len(self.bindings) - 1

    def set_texcoords(self, geometry: int, corners) -> None:
        """Per-corner texture coordinates of a geometry, 3 per triangle."""
        c = np.ascontiguousarray(corners, F32).reshape(-1, 2)
        assert c.shape[0] == 3 * self.geometries[geometry][1].shape[0]
        self.texcoords[geometry] = c
        self._packed = None

    def set_normals(self, geometry: int, corners) -> None:
        """Authored object-space shading normals of a geometry, 3 per triangle."""
        n = np.ascontiguousarray(corners, F32).reshape(-1, 3)
        assert n.shape[0] == 3 * self.geometries[geometry][1].shape[0]
        self.normals[geometry] = n
        self._packed = None

    def set_colors(self, instance: int, colors, constant: bool = False) -> None:
        """The colour set an instance's base colour reads: one colour, or one
        per triangle corner of its geometry."""
        c = np.ascontiguousarray(colors, F32).reshape(-1, 3)
        g = self.instances[instance][0]
        assert c.shape[0] == (1 if constant else 3 * self.geometries[g][1].shape[0])
        self.colors[instance] = (c, bool(constant))
        self._packed = None

    def set_instance_texcoords(self, instance: int, corners) -> None:
        """The texture coordinates an instance's material reads (one set of
        its geometry's), 3 per triangle."""
        c = np.ascontiguousarray(corners, F32).reshape(-1, 2)
        g = self.instances[instance][0]
        assert c.shape[0] == 3 * self.geometries[g][1].shape[0]
        self.instance_texcoords[instance] = c
        self._packed = None

    def add_light(self, record: np.ndarray) -> int:
        self.lights.append(np.array(record, LIGHT_DTYPE))
        self._packed = None
        return len(self.lights) - 1

    def add_material(self, record: np.ndarray) -> int:
        self.materials.append(np.ascontiguousarray(record, np.uint32))
        self._packed = None
        return len(self.materials) - 1

    def copy(self) -> "Scene":
        """An independent copy: editing its arrays leaves this scene as it is."""
        return Scene(geometries=[(p.copy(), t.copy()) for p, t in self.geometries],
                     instances=[(g, t.copy(), m) for g, t, m in self.instances],
                     materials=[m.copy() for m in self.materials], environment=tuple(self.environment),
                     lights=[l.copy() for l in self.lights],
                     textures=[dict(t) for t in self.textures], bindings=[b.copy() for b in self.bindings],
                     texcoords={g: c.copy() for g, c in self.texcoords.items()},
                     instance_texcoords={i: c.copy() for i, c in self.instance_texcoords.items()},
                     normals={g: n.copy() for g, n in self.normals.items()},
                     colors={i: (c.copy(), k) for i, (c, k) in self.colors.items()},
                     environment_map=None if self.environment_map is None else
                     {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                      for k, v in self.environment_map.items()},
                     view=self.view.copy(), projection=self.projection.copy(), name=self.name)

    def set_positions(self, geometry: int, positions) -> None:
        """Moves a geometry's vertices, keeping its vertex count and indices."""
        pos = np.ascontiguousarray(positions, F32).reshape(-1, 3)
        if pos.shape != self.geometries[geometry][0].shape:
            raise ValueError("a moved geometry keeps its vertex count")
        self.geometries[geometry] = (pos, self.geometries[geometry][1])
        self._packed = None

    def set_transform(self, instance: int, transform) -> None:
        g, _, m = self.instances[instance]
        self.instances[instance] = (g, np.ascontiguousarray(transform, F32).reshape(16), m)
        self._packed = None

    def set_material(self, slot: int, record) -> None:
        self.materials[slot] = np.ascontiguousarray(record, np.uint32)
        self._packed = None

    def _packed_corners(self) -> dict:
        """Texture coordinates, normals and colours with their first corners:
        per geometry (coordinates, normals) and per instance (coordinate
        overrides, colour sets); None where the scene has none."""
        none = 0xFFFFFFFF
        out = dict(texcoords=np.zeros(2, F32), geometry_texcoord_offset=None, instance_texcoord_offset=None,
                   normals=np.zeros(3, F32), geometry_normal_offset=None, colors=np.zeros(3, F32),
                   instance_color_offset=None, instance_color_constant=None)
        if self.texcoords or self.instance_texcoords:
            parts, n, goff = [], 0, []
            for g in range(len(self.geometries)):
                if g in self.texcoords:
                    goff.append(n)
                    parts.append(self.texcoords[g])
                    n += self.texcoords[g].shape[0]
                else:
                    goff.append(none)
            ioff = None
            if self.instance_texcoords:
                ioff = []
                for i, (g, _, _) in enumerate(self.instances):
                    if i in self.instance_texcoords:
                        ioff.append(n)
                        parts.append(self.instance_texcoords[i])
                        n += self.instance_texcoords[i].shape[0]
                    else:
                        ioff.append(goff[g])
            out["texcoords"] = np.ascontiguousarray(np.concatenate(parts).astype(F32))
            out["geometry_texcoord_offset"] = np.array(goff, np.uint32)
            out["instance_texcoord_offset"] = None if ioff is None else np.array(ioff, np.uint32)
        if self.normals:
            parts, n, goff = [], 0, []
            for g in range(len(self.geometries)):
                if g in self.normals:
                    goff.append(n)
                    parts.append(self.normals[g])
                    n += self.normals[g].shape[0]
                else:
                    goff.append(none)
            out["normals"] = np.ascontiguousarray(np.concatenate(parts).astype(F32))
            out["geometry_normal_offset"] = np.array(goff, np.uint32)
        if self.colors:
            parts, n, ioff, const = [], 0, [], []
            for i in range(len(self.instances)):
                if i in self.colors:
                    c, k = self.colors[i]
                    ioff.append(n)
                    const.append(1 if k else 0)
                    parts.append(c)
                    n += c.shape[0]
                else:
                    ioff.append(none)
                    const.append(0)
            out["colors"] = np.ascontiguousarray(np.concatenate(parts).astype(F32))
            out["instance_color_offset"] = np.array(ioff, np.uint32)
            out["instance_color_constant"] = np.array(const, np.uint32)
        return out

    @property
    def world_to_clip(self) -> np.ndarray:
        """camera.projection * camera.view (core/render-extraction/extraction.cpp:12-19)."""
        return multiply(self.projection, self.view)

    @property
    def triangle_count(self) -> int:
        return int(sum(self.geometries[g][1].shape[0] for g, _, _ in self.instances))

    def packed(self) -> dict:
        """Contiguous arrays in the layout both the renderer's C ABI and the
        oracle read."""
        if self._packed is not None:
            return self._packed
        vo, to, tc, vc = [], [], [], []
        v = t = 0
        for pos, tri in self.geometries:
            vo.append(v)
            vc.append(pos.shape[0])
            to.append(t)
            tc.append(tri.shape[0])
            v += pos.shape[0]
            t += tri.shape[0]
        pos = [p for p, _ in self.geometries] or [np.zeros((0, 3), F32)]
        tri = [tr for _, tr in self.geometries] or [np.zeros((0, 3), np.uint32)]
        xf = [tr for _, tr, _ in self.instances] or [np.zeros(16, F32)]
        self._packed = dict(
            geometry_count=len(self.geometries),
            positions=np.ascontiguousarray(np.concatenate(pos).astype(F32)),
            triangles=np.ascontiguousarray(np.concatenate(tri).astype(np.uint32)),
            geometry_vertex_offset=np.array(vo, np.uint32),
            geometry_vertex_count=np.array(vc, np.uint32),
            geometry_triangle_offset=np.array(to, np.uint32),
            geometry_triangle_count=np.array(tc, np.uint32),
            instance_count=len(self.instances),
            instance_geometry=np.array([g for g, _, _ in self.instances], np.uint32),
            instance_transform=np.ascontiguousarray(np.stack(xf).astype(F32)),
            instance_material=np.array([m for _, _, m in self.instances], np.uint32),
            material_count=len(self.materials),
            materials=np.ascontiguousarray(np.stack(self.materials).astype(np.uint32)),
            environment=np.array(self.environment, F32),
            lights=np.ascontiguousarray(np.array(self.lights, LIGHT_DTYPE).reshape(-1)),
            textures=list(self.textures),
            bindings=np.ascontiguousarray(np.array(self.bindings, np.uint32).reshape(-1)),
            **self._packed_corners(),
            environment_map=None if self.environment_map is None else dict(
                texels=np.ascontiguousarray(self.environment_map["texels"],
                                            np.uint8 if self.environment_map["texels"].dtype == np.uint8 else F32),
                format=self.environment_map.get("format", TEXTURE_RGBA8_SRGB),
                scale=np.asarray(self.environment_map.get("scale", (1.0, 1.0, 1.0)), F32),
                transform=np.asarray(self.environment_map.get("transform", identity()), F32).reshape(16)),
        )
        return self._packed


# ------------------------------------------------------------ randomness --
def pcg_hash(v: np.ndarray) -> np.ndarray:
    """PCG RXS-M-XS hash (sampling.slang:9-13) over uint32 arrays."""
    v = np.asarray(v, np.uint64) & 0xFFFFFFFF
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return ((word >> 22) ^ word).astype(np.uint32)


def uniform(seed: int, stream: int, index) -> np.ndarray:
    """Seeded uniform numbers in [0, 1) with 24 bits, indexed."""
    key = pcg_hash(np.uint32(seed) ^ pcg_hash(np.uint32(stream)))
    h = pcg_hash(np.asarray(index, np.uint64) ^ np.uint64(key))
    return ((h >> 8).astype(np.float64) * (1.0 / 16777216.0))


# --------------------------------------------------------------- cameras --
def perspective(fovy_degrees: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = 1.0 / math.tan(math.radians(fovy_degrees) / 2.0)
    m = np.zeros(16, F32)
    m[0] = f / aspect
    m[5] = f
    m[10] = (far + near) / (near - far)
    m[11] = -1.0
    m[14] = 2.0 * far * near / (near - far)
    return m


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    eye, target, up = (np.asarray(x, np.float64) for x in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m.astype(F32).reshape(-1, order="F")


# --------------------------------------------------------------- Cornell --
def _quad(corners):
    return np.array(corners, F32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


UNIT_CUBE_POS = np.array([[-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [0.5, 0.5, -0.5], [-0.5, 0.5, -0.5],
                          [-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5]], F32)
UNIT_CUBE_TRI = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                          [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)


def block(size, degrees, centre) -> np.ndarray:
    """reference.cpp:51-62: scale, turn about +Y, move."""
    r = degrees * 3.14159265358979323846 / 180.0
    c, s = F32(math.cos(r)), F32(math.sin(r))
    sx, sy, sz = (F32(v) for v in size)
    return np.array([c * sx, 0, -s * sx, 0, 0, sy, 0, 0, s * sz, 0, c * sz, 0,
                     centre[0], centre[1], centre[2], 1], F32)


def cornell_camera():
    """reference.cpp:222-234: 3 units in front, 45 degree vertical field."""
    focal, near, far = F32(2.41421356), F32(1.0), F32(10.0)
    view = identity()
    view[14] = -3.0
    proj = np.array([focal, 0, 0, 0, 0, focal, 0, 0, 0, 0, (far + near) / (near - far), -1,
                     0, 0, F32(2.0) * far * near / (near - far), 0], F32)
    return view, proj


def cornell_v1(red_scale: float = 1.0) -> Scene:
    """The reference scene (reference.cpp:221-280): coated-diffuse walls, an
    emissive panel (15, 13, 10), a GGX metal block and a white block."""
    sc = Scene(name="cornell-v1")
    sc.view, sc.projection = cornell_camera()
    white, red, green = (0.73, 0.73, 0.73), (0.63, 0.065, 0.05), (0.14, 0.45, 0.091)
    red = tuple(float(F32(c) * F32(red_scale)) for c in red)
    # material table: default, then keys in order: green, light, metal, red, white
    m_green = sc.add_material(preview_material(base_color=green))
    m_light = sc.add_material(preview_material(base_color=(0, 0, 0), emission=(15.0, 13.0, 10.0)))
    m_metal = sc.add_material(preview_material(base_color=(0.95, 0.85, 0.6), roughness=0.35, metallic=1.0))
    m_red = sc.add_material(preview_material(base_color=red))
    m_white = sc.add_material(preview_material(base_color=white))
    # meshes in key order: back, ceiling, floor, left, light, right, short, tall
    sc.add_mesh(*_quad([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]]), m_white)
    sc.add_mesh(*_quad([[-1, 1, 1], [-1, 1, -1], [1, 1, -1], [1, 1, 1]]), m_white)
    sc.add_mesh(*_quad([[-1, -1, 1], [1, -1, 1], [1, -1, -1], [-1, -1, -1]]), m_white)
    sc.add_mesh(*_quad([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1]]), m_red)
    sc.add_mesh(*_quad([[-0.35, 0.98, 0.35], [-0.35, 0.98, -0.35], [0.35, 0.98, -0.35], [0.35, 0.98, 0.35]]),
                m_light)
    sc.add_mesh(*_quad([[1, -1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]]), m_green)
    sc.add_mesh(UNIT_CUBE_POS, UNIT_CUBE_TRI, m_white, block((0.6, 0.6, 0.6), -18.0, (0.38, -0.699, 0.3)))
    sc.add_mesh(UNIT_CUBE_POS, UNIT_CUBE_TRI, m_metal, block((0.6, 1.2, 0.6), 17.0, (-0.35, -0.399, -0.3)))
    return sc


def cornell_lambert() -> Scene:
    """The CPU plain-PT workload: the reference box's five walls and two
    blocks, Lambertian (PreviewSurface with ior 1, metallic 0), lit by a
    0.25 x 0.25 emissive ceiling quad of radiance 15. Walls 0.73, left
    (0.65, 0.05, 0.05), right (0.12, 0.45, 0.15); 10 + 24 + 2 triangles."""
    sc = Scene(name="cornell-lambert")
    sc.view, sc.projection = cornell_camera()
    lam = lambda c: preview_material(base_color=c, ior=1.0)
    m_white = sc.add_material(lam((0.73, 0.73, 0.73)))
    m_left = sc.add_material(lam((0.65, 0.05, 0.05)))
    m_right = sc.add_material(lam((0.12, 0.45, 0.15)))
    m_light = sc.add_material(preview_material(base_color=(0, 0, 0), ior=1.0, emission=(15.0, 15.0, 15.0)))
    sc.add_mesh(*_quad([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]]), m_white)
    sc.add_mesh(*_quad([[-1, 1, 1], [-1, 1, -1], [1, 1, -1], [1, 1, 1]]), m_white)
    sc.add_mesh(*_quad([[-1, -1, 1], [1, -1, 1], [1, -1, -1], [-1, -1, -1]]), m_white)
    sc.add_mesh(*_quad([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1]]), m_left)
    h = 0.125
    sc.add_mesh(*_quad([[-h, 0.98, h], [-h, 0.98, -h], [h, 0.98, -h], [h, 0.98, h]]), m_light)
    sc.add_mesh(*_quad([[1, -1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]]), m_right)
    sc.add_mesh(UNIT_CUBE_POS, UNIT_CUBE_TRI, m_white, block((0.6, 0.6, 0.6), -18.0, (0.38, -0.699, 0.3)))
    sc.add_mesh(UNIT_CUBE_POS, UNIT_CUBE_TRI, m_white, block((0.6, 1.2, 0.6), 17.0, (-0.35, -0.399, -0.3)))
    return sc


# ------------------------------------------------------- closed-form scenes --
def cornell_many_lights(count: int = 100_000, seed: int = 2027, buckets: int = 64,
                        camera_aspect: float = 16.0 / 9.0) -> Scene:
    """The ReSTIR + SVGF workload: the config-1 box and blocks (Lambertian),
    the panel light removed, and `count` seeded emissive triangles on a
    ceiling plane just below the ceiling, facing down, each of the same
    area, with power log-uniform over [1e-2, 1e2] (radiance bucketed into
    `buckets` log-spaced levels, one mesh per level)."""
    sc = Scene(name=f"cornell-lights-{count}")
    view, _ = cornell_camera()
    sc.view = view
    sc.projection = perspective(45.0, camera_aspect, 1.0, 10.0)
    lam = lambda c: preview_material(base_color=c, ior=1.0)
    m_white = sc.add_material(lam((0.73, 0.73, 0.73)))
    m_left = sc.add_material(lam((0.65, 0.05, 0.05)))
    m_right = sc.add_material(lam((0.12, 0.45, 0.15)))
    sc.add_mesh(*_quad([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]]), m_white)
    sc.add_mesh(*_quad([[-1, 1, 1], [-1, 1, -1], [1, 1, -1], [1, 1, 1]]), m_white)
    sc.add_mesh(*_quad([[-1, -1, 1], [1, -1, 1], [1, -1, -1], [-1, -1, -1]]), m_white)
    sc.add_mesh(*_quad([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1]]), m_left)
    sc.add_mesh(*_quad([[1, -1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]]), m_right)
    sc.add_mesh(UNIT_CUBE_POS, UNIT_CUBE_TRI, m_white, block((0.6, 0.6, 0.6), -18.0, (0.38, -0.699, 0.3)))
    sc.add_mesh(UNIT_CUBE_POS, UNIT_CUBE_TRI, m_white, block((0.6, 1.2, 0.6), 17.0, (-0.35, -0.399, -0.3)))
    i = np.arange(count)
    cx = uniform(seed, 1, i) * 1.9 - 0.95
    cz = uniform(seed, 2, i) * 1.9 - 0.95
    ang = uniform(seed, 3, i) * 2.0 * math.pi
    size = 0.004  # circumradius: every triangle is equilateral with the same area
    area = 3.0 * math.sqrt(3.0) / 4.0 * size * size
    corners = []
    for k in (0, 2, 1):  # wound so the face normal points down, into the box
        a = ang + k * 2.0 * math.pi / 3.0
        corners.append(np.stack([cx + size * np.cos(a), np.full(count, 0.995), cz - size * np.sin(a)], -1))
    pos = np.stack(corners, 1).astype(F32)  # [count, 3, 3]
    power = 10.0 ** (uniform(seed, 4, i) * 4.0 - 2.0)
    level = np.minimum((np.log10(power) + 2.0) / 4.0 * buckets, buckets - 1).astype(np.int64)
    for b in range(buckets):
        sel = np.nonzero(level == b)[0]
        if sel.size == 0:
            continue
        p = 10.0 ** ((b + 0.5) / buckets * 4.0 - 2.0)  # the bucket's power
        radiance = p / (area * math.pi)
        m = sc.add_material(preview_material(base_color=(0.0, 0.0, 0.0), ior=1.0, emission=(radiance,) * 3))
        tri = np.arange(sel.size * 3, dtype=np.uint32).reshape(-1, 3)
        sc.add_mesh(pos[sel].reshape(-1, 3), tri, m)
    return sc


def camera_spline(frame: int, frames: int = 120, seed: int = 7):
    """A seeded smooth camera path in front of the box: the eye drifts on a
    Lissajous curve around (0, 0, 3), looking at the box centre."""
    t = frame / max(frames - 1, 1)
    ph = [uniform(seed, 9, np.arange(3)) * 2 * math.pi][0]
    eye = (0.35 * math.sin(2 * math.pi * t + ph[0]), 0.2 * math.sin(4 * math.pi * t + ph[1]),
           3.0 + 0.25 * math.sin(2 * math.pi * t + ph[2]))
    return look_at(eye, (0.0, 0.0, -0.2))


def _icosphere(level: int):
    """Unit icosphere: positions and triangles after `level` midpoint splits."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        m = inv.reshape(3, -1).T + base  # midpoints of edges 01, 12, 20
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def icosphere_rig(triangles: int, seed: int) -> Scene:
    """A batched-view scene: displaced icospheres (about `triangles` in all)
    inside a closed box lit by two emissive ceiling panels; seeded
    PreviewSurface materials."""
    sc = Scene(name=f"rig-{triangles}-{seed}")
    u = lambda stream, i: float(uniform(seed, stream, i))
    m_wall = sc.add_material(preview_material(base_color=(0.7, 0.7, 0.68), ior=1.0))
    m_light = sc.add_material(preview_material(base_color=(0.0, 0.0, 0.0), ior=1.0, emission=(12.0, 11.0, 9.0)))
    e = 2.0
    sc.add_mesh(*_quad([[-e, -e, -e], [e, -e, -e], [e, e, -e], [-e, e, -e]]), m_wall)
    sc.add_mesh(*_quad([[-e, e, e], [-e, e, -e], [e, e, -e], [e, e, e]]), m_wall)
    sc.add_mesh(*_quad([[-e, -e, e], [e, -e, e], [e, -e, -e], [-e, -e, -e]]), m_wall)
    sc.add_mesh(*_quad([[-e, -e, e], [-e, -e, -e], [-e, e, -e], [-e, e, e]]), m_wall)
    sc.add_mesh(*_quad([[e, -e, -e], [e, -e, e], [e, e, e], [e, e, -e]]), m_wall)
    sc.add_mesh(*_quad([[-e, -e, e], [-e, e, e], [e, e, e], [e, -e, e]]), m_wall)
    for k, x0 in enumerate((-0.9, 0.5)):
        h = 0.35
        sc.add_mesh(*_quad([[x0 - h, e - 0.01, h], [x0 - h, e - 0.01, -h], [x0 + h, e - 0.01, -h],
                            [x0 + h, e - 0.01, h]]), m_light)
    spheres = 1 + int(u(1, 0) * 6)
    per = max(triangles // spheres, 80)
    level = max(1, min(7, int(round(math.log(per / 20.0, 4)))))
    base_v, base_f = _icosphere(level)
    for s_ in range(spheres):
        rough = 0.05 + 0.9 * u(2, s_)
        metal = 1.0 if u(3, s_) < 0.25 else 0.0
        col = tuple(0.2 + 0.7 * u(4 + c, s_) for c in range(3))
        m = sc.add_material(preview_material(base_color=col, roughness=rough, metallic=metal, ior=1.5))
        radius = 0.25 + 0.35 * u(8, s_)
        centre = np.array([u(9, s_) * 2.4 - 1.2, u(10, s_) * 1.6 - 1.4, u(11, s_) * 2.4 - 1.2])
        d = base_v
        bump = np.sin(5.0 * d[:, 0] + 7.0 * u(12, s_)) * np.sin(4.0 * d[:, 1] + 3.0 * u(13, s_)) * \
            np.sin(6.0 * d[:, 2] + 5.0 * u(14, s_))
        pos = centre + d * (radius * (1.0 + 0.25 * bump))[:, None]
        sc.add_mesh(pos.astype(F32), base_f.astype(np.uint32), m)
    return sc


def batch_views(count: int = 64, seed: int = 4, min_tris: int = 1000, max_tris: int = 200_000,
                min_size: int = 64, max_size: int = 512):
    """The batched-frame workload: `count` seeded views, each over its own
    icosphere rig of log-uniform triangle count, log-uniform square size,
    depth uniform in 2..12 and a camera looking into the rig."""
    views = []
    for i in range(count):
        u = lambda stream: float(uniform(seed, stream, i))
        tris = int(math.exp(math.log(min_tris) + u(1) * (math.log(max_tris) - math.log(min_tris))))
        size = int(math.exp(math.log(min_size) + u(2) * (math.log(max_size) - math.log(min_size))))
        depth = 2 + int(u(3) * 11)
        ang = u(4) * 2 * math.pi
        eye = (1.7 * math.cos(ang), 0.3 + 0.8 * u(5), 1.7 * math.sin(ang))
        view = look_at(eye, (0.0, -0.5, 0.0))
        proj = perspective(60.0, 1.0, 0.05, 20.0)
        views.append(dict(scene_triangles=tris, scene_seed=seed * 1000 + i, size=size, depth=depth,
                          view=view, world_to_clip=multiply(proj, view)))
    return views


def lambert_under_sky() -> Scene:
    """adapters/headless/main.cpp:1525-1545: a lone Lambert triangle (ior 1)
    under a constant environment reflects albedo * environment + emission."""
    sc = Scene(name="lambert-sky")
    sc.view, sc.projection = cornell_camera()
    m = sc.add_material(preview_material(base_color=(0.5, 0.25, 0.75), ior=1.0, emission=(0.1, 0.05, 0.0)))
    sc.add_mesh([[-0.65, -0.35, 0], [0.55, -0.45, 0], [0.15, 0.70, 0]], [[0, 1, 2]], m)
    sc.environment = (0.8, 0.6, 0.4)
    return sc


def sky_image(width: int = 64, height: int = 32, sun=(0.3, 0.25), sun_radiance=60.0, seed: int = 5) -> dict:
    """A seeded latitude-longitude sky: a blue gradient with noise and a
    small bright sun disc around image coordinates `sun` (u, v)."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    uu, vv = np.meshgrid(u, v)
    base = np.stack([0.3 + 0.4 * vv, 0.5 + 0.3 * vv, 0.9 - 0.2 * vv], -1)
    base = base * (0.8 + 0.4 * uniform(seed, 21, np.arange(width * height)).reshape(height, width, 1))
    d = np.hypot((uu - sun[0]) * 2.0, vv - sun[1])
    base = base + (d < 0.05)[..., None] * sun_radiance
    texels = np.concatenate([base, np.ones((height, width, 1))], -1).astype(F32)
    return dict(texels=texels, scale=(1.0, 1.0, 1.0), transform=identity())


def lit_floor(light: np.ndarray, albedo: float = 0.5, height: float = 2.0) -> Scene:
    """A Lambert floor (ior 1) of half-size 50 in y = 0, seen straight down
    through a 1-degree field of view from `height`, lit only by `light`: the
    closed-form scenes of the analytic lights (each pixel sees a point within
    0.02 of the origin, where the light's irradiance is known)."""
    sc = Scene(name="lit-floor")
    m = sc.add_material(preview_material(base_color=(albedo, albedo, albedo), ior=1.0))
    h = 50.0
    sc.add_mesh([[-h, 0.0, -h], [h, 0.0, -h], [h, 0.0, h], [-h, 0.0, h]], [[0, 2, 1], [0, 3, 2]], m)
    sc.view = look_at((0.0, height, 0.0), (0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0))
    sc.projection = perspective(1.0, 1.0, 0.05, 10.0)
    if light is not None:
        sc.add_light(light)
    return sc


def translation(x: float, y: float, z: float, scale: float = 1.0) -> np.ndarray:
    m = identity() * F32(scale)
    m[15] = 1.0
    m[12:15] = (x, y, z)
    return m


def facing_down(x: float, y: float, z: float) -> np.ndarray:
    """A light frame at (x, y, z) whose -Z axis points down (-Y): a rect or
    disk there emits towards the floor, a distant light shines down."""
    m = np.zeros(16, F32)
    m[0:3] = (1.0, 0.0, 0.0)    # X
    m[4:7] = (0.0, 0.0, -1.0)   # Y
    m[8:11] = (0.0, 1.0, 0.0)   # Z: up, so -Z is down
    m[12:16] = (x, y, z, 1.0)
    return m


def closed_box() -> Scene:
    """adapters/headless/main.cpp:3073-3136: a closed Lambert box seen from
    inside; L = Le (1 + a + ... + a^n), unlimited mean Le / (1 - a)."""
    sc = Scene(name="closed-box")
    sc.view = identity()
    sc.projection = perspective(90.0, 1.0, 0.05, 10.0)
    m = sc.add_material(preview_material(base_color=(0.5, 0.25, 0.75), ior=1.0, emission=(0.1, 0.15, 0.025)))
    pos = UNIT_CUBE_POS * F32(2.0)
    sc.add_mesh(pos, UNIT_CUBE_TRI, m)
    sc.environment = (1.0, 1.0, 1.0)
    return sc


# ------------------------------------------------------------ heightfield --
def value_noise(x: np.ndarray, z: np.ndarray, seed: int, octaves: int = 6,
                base_frequency: float = 4.0, gain: float = 0.5, lacunarity: float = 2.0) -> np.ndarray:
    """Seeded fractal value noise in [-1, 1]: per octave, lattice values from
    the PCG hash interpolated with the cubic smoothstep; amplitudes scale by
    `gain`, frequencies by `lacunarity`, normalised by the amplitude sum."""
    total = np.zeros_like(x, np.float64)
    norm = 0.0
    amp, freq = 1.0, base_frequency
    for o in range(octaves):
        fx, fz = (x + 1.0) * 0.5 * freq, (z + 1.0) * 0.5 * freq
        ix, iz = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)
        tx, tz = fx - ix, fz - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)

        def lattice(a, b):
            key = (a & 0xFFFF) | ((b & 0xFFFF) << 16)
            return uniform(seed, 100 + o, key) * 2.0 - 1.0

        v00, v10 = lattice(ix, iz), lattice(ix + 1, iz)
        v01, v11 = lattice(ix, iz + 1), lattice(ix + 1, iz + 1)
        v0 = v00 + (v10 - v00) * sx
        v1 = v01 + (v11 - v01) * sx
        total += amp * (v0 + (v1 - v0) * sz)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def heightfield(cells: int = 1024, seed: int = 2026, amplitude: float = 0.2, tiles: int = 8,
                emissive: int = 64, camera_aspect: float = 16.0 / 9.0) -> Scene:
    """The wavefront workload: a cells x cells heightfield over [-1, 1]^2 in
    x-z (2 triangles per cell, workloads.cpp:18-27 indexing), heights 0.2 x
    6-octave value noise. The grid is split into tiles x tiles meshes, each
    bound to its own seeded PreviewSurface: roughness U[0.05, 1], 20 %
    conductors (metallic 1), 80 % dielectric-coated diffuse (ior 1.5).
    `emissive` seeded triangles become single-triangle meshes with emission
    U[5, 20] (grey). Triangle total: 2 * cells^2."""
    n = cells + 1
    lin = np.linspace(-1.0, 1.0, n)
    gx, gz = np.meshgrid(lin, lin)  # row = z index
    h = amplitude * value_noise(gx, gz, seed)
    pos = np.stack([gx, h, gz], -1).reshape(-1, 3).astype(F32)
    sc = Scene(name=f"heightfield-{cells}")
    sc.view = look_at((0.0, 1.05, 1.75), (0.0, -0.05, 0.0))
    sc.projection = perspective(45.0, camera_aspect, 0.05, 10.0)
    tile_cells = cells // tiles
    # per-tile materials
    u = lambda stream, i: float(uniform(seed, stream, i))
    tile_mat = []
    for t in range(tiles * tiles):
        rough = 0.05 + 0.95 * u(1, t)
        if u(2, t) < 0.2:
            bc = tuple(0.5 + 0.5 * u(3 + c, t) for c in range(3))
            rec = preview_material(base_color=bc, roughness=rough, metallic=1.0)
        else:
            bc = tuple(0.1 + 0.8 * u(3 + c, t) for c in range(3))
            rec = preview_material(base_color=bc, roughness=rough, metallic=0.0, ior=1.5)
        tile_mat.append(sc.add_material(rec))
    # emissive triangles: seeded distinct picks
    total_tris = 2 * cells * cells
    picks = []
    k = 0
    while len(picks) < emissive:
        cand = int(uniform(seed, 7, k) * total_tris)
        k += 1
        if cand not in picks:
            picks.append(cand)
    pick_set = set(picks)
    # tile meshes (each with its own vertex block)
    for ty in range(tiles):
        for tx in range(tiles):
            y0, x0 = ty * tile_cells, tx * tile_cells
            ys, xs = np.meshgrid(np.arange(y0, y0 + tile_cells), np.arange(x0, x0 + tile_cells), indexing="ij")
            a = (ys * n + xs).reshape(-1)
            b = a + 1
            d = a + n
            c = d + 1
            tri = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], 1).reshape(-1, 3)
            cell_index = (ys * cells + xs).reshape(-1)
            gid = np.stack([2 * cell_index, 2 * cell_index + 1], 1).reshape(-1)
            keep = np.array([g not in pick_set for g in gid]) if pick_set else np.ones(len(gid), bool)
            tri = tri[keep]
            used, inv = np.unique(tri.reshape(-1), return_inverse=True)
            sc.add_mesh(pos[used], inv.reshape(-1, 3).astype(np.uint32), tile_mat[ty * tiles + tx])
    for i, g in enumerate(picks):
        cell, half = divmod(g, 2)
        y, x = divmod(cell, cells)
        a = y * n + x
        b, d = a + 1, a + n
        c = d + 1
        tri = (a, b, c) if half == 0 else (a, c, d)
        radiance = 5.0 + 15.0 * u(8, i)
        m = sc.add_material(preview_material(base_color=(0.0, 0.0, 0.0), emission=(radiance,) * 3))
        sc.add_mesh(pos[list(tri)], [[0, 1, 2]], m)
    return sc


def heightfield_wave(sc: Scene, time: float, amplitude: float = 0.05) -> Scene:
    """The heightfield animated: every vertex raised by amplitude x
    sin(4x + 3z + time), the same vertices, triangles and instances (a
    deformation that keeps each mesh's topology)."""
    out = sc.copy()
    for g, (pos, _) in enumerate(out.geometries):
        moved = pos.copy()
        moved[:, 1] += F32(amplitude) * np.sin(F32(4.0) * pos[:, 0] + F32(3.0) * pos[:, 2] + F32(time)).astype(F32)
        out.set_positions(g, moved)
    return out


def random_triangles(count: int, seed: int = 1, extent: float = 1.0, size: float = 0.2) -> Scene:
    """Seeded random triangles in a cube, for traversal parity."""
    idx = np.arange(count * 9)
    centre = (uniform(seed, 1, np.arange(count * 3)).reshape(count, 1, 3) * 2 - 1) * extent
    off = (uniform(seed, 2, idx).reshape(count, 3, 3) * 2 - 1) * size
    pos = (centre + off).astype(F32).reshape(-1, 3)
    sc = Scene(name=f"random-{count}")
    m = sc.add_material(preview_material())
    sc.add_mesh(pos, np.arange(count * 3, dtype=np.uint32).reshape(-1, 3), m)
    sc.view = look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    sc.projection = perspective(45.0, 1.0, 0.05, 10.0)
    return sc


def grid(divisions: int = 64, extent: float = 1.0) -> Scene:
    """The reference's Grid geometry (workloads.cpp:10-37): a divisions^2
    quad grid with a small deterministic z ripple, seen from its +z side.
    Every interior edge is shared: watertightness cases."""
    n = divisions + 1
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    px = extent * (F32(2.0) * x.astype(F32) / F32(divisions) - F32(1.0))
    py = extent * (F32(2.0) * y.astype(F32) / F32(divisions) - F32(1.0))
    pz = F32(0.002) * ((x * 17 + y * 13) % 11).astype(F32)
    pos = np.stack([px, py, pz], -1).reshape(-1, 3).astype(F32)
    ys, xs = np.meshgrid(np.arange(divisions), np.arange(divisions), indexing="ij")
    a = (ys * n + xs).reshape(-1)
    b, d = a + 1, a + n
    c = d + 1
    tri = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], 1).reshape(-1, 3).astype(np.uint32)
    sc = Scene(name=f"grid-{divisions}")
    m = sc.add_material(preview_material())
    sc.add_mesh(pos, tri, m)
    sc.view = look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    sc.projection = perspective(45.0, 1.0, 0.05, 10.0)
    return sc


def restir_plane(colors=((2.0, 0.5, 1.0), (0.2, 2.0, 1.5), (1.0, 1.0, 1.0)), zoom: float = 1.0) -> Scene:
    """The reservoir-arithmetic scene of validation/restir-gpu-test.cpp:30-51,
    124-134: a 24 x 24 diffuse plane (base colour (0.5, 0.75, 1), ior 1) at
    z = -4 under the camera's identity view, lit by parallel lights of the
    given radiances shining along -Z; the third shines along +Z, from below
    the plane, so its target is zero. Every light's unshadowed target is a
    closed form: max over channels of base colour x radiance / pi."""
    sc = Scene(name="restir-plane")
    sc.view = identity()
    sc.projection = np.array([zoom, 0, 0, 0, 0, zoom, 0, 0, 0, 0, -1.2, -1, 0, 0, -2.2, 0], F32)
    m = sc.add_material(preview_material(base_color=(0.5, 0.75, 1.0), ior=1.0))
    sc.add_mesh([[-12, -12, -4], [12, -12, -4], [12, 12, -4], [-12, 12, -4]], [[0, 1, 2], [0, 2, 3]], m)
    for i, c in enumerate(colors):
        t = identity()
        if i == 2:
            t[10] = -1.0
        sc.add_light(analytic_light(LIGHT_DISTANT, transform=t, radiance=c, angle=0.0))
    return sc
