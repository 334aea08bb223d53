"""The reference's material fixtures (adapters/common/material_fixtures.cpp,
MaterialFixtures at :2402-2435) as scenes of this renderer: the SceneUpdate
the reference's SceneExtraction makes of each (written by
tests/golden/make_material_fixtures.py through oracle/_ref into
tests/golden/material-*.npz), turned into a plan as GpuScene::Apply fills its
records (backend/vulkan/gpu_scene.cpp:1970-2118): per-corner normals per
geometry, each instance's texture-coordinate and colour set, the material
table with its texture bindings. Reference images are the committed
validation/reference/material-*-{mean,variance}.pfm (copied into
tests/golden/), 4096 spp BSDF-sampled (adapters/headless/main.cpp:3577-3604)."""
import os

import numpy as np

from lotus_amd import scenes
from oracle import desq5

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NONE = 0xFFFFFFFF
REFERENCE_SAMPLES = 4096  # kMaterialReferenceSamples, main.cpp:3577
LEVELS = (16, 256)        # kMaterialLevels, main.cpp:3579
COMPARED_SAMPLE_INDEX = 1 << 20


class _Words:
    def __init__(self, words):
        self.w = np.ascontiguousarray(words, np.uint32)
        self.i = 0

    def u(self, n=None):
        if n is None:
            v = int(self.w[self.i])
            self.i += 1
            return v
        v = self.w[self.i:self.i + n].copy()
        self.i += n
        return v

    def f(self, n=None):
        if n is None:
            return float(self.u(1).view(np.float32)[0])
        return self.u(n).view(np.float32)


def scene_from_words(words, name="material-fixture"):
    """A scenes.Scene and the fixture's image size from the driver's words."""
    r = _Words(words)
    size = r.u()
    sc = scenes.Scene(name=name)
    sc.view = r.f(16)
    sc.projection = r.f(16)
    sc.environment = tuple(float(x) for x in r.f(3))
    for _ in range(r.u()):
        shape, normalize = r.u(), r.u()
        xf = r.f(16)
        radiance = r.f(3)
        radius, width, height, angle = r.f(), r.f(), r.f(), r.f()
        sc.add_light(scenes.analytic_light(shape, xf, tuple(radiance), radius, width, height, angle,
                                           bool(normalize)))
    for _ in range(r.u()):
        w, h, fmt, nbytes = r.u(), r.u(), r.u(), r.u()
        raw = r.u((nbytes + 3) // 4).view(np.uint8)[:nbytes]
        texels = raw.view(np.float32).reshape(h, w, 4) if fmt == scenes.TEXTURE_RGBA32F else raw.reshape(h, w, 4)
        sc.add_texture(texels.copy(), fmt)
    geometries = []
    for _ in range(r.u()):
        nv, nt = r.u(), r.u()
        pos = r.f(3 * nv).reshape(-1, 3)
        tri = r.u(3 * nt).reshape(-1, 3)
        nn = r.u()
        normals = r.f(3 * nn).reshape(-1, 3) if nn else None
        tcs = [r.f(2 * r.u()).reshape(-1, 2) for _ in range(r.u())]
        cols = [r.f(3 * r.u()).reshape(-1, 3) for _ in range(r.u())]
        geometries.append((pos, tri, normals, tcs, cols))
    for g, (pos, tri, normals, _, _) in enumerate(geometries):
        sc.geometries.append((np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(tri, np.uint32)))
        if normals is not None and len(normals):
            sc.set_normals(g, normals)
    materials = []
    for _ in range(r.u()):
        model = r.u()
        base = r.f(3)
        roughness, metallic, ior = r.f(), r.f(), r.f()
        specular_workflow = r.u()
        specular, emission = r.f(3), r.f(3)
        transmission = r.f()
        transmission_color = r.f(3)
        thin = r.u()
        fallback_st = r.f(2)
        opacity, threshold = r.f(), r.f()
        normal = r.f(3)
        g_base = r.f(3)
        g_roughness, g_metallic = r.f(), r.f()
        g_emissive = r.f(3)
        g_strength, g_ior, g_specular = r.f(), r.f(), r.f()
        g_specular_color = r.f(3)
        g_transmission, g_clearcoat, g_cc_roughness = r.f(), r.f(), r.f()
        g_sheen_color = r.f(3)
        g_sheen_roughness = r.f()
        if model == 1:
            rec = scenes.gltf_material(tuple(g_base), g_roughness, g_metallic, tuple(g_emissive), g_strength, g_ior,
                                       g_specular, tuple(g_specular_color), g_transmission, g_clearcoat,
                                       g_cc_roughness, tuple(g_sheen_color), g_sheen_roughness, opacity=opacity,
                                       opacity_threshold=threshold, normal=tuple(normal))
            rec.view(np.float32)[4:6] = fallback_st
            model = 0  # the record holds its model
        else:
            rec = None
        rec = rec if rec is not None else scenes.preview_material(base_color=tuple(base), roughness=roughness, metallic=metallic, ior=ior,
                                      emission=tuple(emission), specular_workflow=bool(specular_workflow),
                                      specular_color=tuple(specular), transmission=transmission,
                                      transmission_color=tuple(transmission_color), thin_walled=bool(thin),
                                      opacity=opacity, opacity_threshold=threshold, normal=tuple(normal))
        rec[0] |= model
        rec.view(np.float32)[4:6] = fallback_st
        slots = []
        for slot in range(14):
            if not r.u():
                continue
            tex, channel, ws, wt = r.u(), r.u(), r.u(), r.u()
            scale, bias, fallback, st = r.f(4), r.f(4), r.f(4), r.f(6)
            b = scenes.texture_binding(0 if tex == NONE else tex, (ws, wt), channel, tuple(scale), tuple(bias),
                                       tuple(fallback), tuple(st[:3]), tuple(st[3:]), fallback_only=tex == NONE)
            slots.append((slot, b))
        if slots:
            first = len(sc.bindings)
            for _, b in slots:
                sc.add_binding(b)
            rec = scenes.with_lookups(rec, first, [slot for slot, _ in slots])
        materials.append(rec)
    sc.materials = [np.ascontiguousarray(m, np.uint32) for m in materials]
    for i in range(r.u()):
        g = r.u()
        xf = r.f(16)
        mat, tset, cset = r.u(), r.u(), r.u()
        sc.instances.append((g, np.ascontiguousarray(xf, np.float32), mat))
        _, _, _, tcs, cols = geometries[g]
        if tset != NONE:
            sc.set_instance_texcoords(i, tcs[tset])
        if cset != NONE:
            sc.set_colors(i, cols[cset], constant=cols[cset].shape[0] == 1)
    assert r.i == len(r.w)
    sc._packed = None
    return sc, size


def load(fixture, control=False):
    """(scene, size) of a committed fixture, or its negative control."""
    data = np.load(os.path.join(GOLDEN, f"{fixture}.npz"))
    return scene_from_words(data["control" if control else "scene"], fixture)


def reference_images(fixture):
    """(mean, variance) as float32 [size, size, 3], rows from the top."""
    return tuple(desq5.read_pfm(os.path.join(GOLDEN, f"{fixture}-{kind}.pfm")) for kind in ("mean", "variance"))
