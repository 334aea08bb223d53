"""Material texture lookups on the GPU against the oracle (scene.slang:318-374,
572-577, 1184-1221; gpu_scene.cpp:144-179, 612-629).

Bars, as for the untextured scenes: the light table bitwise (textured
emitters weigh scale + bias); per-sample queue counts and shadow-ray counts
exact; one-sample radiance bit-exact for >= 99.9 % of pixels under BSDF,
MIS and ReSTIR sampling; accumulations to 1e-4.
"""
import sys

import numpy as np
import pytest

import oracle
from lotus_amd import hip, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    status, detail = hip.probe()
    assert status == hip.PASS, detail
    r = hip.Renderer(0)
    yield r
    r.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def textured_cornell():
    """The Lambert Cornell box with an sRGB checker on the back wall and
    floor (repeat / mirror), a float emission texture on the ceiling light
    (clamp), and a glossy block whose roughness, metallic and specular colour
    come from one unorm texture through its material's fallback coordinates."""
    rng = np.random.default_rng(11)
    sc = scenes.cornell_lambert()
    checker = np.zeros((8, 8, 4), np.uint8)
    checker[..., 3] = 255
    checker[::2, ::2, :3] = (230, 200, 120)
    checker[1::2, 1::2, :3] = (40, 90, 200)
    checker[::2, 1::2, :3] = (180, 180, 180)
    checker[1::2, ::2, :3] = (90, 30, 30)
    t_checker = sc.add_texture(checker, scenes.TEXTURE_RGBA8_SRGB)
    glow = np.ones((4, 6, 4), np.float32)
    glow[..., :3] = rng.uniform(0.2, 1.5, (4, 6, 3))
    t_glow = sc.add_texture(glow, scenes.TEXTURE_RGBA32F)
    t_gloss = sc.add_texture(rng.integers(0, 256, (5, 3, 4), dtype=np.uint8), scenes.TEXTURE_RGBA8_UNORM)

    wall = sc.add_binding(scenes.texture_binding(t_checker, wrap=(scenes.WRAP_REPEAT, scenes.WRAP_MIRROR),
                                                 st_s=(2.0, 0.0, 0.1), st_t=(0.0, 2.0, 0.0)))
    sc.materials.append(scenes.with_lookups(scenes.preview_material(base_color=(1, 1, 1), ior=1.0), wall,
                                            [scenes.SLOT_BASE_COLOR]))
    m_wall = len(sc.materials) - 1
    light = sc.add_binding(scenes.texture_binding(t_glow, wrap=(scenes.WRAP_CLAMP, scenes.WRAP_CLAMP),
                                                  scale=(12.0, 12.0, 12.0, 1.0), bias=(1.0, 1.0, 1.0, 0.0)))
    sc.materials.append(scenes.with_lookups(scenes.preview_material(base_color=(0, 0, 0), ior=1.0), light,
                                            [scenes.SLOT_EMISSION]))
    m_light = len(sc.materials) - 1
    g0 = sc.add_binding(scenes.texture_binding(t_gloss, channel=1, scale=(0.6,) * 4, bias=(0.2,) * 4))
    sc.add_binding(scenes.texture_binding(t_gloss, channel=2, wrap=(scenes.WRAP_BLACK, scenes.WRAP_CLAMP)))
    sc.add_binding(scenes.texture_binding(t_gloss, fallback_only=True, fallback=(0.9, 0.8, 0.3, 1.0)))
    gloss = scenes.with_lookups(scenes.preview_material(base_color=(0.6, 0.6, 0.6), ior=1.5), g0,
                                [scenes.SLOT_ROUGHNESS, scenes.SLOT_METALLIC, scenes.SLOT_SPECULAR_COLOR])
    gloss.view(np.float32)[4:6] = (0.3, 0.7)
    sc.materials.append(gloss)
    m_gloss = len(sc.materials) - 1
    # instances: 0 back, 1 ceiling, 2 floor, 3 left, 4 light, 5 right, 6-7 blocks
    retint = {0: m_wall, 2: m_wall, 4: m_light, 6: m_gloss}
    sc.instances = [(g, t, retint.get(i, m)) for i, (g, t, m) in enumerate(sc.instances)]
    for geo in (0, 2, 4):
        pos, tri = sc.geometries[geo]
        corners = pos[tri.reshape(-1)]
        axes = [a for a in range(3) if np.ptp(pos[:, a]) > 0]
        sc.set_texcoords(geo, (corners[:, axes] - corners[:, axes].min(0)) / np.ptp(corners[:, axes], 0))
    sc._packed = None
    return sc


def test_textured_light_table_matches_oracle(renderer):
    sc = textured_cornell()
    renderer.apply(sc)
    off_g, rec_g = renderer.light_table()
    off_c, rec_c = oracle.OracleScene(sc).light_table()
    assert np.array_equal(off_g, off_c) and np.array_equal(rec_g, rec_c)


@pytest.mark.parametrize("light", [hip.LIGHT_BSDF, hip.LIGHT_MIS, hip.LIGHT_RESTIR])
def test_textured_cornell_counts_and_radiance(renderer, light):
    _say(f"textured cornell light={light}")
    sc = textured_cornell()
    renderer.apply(sc)
    w = h = 96
    spp = 4
    o = oracle.OracleScene(sc)
    acc, counts, shadow = o.render(oracle.make_frame(sc.world_to_clip, w, h, spp=spp, max_bounces=8,
                                                     light_sampling=light))
    first = None
    for k in range(spp):
        out = renderer.render(sc.world_to_clip, w, h, frame_count=1, max_bounces=8, light_sampling=light)
        n = len(out["counts"])
        assert np.array_equal(out["counts"], counts[k][:n]) and not counts[k][n:].any(), (k, out["counts"], counts[k])
        assert out["shadow_rays"] == shadow[k], (k, out["shadow_rays"], shadow[k])
        if first is None:
            first = out["color"]
    acc1, _, _ = o.render(oracle.make_frame(sc.world_to_clip, w, h, spp=1, max_bounces=8, light_sampling=light))
    ref1 = acc1[..., :3].astype(np.float32)
    hit = acc1[..., 3] > 0
    exact = np.mean(np.all(first[..., :3][hit] == ref1[hit], -1))
    _say(f"  1-spp radiance bit-exact fraction {exact:.6f}")
    assert exact > 0.999
    assert np.allclose(first[..., :3], ref1, rtol=1e-4, atol=1e-5)
    sums, n = renderer.accumulation(w, h)
    assert n == spp
    assert np.allclose(sums[..., :3], acc[..., :3], rtol=1e-4, atol=1e-4)


def test_texture_edits_and_rejected_slots(renderer):
    sc = textured_cornell()
    renderer.apply(sc)
    # a new binding value changes the tables, not the geometry
    brighter = sc.copy()
    brighter.bindings[1] = scenes.texture_binding(1, wrap=(scenes.WRAP_CLAMP, scenes.WRAP_CLAMP),
                                                  scale=(20.0, 20.0, 20.0, 1.0))
    brighter._packed = None
    ev = renderer.apply(brighter)
    assert ev.update in (hip.SCENE_TABLES, hip.SCENE_BUILD) and ev.bvh_refits == 0
    assert np.array_equal(renderer.light_table()[1], oracle.OracleScene(brighter).light_table()[1])
    # a lookup in a slot the preview surface does not read is refused
    bad = sc.copy()
    rec = np.array(bad.materials[1], np.uint32)
    rec[2] |= 1 << 7
    bad.set_material(1, rec)
    with pytest.raises(hip.LotusError, match="a slot its model does not read"):
        renderer.apply(bad)
    # opacity and its threshold lie in [0, 1] (ValidGeometry, gpu_scene.cpp:821-831)
    bad = sc.copy()
    rec = np.array(bad.materials[1], np.uint32)
    rec.view(np.float32)[6] = 1.5
    bad.set_material(1, rec)
    with pytest.raises(hip.LotusError, match="opacity and its threshold"):
        renderer.apply(bad)
