"""Pins the oracle's material texture lookups (scene.slang:318-374, 572-577,
1184-1221; gpu_scene.cpp:144-179, 612-629) against an independent numpy
restatement in float64: bilinear filtering at (s', 1 - t') of the binding's
coordinate rows, the four wrap modes, scale and bias, the fallback mode,
sRGB decoding, the channel a scalar input reads, per-corner texture
coordinates and the material's fallback coordinates; and by closed forms
(a constant texture is the constant material, a textured emitter's light
weight is scale + bias)."""
import math

import numpy as np
import pytest

import oracle
from lotus_amd import scenes


def _np_wrap(i, n, mode):
    if mode == scenes.WRAP_REPEAT:
        return i % n, False
    if mode == scenes.WRAP_MIRROR:
        r = i % (2 * n)
        return (r if r < n else 2 * n - 1 - r), False
    if i < 0 or i >= n:
        return min(max(i, 0), n - 1), mode == scenes.WRAP_BLACK
    return i, False


def _np_lookup(texels, binding, s, t):
    f = binding.view(np.float32).astype(np.float64)
    if binding[3] == 2:
        return f[12:16]
    ss = f[16] * s + f[17] * t + f[18]
    tt = 1.0 - (f[20] * s + f[21] * t + f[22])
    H, W = texels.shape[:2]
    x, y = ss * W - 0.5, tt * H - 0.5
    x0, y0 = math.floor(x), math.floor(y)
    fx, fy = x - x0, y - y0
    ws, wt = binding[1] >> 2, binding[1] & 3

    def tex(i, j):
        xi, bx = _np_wrap(i, W, ws)
        yj, by = _np_wrap(j, H, wt)
        return np.zeros(4) if (bx or by) else texels[yj, xi].astype(np.float64)
    v = ((1 - fx) * tex(x0, y0) + fx * tex(x0 + 1, y0)) * (1 - fy) + \
        ((1 - fx) * tex(x0, y0 + 1) + fx * tex(x0 + 1, y0 + 1)) * fy
    return v * f[4:8] + f[8:12]


def _srgb(b):
    v = b / 255.0
    return np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)


def _textured_quad(wrap, fmt, st_rows, channel=1, fallback_only=False):
    rng = np.random.default_rng(9)
    sc = scenes.Scene()
    if fmt == scenes.TEXTURE_RGBA32F:
        texels = rng.uniform(0.0, 2.0, (5, 7, 4)).astype(np.float32)
        linear = texels.astype(np.float64)
    else:
        texels = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)
        linear = texels / 255.0
        if fmt == scenes.TEXTURE_RGBA8_SRGB:
            linear = linear.copy()
            linear[..., :3] = _srgb(texels[..., :3].astype(np.float64))
    tex = sc.add_texture(texels, fmt)
    binds = [scenes.texture_binding(tex, wrap=wrap, channel=channel, scale=(0.9, 0.8, 0.7, 0.6),
                                    bias=(0.05, 0.0, 0.1, 0.02), st_s=st_rows[0], st_t=st_rows[1],
                                    fallback=(0.3, 0.2, 0.1, 0.4), fallback_only=fallback_only)
             for _ in range(5)]
    first = sc.add_binding(binds[0])
    for b in binds[1:]:
        sc.add_binding(b)
    mat = scenes.with_lookups(scenes.preview_material(base_color=(0.5, 0.5, 0.5)), first,
                              [scenes.SLOT_BASE_COLOR, scenes.SLOT_ROUGHNESS, scenes.SLOT_METALLIC,
                               scenes.SLOT_EMISSION, scenes.SLOT_SPECULAR_COLOR])
    m = sc.add_material(mat)
    sc.add_mesh([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], [[0, 1, 2], [0, 2, 3]], m)
    corners = rng.uniform(-1.5, 2.5, (6, 2)).astype(np.float32)
    sc.set_texcoords(0, corners)
    return sc, linear, binds[0], corners


@pytest.mark.parametrize("wrap", [(scenes.WRAP_REPEAT, scenes.WRAP_CLAMP), (scenes.WRAP_MIRROR, scenes.WRAP_BLACK),
                                  (scenes.WRAP_CLAMP, scenes.WRAP_REPEAT), (scenes.WRAP_BLACK, scenes.WRAP_MIRROR)])
@pytest.mark.parametrize("fmt", [scenes.TEXTURE_RGBA8_UNORM, scenes.TEXTURE_RGBA8_SRGB, scenes.TEXTURE_RGBA32F])
def test_lookups_match_numpy(wrap, fmt):
    st_rows = ((1.3, 0.2, 0.1), (-0.4, 0.9, 0.25))
    sc, linear, binding, corners = _textured_quad(wrap, fmt, st_rows)
    o = oracle.OracleScene(sc)
    rng = np.random.default_rng(1)
    for _ in range(200):
        prim = int(rng.integers(0, 2))
        u, v = rng.uniform(0, 1, 2)
        if u + v > 1:
            u, v = 1 - u, 1 - v
        u, v = np.float32(u), np.float32(v)
        c = corners[3 * prim:3 * prim + 3].astype(np.float64)
        s, t = c[0] + u * (c[1] - c[0]) + v * (c[2] - c[0])
        look = _np_lookup(linear, binding, s, t)
        got = o.material_inputs(0, prim, float(u), float(v))
        sat = np.clip(look, 0, 1)
        assert np.allclose(got[0:3], sat[:3], atol=2e-5), (got[:3], sat[:3])          # base colour
        assert np.isclose(got[3], sat[1], atol=2e-5)                                  # roughness: channel 1
        assert np.allclose(got[4:7], np.maximum(look[:3], 0), atol=2e-5)              # emission
        assert np.isclose(got[7], sat[1], atol=2e-5)                                  # metallic
        assert np.allclose(got[8:11], sat[:3], atol=2e-5)                             # specular colour


def test_fallback_lookups_and_fallback_coordinates():
    # a binding in fallback mode returns its fallback unscaled; a geometry
    # without coordinates reads the material's (words 4-5)
    sc, linear, binding, _ = _textured_quad((2, 2), scenes.TEXTURE_RGBA32F, ((1, 0, 0), (0, 1, 0)),
                                            fallback_only=True)
    got = oracle.OracleScene(sc).material_inputs(0, 0, 0.2, 0.3)
    assert np.allclose(got[0:3], [0.3, 0.2, 0.1]) and np.isclose(got[3], 0.2)
    sc2, linear2, binding2, _ = _textured_quad((2, 2), scenes.TEXTURE_RGBA32F, ((1, 0, 0), (0, 1, 0)))
    sc2.texcoords = {}
    rec = np.array(sc2.materials[1])
    rec.view(np.float32)[4:6] = (0.37, 0.61)
    sc2.set_material(1, rec)
    got = oracle.OracleScene(sc2).material_inputs(0, 1, 0.5, 0.25)
    look = _np_lookup(linear2, binding2, 0.37, 0.61)
    assert np.allclose(got[0:3], np.clip(look[:3], 0, 1), atol=2e-5)


def test_constant_texture_is_the_constant_material():
    # main.cpp:1525-1545's Lambert triangle under a constant sky, its base
    # colour now a lookup of a constant texture: the same image to 1 ulp
    sc = scenes.lambert_under_sky()
    tex = np.zeros((4, 4, 4), np.float32)
    tex[...] = (0.5, 0.25, 0.75, 1.0)
    t = sc.add_texture(tex, scenes.TEXTURE_RGBA32F)
    b = sc.add_binding(scenes.texture_binding(t))
    sc.set_material(1, scenes.with_lookups(sc.materials[1], b, [scenes.SLOT_BASE_COLOR]))
    o = oracle.OracleScene(sc)
    f = oracle.make_frame(sc.world_to_clip, 48, 48, spp=2, max_bounces=64, light_sampling=0)
    acc, _, _ = o.render(f)
    hit = acc[..., 3] > 0
    img = acc[..., :3][hit] / acc[..., 3:4][hit]
    assert np.allclose(img, [0.5 * 0.8 + 0.1, 0.25 * 0.6 + 0.05, 0.75 * 0.4], rtol=1e-6)


def test_textured_emitter_light_weight():
    # gpu_scene.cpp:144-179: a textured emission's weight is max(scale, 0) +
    # max(bias, 0) per channel (the texture's values do not enter)
    sc = scenes.lit_floor(None)
    tex = np.full((2, 2, 4), 7.0, np.float32)
    t = sc.add_texture(tex, scenes.TEXTURE_RGBA32F)
    b = sc.add_binding(scenes.texture_binding(t, scale=(2.0, 1.0, 0.0, 1.0), bias=(0.5, -1.0, 1.0, 0.0)))
    glow = sc.add_material(scenes.with_lookups(scenes.preview_material(base_color=(0, 0, 0), ior=1.0), b,
                                               [scenes.SLOT_EMISSION]))
    sc.add_mesh([[0, 0.5, 0], [1, 0.5, 0], [0, 0.5, 2]], [[0, 1, 2]], glow)
    emissive = scenes.preview_material(base_color=(0, 0, 0), ior=1.0, emission=(3.0, 3.0, 3.0))
    e2 = sc.add_material(emissive)
    sc.add_mesh([[0, 0.6, 0], [1, 0.6, 0], [0, 0.6, 2]], [[0, 1, 2]], e2)
    offsets, records = oracle.OracleScene(sc).light_table()
    pdf = records[:, 0].view(np.float32)
    w1 = (2.5 + 1.0 + 1.0) / 3.0  # (2 + 0.5, 1 + 0, 0 + 1) over 3, area 1
    assert np.allclose(pdf, np.array([w1, 3.0]) / (w1 + 3.0), rtol=2e-6)
