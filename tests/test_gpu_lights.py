"""Analytic lights on the GPU against the oracle (scene.slang:682-1182,
transport.slang:82-247, gpu_scene.cpp:271-609).

Bars: the light table bitwise; per-sample queue counts and shadow-ray counts
exact; one-sample radiance bit-exact for >= 99.9 % of pixels under BSDF, MIS
and ReSTIR sampling; accumulations to 1e-4; the closed-form floor scenes the
oracle is pinned by, reached by the GPU's own estimates.
"""
import math
import sys

import numpy as np
import pytest

import oracle
from lotus_amd import hip, scenes

from test_oracle_lights import CASES

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


def lit_cornell(parallel=1):
    """The Lambert Cornell box with every analytic shape: a sphere and a disk
    inside, a rect on the back wall, a distant light and `parallel` parallel
    lights shining in through the open front."""
    sc = scenes.cornell_lambert()
    sc.add_light(scenes.analytic_light(scenes.LIGHT_SPHERE, scenes.translation(0.35, 0.45, 0.25), radiance=(6, 5, 4),
                                       radius=0.12))
    back = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, -0.3, 0.2, -0.97, 1], np.float32)  # emits towards +z
    sc.add_light(scenes.analytic_light(scenes.LIGHT_RECT, back, radiance=(2, 3, 4), width=0.4, height=0.3))
    sc.add_light(scenes.analytic_light(scenes.LIGHT_DISK, scenes.facing_down(-0.4, 0.9, -0.2), radiance=(5, 5, 2),
                                       radius=0.1, normalize=True))
    sc.add_light(scenes.analytic_light(scenes.LIGHT_DISTANT, None, radiance=(1, 1, 1.5), angle=25.0))
    for k in range(parallel):
        tilt = scenes.identity()
        tilt[8:11] = (0.3 * (k + 1), 0.4, 1.0)  # shines along -Z: into the box
        sc.add_light(scenes.analytic_light(scenes.LIGHT_DISTANT, tilt, radiance=(0.5, 0.4, 0.3), angle=0.0))
    return sc


def test_light_table_matches_oracle(renderer):
    sc = lit_cornell()
    ev = renderer.apply(sc)
    off_g, rec_g = renderer.light_table()
    off_c, rec_c = oracle.OracleScene(sc).light_table()
    assert ev.light_count == len(rec_c) == 2 + 5
    assert np.array_equal(off_g, off_c) and np.array_equal(rec_g, rec_c)
    # a light edit rewrites the tables only
    moved = sc.copy()
    moved.lights[0]["world_from_light"][13] = 0.3
    moved._packed = None
    ev = renderer.apply(moved)
    assert ev.update == hip.SCENE_TABLES
    off_g, rec_g = renderer.light_table()
    off_c, rec_c = oracle.OracleScene(moved).light_table()
    assert np.array_equal(off_g, off_c) and np.array_equal(rec_g, rec_c)


@pytest.mark.parametrize("light", [hip.LIGHT_BSDF, hip.LIGHT_MIS, hip.LIGHT_RESTIR])
def test_lit_cornell_counts_and_radiance(renderer, light):
    _say(f"lit cornell light={light}")
    sc = lit_cornell()
    renderer.apply(sc)
    w = h = 96
    spp = 6
    o = oracle.OracleScene(sc)
    f = oracle.make_frame(sc.world_to_clip, w, h, spp=spp, max_bounces=8, light_sampling=light)
    acc, counts, shadow = o.render(f)
    first = None
    for k in range(spp):
        out = renderer.render(sc.world_to_clip, w, h, frame_count=1, max_bounces=8,
                              light_sampling=light)
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


@pytest.mark.parametrize("light_sampling", [hip.LIGHT_BSDF, hip.LIGHT_MIS])
@pytest.mark.parametrize("shape", ["sphere", "rect", "distant", "parallel"])
def test_floor_closed_forms_on_gpu(renderer, shape, light_sampling):
    make, expect = CASES[shape]
    sc = scenes.lit_floor(make())
    renderer.apply(sc)
    frames = 512
    out = renderer.render(sc.world_to_clip, 16, 16, frame_count=frames, max_bounces=1, light_sampling=light_sampling)
    px = out["color"][..., :3].reshape(-1, 3).astype(np.float64)
    se = px.std(0) / math.sqrt(px.shape[0])
    assert np.all(np.abs(px.mean(0) - expect) <= 4 * se + 1e-4 * expect), (shape, px.mean(0), expect, se)


@pytest.mark.parametrize("parallel", [2, 3])
def test_bsdf_sampling_walks_every_parallel_light(renderer, parallel):
    # ParallelLights (transport.slang:139-146) with several parallel lights:
    # the shading lanes (k_shade and, past the rounds, k_tail) walk each
    # light's shadow ray themselves. Paths of up to 24 bounces reach the tail.
    _say(f"bsdf sampling, {parallel} parallel lights")
    sc = lit_cornell(parallel=parallel)
    renderer.apply(sc)
    w = h = 64
    spp, bounces = 4, 24
    o = oracle.OracleScene(sc)
    acc, counts, shadow = o.render(oracle.make_frame(sc.world_to_clip, w, h, spp=spp, max_bounces=bounces,
                                                     light_sampling=0))
    first, tails = None, 0
    for k in range(spp):
        out = renderer.render(sc.world_to_clip, w, h, frame_count=1, max_bounces=bounces,
                              light_sampling=hip.LIGHT_BSDF)
        n = len(out["counts"])
        assert np.array_equal(out["counts"], counts[k][:n]) and not counts[k][n:].any(), (k, out["counts"], counts[k])
        assert out["shadow_rays"] == shadow[k], (k, out["shadow_rays"], shadow[k])
        tails += out["rounds"] < n
        if first is None:
            first = out["color"]
    assert tails > 0  # the tail shaded some of the paths
    acc1, _, _ = o.render(oracle.make_frame(sc.world_to_clip, w, h, spp=1, max_bounces=bounces, light_sampling=0))
    ref1 = acc1[..., :3].astype(np.float32)
    hit = acc1[..., 3] > 0
    exact = np.mean(np.all(first[..., :3][hit] == ref1[hit], -1))
    _say(f"  1-spp radiance bit-exact fraction {exact:.6f}")
    assert exact > 0.999
    assert np.allclose(first[..., :3], ref1, rtol=1e-4, atol=1e-5)


def sky_cornell():
    """The lit Cornell box under a sky image: every escaping ray looks it up,
    and light sampling chooses it from the light table."""
    sc = lit_cornell()
    sc.environment_map = scenes.sky_image()
    sc.environment = (0.05, 0.05, 0.05)
    return sc


def test_environment_light_table_matches_oracle(renderer):
    sc = sky_cornell()
    renderer.apply(sc)
    off_g, rec_g = renderer.light_table()
    off_c, rec_c = oracle.OracleScene(sc).light_table()
    assert rec_c[-1, 3] == 0xFFFFFFFF
    assert np.array_equal(off_g, off_c) and np.array_equal(rec_g, rec_c)
    # a new image is a table rewrite; removing it too
    brighter = sc.copy()
    brighter.environment_map["scale"] = (2.0, 2.0, 2.0)
    brighter._packed = None
    assert renderer.apply(brighter).update == hip.SCENE_TABLES
    assert np.array_equal(renderer.light_table()[1], oracle.OracleScene(brighter).light_table()[1])
    none = sc.copy()
    none.environment_map = None
    none._packed = None
    assert renderer.apply(none).update == hip.SCENE_TABLES
    assert np.array_equal(renderer.light_table()[1], oracle.OracleScene(none).light_table()[1])


@pytest.mark.parametrize("light", [hip.LIGHT_BSDF, hip.LIGHT_MIS, hip.LIGHT_RESTIR])
def test_sky_cornell_counts_and_radiance(renderer, light):
    _say(f"sky cornell light={light}")
    sc = sky_cornell()
    renderer.apply(sc)
    w = h = 96
    o = oracle.OracleScene(sc)
    acc, counts, shadow = o.render(oracle.make_frame(sc.world_to_clip, w, h, spp=3, max_bounces=8,
                                                     light_sampling=light))
    first = None
    for k in range(3):
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
    assert n == 3
    assert np.allclose(sums[..., :3], acc[..., :3], rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("light_sampling", [hip.LIGHT_BSDF, hip.LIGHT_MIS])
def test_floor_under_constant_environment_on_gpu(renderer, light_sampling):
    sc = scenes.lit_floor(None)
    tex = np.ones((16, 32, 4), np.float32)
    tex[..., :3] = (0.7, 0.5, 0.3)
    sc.environment_map = dict(texels=tex, scale=(2.0, 1.0, 0.5), transform=scenes.identity())
    renderer.apply(sc)
    expect = 0.5 * np.array([0.7, 0.5, 0.3]) * np.array([2.0, 1.0, 0.5])
    out = renderer.render(sc.world_to_clip, 16, 16, frame_count=256, max_bounces=1, light_sampling=light_sampling)
    px = out["color"][..., :3].reshape(-1, 3).astype(np.float64)
    if light_sampling == hip.LIGHT_BSDF:
        assert np.allclose(px, expect, rtol=1e-5)
    else:
        se = px.std(0) / math.sqrt(px.shape[0])
        assert np.all(np.abs(px.mean(0) - expect) <= 4 * se + 1e-6), (px.mean(0), expect, se)


def _srgb_bytes(linear):
    """IEC 61966-2-1 encoding of linear values in [0, 1] to bytes."""
    v = np.clip(linear, 0.0, 1.0)
    e = np.where(v <= 0.0031308, v * 12.92, 1.055 * v ** (1.0 / 2.4) - 0.055)
    return np.round(e * 255.0).astype(np.uint8)


def test_srgb_environment_image_matches_oracle(renderer):
    # an 8-bit sRGB sky (EnvironmentMap over an Rgba8Srgb Texture): decoded as
    # TexelRgb decodes it; its light table, counts and radiance against the
    # oracle, and the same scene as its decoded RGBA32F image
    sc = sky_cornell()
    tex = sc.environment_map["texels"]
    texels8 = np.full(tex.shape, 255, np.uint8)
    texels8[..., :3] = _srgb_bytes(tex[..., :3] / 64.0)
    sc.environment_map = dict(texels=texels8, format=scenes.TEXTURE_RGBA8_SRGB, scale=(64.0, 64.0, 64.0),
                              transform=scenes.identity())
    renderer.apply(sc)
    o = oracle.OracleScene(sc)
    assert np.array_equal(renderer.light_table()[1], o.light_table()[1])
    w = h = 96
    out = renderer.render(sc.world_to_clip, w, h, frame_count=1, max_bounces=8, light_sampling=hip.LIGHT_MIS)
    acc, counts, shadow = o.render(oracle.make_frame(sc.world_to_clip, w, h, spp=1, max_bounces=8,
                                                     light_sampling=hip.LIGHT_MIS))
    n = len(out["counts"])
    assert np.array_equal(out["counts"], counts[0][:n]) and out["shadow_rays"] == shadow[0]
    hit = acc[..., 3] > 0
    exact = np.mean(np.all(out["color"][..., :3][hit] == acc[..., :3].astype(np.float32)[hit], -1))
    _say(f"  sRGB sky 1-spp radiance bit-exact fraction {exact:.6f}")
    assert exact > 0.999
    decoded = sc.copy()
    decoded.environment_map = dict(sc.environment_map, texels=oracle.decode_environment(texels8, 1))
    decoded._packed = None
    assert renderer.apply(decoded).update == hip.SCENE_UNCHANGED
    unorm = sc.copy()
    unorm.environment_map = dict(sc.environment_map, format=scenes.TEXTURE_RGBA8_UNORM)
    unorm._packed = None
    assert renderer.apply(unorm).update == hip.SCENE_TABLES
    assert np.array_equal(renderer.light_table()[1], oracle.OracleScene(unorm).light_table()[1])
