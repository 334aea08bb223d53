"""Pins the oracle's analytic lights (scene.slang:682-1182, transport.slang:
82-247, gpu_scene.cpp:271-609) by closed forms and by estimator agreement.

A Lambert floor of albedo a seen straight down, lit by one light above it:
its outgoing radiance at the origin is a/pi times the light's irradiance
there, which the mathematics fixes for each shape:
* sphere of radius r at height h: E = pi L r^2 / h^2;
* disk of radius R facing down at height h: E = pi L R^2 / (h^2 + R^2);
* rect of half sizes A, B facing down at height c: E = 2 L (X atan(Y') + Y
  atan(X')) with X = A / sqrt(A^2 + c^2) ... (the form factor of a
  rectangle from a point below its centre, summed over its four quadrants);
* distant light with UsdLux normalize: E = L (any angle up to 180 deg);
* parallel light: E = L, and every sample carries it exactly.
Both estimators must reach the closed form: BSDF sampling alone finds the
light only through CrossedLights, MIS mostly through SampleAnalytic; a
wrong pdf, a wrong MIS weight or a missed crossing fails one of them.
"""
import math

import numpy as np
import pytest

import oracle
from lotus_amd import scenes

L = np.array([4.0, 2.0, 1.0])
ALBEDO = 0.5
H = 2.0


def _rect_irradiance(a, b, c):
    x, y = a / c, b / c
    return 2.0 * (x / math.sqrt(1 + x * x) * math.atan(y / math.sqrt(1 + x * x)) +
                  y / math.sqrt(1 + y * y) * math.atan(x / math.sqrt(1 + y * y)))


CASES = {
    "sphere": (lambda: scenes.analytic_light(scenes.LIGHT_SPHERE, scenes.translation(0, H, 0), radiance=L,
                                             radius=0.8), ALBEDO * L * 0.8 ** 2 / H ** 2),
    "disk": (lambda: scenes.analytic_light(scenes.LIGHT_DISK, scenes.facing_down(0, H, 0), radiance=L, radius=0.7),
             ALBEDO * L * 0.7 ** 2 / (H * H + 0.7 ** 2)),
    "rect": (lambda: scenes.analytic_light(scenes.LIGHT_RECT, scenes.facing_down(0, H, 0), radiance=L, width=1.2,
                                           height=0.8), ALBEDO / math.pi * L * _rect_irradiance(0.6, 0.4, H)),
    "distant": (lambda: scenes.analytic_light(scenes.LIGHT_DISTANT, scenes.facing_down(0, 0, 0), radiance=L,
                                              angle=40.0, normalize=True), ALBEDO * L / math.pi),
    "parallel": (lambda: scenes.analytic_light(scenes.LIGHT_DISTANT, scenes.facing_down(0, 0, 0), radiance=L,
                                               angle=0.0), ALBEDO * L / math.pi),
}


def _pixel_means(o, sc, light_sampling, spp=512, size=16):
    f = oracle.make_frame(sc.world_to_clip, size, size, spp=spp, max_bounces=1, light_sampling=light_sampling)
    acc, counts, shadow = o.render(f)
    assert np.all(acc[..., 3] == spp)  # every camera ray hits the floor
    return (acc[..., :3] / spp).reshape(-1, 3), shadow


@pytest.mark.parametrize("light_sampling", [0, 1])
@pytest.mark.parametrize("shape", list(CASES))
def test_floor_under_analytic_light_closed_form(shape, light_sampling):
    make, expect = CASES[shape]
    sc = scenes.lit_floor(make())
    o = oracle.OracleScene(sc)
    px, shadow = _pixel_means(o, sc, light_sampling)
    mean = px.mean(0)
    # four standard errors of the pixel means, plus 1e-4 for the pixels'
    # spread around the origin (under 0.02 of it at height 2)
    se = px.std(0) / math.sqrt(px.shape[0])
    assert np.all(np.abs(mean - expect) <= 4 * se + 1e-4 * expect), (shape, light_sampling, mean, expect, se)
    if shape == "parallel":
        assert np.allclose(px, expect, rtol=1e-5)  # a delta light: every sample is exact
        assert shadow.sum() == px.shape[0] * 512  # one shadow ray per hit, either estimator
    elif light_sampling == 0:
        assert shadow.sum() == 0


def _parallel(tilt_degrees, radiance):
    t = math.radians(tilt_degrees)
    frame = scenes.identity()
    frame[8:11] = (math.sin(t), math.cos(t), 0.0)  # towards the light; it shines along -Z
    return scenes.analytic_light(scenes.LIGHT_DISTANT, frame, radiance=radiance, angle=0.0)


def test_floor_under_several_parallel_lights_bsdf_sampling():
    # BSDF sampling finds no parallel light: ParallelLights (transport.slang:
    # 139-146) adds each one's exact term through a shadow ray of its own, so
    # every sample is a/pi * sum_j L_j cos(theta_j) and each hit sends one
    # shadow ray per light.
    lights = [(0.0, L), (50.0, np.array([1.0, 3.0, 0.5])), (-70.0, np.array([0.5, 0.5, 2.0]))]
    sc = scenes.lit_floor(_parallel(*lights[0]))
    for tilt, radiance in lights[1:]:
        sc.add_light(_parallel(tilt, radiance))
    o = oracle.OracleScene(sc)
    px, shadow = _pixel_means(o, sc, 0, spp=16)
    expect = ALBEDO / math.pi * sum(r * math.cos(math.radians(t)) for t, r in lights)
    assert np.allclose(px, expect, rtol=1e-5)
    assert shadow.sum() == len(lights) * px.shape[0] * 16


def test_many_sphere_lights_estimators_agree():
    # Seven spheres over the floor: the crossing hierarchy (three levels) must
    # find every one that BSDF sampling hits, as light sampling does.
    rng = np.random.default_rng(3)
    sc = scenes.lit_floor(scenes.analytic_light(scenes.LIGHT_SPHERE, scenes.translation(0.0, 1.5, 0.0),
                                                radiance=L, radius=0.3))
    for k in range(6):
        x, z = rng.uniform(-1.5, 1.5, 2)
        sc.add_light(scenes.analytic_light(scenes.LIGHT_SPHERE, scenes.translation(x, rng.uniform(1, 2.5), z),
                                           radiance=L * rng.uniform(0.5, 2.0), radius=rng.uniform(0.2, 0.5)))
    o = oracle.OracleScene(sc)
    bsdf, _ = _pixel_means(o, sc, 0, spp=1024)
    mis, _ = _pixel_means(o, sc, 1, spp=256)
    se = math.hypot(bsdf.std(0)[0] / math.sqrt(bsdf.shape[0]), mis.std(0)[0] / math.sqrt(mis.shape[0]))
    assert abs(bsdf.mean(0)[0] - mis.mean(0)[0]) <= 4 * se, (bsdf.mean(0), mis.mean(0), se)


def test_light_table_weights_closed_form():
    # gpu_scene.cpp:271-369, 512-609: an emissive triangle's weight is its
    # mean emission times area; an analytic light's its mean radiance times
    # its area (a rect's w h, a disk's pi R^2, a sphere's 4 pi r^2) over 2;
    # a normalized distant light's its irradiance over 2 times the scene
    # radius squared. Entries follow the triangles in the lights' order.
    sc = scenes.lit_floor(scenes.analytic_light(scenes.LIGHT_SPHERE, scenes.translation(0, 2, 0), radiance=L,
                                                radius=0.8))
    sc.add_light(scenes.analytic_light(scenes.LIGHT_RECT, scenes.facing_down(1, 2, 0), radiance=2 * L, width=1.2,
                                       height=0.8))
    sc.add_light(scenes.analytic_light(scenes.LIGHT_DISK, scenes.facing_down(-1, 2, 0), radiance=L, radius=0.5))
    sc.add_light(scenes.analytic_light(scenes.LIGHT_DISTANT, scenes.facing_down(0, 0, 0), radiance=L, angle=10.0,
                                       normalize=True))
    sc.add_light(scenes.analytic_light(scenes.LIGHT_SPHERE, scenes.translation(0, 1, 0), radiance=(0, 0, 0)))
    glow = sc.add_material(scenes.preview_material(base_color=(0, 0, 0), ior=1.0, emission=(3.0, 3.0, 3.0)))
    sc.add_mesh([[0, 0.5, 0], [1, 0.5, 0], [0, 0.5, 2]], [[0, 1, 2]], glow)
    o = oracle.OracleScene(sc)
    offsets, records = o.light_table()
    mean = L.mean()
    radius = 0.5 * math.sqrt(100.0 ** 2 + 0.5 ** 2 + 100.0 ** 2)  # the floor's and the triangle's world box
    weights = np.array([mean * 0.8 ** 2 * 4 * math.pi / 2, 2 * mean * 1.2 * 0.8 / 2, mean * math.pi * 0.25 / 2,
                        mean / 2 * radius ** 2, 3.0 * 1.0])
    # the black sphere emits nothing: no record, no entry
    assert list(offsets) == [0xFFFFFFFF, 0]
    assert len(records) == 5
    assert records[0, 3] == 1  # the triangle's entry names its instance
    assert list(records[1:, 3]) == [0x80000000, 0x80000001, 0x80000002, 0x80000003]
    order = [4, 0, 1, 2, 3]
    pdf = records[:, 0].view(np.float32)
    assert np.allclose(pdf, (weights / weights.sum())[order], rtol=2e-6)
    # the alias table: each entry's selection probability is its pdf
    n = len(records)
    thr, alias = records[:, 1].view(np.float32).astype(np.float64), records[:, 2]
    p = thr / n
    for i in range(n):
        p[alias[i]] += (1 - thr[i]) / n
    assert np.allclose(p, pdf, atol=1e-6)


def _constant_sky(c=(0.7, 0.5, 0.3), scale=(2.0, 1.0, 0.5)):
    tex = np.ones((16, 32, 4), np.float32)
    tex[..., :3] = c
    return dict(texels=tex, scale=scale, transform=scenes.identity())


def test_floor_under_constant_environment_image():
    # A constant image is a constant sky: the floor reflects albedo * c *
    # scale, exactly under BSDF sampling (every escaping direction sees c)
    # and within four standard errors under MIS, whose light-table half
    # samples the image's distribution (density ~ sin(theta), the uniform
    # sphere) and weighs both halves by the power heuristic.
    sc = scenes.lit_floor(None)
    sc.environment_map = _constant_sky()
    expect = ALBEDO * np.array([0.7, 0.5, 0.3]) * np.array([2.0, 1.0, 0.5])
    o = oracle.OracleScene(sc)
    px, shadow = _pixel_means(o, sc, 0, spp=64)
    assert np.allclose(px, expect, rtol=1e-6)
    assert shadow.sum() == 0
    px, shadow = _pixel_means(o, sc, 1, spp=256)
    se = px.std(0) / math.sqrt(px.shape[0])
    assert np.all(np.abs(px.mean(0) - expect) <= 4 * se), (px.mean(0), expect, se)
    # light sampling sends a shadow ray for the upper half of the sphere only
    assert 0.45 < shadow.sum() / (256 * 256) < 0.55


def test_environment_image_estimators_agree():
    # a sky with a sun 60 times brighter than the rest: BSDF sampling and MIS
    # must agree, which a wrong direction pdf, uv mapping or frame fails
    sc = scenes.lit_floor(None)
    sc.environment_map = scenes.sky_image()
    o = oracle.OracleScene(sc)
    bsdf, _ = _pixel_means(o, sc, 0, spp=2048)
    mis, _ = _pixel_means(o, sc, 1, spp=256)
    se = np.hypot(bsdf.std(0) / math.sqrt(bsdf.shape[0]), mis.std(0) / math.sqrt(mis.shape[0]))
    assert np.all(np.abs(bsdf.mean(0) - mis.mean(0)) <= 4 * se), (bsdf.mean(0), mis.mean(0), se)
    # and a tilted frame moves the sun's elevation: the estimate changes (a
    # turn about +Y would not: the floor's irradiance ignores azimuth)
    turned = sc.copy()
    t = scenes.identity()
    t[4:7], t[8:11] = (0.0, 0.0, 1.0), (0.0, -1.0, 0.0)  # a quarter turn about +X
    turned.environment_map["transform"] = t
    turned._packed = None
    mis2, _ = _pixel_means(oracle.OracleScene(turned), turned, 1, spp=256)
    assert not np.allclose(mis2.mean(0), mis.mean(0), rtol=0.02)


def test_environment_light_table_entry():
    # gpu_scene.cpp:502-510, 631-704: the image's entry comes last, weighted
    # pi^2 R^2 times the mean cell value (cells of mean radiance x sin theta
    # at the row centre), here against an emissive triangle of weight 3.
    sc = scenes.lit_floor(None)
    env = _constant_sky(c=(1.0, 1.0, 1.0), scale=(1.0, 1.0, 1.0))
    sc.environment_map = env
    glow = sc.add_material(scenes.preview_material(base_color=(0, 0, 0), ior=1.0, emission=(3.0, 3.0, 3.0)))
    sc.add_mesh([[0, 0.5, 0], [1, 0.5, 0], [0, 0.5, 2]], [[0, 1, 2]], glow)
    offsets, records = oracle.OracleScene(sc).light_table()
    rows = 16
    mean_cell = np.mean(np.sin(math.pi * (np.arange(rows) + 0.5) / rows))
    radius = 0.5 * math.sqrt(100.0 ** 2 + 0.5 ** 2 + 100.0 ** 2)
    w_env = math.pi ** 2 * radius ** 2 * mean_cell
    assert records[-1, 3] == 0xFFFFFFFF
    pdf = records[:, 0].view(np.float32)
    assert np.allclose(pdf, np.array([3.0, w_env]) / (3.0 + w_env), rtol=2e-6)


def test_environment_texel_decode_known_answers():
    # TexelRgb (gpu_scene.cpp:612-629): channel / 255; sRGB colour channels
    # through IEC 61966-2-1's decoding, alpha linear; UNorm linear
    b = np.array([[[0, 10, 128, 128], [255, 11, 200, 0]]], np.uint8)
    s = oracle.decode_environment(b, 1)
    assert s[0, 0, 0] == 0.0 and s[0, 1, 0] == 1.0
    assert s[0, 0, 1] == np.float32(10 / 255 / 12.92)  # the linear segment (10/255 <= 0.04045)
    assert abs(float(s[0, 0, 2]) - 0.21586050011389926) < 1e-7  # ((128/255 + 0.055) / 1.055)^2.4
    assert s[0, 0, 3] == np.float32(128 / 255) and s[0, 1, 3] == 0.0
    assert s[0, 1, 1] == np.float32(((11 / 255 + 0.055) / 1.055) ** 2.4)  # 11/255 > 0.04045
    u = oracle.decode_environment(b, 0)
    assert np.array_equal(u, (b / 255.0).astype(np.float32))
