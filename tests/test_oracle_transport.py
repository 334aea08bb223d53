"""Pins the oracle's transport: closed forms of the reference's evidence
scenes and the reference's committed Cornell images (DES-Q5)."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from oracle import desq5
from lotus_amd import scenes

from conftest import GOLDEN


def _mean_image(acc, background=(0.0, 0.0, 0.0)):
    """path_trace.slang:122-136: missed samples count as the background."""
    n = acc[..., 3:4]
    return acc[..., :3] / np.maximum(n, 1)


def test_lambert_under_sky_exact():
    # main.cpp:1525-1545: every hit sample is albedo * environment + emission.
    sc = scenes.lambert_under_sky()
    o = oracle.OracleScene(sc)
    f = oracle.make_frame(sc.world_to_clip, 48, 48, spp=4, max_bounces=64, light_sampling=0)
    acc, counts, _ = o.render(f)
    hit = acc[..., 3] > 0
    assert hit.sum() > 100
    img = _mean_image(acc)[hit]
    expect = np.array([0.5 * 0.8 + 0.1, 0.25 * 0.6 + 0.05, 0.75 * 0.4])
    assert np.allclose(img, expect, rtol=2e-6)
    # bounces = 0 keeps only the emission (main.cpp:1540)
    f0 = oracle.make_frame(sc.world_to_clip, 48, 48, spp=1, max_bounces=0, light_sampling=0)
    acc0, _, _ = o.render(f0)
    assert np.allclose(_mean_image(acc0)[acc0[..., 3] > 0], [0.1, 0.05, 0.0], atol=1e-7)


@pytest.mark.parametrize("bounces", [0, 1, 3])
def test_closed_box_bounce_limited_exact(bounces):
    # main.cpp:3073-3100: L = Le (1 + a + ... + a^n), BSDF sampling.
    sc = scenes.closed_box()
    o = oracle.OracleScene(sc)
    f = oracle.make_frame(sc.world_to_clip, 32, 32, spp=2, max_bounces=bounces, light_sampling=0)
    acc, counts, _ = o.render(f)
    assert np.all(acc[..., 3] == 2)  # every camera ray hits the inside
    a = np.array([0.5, 0.25, 0.75])
    e = np.array([0.1, 0.15, 0.025])
    expect = sum(e * a ** k for k in range(bounces + 1))
    assert np.allclose(_mean_image(acc), expect, rtol=1e-5)
    # ray-queue occupancy: every path continues until the bounce limit
    assert list(counts[0][: bounces + 1]) == [32 * 32] * (bounces + 1)


def test_closed_box_roulette_mean():
    # main.cpp:3102-3107: the unlimited mean is Le / (1 - a).
    sc = scenes.closed_box()
    o = oracle.OracleScene(sc)
    f = oracle.make_frame(sc.world_to_clip, 32, 32, spp=64, max_bounces=1024, light_sampling=0)
    acc, _, _ = o.render(f)
    img = _mean_image(acc).mean(axis=(0, 1))
    expect = np.array([0.1, 0.15, 0.025]) / (1 - np.array([0.5, 0.25, 0.75]))
    # standard error of the image mean: per-pixel variance estimate / pixels
    n = acc[..., 3].sum()
    var = (acc[..., 4:7].sum(axis=(0, 1)) / n - img ** 2) / n
    assert np.all(np.abs(img - expect) < 5 * np.sqrt(var) + 1e-4)


def test_estimators_agree_on_closed_box():
    # NEE + MIS converges to the BSDF estimator's (and the closed form's) mean.
    sc = scenes.closed_box()
    o = oracle.OracleScene(sc)
    f = oracle.make_frame(sc.world_to_clip, 32, 32, spp=16, max_bounces=3, light_sampling=1)
    acc, _, shadow = o.render(f)
    assert shadow.sum() > 0
    a = np.array([0.5, 0.25, 0.75])
    e = np.array([0.1, 0.15, 0.025])
    expect = sum(e * a ** k for k in range(4))
    img = _mean_image(acc).mean(axis=(0, 1))
    assert np.allclose(img, expect, rtol=0.02)


def _cornell_reference():
    mean = desq5.read_pfm(os.path.join(GOLDEN, "cornell-box-mean.pfm"))
    var = desq5.read_pfm(os.path.join(GOLDEN, "cornell-box-variance.pfm"))
    return mean, var


@pytest.mark.parametrize("light_sampling", [0, 1, 2])
def test_cornell_matches_reference_images(light_sampling):
    # light_sampling 2: ReSTIR's initial reservoir (8 candidates) for the
    # first hit's direct light, partitioned against BSDF sampling.
    # validation/reference/cornell-box-{mean,variance}.pfm: 128^2, 1024 spp,
    # sample index 0 (docs/reference/SCENE.md:1014-1024); DES-Q5 at 64 spp.
    mean, var = _cornell_reference()
    sc = scenes.cornell_v1()
    o = oracle.OracleScene(sc)
    f = oracle.make_frame(sc.world_to_clip, 128, 128, spp=64, max_bounces=64,
                          light_sampling=light_sampling)
    acc, counts, _ = o.render(f)
    assert np.all(acc[..., 3] == 64)  # every camera ray enters the box
    img = _mean_image(acc).astype(np.float32)
    c = desq5.compare(img, 64, mean, var, 1024)
    assert c.failure == "", c.failure


def test_cornell_negative_control_fails():
    # main.cpp:3487-3573: a 10 % brighter red wall must not match.
    mean, var = _cornell_reference()
    sc = scenes.cornell_v1(red_scale=1.1)
    o = oracle.OracleScene(sc)
    f = oracle.make_frame(sc.world_to_clip, 128, 128, spp=256, max_bounces=64, light_sampling=1)
    acc, _, _ = o.render(f)
    c = desq5.compare(_mean_image(acc).astype(np.float32), 256, mean, var, 1024)
    assert c.failure != ""


# ------------------------------------------------ pins against oracle/_ref --
needs_ref = pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built (no reference sources)")


@needs_ref
def test_desq5_restatement_matches_reference_compare():
    mean, var = _cornell_reference()
    R = oracle.ref()
    rng = np.random.default_rng(0)
    for scale, spp in [(1.0, 64), (1.0, 1), (1.03, 256), (0.9, 16)]:
        img = (mean * scale + rng.normal(0, 1, mean.shape) * np.sqrt(var / spp)).astype(np.float32)
        c = desq5.compare(img, spp, mean, var, 1024)
        z = np.zeros(3)
        mz = np.zeros(1)
        mz2 = np.zeros(1)
        tile = np.zeros(1, np.uint32)
        fail = C.create_string_buffer(512)
        rc = R.ref_compare(oracle.ptr(img, C.c_float), 128, 128, spp, oracle.ptr(mean, C.c_float),
                           oracle.ptr(var, C.c_float), 1024, oracle.ptr(z, C.c_double),
                           oracle.ptr(mz, C.c_double), oracle.ptr(mz2, C.c_double),
                           oracle.ptr(tile, C.c_uint32), fail, 512)
        assert (rc == 0) == (c.failure == "")
        assert np.allclose(z, c.image_z, rtol=1e-9)
        assert abs(mz[0] - c.max_tile_z) < 1e-9 * max(1, mz[0])
        assert tile[0] == c.tile


@needs_ref
def test_pfm_reader_matches_reference():
    R = oracle.ref()
    path = os.path.join(GOLDEN, "cornell-box-mean.pfm").encode()
    w = np.zeros(1, np.uint32)
    h = np.zeros(1, np.uint32)
    assert R.ref_read_pfm(path, oracle.ptr(w, C.c_uint32), oracle.ptr(h, C.c_uint32), None) == 0
    out = np.zeros((h[0], w[0], 3), np.float32)
    R.ref_read_pfm(path, oracle.ptr(w, C.c_uint32), oracle.ptr(h, C.c_uint32), oracle.ptr(out, C.c_float))
    assert np.array_equal(out, desq5.read_pfm(path.decode()))


@needs_ref
@pytest.mark.parametrize("red_scale", [1.0, 1.1])
def test_cornell_fixture_bit_identical_to_reference(red_scale):
    R = oracle.ref()
    n = R.ref_cornell_scene(red_scale, None, 0, None)
    buf = np.zeros(n, np.uint32)
    cam = np.zeros(32, np.float32)
    R.ref_cornell_scene(red_scale, oracle.ptr(buf, C.c_uint32), n, oracle.ptr(cam, C.c_float))
    sc = scenes.cornell_v1(red_scale=red_scale)
    k = 0
    for g, t, m in sc.instances:
        pos, tri = sc.geometries[g]
        nv, nt = int(buf[k]), int(buf[k + 1])
        k += 2
        assert np.array_equal(buf[k:k + 3 * nv], pos.reshape(-1).view(np.uint32))
        k += 3 * nv
        assert np.array_equal(buf[k:k + 3 * nt], tri.reshape(-1))
        k += 3 * nt
        assert np.array_equal(buf[k:k + 16], t.view(np.uint32))
        k += 16
        f = sc.materials[m].view(np.float32)
        mine = np.array([*f[12:15], f[15], f[19], f[23], *f[16:19]], np.float32)
        assert np.array_equal(buf[k:k + 9], mine.view(np.uint32))
        k += 9
    assert k == n
    assert np.array_equal(cam[:16], sc.view) and np.array_equal(cam[16:], sc.projection)
