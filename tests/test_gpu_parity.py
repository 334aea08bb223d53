"""GPU parity of the C-ABI path tracer against the CPU oracle.

Bars: closest-hit instance/primitive ids and integer ray-queue counts match
exactly; hit t and barycentrics are compared bit for bit (both sides run the
same watertight FP32 test without contraction) and must at least agree to
1e-5 relative; converged radiance passes DES-Q5 against the reference's
committed Cornell images.
"""
import sys

import numpy as np
import pytest

import oracle
from oracle import desq5
from lotus_amd import hip, scenes

from conftest import GOLDEN

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


def _random_rays(n, seed, spread=2.5, target=0.8):
    u = scenes.uniform(seed, 11, np.arange(n * 6)).reshape(n, 6)
    o = (u[:, :3] * 2 - 1) * spread
    t = (u[:, 3:] * 2 - 1) * target
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, 3.0e38
    return rays


def _compare_hits(gpu, cpu):
    assert np.array_equal(gpu[:, :2], cpu[:, :2]), f"{np.sum(np.any(gpu[:, :2] != cpu[:, :2], 1))} id mismatches"
    hit = cpu[:, 0] != hip.MISS
    g = gpu[hit, 2:].view(np.float32)
    c = cpu[hit, 2:].view(np.float32)
    assert np.allclose(g, c, rtol=1e-5, atol=1e-7)
    return np.mean(np.all(gpu[hit, 2:] == cpu[hit, 2:], 1)) if hit.any() else 1.0


def test_probe_reports_gfx950():
    status, detail = hip.probe()
    assert status == hip.PASS and "gfx950" in detail


@pytest.mark.parametrize("builder", [hip.BVH_LBVH, hip.BVH_SAH])
@pytest.mark.parametrize("count", [1, 7, 300, 20000])
def test_trace_random_triangles(renderer, count, builder):
    # closest hits and any-hit results do not depend on the tree: both
    # topologies give the oracle's ids, barycentrics and t bit for bit
    _say(f"trace random triangles {count}, builder {builder}")
    sc = scenes.random_triangles(count, seed=count)
    ev = renderer.apply(sc, builder=builder)
    assert ev.triangle_count == count
    o = oracle.OracleScene(sc)
    rays = _random_rays(20000, seed=count)
    g = renderer.trace(rays)
    c = o.trace(rays)
    exact = _compare_hits(g, c)
    assert exact == 1.0
    occ_g = renderer.trace(rays, any_hit=True)
    occ_c = o.occluded(rays)
    assert np.array_equal(occ_g, occ_c)


def test_trace_grid_shared_edges(renderer):
    _say("trace grid shared edges")
    sc = scenes.grid(divisions=64)
    renderer.apply(sc)
    o = oracle.OracleScene(sc)
    xs = np.linspace(-1, 1, 65).astype(np.float32)
    pts = np.array([(x, y) for y in xs[1:-1] for x in xs[1:-1]], np.float32)
    pts = np.concatenate([pts, pts + np.float32(1.0 / 64)])
    rays = np.zeros((len(pts), 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2] = pts[:, 0], pts[:, 1], 2.0
    rays[:, 6], rays[:, 7] = -1.0, 3.0e38
    g = renderer.trace(rays)
    assert np.all(g[:, 0] != hip.MISS)  # no cracks at shared edges and vertices
    _compare_hits(g, o.trace(rays))


def test_empty_scene_misses(renderer):
    sc = scenes.Scene()
    renderer.apply(sc)
    g = renderer.trace(_random_rays(100, 3))
    assert np.all(g[:, 0] == hip.MISS)


def test_invalid_plan_fails_without_latching(renderer):
    sc = scenes.Scene()
    m = sc.add_material(scenes.preview_material())
    sc.add_mesh([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 5]], m)
    with pytest.raises(hip.LotusError, match="missing vertex"):
        renderer.apply(sc)
    renderer.apply(scenes.random_triangles(4))  # still usable


def test_light_table_matches_oracle(renderer):
    sc = scenes.heightfield(cells=64, emissive=16, tiles=4)
    renderer.apply(sc)
    off_g, rec_g = renderer.light_table()
    off_c, rec_c = oracle.OracleScene(sc).light_table()
    assert np.array_equal(off_g, off_c)
    assert np.array_equal(rec_g, rec_c)


def _oracle_frame(sc, w, h, spp, bounces, light):
    return oracle.make_frame(sc.world_to_clip, w, h, spp=spp, max_bounces=bounces, light_sampling=light)


@pytest.mark.parametrize("light", [hip.LIGHT_BSDF, hip.LIGHT_MIS, hip.LIGHT_RESTIR])
def test_cornell_lambert_counts_and_radiance(renderer, light):
    # The CPU plain-PT workload: 256x256, 16 spp, max depth 8.
    _say(f"cornell lambert light={light}")
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    w = h = 256
    spp = 16
    acc, counts, shadow = oracle.OracleScene(sc).render(_oracle_frame(sc, w, h, spp, 8, light))
    first = None
    for k in range(spp):
        out = renderer.render(sc.world_to_clip, w, h, frame_count=1, max_bounces=8, light_sampling=light)
        n = len(out["counts"])
        assert np.array_equal(out["counts"], counts[k][:n]) and not counts[k][n:].any(), (k, out["counts"], counts[k])
        assert out["shadow_rays"] == shadow[k]
        if first is None:
            first = out["color"]
    # one-sample radiance per pixel: the same float operations on both sides
    acc1, _, _ = oracle.OracleScene(sc).render(_oracle_frame(sc, w, h, 1, 8, light))
    ref1 = acc1[..., :3].astype(np.float32)
    hitmask = acc1[..., 3] > 0
    exact = np.mean(np.all(first[..., :3][hitmask] == ref1[hitmask], -1))
    _say(f"  1-spp radiance bit-exact fraction {exact:.6f}")
    assert exact > 0.999
    assert np.allclose(first[..., :3], ref1, rtol=1e-4, atol=1e-5)
    sums, n = renderer.accumulation(w, h)
    assert n == spp
    assert np.allclose(sums[..., :3], acc[..., :3], rtol=1e-4, atol=1e-4)
    assert np.array_equal(sums[..., 3], acc[..., 3].astype(np.float32))


@pytest.mark.parametrize("light", [hip.LIGHT_BSDF, hip.LIGHT_MIS, hip.LIGHT_RESTIR])
def test_cornell_v1_matches_reference_images(renderer, light):
    _say(f"cornell-v1 DES-Q5 light={light}")
    mean = desq5.read_pfm(f"{GOLDEN}/cornell-box-mean.pfm")
    var = desq5.read_pfm(f"{GOLDEN}/cornell-box-variance.pfm")
    sc = scenes.cornell_v1()
    renderer.apply(sc)
    out = renderer.render(sc.world_to_clip, 128, 128, frame_count=64, max_bounces=64, light_sampling=light)
    assert out["spp"] == 64
    c = desq5.compare(out["color"][..., :3].copy(), 64, mean, var, 1024)
    _say(f"  image z {c.image_z}, max tile z {c.max_tile_z:.2f}")
    assert c.failure == "", c.failure


def test_cornell_red_wall_negative_control(renderer):
    mean = desq5.read_pfm(f"{GOLDEN}/cornell-box-mean.pfm")
    var = desq5.read_pfm(f"{GOLDEN}/cornell-box-variance.pfm")
    sc = scenes.cornell_v1(red_scale=1.1)
    renderer.apply(sc)
    out = renderer.render(sc.world_to_clip, 128, 128, frame_count=256, max_bounces=64, light_sampling=hip.LIGHT_MIS)
    assert desq5.compare(out["color"][..., :3].copy(), 256, mean, var, 1024).failure != ""


def test_closed_forms(renderer):
    sc = scenes.lambert_under_sky()
    renderer.apply(sc)
    out = renderer.render(sc.world_to_clip, 48, 48, frame_count=4, max_bounces=64, light_sampling=hip.LIGHT_BSDF,
                          clear_color=(0, 0, 0, 0))
    hit = out["color"][..., 3] == 1.0  # every sample of the pixel hit
    assert hit.sum() > 100
    assert np.allclose(out["color"][hit][:, :3], [0.5, 0.2, 0.3], rtol=2e-6)
    sc = scenes.closed_box()
    renderer.apply(sc)
    a = np.array([0.5, 0.25, 0.75])
    e = np.array([0.1, 0.15, 0.025])
    for bounces in (0, 1, 3):
        out = renderer.render(sc.world_to_clip, 32, 32, frame_count=2, max_bounces=bounces,
                              light_sampling=hip.LIGHT_BSDF)
        expect = sum(e * a ** k for k in range(bounces + 1))
        assert np.allclose(out["color"][..., :3], expect, rtol=1e-5)
        assert list(out["counts"]) == [1024] * (bounces + 1)


@pytest.mark.parametrize("builder", [hip.BVH_LBVH, hip.BVH_SAH])
def test_heightfield_counts_and_primary_hits(renderer, builder):
    # The 2.1M-triangle heightfield at reduced resolution: queue counts per
    # bounce exact, per-pixel one-sample radiance, pixel-centre hits exact,
    # on the GPU-built tree and on the SAH topology.
    _say(f"heightfield 1024 parity, builder {builder}")
    sc = scenes.heightfield(cells=1024)
    ev = renderer.apply(sc, builder=builder)
    assert ev.triangle_count == 2 * 1024 * 1024
    o = oracle.OracleScene(sc)
    w, h = 192, 108
    out = renderer.render(sc.world_to_clip, w, h, frame_count=1, max_bounces=6, light_sampling=hip.LIGHT_MIS)
    acc, counts, shadow = o.render(_oracle_frame(sc, w, h, 1, 6, hip.LIGHT_MIS))
    n = len(out["counts"])
    assert np.array_equal(out["counts"], counts[0][:n]) and not counts[0][n:].any()
    assert out["shadow_rays"] == shadow[0]
    # every ray walked by exactly one traversal kernel: camera, the rounds, the tail
    walked = out["walked"]
    assert walked[0] == out["camera_rays"]
    assert walked[1] + walked[2] + walked[3] == out["extension_rays"] + out["shadow_rays"]
    ref1 = acc[..., :3].astype(np.float32)
    assert np.mean(np.all(out["color"][..., :3] == ref1, -1)) > 0.999
    # camera-pass barycentrics vs the oracle's pixel-centre rays
    bary = renderer.render(sc.world_to_clip, w, h, output=hip.OUTPUT_BARYCENTRICS, want_depth=True)
    f = _oracle_frame(sc, w, h, 1, 6, 1)
    rays = np.zeros((w * h, 8), np.float32)
    ys, xs = np.mgrid[0:h, 0:w]
    nx = (xs.reshape(-1).astype(np.float32) + np.float32(0.5)) * np.float32(f.ndc_from_pixel[0]) + np.float32(-1.0)
    ny = (ys.reshape(-1).astype(np.float32) + np.float32(0.5)) * np.float32(f.ndc_from_pixel[1]) + np.float32(-1.0)
    import ctypes as C
    c2w = np.array(f.clip_to_world[:], np.float32)
    for i in range(w * h):
        oracle.lib().oracle_camera_ray(oracle.ptr(c2w, C.c_float), float(nx[i]), float(ny[i]),
                                       oracle.ptr(rays[i], C.c_float))
    hits = o.trace(rays)
    g = renderer.trace(rays)
    _compare_hits(g, hits)
    hit = hits[:, 0] != hip.MISS
    b = bary["color"].reshape(-1, 4)
    uv = hits[hit, 2:4].view(np.float32)
    assert np.allclose(b[hit, 1:3], uv, atol=1e-6)
    assert np.all(bary["depth"].reshape(-1)[~hit] == 1.0)
    assert np.all(bary["depth"].reshape(-1)[hit] < 1.0)


def test_restir_initial_records_match_oracle(renderer):
    # restir.slang:163-233 on the reference Cornell scene (GGX metal block,
    # coated walls): every pixel's initial reservoir, 16 candidates, the
    # power-table proposal and the mixture.
    sc = scenes.cornell_v1()
    renderer.apply(sc)
    o = oracle.OracleScene(sc)
    w = h = 64
    for proposal in (0, 2):
        _say(f"restir initial records proposal={proposal}")
        out = renderer.render(sc.world_to_clip, w, h, frame_count=1, sample_index=11, max_bounces=8,
                              light_sampling=hip.LIGHT_RESTIR, restir_candidates=16, restir_proposal=proposal,
                              want_restir=True)
        # visibility rays are the frame's only bounce-0 shadow rays
        f = oracle.make_frame(sc.world_to_clip, w, h, spp=1, sample_index=11, max_bounces=8, light_sampling=2,
                              restir_candidates=16, restir_proposal=proposal)
        g = out["restir"]
        exact = 0
        for y in range(h):
            for x in range(w):
                ref = o.restir_initial(f, x, y, 11)
                rec = g[y, x]
                assert ref is not None
                assert rec[0] == ref[0] and rec[3] == ref[3] and rec[7] == ref[7], (x, y, rec, ref)
                exact += np.array_equal(rec, ref)
                fw = [1, 2, 4, 5, 6, 8, 9, 10, 11]  # the float words
                assert np.allclose(rec.view(np.float32)[fw], ref.view(np.float32)[fw], rtol=1e-5, atol=1e-7)
        _say(f"  bit-exact records {exact}/{w * h}")
        assert exact >= 0.999 * w * h
        acc, counts, shadow = o.render(f)
        assert out["shadow_rays"] == shadow[0]


def test_settings_errors_do_not_latch(renderer):
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    with pytest.raises(hip.LotusError, match=r"candidates must be in \[1, 64\]"):
        renderer.render(sc.world_to_clip, 16, 16, light_sampling=hip.LIGHT_RESTIR, restir_candidates=0)
    with pytest.raises(hip.LotusError, match="unknown reservoir proposal"):
        renderer.render(sc.world_to_clip, 16, 16, light_sampling=hip.LIGHT_RESTIR, restir_proposal=3)
    with pytest.raises(hip.LotusError, match="at most 1024 bounces"):
        renderer.render(sc.world_to_clip, 16, 16, max_bounces=1025)
    with pytest.raises(hip.LotusError, match="singular"):
        renderer.render(np.zeros(16, np.float32), 16, 16)
    out = renderer.render(sc.world_to_clip, 16, 16, max_bounces=4)
    assert out["spp"] == 1


def test_accumulation_restart_and_max_samples(renderer):
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    a = renderer.render(sc.world_to_clip, 32, 32, frame_count=3, max_bounces=4)
    assert a["spp"] == 3
    b = renderer.render(sc.world_to_clip, 32, 32, frame_count=2, max_bounces=4)
    assert b["spp"] == 5  # continues
    c = renderer.render(sc.world_to_clip, 32, 32, frame_count=2, max_bounces=5)
    assert c["spp"] == 2  # max_bounces restarts
    d = renderer.render(sc.world_to_clip, 32, 32, frame_count=9, max_bounces=5, max_samples=4)
    assert d["spp"] == 4
    # deterministic: same first sample index, same image
    e1 = renderer.render(sc.world_to_clip, 32, 32, frame_count=1, max_bounces=4, sample_index=7)
    renderer.render(sc.world_to_clip, 32, 32, frame_count=1, max_bounces=4, sample_index=8)
    e2 = renderer.render(sc.world_to_clip, 32, 32, frame_count=1, max_bounces=4, sample_index=7)
    assert np.array_equal(e1["color"], e2["color"])


def test_deep_paths_use_the_tail(renderer):
    # max_bounces beyond the recorded rounds: the tail kernel finishes paths,
    # counts still exact against the oracle. The second sample's rounds follow
    # the first sample's counts, so its later bounces go to the tail.
    sc = scenes.closed_box()
    renderer.apply(sc)
    out = renderer.render(sc.world_to_clip, 32, 32, frame_count=2, max_bounces=200, light_sampling=hip.LIGHT_MIS)
    acc, counts, shadow = oracle.OracleScene(sc).render(_oracle_frame(sc, 32, 32, 2, 200, hip.LIGHT_MIS))
    n = len(out["counts"])
    assert np.array_equal(out["counts"], counts[1][:n]) and not counts[1][n:].any()
    assert out["shadow_rays"] == shadow[0] + shadow[1]
    assert out["rounds"] < n
    walked = out["walked"]
    assert walked[0] == out["camera_rays"] and walked[3] > 0
    assert walked[1] + walked[2] + walked[3] == out["extension_rays"] + out["shadow_rays"]
    # a round launches the traversal once, except the last round before the
    # tail (the tail walks its shadow rays): 31 for the first sample's 32
    # rounds, rounds - 1 for the second's
    assert out["trace_launches"] == 31 + out["rounds"] - 1
    # every ray walked by exactly one traversal kernel: camera, the rounds, the tail
    walked = out["walked"]
    assert walked[0] == out["camera_rays"] and walked[3] > 0
    assert walked[1] + walked[2] + walked[3] == out["extension_rays"] + out["shadow_rays"]
    # a round launches the traversal once, except the last round before the
    # tail (the tail walks its shadow rays): 31 for the first sample's 32
    # rounds, rounds - 1 for the second's
    assert out["trace_launches"] == 31 + out["rounds"] - 1


def test_split_kind_launches_give_the_same_frame(renderer):
    # extension and shadow rays in one launch or in launches of their own:
    # the same counts and the same radiance, bit for bit
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    a = renderer.render(sc.world_to_clip, 64, 64, frame_count=1, max_bounces=8, light_sampling=hip.LIGHT_MIS,
                        sample_index=3)
    # another first sample index restarts the accumulation, so b starts over at 3
    renderer.render(sc.world_to_clip, 64, 64, frame_count=1, max_bounces=8, light_sampling=hip.LIGHT_MIS,
                    sample_index=4)
    b = renderer.render(sc.world_to_clip, 64, 64, frame_count=1, max_bounces=8, light_sampling=hip.LIGHT_MIS,
                        sample_index=3, split_kinds=True)
    assert np.array_equal(a["counts"], b["counts"])
    assert a["shadow_rays"] == b["shadow_rays"]
    assert np.array_equal(a["color"], b["color"])
    assert a["trace_ms"] > 0 and b["trace_ms"] == 0 and b["intersect_ms"] > 0 and b["shadow_ms"] > 0


def test_tuning_does_not_change_results(renderer):
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    outs = []
    # drivers, the tail taking every bounce from the first that held fewer
    # rays than the camera's (1), the reference's rule (16), or none (0), and
    # the persistent kernels' resident workgroups (percent of the occupancy's)
    for cam, tr, div, pct, tile in [(1, 1, 16, 0, 0), (0, 0, 0, 50, 1), (0, 1, 1, 25, 2), (1, 0, 4, 1, 1)]:
        renderer.set_tuning(camera_dynamic=cam, trace_dynamic=tr, tail_divisor=div, resident_percent=pct,
                            camera_tile=tile)
        renderer.render(sc.world_to_clip, 64, 64, frame_count=1, max_bounces=8, sample_index=50)  # restart
        outs.append(renderer.render(sc.world_to_clip, 64, 64, frame_count=1, max_bounces=8, sample_index=5,
                                    light_sampling=hip.LIGHT_MIS))
    with pytest.raises(hip.LotusError, match="resident_percent"):
        renderer.set_tuning(resident_percent=101)
    with pytest.raises(hip.LotusError, match="camera_tile"):
        renderer.set_tuning(camera_tile=3)
    renderer.set_tuning(camera_dynamic=0, trace_dynamic=0, tail_divisor=16, resident_percent=0, camera_tile=0)
    for o in outs[1:]:
        assert np.array_equal(o["counts"], outs[0]["counts"])
        assert np.array_equal(o["color"], outs[0]["color"])
    assert outs[2]["rounds"] <= outs[0]["rounds"] <= outs[1]["rounds"] and outs[2]["rounds"] < outs[1]["rounds"]


def _data_scissor(data, w, h):
    # DataScissor (vulkan_backend.cpp:660-677)
    if data[2] <= 0 or data[3] <= 0:
        return (0, 0, w, h)
    x0, y0 = max(data[0], 0), max(data[1], 0)
    x1, y1 = min(data[0] + data[2], w), min(data[1] + data[3], h)
    return (0, 0, 0, 0) if x1 <= x0 or y1 <= y0 else (x0, y0, x1 - x0, y1 - y0)


@pytest.mark.parametrize("display,data", [
    ((-8.5, 4.25, 80.0, 60.0), (5, 7, 40, 30)),   # an offset, scaled display window; an inner data window
    ((0.0, 0.0, 0.0, 0.0), (50, -4, 40, 30)),     # the data window clipped at the right and the top
    ((3.0, -2.0, 58.0, 44.0), (0, 0, -1, 5)),     # no data window (a size that is not positive)
    ((0.0, 0.0, 0.0, 0.0), (100, 0, 10, 10)),     # a data window outside the image: nothing renders
])
def test_display_and_data_windows(renderer, display, data):
    # OffscreenTarget's windows (DisplayViewport, DataScissor, the
    # ndc_from_pixel of vulkan_backend.cpp:2408-2411) against the oracle
    # given the same windows: counts and shadow rays exact, one-sample
    # radiance bit-exact inside the data window, the clear colour outside.
    _say(f"windows display={display} data={data}")
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    w, h = 64, 48
    clear = (0.25, 0.5, 0.75, 1.0)
    out = renderer.render(sc.world_to_clip, w, h, frame_count=1, sample_index=3, max_bounces=6,
                          light_sampling=hip.LIGHT_MIS, display_window=display, data_window=data, clear_color=clear)
    rect = _data_scissor(data, w, h)
    inside = np.zeros((h, w), bool)
    inside[rect[1]:rect[1] + rect[3], rect[0]:rect[0] + rect[2]] = True
    color = out["color"]
    assert np.all(color[~inside] == np.array(clear, np.float32))
    if rect[2] == 0:
        assert out["camera_rays"] == 0 and len(out["counts"]) == 0
        return
    assert out["camera_rays"] == rect[2] * rect[3]
    disp = None if display[2] <= 0 or display[3] <= 0 else display
    fr = oracle.make_frame(sc.world_to_clip, w, h, spp=1, sample_index=3, max_bounces=6, light_sampling=1, rect=rect,
                           display=disp)
    acc, counts, shadow = oracle.OracleScene(sc).render(fr)
    n = len(out["counts"])
    assert np.array_equal(out["counts"], counts[0][:n]) and not counts[0][n:].any(), (out["counts"], counts[0])
    assert out["shadow_rays"] == shadow[0]
    hit = inside & (acc[..., 3] > 0)
    miss = inside & (acc[..., 3] == 0)
    assert np.all(color[miss] == np.array(clear, np.float32))
    exact = np.mean(np.all(color[hit][:, :3] == acc[hit][:, :3].astype(np.float32), -1))
    _say(f"  hits {hit.sum()}, misses {miss.sum()}, bit-exact {exact:.6f}")
    assert hit.sum() > 100 and exact > 0.999


@pytest.mark.parametrize("light", [hip.LIGHT_BSDF, hip.LIGHT_MIS])
@pytest.mark.parametrize("which", ["lambert", "lit_sky", "covered"])
def test_path_megakernel_equals_the_rounds(renderer, light, which):
    # each path from its camera ray to its end in one launch (lane pairs, as
    # the tail walks them) gives the rounds' frame: counts, shadow rays and
    # colour bit for bit, in the plain, extras and covered-surface variants
    if which == "lambert":
        sc = scenes.cornell_lambert()
    elif which == "lit_sky":
        from test_gpu_lights import lit_cornell
        sc = lit_cornell()
        sc.environment_map = scenes.sky_image()
    else:
        import material_fixtures as mf
        sc, _ = mf.load("material-alpha-v1")
    renderer.apply(sc)
    outs = []
    for mega in (0, 1):
        renderer.set_tuning(megakernel=mega)
        renderer.render(sc.world_to_clip, 80, 64, frame_count=1, max_bounces=8, sample_index=90)  # restart
        outs.append(renderer.render(sc.world_to_clip, 80, 64, frame_count=2, max_bounces=8, sample_index=4,
                                    light_sampling=light))
    renderer.set_tuning(megakernel=0)
    assert outs[1]["rounds"] == 0 and outs[0]["rounds"] > 0
    assert np.array_equal(outs[0]["counts"], outs[1]["counts"]), (outs[0]["counts"], outs[1]["counts"])
    assert outs[0]["shadow_rays"] == outs[1]["shadow_rays"]
    assert np.array_equal(outs[0]["color"], outs[1]["color"])
