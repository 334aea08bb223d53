"""ReSTIR DI reuse on the GPU, re-derived pixel by pixel: every temporal
record (restir.slang:302-387) and every spatial record with its accepted
neighbours (:392-491) against the oracle's restatement of the same passes,
fed the renderer's own inputs (the previous frame's published records and
receiver surfaces, the history reprojection and the pixel-centre ids). The
oracle is pinned against the reference's reservoir.cpp in
test_oracle_restir_reuse.py. Discrete fields (light, uv, M, flags,
supporting candidates, previous index, age) must match exactly; weights,
normalizations and contributions within 4e-6 relative, the reference's own
bar (validation/restir-gpu-test.cpp:497-503). The receiver surfaces the GPU
stores must equal the oracle's."""
import numpy as np
import pytest

import oracle
from lotus_amd import hip, scenes

import material_fixtures as mf
import restir_reuse as rr

pytestmark = [pytest.mark.gpu]

_FRAME = iter(range(5000, 1 << 30))


@pytest.fixture(scope="module")
def renderer():
    r = hip.Renderer(0)
    yield r
    r.close()


def _compare_surfaces(got, want, hits, what):
    g = got.reshape(-1, 40)[hits.reshape(-1)]
    w = want.reshape(-1, 40)[hits.reshape(-1)]
    gb, wb = g.view(np.uint32), w.view(np.uint32)
    ids = [7, 11, 30, 31]  # instance and material id words
    assert np.array_equal(gb[:, ids], wb[:, ids]), what
    err = np.abs(g.astype(np.float64) - w) / np.maximum(np.abs(w), 1e-6)
    err[:, ids] = 0
    assert err.max() <= 4e-6, (what, err.max(), np.unravel_index(np.argmax(err), err.shape))
    return float(np.mean(np.all(gb == wb, axis=1)))


def _run(renderer, sc, w, h, frames, views, *, radius, cap, candidates, max_bounces, spatial, first_sample,
         disk=False):
    renderer.apply(sc)
    o = oracle.OracleScene(sc)
    prev_final = prev_surf = None
    bitwise = []
    for f in range(frames):
        view = views(f)
        w2c = scenes.multiply(sc.projection, view)
        print(f"frame {f} {w}x{h} spatial={spatial}", flush=True)
        out = renderer.render_temporal(w2c, view, w, h, frame_index=next(_FRAME), sample_index=first_sample,
                                       camera_cut=f == 0,
                                       restir_temporal=True, restir_spatial=spatial, radius=radius,
                                       max_history_candidates=cap, max_bounces=max_bounces,
                                       light_sampling=hip.LIGHT_RESTIR, restir_candidates=candidates,
                                       want_products=True, spatial_disk=disk)
        sample = out["batch_sample_index"]  # the stream continues from the base by itself
        assert sample == first_sample + f and out["batch_samples"] == 1
        fr = oracle.make_frame(w2c, w, h, max_bounces=max_bounces, light_sampling=2, restir_candidates=candidates)
        vd = out["validity"]
        validity = np.stack([vd["valid"], vd["previous_index"]], -1).astype(np.uint32)
        centre = np.stack([out["surface"]["instance_id"], out["surface"]["material_id"]], -1)
        gtemp, gsurf = out["restir_temporal"], out["restir_surfaces"]
        rec, surf, hits = rr.oracle_temporal_frame(o, fr, sample, view, prev_final, prev_surf, validity, centre,
                                                   f != 0, cap)
        assert hits.mean() > 0.5
        bitwise.append(_compare_surfaces(gsurf, surf, hits, f"surfaces, frame {f}"))
        bitwise.append(rr.compare_records(gtemp, rec, f"temporal records, frame {f}"))
        if f:
            assert (gtemp[..., 14] != rr.NO).mean() > 0.5  # most pixels merged their history
        final = gtemp
        if spatial:
            srec, src = rr.oracle_spatial_frame(o, w, h, sample, radius, cap, gtemp, gsurf, disk)
            gsrc = out["restir_sources"]
            assert np.array_equal(gsrc[..., :5 if disk else 4], src[..., :5 if disk else 4]), f"sources, frame {f}"
            if disk:
                assert np.all(gsrc[..., 5:] == rr.NO)
            bitwise.append(rr.compare_records(out["restir_spatial"], srec, f"spatial records, frame {f}"))
            assert (src != rr.NO).any()
            final = out["restir_spatial"]
        prev_final, prev_surf = final, gsurf
    print("bitwise shares:", [round(b, 4) for b in bitwise])


def test_reuse_on_the_parallel_light_plane(renderer):
    # validation/restir-gpu-test.cpp:124-134, 435-446: 64 x 64, two candidates,
    # history cap 3, one bounce, a still camera.
    sc = scenes.restir_plane()
    _run(renderer, sc, 64, 64, 8, lambda f: sc.view, radius=1, cap=3, candidates=2, max_bounces=1,
         spatial=True, first_sample=17)


def test_reuse_temporal_only_on_the_cornell_box(renderer):
    # The reference Cornell box (area light, GGX metal, coated walls) under a
    # moving camera: temporal reuse alone, the history found by reprojection.
    sc = scenes.cornell_v1()
    _run(renderer, sc, 64, 48, 6, lambda f: scenes.camera_spline(f, frames=24), radius=4, cap=20, candidates=8,
         max_bounces=4, spatial=False, first_sample=3)


def test_reuse_on_the_many_light_workload(renderer):
    # The ReSTIR + SVGF workload at reduced resolution: the box with 100,000
    # seeded emissive ceiling triangles, 32 candidates per pixel, history cap
    # 20, spatial neighbours at r = 30 px, the seeded camera path.
    sc = scenes.cornell_many_lights()
    _run(renderer, sc, 160, 90, 4, lambda f: scenes.camera_spline(f), radius=30, cap=20, candidates=32,
         max_bounces=8, spatial=True, first_sample=40)


def test_reuse_five_neighbours_within_the_radius(renderer):
    # The many-light workload's spatial mode: five neighbours drawn uniformly
    # within r = 30 px (the reference's four axis neighbours at exactly r stay
    # the default); records, sources and surfaces bitwise as above (the
    # oracle's draws are pinned on the plane in test_oracle_restir_reuse.py),
    # and every accepted neighbour lies within the radius.
    sc = scenes.cornell_many_lights()
    _run(renderer, sc, 160, 90, 3, lambda f: scenes.camera_spline(f), radius=30, cap=20, candidates=32,
         max_bounces=8, spatial=True, first_sample=60, disk=True)
    out = renderer.render_temporal(scenes.multiply(sc.projection, sc.view), sc.view, 160, 90,
                                   frame_index=next(_FRAME), camera_cut=True, restir_spatial=True, radius=30,
                                   max_history_candidates=20, max_bounces=8, light_sampling=hip.LIGHT_RESTIR,
                                   restir_candidates=32, want_products=True, spatial_disk=True)
    src = out["restir_sources"][..., :5].reshape(-1, 5)
    ys, xs = np.divmod(np.arange(160 * 90), 160)
    ok = src != rr.NO
    dx = (src % 160).astype(np.int64) - xs[:, None]
    dy = (src // 160).astype(np.int64) - ys[:, None]
    d2 = (dx * dx + dy * dy)[ok]
    print(f"accepted neighbours {ok.sum()}, mean squared distance {d2.mean():.1f}")
    assert ok.sum() > 2000 and d2.max() <= 31 * 31 and d2.min() >= 1


@pytest.mark.parametrize("fixture", ["material-gltf-v1", "material-gltf-extensions-v1"])
def test_reuse_over_gltf_receivers(renderer, fixture):
    # glTF PBR receivers keep their closure's 22 terms in the stored surface
    # (StoreClosure / LoadClosure, restir.slang:80-141): coat, sheen,
    # specular, thin transmission and metal lobes re-evaluated at the
    # neighbours' and the previous frame's samples.
    sc, _ = mf.load(fixture)
    _run(renderer, sc, 64, 64, 4, lambda f: sc.view, radius=4, cap=20, candidates=8, max_bounces=4,
         spatial=True, first_sample=9)
