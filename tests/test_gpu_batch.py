"""Batched multi-view frames: every view of a batch (its own scene, size and
depth) accumulates exactly what it accumulates rendered alone, eagerly and
when the whole frame is one hipGraph, with graph replays interleaved with
eager frames; the autotuner leaves results unchanged."""
import numpy as np
import pytest

from lotus_amd import hip, scenes

pytestmark = pytest.mark.gpu


def _views(n=6):
    vs = scenes.batch_views(count=n, seed=11, min_tris=800, max_tris=20000, min_size=40, max_size=96)
    for v in vs:
        v["depth"] = min(v["depth"], 8)
    return vs


def _renderers(vs):
    out = []
    for v in vs:
        r = hip.Renderer(0)
        r.apply(scenes.icosphere_rig(v["scene_triangles"], v["scene_seed"]))
        out.append(r)
    return out


def _kw(v, sample_index=0, nudge=0.0):
    w2c = np.array(v["world_to_clip"], np.float32).copy()
    w2c[12] += np.float32(nudge)  # a camera moved sideways
    return dict(world_to_clip=w2c, width=v["size"], height=v["size"], max_bounces=v["depth"],
                sample_index=sample_index, light_sampling=hip.LIGHT_MIS)


def test_batched_views_match_unbatched():
    vs = _views()
    alone, eager, graph = _renderers(vs), _renderers(vs), _renderers(vs)
    kinds = []
    try:
        for frame in range(8):
            # the cameras move at frame 5: every accumulation starts over
            nudge = 0.01 if frame >= 5 else 0.0
            ref = [r.render(**_kw(v, nudge=nudge)) for r, v in zip(alone, vs)]
            out_e, ms_e = hip.render_batch(eager, [_kw(v, nudge=nudge) for v in vs], graph=False)
            # graph frames on 0, 2, 3, ... with an eager frame between: the
            # first is captured and instantiated (batch_graph 1), the next
            # replay it with their frame variables (3); the moved cameras
            # change kernel arguments, so that frame is captured and updates
            # the executable graph in place (2), and later ones replay it
            out_g, ms_g = hip.render_batch(graph, [_kw(v, nudge=nudge) for v in vs], graph=frame != 1)
            kinds.append(out_g[0]["batch_graph"])
            assert all(o["batch_graph"] == kinds[-1] for o in out_g)
            print(f"frame {frame}: eager {ms_e:.3f} ms, {'graph' if frame != 1 else 'eager'} {ms_g:.3f} ms "
                  f"(graph kind {kinds[-1]})", flush=True)
            for i, v in enumerate(vs):
                a, _ = alone[i].accumulation(v["size"], v["size"])
                b, nb = eager[i].accumulation(v["size"], v["size"])
                c, nc = graph[i].accumulation(v["size"], v["size"])
                assert nb == nc == (frame + 1 if frame < 5 else frame - 4)
                assert np.array_equal(a, b), (frame, i)
                assert np.array_equal(a, c), (frame, i)
                assert np.array_equal(ref[i]["color"], out_e[i]["color"])
                assert np.array_equal(out_e[i]["counts"], ref[i]["counts"])
                assert out_g[i]["shadow_rays"] == ref[i]["shadow_rays"]
        assert kinds == [1, 0, 3, 3, 3, 2, 3, 3], kinds
    finally:
        for r in alone + eager + graph:
            r.close()


def test_batch_rejects_unsupported_views():
    vs = _views(2)
    rs = _renderers(vs)
    try:
        with pytest.raises(hip.LotusError, match="own renderer"):
            hip.render_batch([rs[0], rs[0]], [_kw(vs[0]), _kw(vs[0])])
        bad = dict(_kw(vs[0]), frame_count=2)
        with pytest.raises(hip.LotusError, match="one radiance sample"):
            hip.render_batch([rs[0]], [bad])
        out, _ = hip.render_batch(rs, [_kw(v) for v in vs])  # still usable
        assert all(o["spp"] == 1 for o in out)
    finally:
        for r in rs:
            r.close()


def test_autotune_keeps_results():
    vs = _views(1)
    a, b = _renderers(vs + vs)
    try:
        chosen = b.autotune(trials=1, **_kw(vs[0]))
        assert set(chosen) == {"camera_dynamic", "trace_dynamic", "tail_divisor", "resident_percent", "megakernel",
                               "camera_tile"}
        assert chosen["tail_divisor"] in (0, 16) and chosen["resident_percent"] in (50, 100)
        x = a.render(**_kw(vs[0], 5))["color"]
        y = b.render(**_kw(vs[0], 5))["color"]
        assert np.array_equal(x, y)
    finally:
        a.close()
        b.close()


def _lit_view():
    """The lit Cornell box under a sky image (analytic lights and an
    environment: the megakernel's variant with extras) as a batch view."""
    from test_gpu_lights import lit_cornell
    sc = lit_cornell()
    sc.environment_map = scenes.sky_image()
    return sc, dict(world_to_clip=sc.world_to_clip, width=48, height=40, max_bounces=6, sample_index=0,
                    light_sampling=hip.LIGHT_MIS)


def test_fused_batch_matches_unbatched():
    # one path megakernel per kernel variant for every view, fed by one
    # device work queue: each view accumulates what the wavefront rounds
    # accumulate rendered alone (colour, counts, shadow rays bit for bit)
    vs = _views()
    alone, fused = _renderers(vs), _renderers(vs)
    lit, lit_kw = _lit_view()
    alone_lit, fused_lit = hip.Renderer(0), hip.Renderer(0)
    alone_lit.apply(lit)
    fused_lit.apply(lit)
    alone.append(alone_lit)
    fused.append(fused_lit)
    kws = [_kw(v) for v in vs] + [lit_kw]
    sizes = [(v["size"], v["size"]) for v in vs] + [(lit_kw["width"], lit_kw["height"])]
    try:
        for r in fused:
            r.set_tuning(megakernel=1)
        with pytest.raises(hip.LotusError, match="eagerly"):
            hip.render_batch(fused, kws, graph=True, fused=True)
        for frame in range(3):
            ref = [r.render(**kw) for r, kw in zip(alone, kws)]
            out, ms = hip.render_batch(fused, kws, fused=True)
            print(f"fused frame {frame}: {ms:.3f} ms", flush=True)
            for i, (w, h) in enumerate(sizes):
                a, na = alone[i].accumulation(w, h)
                b, nb = fused[i].accumulation(w, h)
                assert na == nb == frame + 1
                assert np.array_equal(a, b), (frame, i)
                assert np.array_equal(ref[i]["color"], out[i]["color"]), (frame, i)
                assert np.array_equal(out[i]["counts"], ref[i]["counts"]), (frame, i, out[i]["counts"], ref[i]["counts"])
                assert out[i]["shadow_rays"] == ref[i]["shadow_rays"]
    finally:
        for r in alone + fused:
            r.close()
