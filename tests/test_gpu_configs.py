"""The benchmark workloads at their own sizes on the GPU:

* the 2.1M-triangle heightfield at 1920x1080, one sample, depth 6, NEE/MIS:
  per-bounce ray-queue counts and the shadow-ray count exact against the
  oracle, one-sample radiance bit-exact for >= 99.9 % of the pixels;
* the radiance cache on that scene and resolution with the 16-level, 2^19
  entry hash grid and 65,536 training paths: a zero network's frame equals
  plain frames bit for bit (training pixels the depth-6 frame, the others the
  depth-1 frame), and the network's forward pass and loss gradients on that
  frame's own training records match the cache oracle's device model
  (oracle/nrc.py) at the same FP16 rounding points;
* the 64-view batch (64^2..512^2 views of 1k..200k-triangle rigs, depth
  2..12): every view of a batched frame, eager and as one hipGraph, equals
  its unbatched render bit for bit, and the smallest views' counts are exact
  against the oracle."""
import sys

import numpy as np
import pytest

import oracle
from lotus_amd import hip, scenes
from oracle import nrc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

W, H = 1920, 1080


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def heightfield():
    return scenes.heightfield(1024)


def test_heightfield_1080p_frame_matches_oracle(heightfield):
    # on the SAH topology the benchmark builds
    sc = heightfield
    r = hip.Renderer(0)
    try:
        ev = r.apply(sc, builder=hip.BVH_SAH)
        _say(f"heightfield: {sc.triangle_count} triangles")
        out = r.render(sc.world_to_clip, W, H, frame_count=1, sample_index=5, max_bounces=6,
                       light_sampling=hip.LIGHT_MIS)
    finally:
        r.close()
    o = oracle.OracleScene(sc)
    acc, counts, shadow = o.render(oracle.make_frame(sc.world_to_clip, W, H, spp=1, sample_index=5, max_bounces=6,
                                                     light_sampling=1), threads=16)
    n = len(out["counts"])
    _say(f"counts {list(out['counts'])} vs {list(counts[0])}; shadow {out['shadow_rays']} vs {shadow[0]}")
    assert np.array_equal(out["counts"], counts[0][:n]) and not counts[0][n:].any()
    assert out["shadow_rays"] == shadow[0]
    ref = acc[..., :3].astype(np.float32)
    exact = np.mean(np.all(out["color"][..., :3] == ref, -1))
    _say(f"one-sample radiance bit-exact fraction {exact:.6f}")
    assert exact > 0.999
    assert np.allclose(out["color"][..., :3], ref, rtol=1e-4, atol=1e-4)


def _mix(x):
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


CACHE = dict(log2_table=19, train_paths=65536, batch_capacity=65536, train_steps=4, learning_rate=1e-2)


def test_cache_zero_network_frame_equals_plain_frames(heightfield):
    sc = heightfield
    r = hip.Renderer(0)
    try:
        r.apply(sc)
        cfg = r.cache_configure(**dict(CACHE, query_depth=1))
        r.cache_parameters(np.zeros(hip.CACHE_PARAMS, np.float32), None)
        out = r.render(sc.world_to_clip, W, H, max_bounces=6, cache=True, want_color=False)
        cache_sums, _ = r.accumulation(W, H)
        r.render(sc.world_to_clip, W, H, max_bounces=1, want_color=False)
        short, _ = r.accumulation(W, H)
        r.render(sc.world_to_clip, W, H, max_bounces=6, want_color=False)
        full, _ = r.accumulation(W, H)
    finally:
        r.close()
    stride = max(1, (W * H) // CACHE["train_paths"])
    offset = _mix((0 * 0x9E3779B9 + 0x632BE5AB) & 0xFFFFFFFF) % stride
    slot = np.arange(W * H).reshape(H, W)
    train = (slot + offset) % stride == 0
    _say(f"queries {out['cache_queries']}, samples {out['cache_samples']}, training pixels {train.sum()}")
    assert cfg["resolution"][-1] > cfg["resolution"][0]
    assert out["cache_queries"] > 100_000 and out["cache_samples"] > 10_000
    assert np.array_equal(cache_sums[train], full[train])
    assert np.array_equal(cache_sums[~train], short[~train])


def test_cache_network_on_the_frame_training_records(heightfield):
    sc = heightfield
    r = hip.Renderer(0)
    try:
        r.apply(sc)
        cfg = r.cache_configure(**CACHE)
        res = cfg["resolution"]
        rng = np.random.default_rng(9)
        T = 1 << CACHE["log2_table"]
        grid = (rng.uniform(-1, 1, (16, T, 2)) * 1e-2).astype(np.float32)
        w = np.concatenate([rng.uniform(-1, 1, a * b) * np.sqrt(6.0 / (a + b)) for a, b in nrc.SHAPES])
        w = (w * 0.5).astype(np.float32)
        r.cache_parameters(w, grid)
        # the frame trains the network: read its training records, then
        # restore the seeded parameters before checking against them
        out = r.render(sc.world_to_clip, W, H, max_bounces=6, cache=True, want_color=False,
                       want_training=2 * CACHE["train_paths"])
        records = out["training"]
        _say(f"training records {records.shape[0]} (samples counted {out['cache_samples']})")
        assert records.shape[0] == min(out["cache_samples"], 2 * CACHE["train_paths"]) > 10_000
        assert np.all(np.isfinite(records))
        r.cache_parameters(w, grid)
        sub = records[np.linspace(0, records.shape[0] - 1, 2048).astype(np.int64)]
        q = torch.tensor(np.ascontiguousarray(sub[:, :16]), device="cuda")
        pred = torch.zeros((sub.shape[0], 3), device="cuda")
        r.cache_infer(q.data_ptr(), sub.shape[0], pred.data_ptr())
        ref, _ = nrc.device_forward(sub[:, :16].astype(np.float64), w.astype(np.float64),
                                    grid.astype(np.float64), res)
        got = pred.cpu().numpy().astype(np.float64)
        scale = np.abs(ref).max()
        err = np.abs(got - ref)
        _say(f"forward on the frame's records: scale {scale:.3g}, max err {err.max():.3g}")
        assert err.max() < 1e-2 * scale and err.mean() < 1e-3 * scale
        sd = torch.tensor(np.ascontiguousarray(sub), device="cuda")
        wg, gg, loss = r.cache_gradients(sd.data_ptr(), sub.shape[0])
        loss_ref, wg_ref, gg_ref, _ = nrc.device_loss_and_grads(sub.astype(np.float64), w.astype(np.float64),
                                                                 grid.astype(np.float64), res)
    finally:
        r.close()
    _say(f"loss {loss:.6g} vs {loss_ref:.6g}")
    assert abs(loss - loss_ref) < 1e-3 * abs(loss_ref)
    off = 0
    for li, (a, b) in enumerate(nrc.SHAPES):
        x, y = wg[off:off + a * b], wg_ref[off:off + a * b]
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        _say(f"layer {li}: gradient rel err {rel:.3g}")
        assert rel < 1e-2, li
        off += a * b
    rel = np.linalg.norm(gg - gg_ref) / np.linalg.norm(gg_ref)
    _say(f"grid gradient rel err {rel:.3g}")
    assert rel < 1e-2
    assert np.all(gg[gg_ref == 0] == 0)


def _kw(v):
    return dict(world_to_clip=v["world_to_clip"], width=v["size"], height=v["size"], max_bounces=v["depth"],
                sample_index=0, light_sampling=hip.LIGHT_MIS)


def test_64_view_batch_matches_unbatched_views():
    views = scenes.batch_views()
    assert len(views) == 64
    rigs = [scenes.icosphere_rig(v["scene_triangles"], v["scene_seed"]) for v in views]
    _say(f"views {min(v['size'] for v in views)}..{max(v['size'] for v in views)} px, "
         f"{min(s.triangle_count for s in rigs)}..{max(s.triangle_count for s in rigs)} triangles")
    sets = []
    try:
        for _ in range(3):  # alone, eager batch, graph batch
            rs = []
            for sc in rigs:
                rr = hip.Renderer(0)
                rr.apply(sc)
                rs.append(rr)
            sets.append(rs)
        alone, eager, graph = sets
        for frame in range(2):
            ref = [rr.render(**_kw(v)) for rr, v in zip(alone, views)]
            out_e, ms_e = hip.render_batch(eager, [_kw(v) for v in views], graph=False)
            out_g, ms_g = hip.render_batch(graph, [_kw(v) for v in views], graph=True)
            _say(f"frame {frame}: eager batch {ms_e:.2f} ms, graph {ms_g:.2f} ms")
            for i, v in enumerate(views):
                assert np.array_equal(out_e[i]["color"], ref[i]["color"]), (frame, i)
                assert np.array_equal(out_e[i]["counts"], ref[i]["counts"]), (frame, i)
                a, _ = alone[i].accumulation(v["size"], v["size"])
                c, nc = graph[i].accumulation(v["size"], v["size"])
                assert nc == frame + 1 and np.array_equal(a, c), (frame, i)
                assert out_g[i]["shadow_rays"] == ref[i]["shadow_rays"], (frame, i)
            if frame == 0:
                first = ref
    finally:
        for rs in sets:
            for rr in rs:
                rr.close()
    for i in np.argsort([v["size"] for v in views])[:3]:
        v = views[i]
        o = oracle.OracleScene(rigs[i])
        acc, counts, shadow = o.render(oracle.make_frame(v["world_to_clip"], v["size"], v["size"], spp=1,
                                                         max_bounces=v["depth"], light_sampling=1))
        n = len(first[i]["counts"])
        assert np.array_equal(first[i]["counts"], counts[0][:n]) and not counts[0][n:].any(), i
        assert first[i]["shadow_rays"] == shadow[0], i
