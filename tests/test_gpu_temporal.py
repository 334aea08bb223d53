"""Temporal frames on the GPU: SVGF's variance estimate and a-trous filter
and the history reprojection against the reference's own CPU oracles
(backend/vulkan/temporal.cpp, built into oracle/_ref), temporal
accumulation against progressive accumulation, and ReSTIR DI reuse
(temporal + spatial) against plain next event estimation."""
import numpy as np
import pytest

import oracle
from lotus_amd import hip, scenes

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built (no reference sources)")]

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def renderer():
    r = hip.Renderer(0)
    yield r
    r.close()


def _pack_normal(n):
    """Octahedral snorm16 packing (the surface product's shading normal)."""
    s = np.abs(n).sum(-1, keepdims=True)
    e = n[..., :2] / s
    neg = n[..., 2] < 0
    sign = np.where(e >= 0, 1.0, -1.0)
    fold = (1.0 - np.abs(e[..., ::-1])) * sign
    e = np.where(neg[..., None], fold, e)
    q = np.rint(np.clip(e, -1, 1) * 32767).astype(np.int64)
    return ((q[..., 0] & 0xFFFF) | ((q[..., 1] & 0xFFFF) << 16)).astype(np.uint32)


def _synthetic(w, h, seed, empty=0.1):
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:h, 0:w]
    surf = np.zeros((h, w), hip.TEMPORAL_PIXEL)
    inst = np.where(xs < w // 3, 1, np.where(xs < 2 * w // 3, 2, 3)) + np.where(ys > h // 2, 10, 0)
    depth = 2.0 + 0.01 * xs + 0.02 * ys + 0.5 * (inst % 2)
    n = np.stack([0.1 * np.sin(xs / 5.0), 0.1 * np.cos(ys / 7.0), np.ones_like(depth)], -1)
    n += rng.normal(0, 0.02, n.shape)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    surf["normal_depth"][..., :3] = n
    surf["normal_depth"][..., 3] = depth
    surf["instance_id"] = inst
    surf["material_id"] = inst
    surf["flags"] = 7
    surf["albedo"] = rng.integers(0, 1 << 24, (h, w), dtype=np.uint32)
    surf["albedo"][rng.random((h, w)) < 0.05] &= 0xFFFF00  # zero albedo channels
    sn = n + rng.normal(0, 0.05, n.shape)
    surf["shading_normal"] = _pack_normal(sn / np.linalg.norm(sn, axis=-1, keepdims=True))
    rad = np.zeros((h, w), hip.TEMPORAL_RADIANCE)
    count = rng.integers(1, 60, (h, w)).astype(np.uint32)
    count[rng.random((h, w)) < empty] = 0
    rad["sample_count"] = count
    rad["color"][..., :3] = rng.lognormal(-1, 1, (h, w, 3))
    rad["color"][..., 3] = 1.0
    m1 = rng.lognormal(-1, 0.7, (h, w))
    rad["first_moment"] = m1
    rad["second_moment"] = m1 * m1 + rng.lognormal(-3, 1, (h, w))
    return surf, rad


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


def _host(t, dtype, shape):
    return t.cpu().numpy().view(dtype).reshape(shape)


@pytest.mark.parametrize("batch,cap", [(1, 32), (4, 16)])
def test_svgf_variance_and_filter_match_reference(renderer, batch, cap):
    R = oracle.ref()
    w, h = 61, 47
    surf, rad = _synthetic(w, h, batch * 7 + cap)
    ds, dr = _dev(surf), _dev(rad)
    color = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    renderer.svgf(ds.data_ptr(), dr.data_ptr(), w, h, batch, cap, 5, color.data_ptr())
    out = _host(dr, hip.TEMPORAL_RADIANCE, (h, w))
    ref_var = np.array([[R.ref_estimate_variance(surf.ctypes.data, rad.ctypes.data, x, y, w, h, batch, cap)
                         for x in range(w)] for y in range(h)], np.float32)
    err = np.abs(out["variance"] - ref_var)
    print(f"variance: max rel err {np.max(err / np.maximum(ref_var, 1e-12)):.3g}")
    assert np.all(err <= 2e-5 * np.abs(ref_var) + 1e-9)
    # the filter against the reference's on the same variance words
    rad2 = rad.copy()
    rad2["variance"] = out["variance"]
    ref_col = np.zeros((h, w, 4), np.float32)
    assert R.ref_filter_radiance(surf.ctypes.data, rad2.ctypes.data, w, h, batch, cap, 5,
                                 ref_col.ctypes.data_as(oracle.C.POINTER(oracle.C.c_float)))
    got = color.cpu().numpy().reshape(h, w, 4)
    rel = np.abs(got - ref_col) / np.maximum(np.abs(ref_col), 1e-6)
    print(f"filter: max rel err {rel.max():.3g}, mean {rel.mean():.3g}")
    assert np.all(np.isfinite(got))
    assert rel.max() < 1e-3


def test_reprojection_matches_reference(renderer):
    R = oracle.ref()
    w, h = 53, 41
    rng = np.random.default_rng(5)
    cur, _ = _synthetic(w, h, 1)
    prev, _ = _synthetic(w, h, 2)
    cur["motion_depth"][..., 0] = rng.uniform(-45, 45, (h, w)).astype(np.float32)
    cur["motion_depth"][..., 1] = rng.uniform(-45, 45, (h, w)).astype(np.float32)
    cur["motion_depth"][..., 2] = prev["normal_depth"][..., 3].mean() * rng.choice([1.0, 1.001, 1.5], (h, w))
    cur["motion_depth"][..., 3] = rng.choice([1.0, 0.0], (h, w), p=[0.9, 0.1])
    cur["flags"] = rng.choice([7, 7, 7, 5, 3, 0], (h, w))
    cur["instance_id"] = rng.choice([1, 2], (h, w))
    prev["instance_id"] = rng.choice([1, 2], (h, w))
    prev["normal_depth"][..., 3] = prev["normal_depth"][..., 3].mean()
    prev["flags"] = rng.choice([7, 0], (h, w), p=[0.95, 0.05])
    val = np.zeros((h, w), hip.TEMPORAL_VALIDITY)
    dc, dp, dv = _dev(cur), _dev(prev), _dev(val)
    settings = np.array([0.01, 0.01, 0.9, 32.0], np.float32)
    renderer.reproject(dc.data_ptr(), dp.data_ptr(), w, h, dv.data_ptr(), settings)
    got = _host(dv, hip.TEMPORAL_VALIDITY, (h, w))
    ref = np.zeros(4, np.uint32)
    codes = np.zeros(11, int)
    for y in range(h):
        for x in range(w):
            R.ref_reproject(cur[y, x:x + 1].ctypes.data, x, y, w, h, prev.ctypes.data,
                            settings.ctypes.data_as(oracle.C.POINTER(oracle.C.c_float)),
                            ref.ctypes.data_as(oracle.C.POINTER(oracle.C.c_uint32)))
            g = got[y, x]
            assert (g["valid"], g["rejection"], g["previous_index"]) == tuple(ref[:3]), (x, y, g, ref)
            codes[ref[1]] += 1
    print("rejection codes", codes)
    assert np.count_nonzero(codes) >= 6  # the cases are exercised
    # without history every hit is HistoryUnavailable
    renderer.reproject(dc.data_ptr(), 0, w, h, dv.data_ptr(), settings)
    got = _host(dv, hip.TEMPORAL_VALIDITY, (h, w))
    assert np.all(got["rejection"][(cur["flags"] & 1) != 0] == 8)


def _cornell(renderer):
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    return sc


def test_static_temporal_accumulation_is_the_progressive_mean(renderer):
    # A still camera reprojects every hit onto itself; with a cap above the
    # frame count the accumulated colour is the mean of the fresh batches,
    # i.e. the progressive accumulation of the same samples.
    sc = _cornell(renderer)
    w = h = 64
    n = 12
    for f in range(n):
        out = renderer.render_temporal(sc.world_to_clip, sc.view, w, h, frame_index=100 + f,
                                       accumulate=True, max_history=64, max_bounces=8, want_products=True)
        assert out["reset_reason"] == (1 if f == 0 else 0)
        assert out["batch_sample_index"] == f and out["batch_samples"] == 1  # the stream continues
    hit = (out["surface"]["flags"] & 1) != 0
    assert hit.mean() > 0.9
    assert np.all(out["validity"]["valid"][hit] == 1)
    assert np.all(np.abs(out["surface"]["motion_depth"][hit][:, :2]) < 1e-3)  # projection rounding only
    assert np.all(out["radiance"]["sample_count"][hit] == n)
    prog = renderer.render(sc.world_to_clip, w, h, frame_count=n, sample_index=0, max_bounces=8)["color"]
    assert np.allclose(out["color"][hit], prog[hit], rtol=2e-5, atol=1e-6)


def test_moving_camera_validity_matches_reference(renderer):
    R = oracle.ref()
    sc = _cornell(renderer)
    w = h = 96
    surfs, vals = [], []
    for f in range(3):
        view = scenes.look_at((0.02 * f, 0.0, 3.4 - 0.03 * f), (0.0, 0.0, 0.0))
        w2c = scenes.multiply(sc.projection, view)
        out = renderer.render_temporal(w2c, view, w, h, frame_index=200 + f, max_bounces=4, want_products=True)
        surfs.append(out["surface"])
        vals.append(out["validity"])
    cur, prev, val = surfs[2], surfs[1], vals[2]
    hit = (cur["flags"] & 1) != 0
    moved = np.abs(cur["motion_depth"][..., :2]).sum(-1)
    assert np.mean(moved[hit] > 0) > 0.9
    assert np.mean(val["valid"][hit]) > 0.8
    settings = np.array([0.01, 0.01, 0.9, 32.0], np.float32)
    ref = np.zeros(4, np.uint32)
    for y in range(0, h, 2):
        for x in range(0, w, 2):
            R.ref_reproject(cur[y, x:x + 1].ctypes.data, x, y, w, h, prev.ctypes.data,
                            settings.ctypes.data_as(oracle.C.POINTER(oracle.C.c_float)),
                            ref.ctypes.data_as(oracle.C.POINTER(oracle.C.c_uint32)))
            g = val[y, x]
            assert (g["valid"], g["rejection"], g["previous_index"]) == tuple(ref[:3]), (x, y)


def _relmse(img, ref):
    return float(np.mean((img[..., :3] - ref) ** 2 / (ref ** 2 + 1e-2)))


def test_restir_reuse_is_consistent_and_reduces_error(renderer):
    # Temporal + spatial reservoir reuse on the reference Cornell scene: no
    # failed records, reuse grows the represented candidate counts, the
    # image mean stays that of plain next event estimation, and one frame's
    # error against a converged image falls below the initial-only frame's.
    sc = scenes.cornell_v1()
    renderer.apply(sc)
    w = h = 96
    ref = renderer.render(sc.world_to_clip, w, h, frame_count=1024, sample_index=1 << 20, max_bounces=8)["color"]
    ref = ref[..., :3].astype(np.float64)
    kw = dict(max_bounces=8, light_sampling=hip.LIGHT_RESTIR, restir_candidates=8, want_restir=True)
    initial_err, reuse_err, means = [], [], []
    for f in range(24):
        out = renderer.render_temporal(sc.world_to_clip, sc.view, w, h, frame_index=300 + f,
                                       restir_temporal=True, restir_spatial=True, radius=8,
                                       max_history_candidates=20, **kw)
        rec = out["restir"]
        assert np.all(rec[..., 7] & 2 == 0)
        means.append(out["color"][..., :3].mean())
        if f >= 8:
            reuse_err.append(_relmse(out["color"], ref))
    assert rec[..., 3].max() > 8  # reuse represents more than the initial candidates
    for f in range(16):
        out = renderer.render(sc.world_to_clip, w, h, sample_index=500 + f, **kw)
        initial_err.append(_relmse(out["color"], ref))
    m_reuse, m_ref = np.mean(means[8:]), ref.mean()
    print(f"mean reuse {m_reuse:.5f} ref {m_ref:.5f}; relMSE initial {np.mean(initial_err):.4g} "
          f"reuse {np.mean(reuse_err):.4g}")
    assert abs(m_reuse - m_ref) < 0.03 * m_ref
    assert np.mean(reuse_err) < np.mean(initial_err)


def test_restir_reuse_products(renderer):
    # RestirProducts (vulkan_backend.hpp:175-200): the temporal and spatial
    # passes' records as the frame leaves them (restir.slang:21-28). Each
    # record's stored normalization is ReservoirReuseNormalization
    # (reservoir.cpp:102-112, built from the reference) of its reservoir and
    # supporting count; no stream failed; previous indices name pixels; the
    # spatial pass's accepted neighbours lie exactly `radius` pixels away
    # along an axis (restir.slang:408-410).
    R = oracle.ref()
    sc = scenes.cornell_v1()
    renderer.apply(sc)
    w, h, radius = 48, 40, 6
    for f in range(4):
        out = renderer.render_temporal(sc.world_to_clip, sc.view, w, h, frame_index=700 + f,
                                       restir_temporal=True, restir_spatial=True, radius=radius,
                                       max_history_candidates=20, max_bounces=8, light_sampling=hip.LIGHT_RESTIR,
                                       restir_candidates=8, want_products=True)
    checked = 0
    for name in ("restir_temporal", "restir_spatial"):
        words = out[name].reshape(-1, 16)
        assert np.all(words[:, 7] & 2 == 0), name
        prev = words[:, 14]
        assert np.all((prev == 0xFFFFFFFF) | (prev < w * h)), name
        live = words[:, 1].view(np.float32) > 0
        assert live.mean() > 0.5, name
        for i in np.flatnonzero(live)[::7]:
            res = np.ascontiguousarray(words[i, :4])
            ref = R.ref_reuse_normalization(res.ctypes.data_as(oracle.C.POINTER(oracle.C.c_uint32)), int(words[i, 13]))
            assert np.isclose(words[i, 12].view(np.float32), ref, rtol=1e-6), (name, i, words[i], ref)
            checked += 1
    assert checked > 100
    src = out["restir_sources"].reshape(h, w, 4)
    ys, xs = np.mgrid[0:h, 0:w]
    for k in range(4):
        s = src[..., k]
        ok = s != 0xFFFFFFFF
        assert ok.any()
        dx, dy = np.abs(s[ok] % w - xs[ok]), np.abs(s[ok] // w - ys[ok])
        assert np.all(((dx == radius) & (dy == 0)) | ((dx == 0) & (dy == radius)))


def test_temporal_settings_errors(renderer):
    sc = _cornell(renderer)
    with pytest.raises(hip.LotusError, match="temporal reservoirs require"):
        renderer.render_temporal(sc.world_to_clip, sc.view, 16, 16, frame_index=900, restir_temporal=True)
    with pytest.raises(hip.LotusError, match="spatial reservoirs require"):
        renderer.render_temporal(sc.world_to_clip, sc.view, 16, 16, frame_index=901, restir_spatial=True,
                                 light_sampling=hip.LIGHT_RESTIR, radius=0)
    with pytest.raises(hip.LotusError, match="filter_iterations"):
        renderer.render_temporal(sc.world_to_clip, sc.view, 16, 16, frame_index=902, filter_iterations=3)


@pytest.mark.parametrize("mode", [hip.RECTIFY_NONE, hip.RECTIFY_CLAMP, hip.RECTIFY_ANTI_LAG, hip.RECTIFY_REFLECTION])
def test_accumulation_with_rectification_matches_oracle(renderer, mode):
    # AccumulateTemporalRadiance (temporal.slang:219-267) under each history
    # rectification mode against the oracle's float32 restatement (pinned by
    # closed forms in test_oracle_temporal.py): sample counts exact, colours
    # and moments to 1e-6.
    from oracle import temporal as ot
    w, h = 57, 43
    surf, old = _synthetic(w, h, 21 + mode)
    rng = np.random.default_rng(mode)
    surf["reserved"] = rng.uniform(0, 1, (h, w)).astype(np.float32).view(np.uint32)
    surf["flags"][rng.random((h, w)) < 0.05] = 0  # some misses
    batch = np.ones((h, w, 4), np.float32)
    batch[..., :3] = rng.lognormal(-1, 0.8, (h, w, 3))
    valid = np.zeros((h, w), hip.TEMPORAL_VALIDITY)
    valid["valid"] = rng.random((h, w)) < 0.85
    valid["previous_index"] = rng.integers(0, w * h, (h, w))
    old["color"][rng.random((h, w)) < 0.2, :3] *= 20.0  # some far histories
    out = torch.zeros(w * h * 32, dtype=torch.uint8, device="cuda")
    db = torch.from_numpy(batch.reshape(-1).view(np.uint8).copy()).cuda()
    ds, dv, do = _dev(surf), _dev(valid), _dev(old)
    renderer.accumulate(db.data_ptr(), ds.data_ptr(), dv.data_ptr(), do.data_ptr(), w, h, 4, 24, mode,
                        out.data_ptr())
    got = _host(out, hip.TEMPORAL_RADIANCE, (h, w))
    ref = ot.accumulate(batch, surf, valid, old, 4, 24, mode)
    assert np.array_equal(got["sample_count"], ref["sample_count"])
    for field in ("color", "first_moment", "second_moment"):
        a, b = got[field].astype(np.float64), ref[field].astype(np.float64)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-12)
        assert rel.max() < 1e-6, (field, rel.max())
    print(f"mode {mode}: kept histories {np.mean(got['sample_count'] > 4):.3f}, "
          f"bit-exact colours {np.mean(np.all(got['color'] == ref['color'], -1)):.4f}")


def test_reflection_weight_of_a_still_camera_keeps_everything(renderer):
    # ReflectionWeight (temporal.slang:611-622): theta = 0 under a still
    # camera, so 1 - s + s / (1 + 0) = 1 for every surface with a valid
    # projection (up to the cosine's rounding, dot(v, v) against |v| |v|);
    # a moved camera keeps less of the glossy metal's history.
    sc = scenes.cornell_v1()
    renderer.apply(sc)
    kw = dict(accumulate=True, max_history=16, want_products=True, rectification=hip.RECTIFY_REFLECTION,
              max_bounces=2)
    renderer.render_temporal(sc.world_to_clip, sc.view, 64, 64, frame_index=9000, camera_cut=True, **kw)
    out = renderer.render_temporal(sc.world_to_clip, sc.view, 64, 64, frame_index=9001, **kw)
    s = out["surface"]
    moving = s["motion_depth"][..., 3] == 1.0
    assert moving.mean() > 0.9
    kept = s["reserved"][moving].view(np.float32)
    print(f"still camera: kept share {kept.min():.7f}..{kept.max():.7f}")
    assert np.all((kept > 1.0 - 1e-4) & (kept <= 1.0))
    view = scenes.look_at((0.4, 0.2, 3.0), (0.0, 0.0, 0.0))
    w2c = scenes.multiply(sc.projection, view)
    out = renderer.render_temporal(w2c, view, 64, 64, frame_index=9002, **kw)
    s = out["surface"]
    keep = s["reserved"].view(np.float32)[s["motion_depth"][..., 3] == 1.0]
    print(f"moved camera: kept share {keep.min():.3f}..{keep.max():.3f}")
    assert keep.min() < 0.9 and keep.max() <= 1.0 and keep.min() > 0.0


def _settle_card():
    # temporal-gpu-test.cpp:918-930: a metal card (GGX, roughness 0.5) with
    # misses around it under a constant environment.
    sc = scenes.Scene(name="settle-card")
    sc.projection = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1.2, -1, 0, 0, -2.2, 0], np.float32)
    m = sc.add_material(scenes.preview_material(base_color=(0.6, 0.4, 0.2), metallic=1.0, roughness=0.5))
    sc.add_mesh([[-2, -2, -4], [2, -2, -4], [2, 2, -4], [-2, 2, -4]], [[0, 1, 2], [0, 2, 3]], m)
    sc.environment = (1.0, 2.0, 3.0)
    return sc


def test_settling_sequence(renderer):
    # SettlingSequence (temporal-gpu-test.cpp:911-1076): while the camera
    # moves, frames accumulate and filter; held still, frames settle: their
    # still samples are the progressive render of the host's sample index
    # bit for bit (a second renderer's), the retained history is untouched,
    # and the output is the reference's SettleTemporalRadiance
    # (temporal.cpp:514-528, oracle/_ref) of the last changing frame's colour
    # and the still mean, bit for bit, until it is the progressive image and
    # stops. Switching settle off and on resets nothing and resumes the
    # stream's samples; a change ends settling at once; a cut drops history.
    from oracle import temporal as ot
    sc = _settle_card()
    renderer.apply(sc)
    progressive = hip.Renderer(0)
    try:
        progressive.apply(sc)
        w, h, max_samples, cap = 64, 48, 40, 32
        frames = iter(range(20000, 30000))
        base = dict(accumulate=True, max_history=cap, filter_iterations=5, max_bounces=1, sample_index=31,
                    max_samples=max_samples, want_products=True)

        def render(view, settle=True, **extra):
            return renderer.render_temporal(scenes.multiply(sc.projection, view), view, w, h,
                                            frame_index=next(frames), settle=settle, **base, **extra)

        last = None
        for k in range(4):
            view = scenes.translation(-0.05 * k, 0.0, 0.0)
            last = render(view, camera_cut=k == 0)
            assert not last["settling"] and last["settled_samples"] == 0, k
            assert k == 0 or last["history_available"], k
            assert last["batch_sample_index"] == 31 + k and last["batch_samples"] == 1, k
        misses = (last["surface"]["flags"] & 1) == 0
        assert 0 < misses.sum() < w * h
        prior = last["color"].copy()
        history = last["radiance"].copy()
        resume = last["batch_sample_index"] + last["batch_samples"]
        w2c = scenes.multiply(sc.projection, view)
        weight = np.where(history["sample_count"] != 0, cap, 0).astype(np.uint32)
        span = np.minimum(weight, max_samples)
        converged = still_ref = None
        crossfaded = 0
        for n in range(1, max_samples + 3):
            out = render(view)
            samples, added = min(n, max_samples), int(n <= max_samples)
            assert out["settling"] and out["history_available"] and out["reset_reason"] == 0, n
            assert out["settled_samples"] == samples and out["spp"] == samples, n
            assert out["batch_samples"] == added and out["batch_sample_index"] == 31 + samples - added, n
            assert out["radiance"].tobytes() == history.tobytes(), f"settling changed the history at {n}"
            still = out["batch"]
            if added:
                if n == 1:
                    progressive.render(w2c, w, h, sample_index=31 ^ 0x80000000, max_bounces=1,
                                       max_samples=max_samples)
                still_ref = progressive.render(w2c, w, h, sample_index=31, max_bounces=1,
                                               max_samples=max_samples)["color"]
            assert still.tobytes() == still_ref.tobytes(), f"still mean differs from the progressive image at {n}"
            expected = ot.ref_settle(prior.reshape(-1, 4), still.reshape(-1, 4), weight.reshape(-1), samples,
                                     max_samples).reshape(h, w, 4)
            got = out["color"]
            assert got.tobytes() == expected.tobytes(), f"crossfade differs from SettleTemporalRadiance at {n}"
            done = samples >= span
            assert got[done].tobytes() == still[done].tobytes()
            crossfaded += int((~done).sum())
            if n == max_samples:
                converged = got.copy()
            elif n > max_samples:
                assert got.tobytes() == converged.tobytes(), "a converged view changed"
        assert crossfaded != 0
        assert converged.tobytes() == still_ref.tobytes(), "the settled image is not the progressive image"

        # settle off: nothing resets, the stream resumes where it stopped
        out = render(view, settle=False)
        assert not out["settling"] and out["history_available"] and out["reset_reason"] == 0
        assert out["batch_sample_index"] == resume
        # settle on again: the first frame leaves the prior, the next settles
        out = render(view)
        assert not out["settling"] and out["history_available"] and out["reset_reason"] == 0
        out = render(view)
        assert out["settling"] and out["settled_samples"] == 1
        kept = out["radiance"].copy()
        # a change ends settling at once and resumes the retained history
        moved = scenes.translation(0.05, 0.0, 0.0)
        out = render(moved)
        assert not out["settling"] and out["history_available"] and out["reset_reason"] == 0
        hit = (out["validity"]["valid"] != 0)
        assert hit.mean() > 0.2  # the card covers a quarter of the target
        assert np.all(out["radiance"]["sample_count"][hit] > 1)  # blended the kept history
        assert kept["sample_count"].max() > 1
        out = render(moved, camera_cut=True)
        assert not out["settling"] and not out["history_available"] and out["reset_reason"] == 3
        with pytest.raises(hip.LotusError, match="settling requires radiance accumulation"):
            renderer.render_temporal(w2c, view, w, h, frame_index=next(frames), settle=True, max_bounces=1)
    finally:
        progressive.close()


def test_reset_reasons_follow_the_reference(renderer):
    # TemporalHistory::Advance (temporal.cpp:114-147, built from the
    # reference into oracle/_ref) decides first frame, camera cut, frame
    # gap, target (extent and both windows) and history key; the renderer's
    # reasons over the same sequence equal its, including the frames it
    # refuses (a non-increasing index, an empty extent) without touching the
    # history. Then vulkan_backend.cpp:1770-1789: with a fresh batch, a
    # change of transport settings (bounces, the sample index, the clear
    # colour) is RenderSettingsChanged; without one (surfaces only) the
    # history carries over them.
    import ctypes as C
    R = oracle.ref()
    sc = _cornell(renderer)
    seq = [  # frame index, cut, width, height, display window, data window, history key
        (40000, 0, 32, 24, (0, 0, 0, 0), (0, 0, 0, 0), 0),
        (40001, 0, 32, 24, (0, 0, 0, 0), (0, 0, 0, 0), 0),
        (40002, 1, 32, 24, (0, 0, 0, 0), (0, 0, 0, 0), 0),      # cut
        (40004, 0, 32, 24, (0, 0, 0, 0), (0, 0, 0, 0), 0),      # gap
        (40005, 0, 32, 24, (0, 0, 0, 0), (0, 0, 0, 0), 0),
        (40006, 0, 40, 24, (0, 0, 0, 0), (0, 0, 0, 0), 0),      # extent
        (40007, 0, 40, 24, (1.5, 0, 40, 24), (0, 0, 0, 0), 0),  # display window
        (40008, 0, 40, 24, (1.5, 0, 40, 24), (2, 3, 30, 20), 0),  # data window
        (40009, 0, 40, 24, (1.5, 0, 40, 24), (2, 3, 30, 20), 7),  # history key
        (40009, 0, 40, 24, (1.5, 0, 40, 24), (2, 3, 30, 20), 7),  # not increasing: refused
        (40010, 0, 0, 24, (1.5, 0, 40, 24), (2, 3, 30, 20), 7),   # empty: refused
        (40010, 0, 40, 24, (1.5, 0, 40, 24), (2, 3, 30, 20), 7),
        (40011, 1, 40, 24, (0, 0, 0, 0), (0, 0, 0, 0), 7),        # a cut with a changed target: the cut
    ]
    n = len(seq)
    fi = np.array([s[0] for s in seq], np.uint64)
    cut = np.array([s[1] for s in seq], np.uint32)
    ext = np.array([(s[2], s[3]) for s in seq], np.uint32)
    disp = np.array([s[4] for s in seq], np.float32)
    data = np.array([s[5] for s in seq], np.int32)
    key = np.array([s[6] for s in seq], np.uint64)
    want = np.zeros(n, np.uint32)
    R.ref_temporal_advance(n, fi.ctypes.data_as(C.POINTER(C.c_uint64)), oracle.ptr(cut, C.c_uint32),
                           oracle.ptr(ext, C.c_uint32), oracle.ptr(disp, C.c_float),
                           data.ctypes.data_as(C.POINTER(C.c_int32)), key.ctypes.data_as(C.POINTER(C.c_uint64)),
                           oracle.ptr(want, C.c_uint32))
    renderer.render_temporal(sc.world_to_clip, sc.view, 8, 8, frame_index=39990, camera_cut=True, max_bounces=2)
    got = []
    for k, (f, c, w, h, dw, dd, hk) in enumerate(seq):
        try:
            out = renderer.render_temporal(sc.world_to_clip, sc.view, w, h, frame_index=f, camera_cut=bool(c),
                                           history_key=hk, display_window=dw, data_window=dd, max_bounces=2)
            got.append(out["reset_reason"])
        except hip.LotusError as err:
            assert ("must increase" in str(err)) or ("extent must be nonempty" in str(err)), err
            got.append(0xFFFFFFFF)
    # the renderer's own history before the sequence makes its first frame a
    # continuation or a gap, not a first frame: compare from the second on
    print("reference", list(want), "renderer", got)
    assert want[0] == 1 and got[0] == 4
    assert list(want[1:]) == got[1:]
    assert set(want) >= {0, 1, 3, 4, 5, 6, 0xFFFFFFFF}
    # transport settings: only frames with a fresh batch drop history over them
    kw = dict(history_key=7, max_bounces=2)
    f = 40100
    renderer.render_temporal(sc.world_to_clip, sc.view, 32, 24, frame_index=f, camera_cut=True, accumulate=True, **kw)
    for change, expect in ((dict(), 0), (dict(max_bounces=3), 7), (dict(sample_index=9), 7),
                           (dict(clear_color=(0.5, 0.0, 0.0, 1.0)), 7), (dict(), 0)):
        f += 1
        args = dict(kw, **change)
        out = renderer.render_temporal(sc.world_to_clip, sc.view, 32, 24, frame_index=f, accumulate=True, **args)
        assert out["reset_reason"] == expect, (change, out["reset_reason"])
        f += 1  # back to the settings before the change (itself a change when the last was one)
        renderer.render_temporal(sc.world_to_clip, sc.view, 32, 24, frame_index=f, accumulate=True, **kw)
    f += 1
    out = renderer.render_temporal(sc.world_to_clip, sc.view, 32, 24, frame_index=f, **dict(kw, max_bounces=5))
    assert out["reset_reason"] == 0  # no fresh batch: the surfaces' history carries over


def test_explicit_instance_identities_reach_the_device(renderer):
    # The per-instance identities are host arrays the call copies to the
    # device: explicit arrays equal to the defaults (instance + 1, material
    # slot + 1, rigid and unchanged, static placement) give the default
    # frames bit for bit, and a placement map moves every motion vector.
    sc = _cornell(renderer)
    n = len(sc.instances)
    ids = np.arange(1, n + 1, dtype=np.uint64)
    mids = np.array([m + 1 for (_, _, m) in sc.instances], np.uint64)
    flags = np.full(n, 6, np.uint32)
    still = np.tile(np.eye(4, dtype=np.float32).reshape(16), n)
    w = h = 48
    outs = []
    for k, explicit in enumerate([False, True]):
        extra = dict(instance_ids=ids, material_ids=mids, instance_flags=flags,
                     previous_from_current=still) if explicit else {}
        for f in range(2):
            print(f"identities explicit={explicit} frame {f}", flush=True)
            out = renderer.render_temporal(sc.world_to_clip, sc.view, w, h, frame_index=50000 + 10 * k + f,
                                           history_key=300 + k, accumulate=True, max_history=8, max_bounces=4,
                                           want_products=True, **extra)
        outs.append(out)
    assert np.array_equal(outs[0]["surface"].view(np.uint8), outs[1]["surface"].view(np.uint8))
    assert np.array_equal(outs[0]["validity"].view(np.uint8), outs[1]["validity"].view(np.uint8))
    assert np.array_equal(outs[0]["color"].view(np.uint32), outs[1]["color"].view(np.uint32))
    moved = still.reshape(n, 16).copy()
    moved[:, 12] = 0.05  # each point's previous position 0.05 further along +x
    # the explicit stream's next frame, whose previous frame has the same view
    out = renderer.render_temporal(sc.world_to_clip, sc.view, w, h, frame_index=50012, history_key=301,
                                   accumulate=True, max_history=8, max_bounces=4, want_products=True,
                                   instance_ids=ids, material_ids=mids, instance_flags=flags,
                                   previous_from_current=moved.reshape(-1))
    hit = (out["surface"]["flags"] & 1) != 0
    assert hit.mean() > 0.9
    mx = np.abs(out["surface"]["motion_depth"][hit][:, 0])
    # 0.05 along x moves the box's points 0.72 to 1.19 px at 48 x 48 (5th to
    # 95th percentile); points that leave the view keep no motion
    assert np.mean(mx > 0.5) > 0.9 and mx.max() < 1.5
