"""Radiance cache on the GPU against the cache oracle (oracle/nrc.py):
inference and gradients at the device's FP16 rounding points, Adam's first
step, online training, and cache frames against plain frames (exact where
the cache cannot contribute)."""
import numpy as np
import pytest

from oracle import nrc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _hip():
    from lotus_amd import hip
    return hip


def _params(cfg, seed, grid_scale=1.0):
    rng = np.random.default_rng(seed)
    T = 1 << cfg["log2_table"]
    grid = (rng.uniform(-1, 1, (16, T, 2)) * grid_scale).astype(np.float32)
    w = np.concatenate([rng.uniform(-1, 1, r * c) * np.sqrt(6.0 / (r + c)) for r, c in nrc.SHAPES])
    return w.astype(np.float32), grid


def _samples(n, seed, target=None):
    rng = np.random.default_rng(seed)
    q = np.zeros((n, 20), np.float32)
    q[:, 0:3] = rng.uniform(0, 1, (n, 3))
    q[:4, 0:3] = [[0, 0, 0], [1, 1, 1], [1, 0, 0.5], [0.25, 1, 0]]  # cube faces and corners
    d = rng.normal(size=(n, 3))
    q[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    q[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q[:, 9:14] = rng.uniform(0, 1, (n, 5))
    q[:, 16:19] = rng.uniform(0, 2, (n, 3)) if target is None else target(q)
    return q


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")


@pytest.fixture(scope="module")
def small():
    hip = _hip()
    r = hip.Renderer(0)
    cfg = r.cache_configure(log2_table=12, min_resolution=4, max_resolution=256, batch_capacity=4096)
    yield r, cfg
    r.close()


def test_cache_levels(small):
    _, cfg = small
    res = cfg["resolution"]
    assert res[0] == 4 and res[-1] <= 256 and all(a <= b for a, b in zip(res, res[1:]))
    for l, rr in enumerate(res):
        assert ((cfg["dense_levels"] >> l) & 1) == int((rr + 1) ** 3 <= 4096)
    assert 0 < cfg["dense_levels"] < (1 << 16)  # both dense and hashed levels are exercised


def test_cache_forward_matches_oracle(small):
    r, cfg = small
    w, grid = _params(cfg, 1)
    r.cache_parameters(w, grid)
    q = _samples(1000, 2)
    out = torch.zeros((1000, 3), device="cuda")
    qd = _dev(q[:, :16])
    r.cache_infer(qd.data_ptr(), 1000, out.data_ptr())
    got = out.cpu().numpy().astype(np.float64)
    ref, _ = nrc.device_forward(q[:, :16].astype(np.float64), w.astype(np.float64), grid.astype(np.float64),
                                cfg["resolution"])
    scale = np.abs(ref).max()
    err = np.abs(got - ref)
    print(f"forward: scale {scale:.3g}, max err {err.max():.3g}, mean err {err.mean():.3g}")
    # FP32 accumulation against float64: an FP16 rounding may flip by one ulp
    assert err.max() < 1e-2 * scale
    assert err.mean() < 1e-3 * scale
    # read-back of what was written
    w2, g2 = r.cache_parameters()
    assert np.array_equal(w2, w) and np.array_equal(g2, grid)


def test_cache_gradients_match_oracle(small):
    r, cfg = small
    w, grid = _params(cfg, 3)
    r.cache_parameters(w, grid)
    n = 777  # not a multiple of the 16-sample tiles or the 32-sample K blocks
    s = _samples(n, 4)
    sd = _dev(s)
    wg, gg, loss = r.cache_gradients(sd.data_ptr(), n)
    loss_ref, wg_ref, gg_ref, _ = nrc.device_loss_and_grads(s.astype(np.float64), w.astype(np.float64),
                                                             grid.astype(np.float64), cfg["resolution"])
    print(f"loss {loss:.6g} vs {loss_ref:.6g}")
    assert abs(loss - loss_ref) < 1e-3 * loss_ref
    off = 0
    for li, (rr, cc) in enumerate(nrc.SHAPES):
        a, b = wg[off:off + rr * cc], wg_ref[off:off + rr * cc]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"layer {li}: rel err {rel:.3g}")
        assert rel < 1e-2, li
        off += rr * cc
    rel = np.linalg.norm(gg - gg_ref) / np.linalg.norm(gg_ref)
    print(f"grid: rel err {rel:.3g}")
    assert rel < 1e-2
    # entries no sample touched stay exactly zero
    assert np.all(gg[gg_ref == 0] == 0)
    # a second call starts from cleared gradients
    wg2, gg2, loss2 = r.cache_gradients(sd.data_ptr(), n)
    assert np.allclose(wg2, wg, rtol=1e-5, atol=1e-7 * np.abs(wg).max())
    # hashed levels sum in packed FP16 atomics, whose order varies: the two
    # calls are a few FP16 roundings of the summed terms apart (an entry whose
    # terms cancel keeps their absolute rounding), far from twice the sums
    assert np.allclose(gg2, gg, rtol=4e-3, atol=1e-3 * np.abs(gg).max())


def test_cache_adam_first_step(small):
    # Adam's first bias-corrected step moves every parameter with a non-zero
    # gradient by the learning rate against the gradient's sign, and leaves
    # untouched grid entries alone.
    r, cfg = small
    w, grid = _params(cfg, 5)
    r.cache_parameters(w, grid)
    n = 512
    sd = _dev(_samples(n, 6))
    wg, gg, _ = r.cache_gradients(sd.data_ptr(), n)
    r.cache_train(sd.data_ptr(), n)
    w1, g1 = r.cache_parameters()
    lr = cfg["learning_rate"]
    big = np.abs(wg) > 1e-4 * np.abs(wg).max()
    assert np.allclose((w - w1)[big], lr * np.sign(wg[big]), rtol=1e-3, atol=1e-7)
    assert np.array_equal(w1[wg == 0], w[wg == 0])
    bigg = np.abs(gg) > 1e-4 * np.abs(gg).max()
    assert np.allclose((grid - g1)[bigg], lr * np.sign(gg[bigg]), rtol=1e-3, atol=1e-7)
    assert np.array_equal(g1[gg == 0], grid[gg == 0])


def test_cache_learns_a_smooth_function():
    hip = _hip()
    r = hip.Renderer(0)
    cfg = r.cache_configure(log2_table=14, min_resolution=8, max_resolution=512, batch_capacity=8192,
                            learning_rate=1e-2)

    def target(q):
        x, y, z = q[:, 0], q[:, 1], q[:, 2]
        return np.stack([0.5 + 0.4 * np.sin(6 * x) * np.cos(4 * z), 0.3 + 0.25 * y + 0.2 * q[:, 5],
                         0.6 + 0.3 * np.cos(5 * (x + z))], 1).astype(np.float32)

    losses = []
    for step in range(300):
        sd = _dev(_samples(8192, 100 + step, target))
        losses.append(r.cache_train(sd.data_ptr(), 8192))
    test = _samples(4096, 7, target)
    out = torch.zeros((4096, 3), device="cuda")
    r.cache_infer(_dev(test[:, :16]).data_ptr(), 4096, out.data_ptr())
    pred = out.cpu().numpy()
    relmse = np.mean((pred - test[:, 16:19]) ** 2 / (test[:, 16:19] ** 2 + 1e-2))
    print(f"loss {losses[0]:.4g} -> {np.mean(losses[-10:]):.4g}; held-out relMSE {relmse:.4g}")
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-10:]) < 0.05 * losses[0]
    assert relmse < 0.01
    r.close()


def _mix(x):
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_cache_frame_with_zero_network_matches_plain_frames():
    # With a zero network the cache adds nothing: training paths are plain
    # depth-6 paths and every other path is a plain depth-1 path (emission and
    # direct light at the first hit, MIS-weighted emission at the second), so
    # both must match plain frames bit for bit.
    hip = _hip()
    from lotus_amd import scenes
    sc = scenes.cornell_lambert()
    W = H = 64
    r = hip.Renderer(0)
    r.apply(sc)
    cfg = r.cache_configure(log2_table=12, min_resolution=4, max_resolution=128, train_paths=256, query_depth=1)
    r.cache_parameters(np.zeros(hip.CACHE_PARAMS, np.float32), None)
    out = r.render(sc.world_to_clip, W, H, max_bounces=6, cache=True, want_color=False)
    cache_sums, _ = r.accumulation(W, H)
    r.render(sc.world_to_clip, W, H, max_bounces=1, want_color=False)
    short, _ = r.accumulation(W, H)
    r.render(sc.world_to_clip, W, H, max_bounces=6, want_color=False)
    full, _ = r.accumulation(W, H)
    stride = max(1, (W * H) // 256)
    offset = _mix((0 * 0x9E3779B9 + 0x632BE5AB) & 0xFFFFFFFF) % stride
    slot = np.arange(W * H).reshape(H, W)
    train = (slot + offset) % stride == 0
    print(f"queries {out['cache_queries']}, samples {out['cache_samples']}, training pixels {train.sum()}")
    assert out["cache_queries"] > 0 and out["cache_samples"] > 0
    assert np.array_equal(cache_sums[train], full[train])
    assert np.array_equal(cache_sums[~train], short[~train])
    r.close()


def _relmse(img, ref):
    return float(np.mean((img[..., :3] - ref) ** 2 / (ref ** 2 + 1e-2)))


def test_trained_cache_beats_plain_and_terminated_paths():
    # What the trained network gains: on the Cornell box (long paths, smooth
    # indirect light) 32 cache frames, trained online from a fresh cache, have
    # a lower relMSE against a 2048-spp plain reference than 32 plain frames,
    # and carry the indirect light that the same paths terminated without a
    # cache (plain depth-1 frames) lack.
    hip = _hip()
    from lotus_amd import scenes
    sc = scenes.cornell_lambert()
    W = H = 128
    r = hip.Renderer(0)
    r.apply(sc)
    for _ in range(2048):
        out = r.render(sc.world_to_clip, W, H, max_bounces=8, sample_index=1 << 20, want_color=False)
    ref = r.render(sc.world_to_clip, W, H, max_bounces=8, sample_index=1 << 20, frame_count=0)["color"][..., :3]
    ref = ref.astype(np.float64)
    for _ in range(32):
        plain = r.render(sc.world_to_clip, W, H, max_bounces=8)["color"]
    for _ in range(32):
        direct = r.render(sc.world_to_clip, W, H, max_bounces=1)["color"]
    r.cache_configure(log2_table=16, train_paths=2048, train_steps=4, batch_capacity=8192)
    losses = []
    for _ in range(32):
        out = r.render(sc.world_to_clip, W, H, max_bounces=8, cache=True)
        losses.append(out["cache_loss"])
    cached = out["color"]
    e_plain, e_direct, e_cache = _relmse(plain, ref), _relmse(direct, ref), _relmse(cached, ref)
    m_ref, m_direct, m_cache = ref.mean(), direct[..., :3].mean(), cached[..., :3].mean()
    print(f"relMSE plain {e_plain:.4g} terminated {e_direct:.4g} cache {e_cache:.4g}; "
          f"mean ref {m_ref:.4g} terminated {m_direct:.4g} cache {m_cache:.4g}; loss {losses[0]:.3g} -> {losses[-1]:.3g}")
    assert np.all(np.isfinite(cached)) and np.all(np.isfinite(losses))
    assert out["spp"] == 32 and out["cache_queries"] > 0 and out["cache_samples"] > 0
    assert e_cache < e_plain
    assert e_cache < e_direct
    assert abs(m_cache - m_ref) < 0.25 * abs(m_direct - m_ref)
    r.close()


def test_cache_frame_samples_do_not_depend_on_the_tail():
    # Training paths record the same vertices whether their later bounces run
    # as rounds or in the tail kernel: the same training samples, loss and
    # frame. The first frame sizes the rounds of the second; the network is
    # set to the same seeded weights before the second (training sums in
    # atomics of varying order, so trained weights differ in their last bits).
    hip = _hip()
    from lotus_amd import scenes
    sc = scenes.cornell_lambert()
    W = H = 64
    outs = []
    for div in (0, 1):
        r = hip.Renderer(0)
        r.apply(sc)
        r.set_tuning(tail_divisor=div)
        cfg = r.cache_configure(log2_table=12, min_resolution=4, max_resolution=128, train_paths=512,
                                query_depth=1, train_steps=1)  # the loss before any of the frame's updates
        r.render(sc.world_to_clip, W, H, max_bounces=8, cache=True, want_color=False, sample_index=3)
        w, grid = _params(cfg, 11, grid_scale=1e-2)
        r.cache_parameters(w * 0.1, grid)
        outs.append(r.render(sc.world_to_clip, W, H, max_bounces=8, cache=True, sample_index=9))
        r.close()
    a, b = outs
    print(f"rounds {a['rounds']} vs {b['rounds']}, samples {a['cache_samples']} vs {b['cache_samples']}, "
          f"loss {a['cache_loss']:.6g} vs {b['cache_loss']:.6g}")
    assert b["rounds"] < a["rounds"]
    assert a["cache_samples"] == b["cache_samples"] > 0
    assert abs(a["cache_loss"] - b["cache_loss"]) <= 1e-5 * abs(a["cache_loss"])
    assert np.array_equal(a["color"], b["color"])
