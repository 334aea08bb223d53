"""lotus_hip_svgf_static, the entry point that allocates nothing, against
lotus_hip_svgf through the C ABI: two calls in a row, the second on a
smaller image, each equal bit for bit to the allocating entry point's
variance words and filtered colour (the same kernels over the same inputs);
an image larger than the static scratch fails without writing."""
import ctypes as C

import numpy as np
import pytest

from lotus_amd import hip

pytestmark = [pytest.mark.gpu]

torch = pytest.importorskip("torch")


def _pack_normal(n):
    s = np.abs(n).sum(-1, keepdims=True)
    e = n[..., :2] / s
    neg = n[..., 2] < 0
    sign = np.where(e >= 0, 1.0, -1.0)
    fold = (1.0 - np.abs(e[..., ::-1])) * sign
    e = np.where(neg[..., None], fold, e)
    q = np.rint(np.clip(e, -1, 1) * 32767).astype(np.int64)
    return ((q[..., 0] & 0xFFFF) | ((q[..., 1] & 0xFFFF) << 16)).astype(np.uint32)


def _records(w, h, seed):
    """Seeded TemporalPixel and TemporalRadiancePixel records: three
    instances, smooth depths and normals, some pixels without history."""
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:h, 0:w]
    surf = np.zeros((h, w), hip.TEMPORAL_PIXEL)
    inst = 1 + (xs * 3 // w) + 10 * (ys > h // 2)
    n = np.stack([0.1 * np.sin(xs / 5.0), 0.1 * np.cos(ys / 7.0), np.ones((h, w))], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    surf["normal_depth"][..., :3] = n
    surf["normal_depth"][..., 3] = 2.0 + 0.01 * xs + 0.02 * ys
    surf["instance_id"] = inst
    surf["material_id"] = inst
    surf["flags"] = 7
    surf["albedo"] = rng.integers(0, 1 << 24, (h, w), dtype=np.uint32)
    surf["shading_normal"] = _pack_normal(n)
    rad = np.zeros((h, w), hip.TEMPORAL_RADIANCE)
    count = rng.integers(1, 60, (h, w)).astype(np.uint32)
    count[rng.random((h, w)) < 0.1] = 0
    rad["sample_count"] = count
    rad["color"][..., :3] = rng.lognormal(-1, 1, (h, w, 3))
    rad["color"][..., 3] = 1.0
    m1 = rng.lognormal(-1, 0.7, (h, w))
    rad["first_moment"] = m1
    rad["second_moment"] = m1 * m1 + rng.lognormal(-3, 1, (h, w))
    return surf, rad


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


def _call(name, r, ds, dr, w, h, batch, cap, iterations, color):
    buf = C.create_string_buffer(256)
    fn = getattr(hip.lib(), name)
    status = fn(r.handle, C.c_void_p(ds.data_ptr()), C.c_void_p(dr.data_ptr()), w, h, batch, cap, iterations,
                C.c_void_p(color.data_ptr()), buf, 256)
    return status, buf.value.decode()


def test_static_scratch_matches_the_allocating_entry_point():
    r = hip.Renderer(0)
    try:
        for w, h, seed in ((333, 217, 1), (97, 61, 2)):  # the second call is smaller
            print(f"svgf {w}x{h} seed {seed}", flush=True)
            surf, rad = _records(w, h, seed)
            outs = {}
            for name in ("lotus_hip_svgf", "lotus_hip_svgf_static"):
                ds, dr = _dev(surf), _dev(rad)
                color = torch.full((w * h * 4,), -1.0, dtype=torch.float32, device="cuda")
                status, detail = _call(name, r, ds, dr, w, h, 2, 24, 5, color)
                assert status == hip.PASS, (name, detail)
                torch.cuda.synchronize()
                outs[name] = (dr.cpu().numpy(), color.cpu().numpy())
            a, b = outs["lotus_hip_svgf"], outs["lotus_hip_svgf_static"]
            assert np.array_equal(a[0], b[0]), "variance words differ"
            assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), "filtered colour differs"
            assert np.isfinite(a[1]).all() and (a[1] != -1.0).all()
        # beyond the scratch: a clear failure before any launch (the size is
        # checked on the host, so the small buffers are never touched), and
        # the renderer goes on as before
        ds, dr = _dev(surf), _dev(rad)
        color = torch.zeros(97 * 61 * 4, dtype=torch.float32, device="cuda")
        status, detail = _call("lotus_hip_svgf_static", r, ds, dr, 2048, 1025, 2, 24, 5, color)
        assert status == hip.FAIL and "at most 2097152 pixels" in detail, detail
        status, detail = _call("lotus_hip_svgf_static", r, ds, dr, 97, 61, 2, 24, 5, color)
        assert status == hip.PASS, detail
        assert np.array_equal(color.cpu().numpy().view(np.uint32), b[1].view(np.uint32))
    finally:
        r.close()
