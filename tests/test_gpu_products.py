"""Render-buffer writes on the GPU against the oracle (adapters/hydra2/src/
adapter.cpp:2544-2596, 2713-2743): the written buffer bitwise for every
format pair at 1080p and on a ragged size, a rendered frame's colour and
depth products written as a host's buffers, the id fill, and the refusals
(which do not latch)."""
import numpy as np
import pytest
import torch

import oracle
from lotus_amd import hip, scenes

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def renderer():
    status, detail = hip.probe()
    assert status == hip.PASS, detail
    r = hip.Renderer(0)
    yield r
    r.close()


def _device(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).cuda()


@pytest.mark.parametrize("size", [(1080, 1920), (7, 333)])
@pytest.mark.parametrize("pair", [(1, 0), (0, 1), (1, 1), (0, 0), (2, 2), (3, 3)])
def test_writes_match_oracle(renderer, size, pair):
    h, w = size
    src_fmt, dst_fmt = pair
    print(f"write {pair} {size}", flush=True)
    rng = np.random.default_rng(h + w + 10 * src_fmt + dst_fmt)
    if src_fmt == 1:
        prod = rng.uniform(-0.5, 1.5, (h, w, 4)).astype(F32)
        prod[rng.uniform(size=(h, w, 4)) < 0.02] = np.nan
        prod[0] = (np.arange(w * 4).reshape(w, 4) % 256 + 0.5).astype(F32) / F32(255)
    elif src_fmt == 0:
        prod = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    elif src_fmt == 2:
        prod = rng.uniform(0, 1, (h, w)).astype(F32)
    else:
        prod = rng.integers(-2**31, 2**31 - 1, (h, w), dtype=np.int32)
    expect = oracle.write_render_buffer(prod, src_fmt, dst_fmt)
    src = _device(prod)
    dst = torch.zeros(expect.size, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()  # the renderer's stream does not wait for torch's
    renderer.write_render_buffer(src.data_ptr(), src_fmt, w, h, dst.data_ptr(), dst_fmt)
    assert np.array_equal(dst.cpu().numpy(), expect)


def test_frame_products_and_ids(renderer):
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    w, h = 160, 96
    out = renderer.render(sc.world_to_clip, w, h, frame_count=2, max_bounces=4, light_sampling=hip.LIGHT_MIS,
                          want_depth=True)
    color = out["color"].astype(F32)
    src = _device(color)
    dst = torch.zeros(w * h * 4, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()  # the renderer's stream does not wait for torch's
    renderer.write_render_buffer(src.data_ptr(), hip.BUFFER_FLOAT32X4, w, h, dst.data_ptr(), hip.BUFFER_UNORM8X4)
    assert np.array_equal(dst.cpu().numpy(), oracle.write_render_buffer(color, 1, 0))
    depth = out["depth"].astype(F32)
    dsrc = _device(depth)
    ddst = torch.zeros(w * h * 4, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()  # the renderer's stream does not wait for torch's
    renderer.write_render_buffer(dsrc.data_ptr(), hip.BUFFER_FLOAT32, w, h, ddst.data_ptr(), hip.BUFFER_FLOAT32)
    assert np.array_equal(ddst.cpu().numpy().view(F32).reshape(h, w), np.flipud(depth))
    ids = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # the renderer's stream does not wait for torch's
    renderer.write_ids(ids.data_ptr(), hip.BUFFER_INT32, w, h, -1)
    assert bool((ids == -1).all())


def test_refusals_do_not_latch(renderer):
    a = torch.zeros(64 * 4, dtype=torch.float32, device="cuda")
    b = torch.zeros(64 * 4, dtype=torch.float32, device="cuda")
    with pytest.raises(hip.LotusError, match="size differs"):
        renderer.write_render_buffer(a.data_ptr(), hip.BUFFER_FLOAT32X4, 8, 8, b.data_ptr(), hip.BUFFER_FLOAT32X4,
                                     buffer_width=8, buffer_height=7)
    with pytest.raises(hip.LotusError, match="does not convert"):
        renderer.write_render_buffer(a.data_ptr(), hip.BUFFER_FLOAT32, 8, 8, b.data_ptr(), hip.BUFFER_INT32)
    with pytest.raises(hip.LotusError, match="Int32"):
        renderer.write_ids(b.data_ptr(), hip.BUFFER_FLOAT32, 8, 8, 3)
    torch.cuda.synchronize()  # the renderer's stream does not wait for torch's
    renderer.write_render_buffer(a.data_ptr(), hip.BUFFER_FLOAT32X4, 8, 8, b.data_ptr(), hip.BUFFER_FLOAT32X4)
