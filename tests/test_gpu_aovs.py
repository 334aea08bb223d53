"""The named temporal and ReSTIR AOVs written on the device
(lotus_hip_write_temporal_aov, lotus_hip_write_restir_aov) against the
oracle's restatement of HdLotusRenderBuffer::WriteTemporal and WriteRestir
(oracle/aovs.py; adapter.cpp:2598-2672), bitwise, from the records of real
frames: temporal ReSTIR frames with radiance accumulation (reuse records
resolved from the spatial pass, then from the temporal pass), an initial-only
ReSTIR frame, no records at all, and a ragged size. Wrong formats and sizes
are refused without writing and without latching."""
import numpy as np
import pytest
import torch

from lotus_amd import hip, scenes
from oracle import aovs

pytestmark = pytest.mark.gpu

_FRAME = iter(range(90_000, 1 << 30))


@pytest.fixture(scope="module")
def renderer():
    status, detail = hip.probe()
    assert status == hip.PASS, detail
    r = hip.Renderer(0)
    yield r
    r.close()


def _device(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).cuda()


def _temporal_aov(renderer, aov, cur, val, rad, w, h):
    name = aovs.TEMPORAL_AOVS[aov]
    expect = aovs.temporal_aov(name, cur, val, rad)
    dst = torch.full((expect.nbytes,), 0xAB, dtype=torch.uint8, device="cuda")
    fmt = hip.BUFFER_FLOAT32X4 if aov == hip.AOV_MOTION else hip.BUFFER_INT32
    bufs = [_device(cur), _device(val), _device(rad) if rad is not None else None]
    torch.cuda.synchronize()  # the renderer's stream does not wait for torch's
    renderer.write_temporal_aov(aov, bufs[0].data_ptr(), bufs[1].data_ptr(),
                                bufs[2].data_ptr() if bufs[2] is not None else 0, w, h, dst.data_ptr(), fmt)
    got = dst.cpu().numpy()
    assert np.array_equal(got, expect.view(np.uint8).reshape(-1)), name


def _restir_aov(renderer, aov, initial, resolved, w, h):
    name = aovs.RESTIR_AOVS[aov]
    expect = aovs.restir_aov(name, initial, resolved, shape=(h, w))
    dst = torch.full((expect.nbytes,), 0xCD, dtype=torch.uint8, device="cuda")
    fmt = hip.BUFFER_FLOAT32X4 if aov in (hip.AOV_RESTIR_WEIGHT, hip.AOV_RESTIR_DIRECT) else hip.BUFFER_INT32
    ini = _device(initial) if initial is not None else None
    res = _device(resolved) if resolved is not None else None
    torch.cuda.synchronize()  # the renderer's stream does not wait for torch's
    renderer.write_restir_aov(aov, ini.data_ptr() if ini is not None else 0, res.data_ptr() if res is not None else 0,
                              w, h, dst.data_ptr(), fmt)
    got = dst.cpu().numpy()
    assert np.array_equal(got, expect.view(np.uint8).reshape(-1)), name


@pytest.mark.parametrize("size", [(96, 64), (37, 23)])
def test_aovs_of_temporal_restir_frames(renderer, size):
    w, h = size
    sc = scenes.cornell_lambert()
    renderer.apply(sc)
    w2c = scenes.multiply(sc.projection, sc.view)
    products = []
    for f in range(3):
        print(f"temporal frame {f} {w}x{h}", flush=True)
        out = renderer.render_temporal(w2c, sc.view, w, h, frame_index=next(_FRAME), camera_cut=f == 0,
                                       accumulate=True, max_history=8, restir_temporal=True, restir_spatial=f == 2,
                                       radius=4, max_history_candidates=20, light_sampling=hip.LIGHT_RESTIR,
                                       restir_candidates=8, want_products=True, want_restir=True)
        products.append(out)
    for f, out in enumerate(products):
        cur = out["surface"].view(aovs.TEMPORAL_PIXEL).reshape(h, w)
        val = out["validity"].view(aovs.VALIDITY_PIXEL).reshape(h, w)
        rad = out["radiance"].view(aovs.RADIANCE_PIXEL).reshape(h, w)
        assert rad["sample_count"].max() > 0
        for aov in range(4):
            _temporal_aov(renderer, aov, cur, val, rad, w, h)
        _temporal_aov(renderer, hip.AOV_HISTORY_LENGTH, cur, val, None, w, h)
        initial = out["restir"].view(aovs.RESTIR_PIXEL).reshape(h, w)
        records = out["restir_spatial"] if f == 2 else out["restir_temporal"]
        resolved = records.view(aovs.RESTIR_REUSE_PIXEL).reshape(h, w)
        assert resolved["reservoir"]["candidate_count"].max() > 0
        for aov in range(6):
            _restir_aov(renderer, aov, initial, resolved, w, h)
            _restir_aov(renderer, aov, initial, None, w, h)
            _restir_aov(renderer, aov, None, None, w, h)


def test_aov_refusals_do_not_write_or_latch(renderer):
    w, h = 8, 4
    cur = np.zeros((h, w), aovs.TEMPORAL_PIXEL)
    val = np.zeros((h, w), aovs.VALIDITY_PIXEL)
    c, v = _device(cur), _device(val)
    dst = torch.full((w * h * 16,), 0x5A, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(hip.LotusError):  # motion is Float32Vec4
        renderer.write_temporal_aov(hip.AOV_MOTION, c.data_ptr(), v.data_ptr(), 0, w, h, dst.data_ptr(),
                                    hip.BUFFER_INT32)
    with pytest.raises(hip.LotusError):  # another size
        renderer.write_temporal_aov(hip.AOV_HISTORY_VALIDITY, c.data_ptr(), v.data_ptr(), 0, w, h, dst.data_ptr(),
                                    hip.BUFFER_INT32, w, h + 1)
    with pytest.raises(hip.LotusError):  # the validity records are missing
        renderer.write_temporal_aov(hip.AOV_HISTORY_VALIDITY, c.data_ptr(), 0, 0, w, h, dst.data_ptr(),
                                    hip.BUFFER_INT32)
    with pytest.raises(hip.LotusError):  # the weight AOV is Float32Vec4
        renderer.write_restir_aov(hip.AOV_RESTIR_WEIGHT, 0, 0, w, h, dst.data_ptr(), hip.BUFFER_INT32)
    assert bool((dst == 0x5A).all())
    _temporal_aov(renderer, hip.AOV_HISTORY_VALIDITY, cur, val, None, w, h)  # still usable
