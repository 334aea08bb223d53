"""Coverage on the GPU (scene.slang:388-460) against the oracle, whose
coverage test_oracle_coverage.py pins by closed forms and
test_oracle_materials.py by the reference's alpha fixture images: closest
hits through covered card stacks (instance, primitive, barycentrics and t
bitwise, every ray), shadow transmittance bitwise (the product of 1 -
coverage in double), the opacity lookup's transmittance bitwise, and
frames of the reference's alpha fixture under ReSTIR initial sampling
(counts exact, one-sample radiance bit-exact for >= 99.9 % of pixels; BSDF
and NEE/MIS frames of the coverage fixtures run in test_gpu_materials.py).
Scenes without covered surfaces run the kernels without the test."""
import sys

import numpy as np
import pytest

import oracle
from lotus_amd import hip

import material_fixtures as mf
from test_oracle_coverage import _cards, _down_rays, lookup_card

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def renderer():
    r = hip.Renderer(0)
    yield r
    r.close()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oblique_rays(n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:2] = rng.uniform(-0.6, 0.6, (n, 2))
    d = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n), -np.ones(n)], -1)
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = 3.0e38
    return rays


def test_closest_hits_and_transmittance_through_cards(renderer):
    for c in ([0.3, 0.5, 0.7, 1.0], [0.25, 0.0, 0.6, 0.9]):
        sc = _cards(c)
        renderer.apply(sc)
        o = oracle.OracleScene(sc)
        for rays in (_down_rays(200_000, 11), _oblique_rays(200_000, 12)):
            _say(f"cards {c}: {len(rays)} rays")
            got = renderer.trace(rays)
            want = o.trace(rays)
            assert np.array_equal(got, want), np.flatnonzero(np.any(got != want, axis=1))[:8]
            for tmax in (2.25, 2.75, 3.0e38):
                r = rays.copy()
                r[:, 7] = tmax
                t_gpu, t_cpu = renderer.transmittance(r), o.transmittance(r)
                assert t_gpu.tobytes() == t_cpu.tobytes(), (tmax, np.abs(t_gpu - t_cpu).max())
                assert np.array_equal(renderer.trace(r, any_hit=True), o.occluded(r))
        hist = np.bincount(got[:, 0][got[:, 0] != 0xFFFFFFFF], minlength=4)
        _say(f"  cards hit {hist}")


def test_opacity_lookup_transmittance(renderer):
    sc, _, _ = lookup_card()
    renderer.apply(sc)
    o = oracle.OracleScene(sc)
    for rays in (_down_rays(100_000, 21, tmax=3.0), _oblique_rays(100_000, 22)):
        t_gpu, t_cpu = renderer.transmittance(rays), o.transmittance(rays)
        assert t_gpu.tobytes() == t_cpu.tobytes(), np.abs(t_gpu - t_cpu).max()
        assert np.array_equal(renderer.trace(rays), o.trace(rays))


def test_alpha_fixture_under_restir(renderer):
    sc, size = mf.load("material-alpha-v1")
    renderer.apply(sc)
    o = oracle.OracleScene(sc)
    for proposal in (0, 2):
        _say(f"alpha fixture, ReSTIR initial, proposal {proposal}")
        fr = oracle.make_frame(sc.world_to_clip, size, size, spp=1, sample_index=5, max_bounces=64, light_sampling=2,
                               restir_candidates=8, restir_proposal=proposal)
        acc, counts, shadow = o.render(fr)
        out = renderer.render(sc.world_to_clip, size, size, frame_count=1, sample_index=5, max_bounces=64,
                              light_sampling=hip.LIGHT_RESTIR, restir_candidates=8, restir_proposal=proposal)
        n = len(out["counts"])
        assert np.array_equal(out["counts"], counts[0][:n]) and not counts[0][n:].any()
        assert out["shadow_rays"] == shadow[0]
        exact = np.mean(np.all(out["color"][..., :3] == acc[..., :3].astype(np.float32), -1))
        _say(f"  bit-exact fraction {exact:.6f}")
        assert exact > 0.999
