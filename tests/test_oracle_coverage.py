"""Pins the oracle's coverage (scene.slang:388-460: Coverage, AcceptSurface,
ShadowTransmittance, and the covered emitter's radiance in SampleTriangle,
:656-657) beyond the reference's alpha fixture images, which
test_oracle_materials.py compares by DES-Q5: through a stack of cards of
coverage c_k a closest-hit ray stops at card k with probability
c_k prod_{j<k} (1 - c_j) (a chi-square test over 200,000 rays), the
decisions do not depend on the order candidates are met in (the BVH walk
and brute force agree ray by ray), a shadow ray carries prod (1 - c_k) of
the cards on its segment (exactly, from the float factors multiplied in
double), 0 at an opaque card, and a thresholded opacity is 0 or 1; an
opacity lookup's coverage is 1 - transmittance = the texel's clamped
channel (an independent numpy bilinear lookup)."""
import numpy as np

import oracle
from lotus_amd import scenes

from test_oracle_textures import _np_lookup

FAR = np.float32(3.0e38)


def _cards(opacities, thresholds=None, gap=0.5):
    sc = scenes.Scene(name="cards")
    thresholds = thresholds or [0.0] * len(opacities)
    for k, (c, th) in enumerate(zip(opacities, thresholds)):
        m = sc.add_material(scenes.preview_material(base_color=(0.5, 0.5, 0.5), opacity=c, opacity_threshold=th))
        z = -2.0 - gap * k
        sc.add_mesh([[-1, -1, z], [1, -1, z], [1, 1, z], [-1, 1, z]], [[0, 1, 2], [0, 2, 3]], m)
    return sc


def _down_rays(n, seed, tmax=FAR, extent=0.9):
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:2] = rng.uniform(-extent, extent, (n, 2))
    rays[:, 6] = -1.0
    rays[:, 7] = tmax
    return rays


def test_acceptance_law_through_a_stack():
    c = [0.3, 0.5, 0.7, 1.0]
    o = oracle.OracleScene(_cards(c))
    rays = _down_rays(200_000, 1)
    hits = o.trace(rays)
    assert np.array_equal(hits, o.trace(rays, brute_force=True))  # order does not matter
    inst = hits[:, 0]
    assert np.all(inst != 0xFFFFFFFF)  # the last card is opaque
    expect, through = [], 1.0
    for ck in c:
        expect.append(through * ck)
        through *= 1 - ck
    counts = np.bincount(inst, minlength=4)[:4]
    e = np.array(expect) * len(rays)
    chi2 = float(np.sum((counts - e) ** 2 / e))
    print("cards hit", counts, "expected", np.round(e), "chi2", round(chi2, 2))
    assert chi2 < 21.1  # 3 degrees of freedom, p = 1e-4
    # the hit is the accepted card's, with t at its plane
    t = hits[:, 4].view(np.float32)
    assert np.allclose(t, 2.0 + 0.5 * inst, rtol=1e-6)


def test_shadow_transmittance_is_the_product():
    c = [0.3, 0.55, 0.8]
    o = oracle.OracleScene(_cards(c))
    factors = [np.float64(np.float32(1.0) - np.float32(x)) for x in c]
    for upto, tmax in ((1, 2.25), (2, 2.75), (3, 3.25)):
        rays = _down_rays(2000, upto, tmax=tmax)
        got = o.transmittance(rays)
        want = np.float32(np.prod(factors[:upto]))
        assert np.all(got == want), (upto, np.unique(got), want)
        assert np.array_equal(got, o.transmittance(rays, brute_force=True))
        assert np.all(o.occluded(rays) == 0)
    # an opaque card anywhere on the segment: 0
    o2 = oracle.OracleScene(_cards([0.3, 1.0, 0.5]))
    rays = _down_rays(2000, 7)
    assert np.all(o2.transmittance(rays) == 0.0) and np.all(o2.occluded(rays) == 1)
    # thresholds: below it the card is not there, at or above it is opaque
    o3 = oracle.OracleScene(_cards([0.4, 0.6], thresholds=[0.5, 0.5]))
    rays = _down_rays(2000, 8, tmax=2.25)
    assert np.all(o3.transmittance(rays) == 1.0)
    hits = o3.trace(_down_rays(2000, 9))
    assert np.all(hits[:, 0] == 1)  # the second card, always
    # coverage 0: never hit, never shadows
    o4 = oracle.OracleScene(_cards([0.0, 1.0]))
    assert np.all(o4.trace(_down_rays(2000, 10))[:, 0] == 1)


def lookup_card():
    """A card whose opacity is looked up (channel 3 of a float texture,
    scaled and biased, repeat and mirror wraps) at (s, t) = ((x + 1) / 2,
    (y + 1) / 2); returns the scene, texels and binding."""
    rng = np.random.default_rng(3)
    texels = rng.uniform(-0.2, 1.3, (4, 5, 4)).astype(np.float32)  # clamped to [0, 1]
    sc = scenes.Scene(name="lookup card")
    tex = sc.add_texture(texels, scenes.TEXTURE_RGBA32F)
    binding = scenes.texture_binding(tex, wrap=(scenes.WRAP_REPEAT, scenes.WRAP_MIRROR), channel=3,
                                     scale=(1, 1, 1, 0.9), bias=(0, 0, 0, 0.05))
    first = sc.add_binding(binding)
    mat = scenes.with_lookups(scenes.preview_material(), first, [0])  # slot 0: opacity
    m = sc.add_material(mat)
    sc.add_mesh([[-1, -1, -2], [1, -1, -2], [1, 1, -2], [-1, 1, -2]], [[0, 1, 2], [0, 2, 3]], m)
    # texture coordinates (s, t) = ((x + 1) / 2, (y + 1) / 2) at the corners
    corners = np.array([[0, 0], [1, 0], [1, 1], [0, 0], [1, 1], [0, 1]], np.float32)
    sc.set_texcoords(0, corners)
    return sc, texels, binding


def test_opacity_lookup():
    sc, texels, binding = lookup_card()
    o = oracle.OracleScene(sc)
    rays = _down_rays(3000, 4, tmax=3.0)
    got = o.transmittance(rays)
    s = (rays[:, 0].astype(np.float64) + 1) / 2
    t = (rays[:, 1].astype(np.float64) + 1) / 2
    want = np.array([1 - np.clip(_np_lookup(texels.astype(np.float64), binding, a, b)[3], 0, 1)
                     for a, b in zip(s, t)])
    assert np.allclose(got, want, atol=2e-5), np.abs(got - want).max()
    assert (want == 0).any() and (want == 1).any() and ((want > 0.05) & (want < 0.95)).any()
