"""The oracle's restatement of the reference's CPU Cornell path tracer
(RenderCornellReference, src/integrators/reference_cpu/CornellReference.cpp)
against the reference's committed golden
(rt/cpu/references/wave1/cornell-256x256-64spp-seed1.json, copied to
tests/golden/cornell/), against the reference's own code compiled into
oracle/_ref, and against what the mathematics fixes (PCG32's published
sequence, Lambert's normalisation and sampling density)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cornell", "cornell-256x256-64spp-seed1.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_scene_hash_matches_golden():
    _, _, scene_hash, _ = ol.cornell_render(8, 8, 1, 1, 1, threads=1)
    assert scene_hash == int(golden()["scene_hash_fnv1a64"], 16)


def test_scene_layout():
    """5 walls, the light quad and two boxes: 36 triangles in primitive order;
    every quad wound toward its stated normal (CornellReference.cpp:159-175)."""
    v, m, order, nodes = ol.cornell_scene()
    assert v.shape[0] == 36
    assert list(m[:10]) == [0, 0, 0, 0, 0, 0, 1, 1, 2, 2] and list(m[10:12]) == [3, 3] and set(m[12:]) == {0}
    assert sorted(order) == list(range(36))
    assert 2 * 36 // 4 - 1 <= nodes <= 2 * 36 - 1
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    # floor faces +y, ceiling -y, back wall +z, left wall (red) +x, right wall (green) -x, light -y
    want = [(1, 1), (1, 1), (1, -1), (1, -1), (2, 1), (2, 1), (0, 1), (0, 1), (0, -1), (0, -1), (1, -1), (1, -1)]
    for i, (axis, sign) in enumerate(want):
        assert np.sign(n[i, axis]) == sign, i


def test_golden_image_and_ray_counts():
    """256x256, 64 spp, 8 bounces, seed 1: the image hash, ray and shadow ray
    counts equal the golden's bit for bit."""
    g = golden()
    px, cnt, scene_hash, _ = ol.cornell_render(g["width"], g["height"], g["spp"], g["max_bounces"], g["seed"])
    assert scene_hash == int(g["scene_hash_fnv1a64"], 16)
    assert int(cnt[0]) == g["ray_count"]
    assert int(cnt[1]) == g["shadow_ray_count"]
    assert int(cnt[2]) == g["non_finite_count"]
    assert ol.linear_rgb_hash(px) == int(g["linear_rgb_hash_fnv1a64"], 16)


def test_thread_count_and_row_bands_do_not_change_the_image():
    """Per-sample generators make the image independent of the thread count
    (tests/unit/L3/DeterminismTests.cpp) and of the row band rendered."""
    a, ca, _, _ = ol.cornell_render(48, 40, 4, 8, 5, threads=1)
    b, cb, _, _ = ol.cornell_render(48, 40, 4, 8, 5, threads=3)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(ca, cb)
    band, _, _, _ = ol.cornell_render(48, 40, 4, 8, 5, threads=2, rows=(13, 29))
    assert np.array_equal(band.view(np.uint32), a[13:29].view(np.uint32))


def test_matches_reference_build():
    """Bit-identical with RenderCornellReference compiled from its sources
    (long double as double, as in the golden's build)."""
    if ol.ref_cornell() is None:
        pytest.skip("oracle/_ref/libref_cornell.so not built (reference tree absent)")
    for (w, h, spp, bounces, seed) in ((96, 64, 16, 8, 7), (33, 17, 3, 2, 11), (20, 20, 8, 1, 1)):
        a, ca, ha, _ = ol.cornell_render(w, h, spp, bounces, seed, threads=4)
        b, cb, hb = ol.ref_cornell_render(w, h, spp, bounces, seed, threads=4)
        assert ha == hb
        assert np.array_equal(ca, cb), (w, h, spp, bounces, seed)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (w, h, spp, bounces, seed)


def test_invalid_options_are_rejected():
    for args in ((0, 8, 1, 1, 1), (8, 0, 1, 1, 1), (8, 8, 0, 1, 1), (8, 8, 1, 0, 1)):
        with pytest.raises(ValueError):
            ol.cornell_render(*args, threads=1)


def test_pcg32_published_sequence():
    """pcg32_srandom_r(42, 54) -> the PCG reference implementation's
    published first outputs (pcg-c-basic's pcg32-demo)."""
    got = [int(x) for x in ol.pcg32_sequence(42, 54, 6)]
    assert got == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_lambert_pdf_normalised_and_sampling_density():
    """Lambert::Pdf integrates to one over the sphere (zero below the
    surface); Sample draws unit directions above it whose reported pdf is
    cos/pi and whose mean cosine is 2/3 (the cosine-weighted hemisphere)."""
    rng = np.random.default_rng(3)
    for n in (np.array([0.0, 0.0, 1.0]), np.array([0.3, -0.5, 0.81]), np.array([0.0, 1.0, 0.0])):
        n = n / np.linalg.norm(n)
        wo = n + np.array([0.1, 0.2, 0.0])
        wo /= np.linalg.norm(wo)
        # midpoint quadrature in (cos theta, phi) about z, rotated to n
        k = 200
        ct = (np.arange(k) + 0.5) / k * 2.0 - 1.0
        ph = (np.arange(k) + 0.5) / k * 2.0 * np.pi
        C, Ph = np.meshgrid(ct, ph, indexing="ij")
        S = np.sqrt(1 - C * C)
        dirs = np.stack([S * np.cos(Ph), S * np.sin(Ph), C], -1).reshape(-1, 3)
        # any orthonormal frame works for the integral
        t = np.cross([0.0, 0.0, 1.0] if abs(n[2]) < 0.9 else [1.0, 0.0, 0.0], n)
        t /= np.linalg.norm(t)
        b = np.cross(n, t)
        world = dirs[:, 0:1] * t + dirs[:, 1:2] * b + dirs[:, 2:3] * n
        total = sum(ol.lambert_pdf(n, wo, w) for w in world) * (2.0 / k) * (2.0 * np.pi / k)
        assert abs(total - 1.0) < 1e-3
        assert ol.lambert_pdf(n, wo, -n) == 0.0
        cos = []
        for i in range(2000):
            out = ol.lambert_sample(n, wo, int(rng.integers(1 << 62)), int(rng.integers(1 << 62)))
            d, pdf = out[:3], out[3]
            assert abs(np.linalg.norm(d) - 1.0) < 1e-12
            c = float(d @ n)
            assert c > 0.0
            assert abs(pdf - c / np.pi) <= 1e-12
            cos.append(c)
        # E[cos] = 2/3, sd = sqrt(1/2 - 4/9) = 0.236
        assert abs(np.mean(cos) - 2.0 / 3.0) < 5 * 0.236 / np.sqrt(len(cos))
