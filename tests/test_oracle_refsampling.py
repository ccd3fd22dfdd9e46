"""The reference's sampling (Q9) restated in the oracle
(or_render_pt_refsampling): per frame, UniformSample's d^2 jittered pixel
points (d = (int)sqrt(spp)) and each area light's RectSample points, every
draw taken by Sampler::Sampling's incremental shuffle of x and y.  Pinned by
the sets' construction (one point per grid cell, RectSample's aspect rule,
SphereSample's spiral), by the exact, noise-free coverage a stratified set
gives a half-covered pixel, by (int)sqrt(spp)^2 samples per pixel, and by
unbiasedness over frames for the point-to-square form factor.  The A/B
against the continuous sampling the renderer uses measures Q9's effect on
rel-MSE at equal spp.

Reference: src/Sample/sampler.cpp:17-66, src/Sample/uniformsample.cpp:18-58,
src/Sample/rectsample.cpp:18-102, src/Sample/spheresample.cpp:20-48,
src/Scene/scene.cpp:206-258, src/Camera/film.cpp:38-44."""
import importlib
import math
import os
import sys

import numpy as np
import pytest

import render_cases as rc

HERE = os.path.dirname(os.path.abspath(__file__))


def test_uniform_set_one_point_per_cell():
    import oracle
    for spp, d in ((16, 4), (20, 4), (32, 5), (64, 8)):
        xy = oracle.reference_point_set("pixel", spp, seed=spp)
        assert xy.shape == (d * d, 2)
        cells = np.floor(xy * d).astype(int)
        assert len({(a, b) for a, b in cells}) == d * d
        assert np.all((xy >= 0) & (xy < 1))
    # one sample: the cell centre, no jitter (UniformSample(1, jitter = false))
    np.testing.assert_array_equal(oracle.reference_point_set("pixel", 1), [[0.5, 0.5]])


def test_rect_set_aspect_rule():
    """RectSample(256, length, width): a 16 x 16 jittered grid whose points
    with y < ratio are kept row by row and stretched to the unit square; a
    2 x 1 light keeps the 8 rows below 1/2 (128 points, x stratified in 16
    columns of 8, y in 8 rows of 16), a 1 x 2 light the same with x and y
    exchanged; a square light keeps all 256."""
    import oracle
    sq = oracle.reference_point_set("rect", 256, 1.0, 1.0, seed=3)
    assert sq.shape == (256, 2)
    wide = oracle.reference_point_set("rect", 256, 2.0, 1.0, seed=3)
    assert wide.shape == (128, 2)
    assert np.all(np.bincount(np.floor(wide[:, 0] * 16).astype(int), minlength=16) == 8)
    assert np.all(np.bincount(np.floor(wide[:, 1] * 8).astype(int), minlength=8) == 16)
    tall = oracle.reference_point_set("rect", 256, 1.0, 2.0, seed=3)
    assert tall.shape == (128, 2)
    assert np.all(np.bincount(np.floor(tall[:, 1] * 16).astype(int), minlength=16) == 8)
    for xy in (sq, wide, tall):
        assert np.all((xy >= 0) & (xy <= 1))


def test_sphere_set_is_the_golden_spiral():
    import oracle
    n = 64
    xy = oracle.reference_point_set("sphere", n)
    i = np.arange(n) + 0.5
    np.testing.assert_allclose(xy[:, 0], np.pi * (1 + math.sqrt(5)) / 2 * i, rtol=1e-12)
    np.testing.assert_allclose(xy[:, 1], np.arccos(1 - 2 * i / n), rtol=1e-12)


def edge_scene(edge):
    """A one-pixel camera (1 degree) looking along +z at an emitter that
    covers the image's right part, x <= edge in pixel-jitter units (the
    camera's x = -(jitter - 1/2) 2 tan(fov/2) mirrors the jitter)."""
    tan = math.tan(math.radians(0.5))
    xe = -(edge - 0.5) * 2 * tan * 10.0
    mats = [dict(kind="emissive", kd=1.0, color=(1.0, 2.0, 3.0))]
    big = 40.0
    rect = dict(kind="rect", p=(xe, -big / 2, 10.0), l=(0.0, big, 0.0), w=(big, 0.0, 0.0),
                material=0, light=True)
    cam = dict(position=(0.0, 0.0, 0.0), basis=np.eye(3).tolist(), fov_deg=1.0, y_ratio=1.0)
    return dict(materials=mats, shapes=[rect], camera=cam, medium=None)


@pytest.mark.parametrize("spp", [16, 20])
def test_stratified_jitter_covers_half_a_pixel_exactly(oracle, spp):
    """The emitter's edge halves the pixel on a stratum boundary: 8 of the 16
    points (two columns of four) see it in every frame, with no noise at 16
    or 20 spp ((int)sqrt(20)^2 = 16 samples, divided by 16)."""
    for seed in range(8):
        m, m2 = oracle.render_pt_refsampling(edge_scene(0.5), 1, 1, spp, 1, seed=seed)
        np.testing.assert_allclose(m[0, 0], 0.5 * np.array([1.0, 2.0, 3.0]), rtol=1e-12)
    m, _ = oracle.render_pt(edge_scene(0.5), 1, 1, 16, 1, seed=1)
    frac = [oracle.render_pt(edge_scene(0.5), 1, 1, 16, 1, seed=s)[0][0, 0, 0] for s in range(8)]
    assert np.std(frac) > 0.02      # continuous jitter: binomial noise


def test_light_set_is_unbiased_over_frames(oracle):
    """256 fixed points per frame on the square light, a fresh set each
    frame: the mean over frames is the point-to-square form factor."""
    sc = rc.square_light_scene()
    want = rc.square_light_radiance(0.5, (1.0, 2.0, 3.0), 0.5, 1.0)
    vals = np.array([oracle.render_pt_refsampling(sc, 1, 1, 4096, 1, light_points=256, seed=s)[0][0, 0]
                     for s in range(24)])
    se = vals.std(0, ddof=1) / np.sqrt(len(vals))
    assert np.all(np.abs(vals.mean(0) - want) < 4 * se + 1e-4 * want), (vals.mean(0), want, se)
    # one light point: every frame renders that point's exact irradiance
    one = [oracle.render_pt_refsampling(sc, 1, 1, 64, 1, light_points=1, seed=s)[0][0, 0, 0]
           for s in range(4)]
    assert np.std(one) > 0.01 * want[0]


@pytest.mark.slow
def test_rel_mse_ab_against_continuous_sampling(oracle):
    """Q9's effect at equal spp on the Cornell box (64^2, 16 spp, depth 5,
    256 light points as the reference's Scene::Build gives its light):
    mean rel-MSE over 96 independent frames of each estimator against the
    4096-spp reference.  Measured: the reference's stratified sets give
    0.90 +- 0.10 of the continuous estimator's rel-MSE, the image means
    agree within 0.2%."""
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "rumi-renderer_amd"))
    scenes = importlib.import_module("scenes")
    sc = scenes.cornell_box(64, 64)
    ref = np.load(os.path.join(HERE, "golden", "cornell64_pt_d5.npz"))
    a = np.array([rc.rel_mse(oracle.render_pt_refsampling(sc, 64, 64, 16, 5, seed=100 + k,
                                                          nthreads=8)[0], ref["mean"])
                  for k in range(96)])
    b = np.array([rc.rel_mse(oracle.render_pt(sc, 64, 64, 16, 5, seed=200 + k, nthreads=8)[0],
                             ref["mean"]) for k in range(96)])
    ratio = a.mean() / b.mean()
    se = ratio * math.sqrt((a.std(ddof=1) / a.mean()) ** 2 / 96 + (b.std(ddof=1) / b.mean()) ** 2 / 96)
    print(f"rel-MSE reference sampling {a.mean():.5f}, continuous {b.mean():.5f}, ratio {ratio:.3f} +- {se:.3f}")
    assert 0.7 < ratio < 1.1, (ratio, se)
