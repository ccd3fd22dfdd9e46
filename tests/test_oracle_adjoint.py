"""The oracle's FP64 adjoint (or_adjoint_pt: forward-mode derivatives carried
through the backward loop of PathTracer::Trace, src/Integrator/pathtracer.cpp:
117-354, plus the score of every parameter-dependent sampling decision)
against a closed form, exact finite differences for texels (no sampling
decision depends on a texel, so renders with one seed differ only through the
texel), and finite differences of the oracle renderer for the medium."""
import numpy as np
import pytest

import render_cases as rc


def test_pure_absorber_closed_form(oracle):
    """sigma_s = 0: a sample carries Le iff the slab is passed, with
    probability P = exp(-avg(sigma_a) (s - 1e-4)), so dI/dsigma_{a,s}k =
    -(s - 1e-4)/3 I for every channel and dI/dg = 0."""
    sc = rc.absorber_slab_scene()
    le = np.array([1.0, 2.0, 3.0])
    spp = 1 << 16
    dimg = np.ones((1, 1, 3))
    img, g, se, _ = oracle.adjoint_pt(sc, dimg, spp, 1, seed=3)
    s = 1.0  # chord of the slab along the (nearly axial) pixel rays
    p = np.exp(-np.mean([0.1, 0.4, 0.7]) * (s - 1e-4))
    # the image is Le times the pass fraction; binomial noise
    frac = img[0, 0] / le
    assert np.allclose(frac, frac[0], rtol=1e-12)
    assert abs(frac[0] - p) < 4 * np.sqrt(p * (1 - p) / spp)
    # per sample the gradient is -(s - 1e-4)/3 times the radiance: exact given the image
    want = -(s - 1e-4) / 3.0 * img[0, 0].sum()
    np.testing.assert_allclose(g[:6], want, rtol=1e-6)
    assert g[6] == 0.0


def test_texel_gradients_equal_exact_differences(oracle):
    sc = rc.textured_scene(12, 12)
    rng = np.random.default_rng(1)
    dimg = rng.uniform(0.5, 1.5, (12, 12, 3))
    spp, depth, seed = 16, 5, 21
    _, _, _, gt = oracle.adjoint_pt(sc, dimg, spp, depth, seed=seed, nthreads=8, textures=True)
    sizes = [t.shape[0] * t.shape[1] for t in sc["textures"]]
    offs = np.cumsum([0] + sizes)
    assert gt.shape == (offs[-1], 3)
    # every texture receives gradient somewhere
    for k in range(3):
        assert np.abs(gt[offs[k]:offs[k + 1]]).sum() > 0
    picks = [(0, 1, 2, 0), (0, 3, 0, 1), (1, 4, 1, 2), (1, 0, 2, 0), (2, 1, 1, 1), (2, 0, 0, 2)]
    for t, y, x, c in picks:
        idx = offs[t] + y * sc["textures"][t].shape[1] + x
        v = float(sc["textures"][t][y, x, c])
        h = 1e-3 * v
        imgs = []
        for sgn in (1, -1):
            s2 = dict(sc)
            s2["textures"] = [a.copy() for a in sc["textures"]]
            s2["textures"][t][y, x, c] = np.float32(v + sgn * h)
            hh = float(s2["textures"][t][y, x, c]) - v
            imgs.append((rc_render(oracle, s2, spp, depth, seed), hh))
        (ip, hp), (im, hm) = imgs
        fd = ((ip - im) * dimg).sum() / (hp - hm)
        np.testing.assert_allclose(gt[idx, c], fd, rtol=2e-5, atol=1e-9, err_msg=str((t, y, x, c)))


def rc_render(oracle, sc, spp, depth, seed):
    h, w = 12, 12
    return oracle.render_pt(sc, w, h, spp, depth, seed=seed, nthreads=8)[0]


def _media(p, w):
    import importlib
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "rumi-renderer_amd"))
    scenes = importlib.import_module("scenes")
    return scenes.cornell_media(w, w, sigma_a=tuple(p[:3]), sigma_s=tuple(p[3:6]), g=float(p[6]))


@pytest.mark.slow
def test_medium_gradient_matches_finite_differences(oracle):
    """Central differences (step 10% of each parameter) of the oracle renderer
    with common random numbers, against the adjoint, 16 x 16 pixels of the
    media Cornell box at depth 16; each side averaged over independent seeds.
    (A 10% step keeps the estimator's jumps -- a sampling decision that flips
    between the two renders -- rare enough for a test of this size; the
    central-difference bias is below 1e-3 relative here.)"""
    theta = np.array([0.05, 0.2, 0.4, 1.0, 1.0, 1.0, 0.7])
    w, spp, depth, runs = 16, 1024, 16, 12
    dimg = np.random.default_rng(0).uniform(0.5, 1.5, (w, w, 3))
    ga, fd = [], []
    for r in range(runs):
        ga.append(oracle.adjoint_pt(_media(theta, w), dimg, spp, depth, seed=10 + r, nthreads=8)[1])
        row = []
        for k in range(7):
            h = 0.1 * theta[k]
            tp, tm = theta.copy(), theta.copy()
            tp[k] += h
            tm[k] -= h
            ip = oracle.render_pt(_media(tp, w), w, w, spp, depth, seed=500 + r, nthreads=8)[0]
            im = oracle.render_pt(_media(tm, w), w, w, spp, depth, seed=500 + r, nthreads=8)[0]
            row.append(((ip - im) * dimg).sum() / (2 * h))
        fd.append(row)
    ga, fd = np.array(ga), np.array(fd)
    se = np.sqrt(ga.var(0, ddof=1) / runs + fd.var(0, ddof=1) / runs)
    z = (ga.mean(0) - fd.mean(0)) / se
    print("adjoint", ga.mean(0), "\nFD", fd.mean(0), "\nz", z)
    assert np.all(np.abs(z) < 4.0), z


def test_image_based_light_texels_get_their_light_sample_gradients(oracle):
    """The image-based light's texels enter the image twice: where the sky is
    seen or reached through a specular chain (the texture of the hit vertex)
    and through every light sample of a matte vertex (ImageBasedLight::L,
    imagebasedlight.cpp:35-64, linear in the texel it reads).  No sampling
    decision depends on a texel, so renders with one seed are linear in each
    texel and central differences along any texel direction are exact: the
    adjoint must equal them, for random directions over the whole sky and
    for single texels lit only through light samples."""
    w = 8
    sc = rc.ibl_scene(w, w, seed=9)
    sky = sc["textures"][0]["rgb"]
    rng = np.random.default_rng(3)
    dimg = rng.uniform(0.5, 1.5, (w, w, 3))
    spp, depth, seed = 64, 4, 17
    _, _, _, gt = oracle.adjoint_pt(sc, dimg, spp, depth, seed=seed, nthreads=8, textures=True)
    assert gt.shape == (sky.shape[0] * sky.shape[1], 3)

    def loss_at(delta):
        s2 = dict(sc)
        s2["textures"] = [dict(kind="hdr", rgb=(sky.astype(np.float64) + delta).astype(np.float32))]
        img = oracle.render_pt(s2, w, w, spp, depth, seed=seed, nthreads=8)[0]
        actual = s2["textures"][0]["rgb"].astype(np.float64) - sky
        return (img * dimg).sum(), actual

    for k in range(2):
        v = rng.uniform(-1.0, 1.0, sky.shape) * 1e-2
        lp, ap = loss_at(v)
        lm, am = loss_at(-v)
        fd = lp - lm
        want = float((gt.reshape(sky.shape) * (ap - am)).sum())
        np.testing.assert_allclose(want, fd, rtol=1e-5)
    # the camera never sees the sky directly here: a texel's whole gradient
    # comes from light samples (and dielectric chains)
    big = np.argsort(-np.abs(gt).sum(1))[:3]
    for i in big:
        y, x = divmod(int(i), sky.shape[1])
        d = np.zeros_like(sky, dtype=np.float64)
        d[y, x, :] = 1e-2 * sky[y, x, :]
        lp, ap = loss_at(d)
        lm, am = loss_at(-d)
        want = float((gt[i] * (ap - am)[y, x]).sum())
        np.testing.assert_allclose(want, lp - lm, rtol=1e-5)
