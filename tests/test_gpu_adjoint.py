"""GPU path-replay adjoint (rumi_adjoint through the C ABI) against a closed
form, exact texel differences, the oracle's FP64 adjoint and finite
differences of the GPU renderer (the estimator the adjoint differentiates is
PathTracer::Trace's, src/Integrator/pathtracer.cpp:117-354)."""
import importlib
import os
import sys

import numpy as np
import pytest

import render_cases as rc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def scenes():
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "rumi-renderer_amd"))
    return importlib.import_module("scenes")


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def adjoint(rumi, ctx, sc, dimg, spp, depth, seed, textures=False, **kw):
    import torch
    scene = rumi.Scene(ctx, sc)
    img, dmed, dtex, st = scene.adjoint(_dev(dimg), spp, depth, seed=seed, textures=textures, **kw)
    torch.cuda.synchronize()
    out = (img.cpu().numpy().astype(np.float64), dmed.cpu().numpy().astype(np.float64),
           None if dtex is None else dtex.cpu().numpy().astype(np.float64), st)
    scene.close()
    return out


def render(rumi, ctx, sc, w, h, spp, depth, seed):
    import torch
    scene = rumi.Scene(ctx, sc)
    img, _ = scene.render(w, h, spp, depth, seed=seed)
    torch.cuda.synchronize()
    out = img.cpu().numpy().astype(np.float64)
    scene.close()
    return out


def test_pure_absorber_closed_form(rumi, ctx):
    sc = rc.absorber_slab_scene()
    spp = 1 << 18
    img, g, _, st = adjoint(rumi, ctx, sc, np.ones((1, 1, 3)), spp, 1, seed=5)
    p = np.exp(-np.mean([0.1, 0.4, 0.7]) * (1.0 - 1e-4))
    frac = img[0, 0] / np.array([1.0, 2.0, 3.0])
    assert abs(frac[0] - p) < 4 * np.sqrt(p * (1 - p) / spp) + 1e-6
    want = -(1.0 - 1e-4) / 3.0 * img[0, 0].sum()
    np.testing.assert_allclose(g[:6], want, rtol=1e-4)
    assert abs(g[6]) < 1e-6 * abs(want)
    assert st.samples == spp


def test_primal_image_equals_render(rumi, ctx, scenes):
    sc = scenes.cornell_media(24, 20)
    img, _, _, _ = adjoint(rumi, ctx, sc, np.ones((20, 24, 3)), 16, 16, seed=77)
    ref = render(rumi, ctx, sc, 24, 20, 16, 16, seed=77)
    np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-7)


def test_texel_gradients_equal_exact_differences(rumi, ctx):
    """No sampling decision depends on a texel, so two renders with one seed
    differ only through it: central differences are exact up to float
    rounding, the order of the film's float atomics and the h^2 term.  The
    texels tested are each texture's most influential ones; the tolerance
    includes the difference two identical renders show."""
    sc = rc.textured_scene(16, 16)
    dimg = np.random.default_rng(2).uniform(0.5, 1.5, (16, 16, 3))
    spp, depth, seed = 64, 5, 31
    _, _, gt, _ = adjoint(rumi, ctx, sc, dimg, spp, depth, seed, textures=True)
    sizes = [t.shape[0] * t.shape[1] for t in sc["textures"]]
    offs = np.cumsum([0] + sizes)
    assert gt.shape == (offs[-1], 3)
    base = render(rumi, ctx, sc, 16, 16, spp, depth, seed)
    again = render(rumi, ctx, sc, 16, 16, spp, depth, seed)
    atomics = abs(((base - again) * dimg).sum())
    picks = []
    for t in range(3):
        g = np.abs(gt[offs[t]:offs[t + 1]])
        for rank in (0, 1):
            flat = np.argsort(g.reshape(-1))[::-1][rank]
            picks.append((t, int(flat // 3), int(flat % 3)))
    for t, texel, c in picks:
        w = sc["textures"][t].shape[1]
        y, x = texel // w, texel % w
        v = float(sc["textures"][t][y, x, c])
        res = []
        for sgn in (1, -1):
            s2 = dict(sc)
            s2["textures"] = [a.copy() for a in sc["textures"]]
            s2["textures"][t][y, x, c] = np.float32(v + sgn * 0.05 * v)
            res.append((render(rumi, ctx, s2, 16, 16, spp, depth, seed),
                        float(s2["textures"][t][y, x, c])))
        (ip, vp), (im, vm) = res
        fd = ((ip - im) * dimg).sum() / (vp - vm)
        got = gt[offs[t] + texel, c]
        assert got != 0.0
        tol = 1e-2 * abs(fd) + 4 * atomics / (vp - vm) + 1e-7
        assert abs(got - fd) <= tol, (t, y, x, c, got, fd, tol)


def test_texel_gradients_match_oracle(rumi, oracle, ctx):
    sc = rc.textured_scene(16, 16)
    dimg = np.random.default_rng(3).uniform(0.5, 1.5, (16, 16, 3))
    runs = 6
    g = np.array([adjoint(rumi, ctx, sc, dimg, 256, 5, seed=40 + k, textures=True)[2]
                  for k in range(runs)])
    c = np.array([oracle.adjoint_pt(sc, dimg, 256, 5, seed=60 + k, nthreads=8, textures=True)[3]
                  for k in range(runs)])
    # texture totals (per texture and channel) are well resolved statistically
    sizes = [t.shape[0] * t.shape[1] for t in sc["textures"]]
    offs = np.cumsum([0] + sizes)
    for t in range(3):
        gs = g[:, offs[t]:offs[t + 1]].sum(1)
        cs = c[:, offs[t]:offs[t + 1]].sum(1)
        se = np.sqrt(gs.var(0, ddof=1) / runs + cs.var(0, ddof=1) / runs)
        z = (gs.mean(0) - cs.mean(0)) / se
        print("texture", t, gs.mean(0), cs.mean(0), z)
        assert np.all(np.abs(z) < 4.0), (t, z)


def _media_params(scenes, p, w):
    return scenes.cornell_media(w, w, sigma_a=tuple(p[:3]), sigma_s=tuple(p[3:6]), g=float(p[6]))


THETA = np.array([0.05, 0.2, 0.4, 1.0, 1.0, 1.0, 0.7])


def test_medium_gradient_matches_oracle_adjoint(rumi, oracle, ctx, scenes):
    w, spp, depth, runs = 32, 1024, 16, 8
    sc = _media_params(scenes, THETA, w)
    dimg = np.random.default_rng(0).uniform(0.5, 1.5, (w, w, 3))
    g = np.array([adjoint(rumi, ctx, sc, dimg, spp, depth, seed=100 + k)[1] for k in range(runs)])
    c = np.array([oracle.adjoint_pt(sc, dimg, spp, depth, seed=200 + k, nthreads=8)[1]
                  for k in range(runs)])
    se = np.sqrt(g.var(0, ddof=1) / runs + c.var(0, ddof=1) / runs)
    z = (g.mean(0) - c.mean(0)) / se
    print("GPU", g.mean(0), "\noracle", c.mean(0), "\nz", z)
    assert np.all(np.abs(z) < 4.0), z


def test_gradients_equal_oracle_on_common_random_numbers(rumi, oracle, ctx, scenes):
    """With one seed the GPU adjoint and the oracle's FP64 forward-mode
    derivative follow the same paths (the renderers share their random-number
    numbering), so their gradients agree to rounding: the 7 medium parameters
    of the media workload at 512 x 512, 32 spp, depth 16 (the benchmark's
    gradient) within 0.1% each, and the texel gradients of a textured scene
    within 0.1% of the largest."""
    w, spp, depth = 512, 32, 16
    sc = _media_params(scenes, THETA, w)
    dimg = np.random.default_rng(4).uniform(0.5, 1.5, (w, w, 3))
    g = adjoint(rumi, ctx, sc, dimg, spp, depth, seed=7)[1]
    c = oracle.adjoint_pt(sc, dimg, spp, depth, seed=7, nthreads=16)[1]
    print("GPU", g, "\noracle", c)
    np.testing.assert_allclose(g, c, rtol=1e-3)
    sc = rc.textured_scene(16, 16)
    dimg = np.random.default_rng(5).uniform(0.5, 1.5, (16, 16, 3))
    gt = adjoint(rumi, ctx, sc, dimg, 256, 5, seed=9, textures=True)[2]
    ct = oracle.adjoint_pt(sc, dimg, 256, 5, seed=9, nthreads=8, textures=True)[3]
    assert np.abs(ct).max() > 0
    np.testing.assert_allclose(gt, ct, rtol=0, atol=1e-3 * np.abs(ct).max())


def test_image_based_light_texel_gradients_equal_oracle(rumi, oracle, ctx):
    """The sky's texels through the light samples of matte vertices (the
    shadow entry carries the texel; the shadow stage adds dLoss x contrib /
    texel when unoccluded) and through dielectric chains: the GPU adjoint
    equals the oracle's FP64 derivative with one seed (the oracle's is pinned
    by exact differences in test_oracle_adjoint.py)."""
    w = 16
    sc = rc.ibl_scene(w, w, seed=9)
    dimg = np.random.default_rng(6).uniform(0.5, 1.5, (w, w, 3))
    gt = adjoint(rumi, ctx, sc, dimg, 128, 4, seed=31, textures=True)[2]
    ct = oracle.adjoint_pt(sc, dimg, 128, 4, seed=31, nthreads=8, textures=True)[3]
    assert np.abs(ct).max() > 0
    np.testing.assert_allclose(gt[:ct.shape[0]], ct, rtol=0, atol=1e-3 * np.abs(ct).max())


def test_medium_gradient_matches_finite_differences(rumi, ctx, scenes):
    """Central differences of the GPU renderer (step 10% of each parameter,
    common random numbers) against the adjoint, over independent seeds."""
    w, spp, depth, runs = 64, 256, 16, 16
    dimg = np.random.default_rng(1).uniform(0.5, 1.5, (w, w, 3))
    ga, fd = [], []
    for r in range(runs):
        ga.append(adjoint(rumi, ctx, _media_params(scenes, THETA, w), dimg, spp, depth,
                          seed=300 + r)[1])
        row = []
        for k in range(7):
            h = 0.1 * THETA[k]
            tp, tm = THETA.copy(), THETA.copy()
            tp[k] += h
            tm[k] -= h
            ip = render(rumi, ctx, _media_params(scenes, tp, w), w, w, spp, depth, seed=900 + r)
            im = render(rumi, ctx, _media_params(scenes, tm, w), w, w, spp, depth, seed=900 + r)
            row.append(((ip - im) * dimg).sum() / (2 * h))
        fd.append(row)
    ga, fd = np.array(ga), np.array(fd)
    se = np.sqrt(ga.var(0, ddof=1) / runs + fd.var(0, ddof=1) / runs)
    z = (ga.mean(0) - fd.mean(0)) / se
    print("adjoint", ga.mean(0), "\nFD", fd.mean(0), "\nz", z)
    assert np.all(np.abs(z) < 4.0), z


def test_adjoint_errors(rumi, ctx, scenes):
    import torch
    sc = scenes.cornell_box(8, 8)
    scene = rumi.Scene(ctx, sc)
    d = torch.ones((8, 8, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(rumi.RumiError):
        scene.adjoint(d, 0, 5)
    with pytest.raises(rumi.RumiError):  # no textures to differentiate
        scene.adjoint(d, 1, 5, textures=True)
    scene.close()
