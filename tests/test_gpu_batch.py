"""Batched rendering (rumi_batch_* through the C ABI): every item's image is
the image of its scene rendered alone with the same seed, the captured
hipGraph replays give the eager result every time, and the batched adjoint's
texel gradients are each scene's rumi_adjoint gradients."""
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


def _items(scenes):
    """Scenes of different kinds and sizes: (scene dict, width, height)."""
    out = [(scenes.cornell_box(40, 30), 40, 30),
           (rc.textured_scene(32, 20), 32, 20),
           (scenes.cornell_media(24, 24), 24, 24),
           (scenes.batch_scene(3, 6, 36, tex_size=16), 36, 36),
           (scenes.batch_scene(7, 12, 20, tex_size=8), 20, 20),
           (rc.ibl_scene(24, 16), 24, 16)]
    for sc, w, h in out:
        sc["camera"]["y_ratio"] = h / w
    return out


def _close(a, b, scale):
    # float atomics accumulate the film in a different order
    np.testing.assert_allclose(a, b, rtol=2e-4, atol=2e-6 * scale)


def test_items_equal_single_renders_and_graph_replays(rumi, ctx, scenes):
    import torch
    items = _items(scenes)
    objs = [rumi.Scene(ctx, sc) for sc, _, _ in items]
    spp, depth, seed = 8, 5, 12
    batch = rumi.Batch(ctx, [(o, w, h) for o, (_, w, h) in zip(objs, items)], spp, depth,
                       seed=seed, queue_capacity=3000)
    eager, _, _, st, ms = batch.run(use_graph=False)
    eager = eager.clone()
    torch.cuda.synchronize()
    assert st.total_samples == sum(w * h for _, w, h in items) * spp
    assert st.generations == -(-st.total_samples // 3000)
    assert all(0.0 <= m <= st.milliseconds + 1.0 for m in ms), (ms, st.milliseconds)
    scale = float(eager.abs().max())
    parts = batch.split(eager)
    for o, (sc, w, h), part in zip(objs, items, parts):
        assert float(part.mean()) > 1e-3  # every item is lit
        single, _ = o.render(w, h, spp, depth, seed=seed)
        torch.cuda.synchronize()
        _close(part.cpu().numpy(), single.cpu().numpy(), scale)
    out = torch.empty_like(eager)
    for k in range(3):
        out.fill_(-1.0)
        batch.run(images=out, use_graph=True)
        torch.cuda.synchronize()
        _close(out.cpu().numpy(), eager.cpu().numpy(), scale)
        # eager work between replays
        objs[0].render(8, 8, 2, 3, seed=k)
        torch.cuda.synchronize()
    batch.close()
    for o in objs:
        o.close()


def test_batched_adjoint_equals_per_scene_adjoint(rumi, ctx, scenes):
    import torch
    items = [(rc.textured_scene(24, 16), 24, 16),
             (scenes.batch_scene(2, 5, 28, tex_size=12), 28, 28),
             (scenes.batch_scene(5, 9, 20, tex_size=8), 20, 20),
             (rc.ibl_scene(20, 16), 20, 16)]
    for sc, w, h in items:
        sc["camera"]["y_ratio"] = h / w
    objs = [rumi.Scene(ctx, sc) for sc, _, _ in items]
    spp, depth, seed = 16, 5, 3
    batch = rumi.Batch(ctx, [(o, w, h) for o, (_, w, h) in zip(objs, items)], spp, depth,
                       seed=seed, adjoint=True)
    rng = np.random.default_rng(5)
    targets = torch.from_numpy(rng.uniform(0.0, 0.3, (batch.total_pixels, 3)).astype(np.float32)).cuda()
    for use_graph in (False, True, True):
        img, tg, loss, st, _ = batch.run(targets=targets, use_graph=use_graph)
        torch.cuda.synchronize()
        img_h, tg_h, loss_h = img.cpu().numpy(), tg.cpu().numpy(), loss.cpu().numpy()
        t_h = targets.cpu().numpy()
        po = to = 0
        for i, (o, (sc, w, h)) in enumerate(zip(objs, items)):
            I = img_h[po:po + w * h]
            T = t_h[po:po + w * h]
            assert I.mean() > 1e-3
            np.testing.assert_allclose(loss_h[i], ((I.astype(np.float64) - T) ** 2).sum(), rtol=1e-4)
            dimg = torch.from_numpy((2.0 * (I - T)).reshape(h, w, 3).astype(np.float32)).cuda()
            _, _, dtex, _ = o.adjoint(dimg.contiguous(), spp, depth, seed=seed + 1, textures=True)
            torch.cuda.synchronize()
            ref = dtex.cpu().numpy()[:o.n_texels]
            got = tg_h[to:to + o.n_texels]
            assert np.abs(ref).max() > 0
            np.testing.assert_allclose(got, ref, rtol=2e-3, atol=2e-4 * float(np.abs(ref).max()))
            po += w * h
            to += o.n_texels
    batch.close()
    for o in objs:
        o.close()


def test_batch_errors(rumi, ctx, scenes):
    o = rumi.Scene(ctx, scenes.cornell_box(8, 8))
    with pytest.raises(rumi.RumiError):
        rumi.Batch(ctx, [], 4, 4)
    with pytest.raises(rumi.RumiError):
        rumi.Batch(ctx, [(o, 0, 8)], 4, 4)
    with pytest.raises(rumi.RumiError):
        rumi.Batch(ctx, [(o, 8, 8)], 0, 4)
    b = rumi.Batch(ctx, [(o, 8, 8)], 2, 3, adjoint=True)
    with pytest.raises(rumi.RumiError):  # the adjoint needs targets
        b.run()
    b.close()
    o.close()
