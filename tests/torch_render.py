"""torch.autograd.Function over rumi_render + rumi_adjoint (test harness
only; the library's boundary is the C ABI).  forward renders the scene that
`build` makes from the medium parameters and the texel pool, and backward
hands the incoming dLoss/dImage to rumi_adjoint with the forward's seed, so
its gradients are those of the expected image at the rendered parameters."""
import numpy as np
import torch


class RumiRender(torch.autograd.Function):
    """RumiRender.apply(medium, texels, build, rumi, ctx, width, height, spp,
    max_depth, seed) -> (height, width, 3) image.

    medium: (7,) tensor sigma_a rgb, sigma_s rgb, g of the homogeneous medium.
    texels: (n_texels, 3) tensor of every image texture's texels in scene
        order, or None.
    build(medium_np, texels_np_or_None) -> scene dict."""

    @staticmethod
    def forward(fctx, medium, texels, build, rumi, ctx, width, height, spp, max_depth, seed):
        m = medium.detach().double().cpu().numpy()
        t = None if texels is None else texels.detach().float().cpu().numpy()
        scene = rumi.Scene(ctx, build(m, t))
        img, _ = scene.render(width, height, spp, max_depth, seed=seed)
        fctx.scene = scene
        fctx.args = (spp, max_depth, seed)
        fctx.has_tex = texels is not None
        fctx.n_texels = 0 if texels is None else int(texels.shape[0])
        fctx.like = (medium.dtype, medium.device,
                     None if texels is None else (texels.dtype, texels.device))
        return img

    @staticmethod
    def backward(fctx, grad_img):
        spp, depth, seed = fctx.args
        _, dmed, dtex, _ = fctx.scene.adjoint(grad_img.contiguous().float(), spp, depth, seed=seed,
                                              textures=fctx.has_tex)
        torch.cuda.synchronize()
        fctx.scene.close()
        m_dtype, m_dev, t_like = fctx.like
        d_tex = dtex[:fctx.n_texels].to(dtype=t_like[0], device=t_like[1]) if fctx.has_tex else None
        return (dmed.to(dtype=m_dtype, device=m_dev), d_tex, None, None, None, None, None, None,
                None, None)


def split_texels(flat: np.ndarray, shapes):
    """The (n_texels, 3) pool back into (h, w, 3) textures."""
    out, o = [], 0
    for h, w in shapes:
        out.append(np.ascontiguousarray(flat[o:o + h * w].reshape(h, w, 3), np.float32))
        o += h * w
    return out
