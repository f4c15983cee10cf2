"""Differentiable rendering over the HIP path: a torch.autograd.Function
whose forward is mrt_render (mean image over n_samples) and whose backward
is the path-replay adjoint mrt_adjoint.  Tensors live on the CPU; the
library owns the GPU side."""
import numpy as np
import torch

from . import HostScene, Renderer, SCENE_LIGHT_EXACT, EST_MIS, OPT_ESTIMATOR


class RenderFunction(torch.autograd.Function):
    """image = render(albedo[n,3], roughness[n], emission[n,3], hdri[h,w,3])."""

    @staticmethod
    def forward(ctx, albedo, roughness, emission, hdri, renderer, n_samples,
                max_depth, first_sample, grad_first_sample):
        renderer.diff_set_materials(albedo.detach().cpu().numpy(),
                                    roughness.detach().cpu().numpy(),
                                    emission.detach().cpu().numpy())
        if hdri is not None:
            renderer.diff_set_hdri(hdri.detach().cpu().numpy())
        renderer.clear()
        renderer.render(first_sample, n_samples, max_depth)
        img = renderer.readback() / np.float32(n_samples)
        ctx.renderer = renderer
        ctx.args = (first_sample if grad_first_sample is None
                    else grad_first_sample, n_samples, max_depth)
        ctx.n_mat = albedo.shape[0]
        ctx.hshape = tuple(hdri.shape[:2]) if hdri is not None else None
        return torch.from_numpy(img)

    @staticmethod
    def backward(ctx, grad_img):
        first, n, depth = ctx.args
        gm, gt = ctx.renderer.adjoint(first, n, depth,
                                      grad_img.detach().cpu().numpy(),
                                      ctx.n_mat, ctx.hshape)
        g_alb = torch.from_numpy(np.ascontiguousarray(gm[:, 0:3]))
        g_rough = torch.from_numpy(np.ascontiguousarray(gm[:, 3]))
        g_emis = torch.from_numpy(np.ascontiguousarray(gm[:, 4:7]))
        g_hdri = torch.from_numpy(gt) if gt is not None else None
        return g_alb, g_rough, g_emis, g_hdri, None, None, None, None, None


def render(renderer, albedo, roughness, emission, hdri=None, n_samples=64,
           max_depth=6, first_sample=0, grad_first_sample=None):
    """Mean image over samples first_sample .. +n_samples-1.  The backward
    pass replays those samples, or the range starting at grad_first_sample:
    with a disjoint range the gradient of an L2 loss is free of the
    correlation between the residual and the derivative."""
    return RenderFunction.apply(albedo, roughness, emission, hdri, renderer,
                                n_samples, max_depth, first_sample,
                                grad_first_sample)


def inverse_scene(seed, width, height, n_materials=16, subdiv=None,
                  floor_cells=8, amp=0.15, freq=100):
    """The inverse-rendering scene: a displaced icosphere (the geodesic
    sphere of 20 * freq^2 triangles, 200,000 by default, or subdiv levels of
    4-way splits) over a flat floor of floor_cells^2 squares, triangles drawn
    uniformly from n_materials materials (lambert, GGX metals and emitters,
    seeded parameters), lit by an HDRI (uploaded by the caller) with ambient
    ratio 1.  Returns the host scene and the material parameters (albedo,
    roughness, emission)."""
    rng = np.random.default_rng(seed)
    hs = HostScene.empty()
    hs.set_resolution(width, height)
    alb = rng.uniform(0.2, 0.9, (n_materials, 3)).astype(np.float32)
    rough = rng.uniform(0.05, 0.6, n_materials).astype(np.float32)
    emis = np.zeros((n_materials, 3), np.float32)
    for k in range(n_materials):
        metal = k % 4 == 1
        if k % 8 == 7:
            emis[k] = rng.uniform(0.5, 2.0, 3)
        hs.add_material(color=tuple(alb[k]), metallic=1.0 if metal else 0.0,
                        roughness=float(rough[k]), emission=1.0,
                        emission_color=tuple(emis[k]))
    if subdiv is not None:
        hs.add_icosphere((0.0, 1.2, 0.0), 1.0, subdiv, amp, seed, mix=2,
                         mat=n_materials)
    else:
        hs.add_icosphere((0.0, 1.2, 0.0), 1.0, 0, amp, seed, mix=2,
                         mat=n_materials, freq=freq)
    if floor_cells:
        # the floor: a flat grid of triangles drawn from the same palette
        n = floor_cells
        xs = np.linspace(-4.0, 4.0, n + 1, dtype=np.float32)
        mats = rng.integers(0, n_materials, (n, n, 2))
        for i in range(n):
            for j in range(n):
                a = (xs[i], 0.0, xs[j])
                b = (xs[i + 1], 0.0, xs[j])
                c = (xs[i + 1], 0.0, xs[j + 1])
                d = (xs[i], 0.0, xs[j + 1])
                hs.add_triangle([a, b, c], int(mats[i, j, 0]))
                hs.add_triangle([a, c, d], int(mats[i, j, 1]))
    hs.set_ambient(1.0, (255, 255, 255))
    hs.set_flags(SCENE_LIGHT_EXACT)
    hs.set_camera((0.0, 1.6, -4.2), (0.0, -0.12, 1.0), fov=50.0)
    return hs, alb, rough, emis


def make_renderer(hs, width, height, hdri, device=0):
    r = Renderer(width, height, device)
    r.upload(hs)
    r.upload_hdri(hdri)
    r.set_option(OPT_ESTIMATOR, EST_MIS)
    return r
