"""Differentiable rendering: the oracle's forward-mode derivative is pinned by
central finite differences (fp64, same RNG stream) and, for roughness, by
the zero mean of the score function and finite differences of the expected
image; the HIP path-replay adjoint is then compared with the oracle's
gradient, with central finite differences of the HIP forward render, and
used through torch.autograd."""
import ctypes as C

import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O
from mrt_hip.diff import inverse_scene

W, H = 24, 16
SEED = 0xD1FF
DEPTH = 6


def grad_scene():
    """Lambert, GGX metals and an emitter on a displaced icosphere and floor,
    a rough glass sphere and a point light (exact sums), under an HDRI."""
    hs, alb, rough, emis = inverse_scene(5, W, H, n_materials=8, subdiv=2,
                                         floor_cells=3)
    g = hs.add_material(color=(0.9, 0.95, 1.0), ior=1.5, transmission=1.0,
                        roughness=0.1, emission=1.0)
    hs.add_sphere((1.4, 0.6, -0.8), 0.5, g)
    hs.add_light((0.0, 4.0, -2.0), 0.7, (255, 255, 255))
    return hs


def grad_env():
    return M.hdri_sky(9, 32, 16, sun_radius_deg=8.0, sun_radiance=20.0)


def oracle_loss(view, bvh, hd, delta, spp, f64=True):
    img, _ = O.render(view, bvh, W, H, 0, spp, DEPTH, SEED, f64=f64,
                      traversal=O.TRAV_ORDERED, estimator=M.EST_MIS, hdri=hd)
    return float(np.sum(delta * img))


def oracle_grad(view, bvh, hd, delta, spp, q, f64=True):
    _, dimg = O.render_grad(view, bvh, W, H, 0, spp, DEPTH, SEED, *q,
                            f64=f64, hdri=hd)
    return float(np.sum(delta * dimg))


def perturb(view, hd, q, h):
    """Adds h to the fp32 parameter; returns its new value (the step that
    actually happened is the difference of two returned values)."""
    kind, mat, chan, texel = q
    m = view.contents.materials[mat]
    if kind == O.P_ALBEDO:
        m.color[chan] += h
        return float(m.color[chan])
    if kind == O.P_EMISSION:
        m.emission_color[chan] += h
        return float(m.emission_color[chan])
    if kind == O.P_ROUGHNESS:
        m.roughness += h
        return float(m.roughness)
    rgb = hd._keep[0]
    idx = (texel // rgb.shape[1], texel % rgb.shape[1], chan)
    rgb[idx] += np.float32(h)
    return float(rgb[idx])


@pytest.fixture(scope="module")
def setup():
    hs = grad_scene()
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    env = grad_env()
    hd = O.make_hdri(env)
    rng = np.random.default_rng(3)
    delta = rng.normal(size=(H, W, 3))
    return hs, view, bvh, env, hd, delta


def test_oracle_grad_image_equals_render(setup):
    hs, view, bvh, env, hd, delta = setup
    img, _ = O.render_grad(view, bvh, W, H, 0, 4, DEPTH, SEED, O.P_ALBEDO,
                           f64=True, hdri=hd)
    ref, _ = O.render(view, bvh, W, H, 0, 4, DEPTH, SEED, f64=True,
                      traversal=O.TRAV_ORDERED, estimator=M.EST_MIS, hdri=hd)
    np.testing.assert_array_equal(img, ref)


def linear_params(view, hd, rng, n):
    nm = view.contents.n_materials
    qs = []
    while len(qs) < n:
        k = rng.integers(3)
        if k == 0:
            qs.append((O.P_ALBEDO, int(rng.integers(nm)), int(rng.integers(3)), 0))
        elif k == 1:
            qs.append((O.P_EMISSION, int(rng.integers(nm)), int(rng.integers(3)), 0))
        else:
            rgb = hd._keep[0]
            qs.append((O.P_TEXEL, 0, int(rng.integers(3)),
                       int(rng.integers(rgb.shape[0] * rgb.shape[1]))))
    return qs


def test_oracle_grad_matches_central_differences(setup):
    """Albedo, emission and texels enter the estimator polynomially with no
    decision depending on them: fp64 central differences on the same RNG
    stream equal the forward-mode derivative."""
    hs, view, bvh, env, hd, delta = setup
    rng = np.random.default_rng(7)
    h = 1e-3
    checked = 0
    for q in linear_params(view, hd, rng, 16):
        g = oracle_grad(view, bvh, hd, delta, 4, q)
        xp = perturb(view, hd, q, h)
        lp = oracle_loss(view, bvh, hd, delta, 4)
        xm = perturb(view, hd, q, -2 * h)
        lm = oracle_loss(view, bvh, hd, delta, 4)
        perturb(view, hd, q, h)
        fd = (lp - lm) / (xp - xm)
        assert abs(g - fd) <= 1e-5 * max(1.0, abs(fd)), (q, g, fd)
        checked += g != 0
    assert checked >= 8


def rough_metal_scene(rough):
    hs = M.HostScene.empty()
    hs.set_resolution(8, 8)
    m = hs.add_material(color=(0.8, 0.8, 0.8), metallic=1.0, roughness=rough)
    hs.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    hs.set_ambient(1.0, (255, 255, 255))
    hs.set_camera((0.0, 0.0, -1.6), (0.0, 0.0, 1.0), fov=40.0)
    return hs


def smooth_env(w=64, h=32):
    y = (np.arange(h)[:, None] + 0.5) / h
    x = (np.arange(w)[None, :] + 0.5) / w
    base = 0.2 + 1.5 * y * y + 0.6 * np.cos(2 * np.pi * x) ** 2
    return np.repeat(base[:, :, None], 3, axis=2).astype(np.float32)


def test_oracle_roughness_score_has_zero_mean_under_constant_env():
    """A rough metal plane under a constant environment: every reflected ray
    escapes (above or below the plane), so each sample is the same
    L * F(cos theta) whatever micro-normal is drawn and the image does not
    depend on the roughness: the score term must average to zero."""
    hs = M.HostScene.empty()
    hs.set_resolution(8, 8)
    m = hs.add_material(color=(0.8, 0.7, 0.6), metallic=1.0, roughness=0.3)
    hs.add_plane((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), m)
    hs.set_ambient(1.0, (255, 255, 255))
    hs.set_camera((0.0, 1.0, -2.0), (0.0, -0.6, 1.0), fov=40.0)
    view = hs.view()
    hd = O.make_hdri(np.full((16, 32, 3), 0.7, np.float32))
    spp = 16384
    img, dimg = O.render_grad(view, None, 8, 8, 0, spp, 2, SEED,
                              O.P_ROUGHNESS, 0, 0, 0, f64=True, hdri=hd)
    g = dimg.sum() / spp
    scale = img.sum() / spp
    # the statistic's standard deviation is about 0.7 % of the scale here
    assert abs(g) < 0.03 * scale, (g, scale)


def test_oracle_roughness_gradient_matches_expected_image_differences():
    """Score-function gradient vs finite differences of the expected image
    (common random numbers, many samples) on a rough metal sphere under a
    smooth environment."""
    spp = 16384
    env = smooth_env()
    hd = O.make_hdri(env)
    r0, h = 0.3, 0.02
    vals = {}
    for r in (r0 - h, r0, r0 + h):
        hs = rough_metal_scene(r)
        view = hs.view()
        bvh = O.make_bvh(M.BuiltBvh(view))
        img, dimg = O.render_grad(view, bvh, 8, 8, 0, spp, 2, SEED,
                                  O.P_ROUGHNESS, 0, 0, 0, f64=True, hdri=hd)
        vals[r] = (img.sum() / spp, dimg.sum() / spp)
    fd = (vals[r0 + h][0] - vals[r0 - h][0]) / (2 * h)
    g = vals[r0][1]
    print(f"roughness: score {g:.4f}, FD of the expectation {fd:.4f}")
    # spread over seeds: about 2 % for the score term, 0.3 % for FD
    assert abs(fd) > 0.05
    assert abs(g - fd) < 0.1 * abs(fd), (g, fd)


# ---- HIP path ---------------------------------------------------------------

def gpu_renderer(hs, env):
    r = M.Renderer(W, H)
    r.upload(hs)
    r.upload_hdri(env)
    r.set_option(M.OPT_ESTIMATOR, M.EST_MIS)
    r.set_seed(SEED)
    return r


def gpu_loss(r, delta, spp):
    r.clear()
    r.render(0, spp, DEPTH)
    return float(np.sum(delta * (r.readback().astype(np.float64) / spp)))


def mat_arrays(view):
    n = view.contents.n_materials
    ms = view.contents.materials
    alb = np.array([[ms[k].color[c] for c in range(3)] for k in range(n)],
                   np.float32)
    rough = np.array([ms[k].roughness for k in range(n)], np.float32)
    emis = np.array([[ms[k].emission_color[c] * ms[k].emission_strength
                      for c in range(3)] for k in range(n)], np.float32)
    return alb, rough, emis


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle_gradient(setup):
    hs, view, bvh, env, hd, delta = setup
    spp = 8
    r = gpu_renderer(hs, env)
    nm = view.contents.n_materials
    gm, gt = r.adjoint(0, spp, DEPTH, delta.astype(np.float32), nm,
                       env.shape[:2])
    r.close()
    rng = np.random.default_rng(11)
    qs = linear_params(view, hd, rng, 24)
    qs += [(O.P_ROUGHNESS, k, 0, 0) for k in range(nm)]
    bad = []
    for q in qs:
        ref = oracle_grad(view, bvh, hd, delta, spp, q, f64=False) / spp
        kind, mat, chan, texel = q
        if kind == O.P_ALBEDO:
            got = gm[mat, chan]
        elif kind == O.P_EMISSION:
            got = gm[mat, 4 + chan]
        elif kind == O.P_ROUGHNESS:
            got = gm[mat, 3]
        else:
            got = gt.reshape(-1, 3)[texel, chan]
        if abs(got - ref) > 2e-3 * max(abs(ref), 1e-3):
            bad.append((q, got, ref))
    print(f"{len(qs) - len(bad)}/{len(qs)} within tolerance", bad[:4])
    assert len(bad) <= 1, bad


@pytest.mark.gpu
def test_gpu_adjoint_matches_central_differences(setup):
    """32 random parameters among those the image depends on: the adjoint
    equals central differences of the HIP forward render (h = 1e-3) within
    1e-2 relative."""
    hs, view, bvh, env, hd, delta = setup
    spp = 16
    r = gpu_renderer(hs, env)
    nm = view.contents.n_materials
    alb, rough, emis = mat_arrays(view)
    r.diff_set_materials(alb, rough, emis)
    gm, gt = r.adjoint(0, spp, DEPTH, delta.astype(np.float32), nm,
                       env.shape[:2])
    cand = [(("a", k, c), gm[k, c]) for k in range(nm) for c in range(3)]
    cand += [(("e", k, c), gm[k, 4 + c]) for k in range(nm) for c in range(3)]
    tflat = gt.reshape(-1)
    for i in np.argsort(-np.abs(tflat))[:64]:
        cand.append((("t", int(i), 0), tflat[i]))
    scale = max(abs(g) for _, g in cand)
    cand = [c for c in cand if abs(c[1]) > 1e-2 * scale]
    rng = np.random.default_rng(5)
    pick = rng.choice(len(cand), size=min(32, len(cand)), replace=False)
    h = 1e-3
    tex = env.copy()
    worst = 0.0
    for i in pick:
        (kind, a, b), g = cand[i]
        vals = []
        for s in (h, -h):
            al, em, tx = alb.copy(), emis.copy(), tex.copy()
            if kind == "a":
                al[a, b] += s
            elif kind == "e":
                em[a, b] += s
            else:
                tx.reshape(-1)[a] += s
            r.diff_set_materials(al, rough, em)
            r.diff_set_hdri(tx)
            vals.append(gpu_loss(r, delta, spp))
        fd = (vals[0] - vals[1]) / (2 * h)
        rel = abs(g - fd) / max(abs(fd), 1e-12)
        worst = max(worst, rel)
        assert rel <= 1e-2, (kind, a, b, g, fd)
    r.close()
    print(f"{len(pick)} parameters, worst relative error {worst:.2e}")


@pytest.mark.gpu
def test_gpu_autograd_inverse_rendering_reduces_the_loss():
    import torch
    from mrt_hip import diff
    hs, alb, rough, emis = inverse_scene(21, 32, 24, n_materials=6, subdiv=3,
                                         floor_cells=4)
    env = grad_env()
    r = diff.make_renderer(hs, 32, 24, env)
    r.set_seed(SEED)
    target = diff.render(r, torch.from_numpy(alb), torch.from_numpy(rough),
                         torch.from_numpy(emis), torch.from_numpy(env),
                         n_samples=16, max_depth=4).detach()
    rng = np.random.default_rng(2)
    a = torch.tensor(np.clip(alb + rng.normal(0, 0.15, alb.shape), 0.05, 0.95),
                     dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([a], lr=0.03)
    losses = []
    for _ in range(40):
        opt.zero_grad()
        img = diff.render(r, a, torch.from_numpy(rough),
                          torch.from_numpy(emis), torch.from_numpy(env),
                          n_samples=16, max_depth=4)
        loss = ((img - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            a.clamp_(0.0, 1.0)
        losses.append(loss.item())
    r.close()
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.1 * losses[0]
