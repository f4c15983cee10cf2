"""Oracle shading pinned by closed-form scenes and sampler distributions."""
import numpy as np
import pytest
from scipy import stats

import mrt_hip as M
import oracle_ffi as O

W = H = 48
SEED = 77


def render(hs, spp=1, depth=8, f64=True, first=0):
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    img, cnt = O.render(view, bvh, W, H, first, spp, depth, SEED, f64=f64,
                        traversal=O.TRAV_REFERENCE if f64 else O.TRAV_ORDERED)
    return img / spp, view, hs


def sphere_scene(mat_kwargs, ambient=0.25):
    hs = M.HostScene.empty()
    hs.set_resolution(W, H)
    m = hs.add_material(**mat_kwargs)
    hs.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    hs.set_ambient(ambient, (255, 255, 255))
    hs.set_camera((0.0, 0.0, -4.0), (0.0, 0.0, 1.0), fov=40.0)
    return hs


def primary(view, f64=True):
    """Primary hit ids / points / normals of sample 0 (fp32 rays)."""
    o, d = O.raygen_f32(view, W, H, 0, SEED)
    bvh = O.make_bvh(M.BuiltBvh(view))
    p, t, _, _ = O.trace_rays(view, bvh, O.TRAV_BRUTE, o, d, f64=True)
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    pts = o + d * t[:, None]
    return p.reshape(H, W), pts.reshape(H, W, 3), d.reshape(H, W, 3)


@pytest.mark.parametrize("f64", [True, False])
def test_lambert_sphere_is_albedo_times_background(f64):
    """Every bounce ray from a convex sphere escapes: pixel = a * B
    (default_mat, material_default.c:85-91; add_skybox, path_trace.c:37-48)."""
    a, B = 0.6, 0.25
    img, view, hs = render(sphere_scene(dict(color=(a, a, a))), spp=3, f64=f64)
    ids, _, _ = primary(view)
    on = ids == M.PRIM_SPHERE
    assert 0.2 < on.mean() < 0.9
    # with per-pixel jitter a border pixel may mix hit and miss samples;
    # interior pixels are exact
    inner = on & np.roll(on, 1, 0) & np.roll(on, -1, 0) & np.roll(on, 1, 1) \
        & np.roll(on, -1, 1)
    np.testing.assert_allclose(img[inner], np.float32(a) * np.float32(B),
                               rtol=1e-6)
    off = ~on & ~np.roll(on, 1, 0) & ~np.roll(on, -1, 0) & \
        ~np.roll(on, 1, 1) & ~np.roll(on, -1, 1)
    np.testing.assert_allclose(img[off], B, rtol=1e-6)


def test_glass_with_ior_one_is_invisible():
    """ior == 1 never reflects (dielectric.c:45-46) and refracts straight
    through with eta = 1, so the sphere leaves the background unchanged."""
    img, _, _ = render(sphere_scene(dict(color=(1, 1, 1), ior=1.0,
                                         transmission=1.0)), spp=2)
    np.testing.assert_allclose(img, 0.25, rtol=1e-9)


def test_smooth_metal_is_schlick_times_background():
    """Roughness 0 mirror (metallic.c:32-36): the reflected ray escapes, so
    pixel = (f0 + (1 - f0)(1 - cos)^5) * B with cos = -d.n."""
    a, B = 0.7, 0.3
    img, view, hs = render(sphere_scene(dict(color=(a, a, a), metallic=1.0),
                                        ambient=B), spp=1)
    ids, pts, d = primary(view)
    on = ids == M.PRIM_SPHERE
    n = pts / np.linalg.norm(pts, axis=-1, keepdims=True)
    cos = -(d * n).sum(-1)
    expect = (a + (1 - a) * (1 - cos) ** 5) * B
    assert on.mean() > 0.2
    np.testing.assert_allclose(img[on][:, 0], expect[on], rtol=1e-5)


def point_light_scene(flags):
    hs = sphere_scene(dict(color=(0.5, 0.5, 0.5)), ambient=0.0)
    hs.add_light((2.0, 3.0, -3.0), brightness=0.8, color=(255, 200, 100))
    hs.set_flags(flags)
    return hs


def lambert_phong(pts, d):
    """Expected default_mat term for one point light and a black sky:
    a * (c b n.l + c b max(r.v, 0)^64) / 255 with r = reflect(-l, n) and v
    toward the camera (p_light.c:67-91, light_manager.c:59-70)."""
    L = np.array([2.0, 3.0, -3.0])
    cam = np.array([0.0, 0.0, -4.0])
    c = np.array([255.0, 200.0, 100.0])
    b = 0.8
    n = pts / np.linalg.norm(pts, axis=-1, keepdims=True)
    ld = L - pts
    ld /= np.linalg.norm(ld, axis=-1, keepdims=True)
    ndl = (n * ld).sum(-1)
    view = cam - pts
    view /= np.linalg.norm(view, axis=-1, keepdims=True)
    r = -ld + 2 * ndl[..., None] * n
    rdv = np.maximum((r * view).sum(-1), 0)
    diff = c * (b * ndl)[..., None]
    spec = (c * b) * (rdv ** 64)[..., None]
    lit = ndl > 0
    return diff, spec, lit


def test_point_light_closed_form_exact_sums():
    img, view, hs = render(point_light_scene(M.SCENE_LIGHT_EXACT), spp=1)
    ids, pts, d = primary(view)
    on = ids == M.PRIM_SPHERE
    diff, spec, lit = lambert_phong(pts, d)
    expect = np.where(lit[..., None], 0.5 * (diff + spec) / 255.0, 0.0)
    assert (on & lit).mean() > 0.1
    err = np.abs(img[on] - expect[on])
    # rays differ from the fp32 rays of this check by ~1e-7 relative
    assert np.max(err) < 2e-4
    np.testing.assert_allclose(img[~on], 0.0, atol=0)


def test_point_light_sums_truncate_like_uint64():
    """t_lcolor is uint64_t (includes/structs.h:26-31): each += truncates,
    so a lit pixel is a * (floor(diff) + floor(spec)) / 255."""
    img, view, hs = render(point_light_scene(0), spp=1)
    ids, pts, d = primary(view)
    on = ids == M.PRIM_SPHERE
    k = img[on] * 255.0 / 0.5
    np.testing.assert_allclose(k, np.round(k), atol=1e-6)
    diff, spec, lit = lambert_phong(pts, d)
    expect = np.where(lit[..., None],
                      0.5 * (np.floor(diff) + np.floor(spec)) / 255.0, 0.0)
    close = np.abs(img[on] - expect[on]) < 1e-9
    assert close.mean() > 0.99
    assert np.max(np.abs(img[on] - expect[on])) <= 2 * 0.5 / 255 + 1e-9


def test_soft_shadow_factor_bounds():
    """Unoccluded soft light (radius, samples) gives factor 1: same image as
    the hard light; fully occluded gives 0 (p_light.c:19-46)."""
    def scene(radius, samples, blocker):
        hs = M.HostScene.empty()
        hs.set_resolution(W, H)
        m = hs.add_material(color=(0.5, 0.5, 0.5))
        hs.add_plane((0, 0, 0), (0, 1, 0), m)
        if blocker:
            hs.add_sphere((0.0, 4.0, 0.0), 2.0, m)
        hs.add_light((0.0, 8.0, 0.0), 1.0, (255, 255, 255), radius, samples)
        hs.set_ambient(0.0, (0, 0, 0))
        hs.set_camera((0.0, 2.0, -6.0), (0.0, -0.3, 1.0), fov=30.0)
        return hs
    hard, _, _ = render(scene(0.0, 0, False), depth=1)
    soft, _, _ = render(scene(0.3, 6, False), depth=1)
    np.testing.assert_allclose(soft, hard, rtol=1e-12)
    assert hard.mean() > 0.05
    # directly below a big occluder nothing is lit
    blocked, _, _ = render(scene(0.3, 6, True), depth=1)
    assert blocked[H // 2 + 6:, W // 2 - 2:W // 2 + 2].max() == 0.0


def test_ggx_micro_normal_distribution():
    """tan^2(theta_m) of ggx_sample_hemisphere follows the GGX NDF law
    P(tan^2 <= x) = x / (alpha^2 + x), alpha = roughness^2 (Walter et al.
    2007); phi is uniform."""
    rough = 0.5
    alpha2 = rough ** 4
    rng = np.random.default_rng(0)
    u = rng.random((4000, 2)).astype(np.float32)
    n = (0.0, 0.0, 1.0)
    m = np.array([O.sample_ggx(n, rough, a, b) for a, b in u])
    cos = m[:, 2]
    tan2 = (1 - cos ** 2) / cos ** 2
    ks = stats.kstest(tan2, lambda x: x / (alpha2 + x))
    assert ks.pvalue > 1e-3
    phi = np.arctan2(m[:, 1], m[:, 0]) % (2 * np.pi)
    assert stats.kstest(phi / (2 * np.pi), "uniform").pvalue > 1e-3
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-12)


def test_cosine_hemisphere_moments():
    """E[cos] = 2/3, E[cos^2] = 1/2 for a cosine-weighted hemisphere."""
    rng = np.random.default_rng(1)
    u = rng.random((20000, 2)).astype(np.float32)
    n = np.array([0.3, 0.8, -0.52])
    n /= np.linalg.norm(n)
    s = np.array([O.sample_cos(n, a, b) for a, b in u[:6000]])
    c = s @ n
    assert np.all(c >= -1e-12)
    assert abs(c.mean() - 2 / 3) < 0.01
    assert abs((c ** 2).mean() - 0.5) < 0.01


def test_unit_sphere_radius_law():
    """random_in_unit_sphere: |p| <= 1 and |p|^3 uniform (inside draws are
    uniform in the ball; outside draws use radius cbrt(U))."""
    rng = np.random.default_rng(2)
    u = rng.random((5000, 4)).astype(np.float32)
    p = np.array([O.unit_sphere(x) for x in u])
    r = np.linalg.norm(p, axis=1)
    assert r.max() <= 1.0 + 1e-12
    assert stats.kstest(r ** 3, "uniform").pvalue > 1e-3


def test_fp32_and_fp64_images_agree():
    hs = M.HostScene.plain(5, W, H)
    img64, _, _ = render(hs, spp=4, f64=True)
    img32, _, _ = render(hs, spp=4, f64=False)
    rel = np.abs(img64 - img32).mean() / img64.mean()
    assert rel < 2e-3
