"""Oracle paths pinned in closed form: refraction and total internal
reflection, directional lights and the thin-lens camera.

Each check derives the expected value from the physics or geometry (Snell's
law, the mirror law, the Lambert/Phong sum of add_dlight, rays through the
focus plane) rather than from the oracle's own formulas.
"""
import numpy as np
import pytest
from scipy import stats

import mrt_hip as M
import oracle_ffi as O

W = H = 48
SEED = 91


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _incident(theta, phi=0.3):
    """Unit direction making angle theta with -n for n = +z (so d.n < 0)."""
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi),
                     -np.cos(theta)])


# ---- refraction (refraction.c:16-76) and its TIR fallback ----------------

@pytest.mark.parametrize("f64", [True, False])
@pytest.mark.parametrize("front", [True, False])
def test_snell_law_on_a_refracting_plane(f64, front):
    """sin(theta_t) = eta sin(theta_i) with eta = ior_all / ior entering
    (front face) and ior / ior_all leaving; the refracted ray stays in the
    plane of incidence, on the far side of the surface, at unit length; the
    power is multiplied by the albedo."""
    ior, ior_all = 1.5, 1.0
    eta = ior_all / ior if front else ior / ior_all
    n = np.array([0.0, 0.0, 1.0])
    crit = np.arcsin(min(1.0, 1.0 / eta))
    col = (0.9, 0.7, 0.5)
    tol = 1e-12 if f64 else 2e-6
    for theta in np.linspace(0.0, crit * 0.999, 23):
        d = _incident(theta)
        out, pw, tir = O.refract(d, n, front, ior, ior_all, col, f64=f64)
        assert not tir
        assert abs(np.linalg.norm(out) - 1.0) < tol
        assert out @ n < 0
        sin_t = np.linalg.norm(np.cross(out, n))
        assert abs(sin_t - eta * np.sin(theta)) < 10 * tol
        # coplanar with d and n
        assert abs(np.cross(d, n) @ out) < 10 * tol
        # the tangential component keeps the incident's direction
        if theta > 0:
            assert (out[:2] @ d[:2]) > 0
        np.testing.assert_allclose(pw, np.float32(col), rtol=1e-7)


@pytest.mark.parametrize("f64", [True, False])
def test_normal_incidence_passes_straight(f64):
    """Through the centre of a glass sphere the ray meets both surfaces
    head on: the refracted direction equals the incident one."""
    n = _unit([0.2, -0.4, 0.9])
    d = -n
    for front in (True, False):
        out, _, tir = O.refract(d, n, front, 1.5, f64=f64)
        assert not tir
        np.testing.assert_allclose(out, d, atol=1e-12 if f64 else 1e-6)


@pytest.mark.parametrize("f64", [True, False])
def test_total_internal_reflection_falls_back_to_the_mirror(f64):
    """Leaving glass (eta = 1.5) beyond the critical angle asin(1/1.5) the
    ray is reflected as by a smooth metal (refraction.c:58-76 ->
    metallic_color): r = d - 2 (d.n) n with power f0 + (1 - f0)(1 - cos)^5,
    f0 the albedo; just below the critical angle it still refracts."""
    n = np.array([0.0, 0.0, 1.0])
    crit = np.arcsin(1.0 / 1.5)
    col = np.array([0.8, 0.6, 0.3])
    tol = 1e-12 if f64 else 2e-6
    for theta in np.linspace(crit * 1.001, np.pi / 2 * 0.999, 17):
        d = _incident(theta, phi=1.1)
        out, pw, tir = O.refract(d, n, False, 1.5, 1.0, col, f64=f64)
        assert tir
        np.testing.assert_allclose(out, d - 2 * (d @ n) * n, atol=10 * tol)
        cos = np.cos(theta)
        f0 = col.astype(np.float32).astype(np.float64)
        expect = f0 + (1 - f0) * (1 - cos) ** 5
        np.testing.assert_allclose(pw, expect, rtol=1e-5)
    out, _, tir = O.refract(_incident(crit * 0.999), n, False, 1.5, f64=f64)
    assert not tir
    # entering the denser medium never reflects totally
    out, _, tir = O.refract(_incident(np.pi / 2 * 0.999), n, True, 1.5,
                            f64=f64)
    assert not tir


def test_ior_all_scales_eta():
    """eta is ior_all / ior on the front face: with ior_all = ior the ray
    goes straight through at any angle."""
    n = np.array([0.0, 0.0, 1.0])
    for theta in (0.2, 0.7, 1.2):
        d = _incident(theta)
        out, _, tir = O.refract(d, n, True, 1.33, 1.33)
        assert not tir
        np.testing.assert_allclose(out, d, atol=1e-12)


# ---- directional light (d_light.c:19-57) ---------------------------------

DL_DIR = _unit([-0.4, -1.0, 0.6])      # orientation: the way the light travels
DL_COL = np.array([255.0, 180.0, 90.0])
DL_B = 0.7
ALB = 0.5


def dl_scene(flags, blocker=False):
    hs = M.HostScene.empty()
    hs.set_resolution(W, H)
    m = hs.add_material(color=(ALB, ALB, ALB))
    hs.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    if blocker:
        # a sphere between the lit side and the light
        hs.add_sphere(tuple(-DL_DIR * 3.0), 0.4, m)
    hs.add_light(tuple(DL_DIR), DL_B, tuple(DL_COL), kind=M.LIGHT_DIRECTIONAL)
    hs.set_ambient(0.0, (0, 0, 0))
    hs.set_camera((0.0, 0.0, -4.0), (0.0, 0.0, 1.0), fov=40.0)
    hs.set_flags(flags)
    return hs


def primary(view):
    o, d = O.raygen_f32(view, W, H, 0, SEED)
    bvh = O.make_bvh(M.BuiltBvh(view))
    p, t, _, _ = O.trace_rays(view, bvh, O.TRAV_BRUTE, o, d, f64=True)
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    return p.reshape(H, W), (o + d * t[:, None]).reshape(H, W, 3)


def dl_terms(pts):
    """c b max(n.l, 0) and c b max(r.v, 0)^64 with l = -orientation,
    r = reflect(-l, n) and v toward the camera."""
    cam = np.array([0.0, 0.0, -4.0])
    n = pts / np.linalg.norm(pts, axis=-1, keepdims=True)
    l = -DL_DIR
    ndl = n @ l
    v = cam - pts
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    r = -l + 2 * ndl[..., None] * n
    rdv = np.maximum((r * v).sum(-1), 0)
    diff = DL_COL * (DL_B * ndl)[..., None]
    spec = DL_COL * DL_B * (rdv ** 64)[..., None]
    return diff, spec, ndl > 0


@pytest.mark.parametrize("f64", [True, False])
def test_directional_light_closed_form(f64):
    """A lambert sphere under one directional light and a black sky: every
    bounce escapes into black, so pixel = a (c b n.l + c b max(r.v,0)^64)
    / 255 on the lit side and 0 elsewhere (add_dlight, d_light.c:35-57;
    default_mat, material_default.c:71-92)."""
    hs = dl_scene(M.SCENE_LIGHT_EXACT)
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    img, _ = O.render(view, bvh, W, H, 0, 1, 8, SEED, f64=f64,
                      traversal=O.TRAV_REFERENCE if f64 else O.TRAV_ORDERED)
    ids, pts = primary(view)
    on = ids == M.PRIM_SPHERE
    diff, spec, lit = dl_terms(pts)
    expect = np.where(lit[..., None], ALB * (diff + spec) / 255.0, 0.0)
    assert (on & lit).mean() > 0.1 and (on & ~lit).mean() > 0.05
    assert np.max(np.abs(img[on] - expect[on])) < (2e-4 if f64 else 5e-4)
    np.testing.assert_array_equal(img[on & ~lit], 0.0)
    np.testing.assert_array_equal(img[~on], 0.0)


def test_directional_light_sums_truncate_and_shadow():
    """Without MRT_SCENE_LIGHT_EXACT each += truncates (t_lcolor is
    uint64_t); a sphere placed toward the light shadows the points whose
    ray along -orientation meets it (check_dlight_hit_dir, d_light.c:19-33)."""
    hs = dl_scene(0)
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    img, _ = O.render(view, bvh, W, H, 0, 1, 1, SEED, f64=True,
                      traversal=O.TRAV_REFERENCE)
    ids, pts = primary(view)
    on = ids == M.PRIM_SPHERE
    diff, spec, lit = dl_terms(pts)
    expect = np.where(lit[..., None],
                      ALB * (np.floor(diff) + np.floor(spec)) / 255.0, 0.0)
    assert (np.abs(img[on] - expect[on]) < 1e-9).mean() > 0.99
    # blocked: points whose shadow ray toward -orientation hits the blocker
    hs2 = dl_scene(M.SCENE_LIGHT_EXACT, blocker=True)
    view2 = hs2.view()
    bvh2 = O.make_bvh(M.BuiltBvh(view2))
    img2, _ = O.render(view2, bvh2, W, H, 0, 1, 1, SEED, f64=True,
                       traversal=O.TRAV_REFERENCE)
    ids2, pts2 = primary(view2)
    c2 = -DL_DIR * 3.0
    on2 = ids2 == M.PRIM_SPHERE
    p = pts2[on2]
    l = -DL_DIR
    # ray p + s l against the blocker sphere |x - c2| = 0.4
    oc = p - c2
    b = oc @ l
    c = (oc * oc).sum(-1) - 0.4 ** 2
    disc = b * b - c
    s_far = -b + np.sqrt(np.maximum(disc, 0))
    shadowed = (disc > 0) & (s_far > 1e-3)
    n = p / np.linalg.norm(p, axis=-1, keepdims=True)
    lit2 = (n @ l) > 0
    # keep away from the shadow's edge
    edge = np.abs(disc) < 0.02
    sel = shadowed & lit2 & ~edge
    assert sel.sum() > 20
    np.testing.assert_array_equal(img2[on2][sel], 0.0)
    free = ~shadowed & lit2 & ~edge
    assert free.sum() > 20
    assert np.all(img2[on2][free].sum(-1) > 0)


# ---- thin-lens camera (focus.c:31-42, viewport.c:80-88) ------------------

def lens_rays(angle, focus, w=W, h=H, sample=0):
    hs = M.HostScene.empty()
    hs.set_resolution(w, h)
    hs.set_camera((0.3, 1.1, -5.0), (0.1, -0.2, 1.0), fov=55.0,
                  defocus=angle, focus=focus)
    return O.raygen_f32(hs.view(), w, h, sample, SEED), hs


@pytest.mark.parametrize("angle,focus", [(2.0, 5.0), (8.0, 3.0)])
def test_lens_rays_meet_the_pixel_point_on_the_focus_plane(angle, focus):
    """defocus_disk_sample: the origin is the camera plus a point of the
    disk of radius focus * tan(angle / 2) spanned by the viewport's u and v;
    the direction aims at the pixel's point on the viewport, which lies in
    the focus plane.  So every lens ray of a pixel passes through the point
    where the pinhole ray with the same jitter meets that plane."""
    (o, d), hs = lens_rays(angle, focus)
    (o0, d0), _ = lens_rays(0.0, focus)
    cam = np.array([0.3, 1.1, -5.0])
    fwd = _unit([0.1, -0.2, 1.0])
    np.testing.assert_allclose(o0, np.broadcast_to(cam, o0.shape), atol=0)
    d0 = d0.astype(np.float64)
    # the pixel's point on the focus plane (distance focus along fwd)
    P = cam + d0 * (focus / (d0 @ fwd))[:, None]
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    # the lens point lies in the plane through cam normal to fwd, inside
    # the disk
    off = o - cam
    R = focus * np.tan(np.radians(angle) / 2)
    assert np.max(np.abs(off @ fwd)) < 1e-5 * max(1.0, R)
    r = np.linalg.norm(off, axis=1)
    assert r.max() <= R * (1 + 1e-5)
    assert r.max() > 0.9 * R
    # the lens ray passes through P
    to_p = P - o
    miss = np.linalg.norm(np.cross(to_p, d), axis=1) / np.linalg.norm(to_p,
                                                                      axis=1)
    assert np.max(miss) < 2e-6
    assert np.all((to_p * d).sum(1) > 0)


def test_lens_points_fill_the_disk_uniformly():
    """Over many samples the lens offsets are uniform on the disk: r^2 / R^2
    is uniform and the angle is uniform."""
    angle, focus = 4.0, 4.0
    R = focus * np.tan(np.radians(angle) / 2)
    cam = np.array([0.3, 1.1, -5.0])
    offs = []
    for s in range(4):
        (o, _), _ = lens_rays(angle, focus, 32, 32, sample=s)
        offs.append(o.astype(np.float64) - cam)
    off = np.concatenate(offs)
    fwd = _unit([0.1, -0.2, 1.0])
    u = _unit(np.cross([0.0, 1.0, 0.0], -fwd))
    v = np.cross(fwd, u)
    x, y = off @ u, off @ v
    r2 = (x * x + y * y) / R ** 2
    assert stats.kstest(r2, "uniform").pvalue > 1e-3
    phi = (np.arctan2(y, x) % (2 * np.pi)) / (2 * np.pi)
    assert stats.kstest(phi, "uniform").pvalue > 1e-3
