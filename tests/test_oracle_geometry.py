"""Oracle intersection and traversal pinned by closed forms and brute force."""
import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O


def scene_with(add):
    hs = M.HostScene.empty()
    m = hs.add_material(color=(0.5, 0.5, 0.5))
    add(hs, m)
    hs.set_ambient(0.1, (255, 255, 255))
    return hs


def trace(hs, o, d, mode, f64):
    view = hs.view()
    built = M.BuiltBvh(view) if (view.contents.n_triangles
                                 + view.contents.n_spheres) else None
    bvh = O.make_bvh(built) if built is not None else None
    if bvh is None:
        mode = O.TRAV_BRUTE
    p, t, _, _ = O.trace_rays(view, bvh, mode, np.array([o], np.float32),
                              np.array([d], np.float32), f64=f64)
    return int(p[0]), float(t[0])


MODES = [(m, f) for m in (O.TRAV_BRUTE, O.TRAV_REFERENCE, O.TRAV_ORDERED)
         for f in (False, True)]


@pytest.mark.parametrize("mode,f64", MODES)
def test_sphere_roots(mode, f64):
    hs = scene_with(lambda h, m: h.add_sphere((0, 0, 0), 1.0, m))
    # from outside: nearer root t = 5 - 1
    assert trace(hs, (0, 0, -5), (0, 0, 1), mode, f64) == (M.PRIM_SPHERE, 4.0)
    # from the centre the near root is negative: the far root t = 1
    assert trace(hs, (0, 0, 0), (1, 0, 0), mode, f64) == (M.PRIM_SPHERE, 1.0)
    # pointing away: both roots negative
    assert trace(hs, (0, 0, -5), (0, 0, -1), mode, f64)[0] == M.PRIM_NONE
    # passing beside: discriminant < 0
    assert trace(hs, (0, 1.01, -5), (0, 0, 1), mode, f64)[0] == M.PRIM_NONE
    # off-axis: t = 5 - sqrt(1 - y^2) for y = 0.6 -> 4.2
    p, t = trace(hs, (0, 0.6, -5), (0, 0, 1), mode, f64)
    assert p == M.PRIM_SPHERE and abs(t - 4.2) < 1e-6


@pytest.mark.parametrize("mode,f64", MODES)
def test_triangle_barycentric_edges(mode, f64):
    tri = [(0, 0, 2), (1, 0, 2), (0, 1, 2)]
    hs = scene_with(lambda h, m: h.add_triangle(tri, m))
    assert trace(hs, (0.25, 0.25, 0), (0, 0, 1), mode, f64) == (0, 2.0)
    # u + v > 1 -> outside
    assert trace(hs, (0.6, 0.6, 0), (0, 0, 1), mode, f64)[0] == M.PRIM_NONE
    # two-sided: hit from behind
    assert trace(hs, (0.2, 0.2, 5), (0, 0, -1), mode, f64) == (0, 3.0)
    # parallel to the plane of the triangle
    assert trace(hs, (0.2, 0.2, 2), (1, 0, 0), mode, f64)[0] == M.PRIM_NONE


def test_one_sided_triangles_cull_back_faces():
    tri = [(0, 0, 2), (1, 0, 2), (0, 1, 2)]
    hs = scene_with(lambda h, m: h.add_triangle(tri, m))
    hs.set_flags(M.SCENE_TRIANGLE_ONE_SIDED)
    # e1 x e2 = +z: a ray along +z sees a = e1.(d x e2) < 0 -> culled
    assert trace(hs, (0.25, 0.25, 0), (0, 0, 1), O.TRAV_BRUTE, True)[0] \
        == M.PRIM_NONE
    assert trace(hs, (0.25, 0.25, 5), (0, 0, -1), O.TRAV_BRUTE, True) == (0, 3.0)


@pytest.mark.parametrize("f64", [False, True])
def test_plane_and_t_range(f64):
    hs = scene_with(lambda h, m: h.add_plane((0, 0, 0), (0, 1, 0), m))
    assert trace(hs, (0, 5, 0), (0, -1, 0), O.TRAV_BRUTE, f64) == \
        (M.PRIM_PLANE, 5.0)
    assert trace(hs, (0, 5, 0), (1, 0, 0), O.TRAV_BRUTE, f64)[0] == M.PRIM_NONE
    # IT_MAX = 1000 inclusive for planes, beyond it a miss
    assert trace(hs, (0, 1000, 0), (0, -1, 0), O.TRAV_BRUTE, f64)[0] == \
        M.PRIM_PLANE
    assert trace(hs, (0, 1000.5, 0), (0, -1, 0), O.TRAV_BRUTE, f64)[0] == \
        M.PRIM_NONE
    # IT_MIN = 0.001: a ray leaving the surface does not hit it again
    assert trace(hs, (0, 0.0005, 0), (0, -1, 0), O.TRAV_BRUTE, f64)[0] == \
        M.PRIM_NONE


def random_soup(seed, n):
    rng = np.random.default_rng(seed)
    hs = M.HostScene.empty()
    m = hs.add_material()
    c = rng.uniform(-2, 2, size=(n, 3))
    for i in range(n):
        e = rng.normal(scale=0.25, size=(2, 3))
        v = np.stack([c[i], c[i] + e[0], c[i] + e[1]])
        hs.add_triangle(v, m)
    for i in range(5):
        hs.add_sphere(rng.uniform(-2, 2, 3), rng.uniform(0.1, 0.5), m)
    hs.add_plane((0, -2.5, 0), (0, 1, 0), m)
    return hs


@pytest.mark.parametrize("f64", [False, True])
def test_bvh_traversals_equal_brute_force(f64):
    hs = random_soup(11, 3000)
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    rng = np.random.default_rng(2)
    n = 4000
    o = rng.uniform(-4, 4, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    pb, tb, _, _ = O.trace_rays(view, None, O.TRAV_BRUTE, o, d, f64=f64)
    for mode in (O.TRAV_REFERENCE, O.TRAV_ORDERED):
        p, t, nodes, prims = O.trace_rays(view, bvh, mode, o, d, f64=f64)
        assert np.mean(p == pb) >= 0.9995
        hit = (p == pb) & (pb != M.PRIM_NONE)
        np.testing.assert_allclose(t[hit], tb[hit], rtol=1e-6)
    assert (pb != M.PRIM_NONE).mean() > 0.5
    # ordered + culled traversal visits fewer nodes than the reference's
    _, _, n_ref, _ = O.trace_rays(view, bvh, O.TRAV_REFERENCE, o, d, f64=f64)
    _, _, n_ord, _ = O.trace_rays(view, bvh, O.TRAV_ORDERED, o, d, f64=f64)
    assert n_ord.sum() < n_ref.sum()


def test_fp32_against_fp64_primary_hits():
    hs = M.HostScene.mesh(3, 320, 180, subdiv=6, floor_cells=64)
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    o, d = O.raygen_f32(view, 320, 180, 0, 99)
    p32, t32, _, _ = O.trace_rays(view, bvh, O.TRAV_ORDERED, o, d)
    p64, t64, _, _ = O.trace_rays(view, bvh, O.TRAV_REFERENCE, o, d, f64=True)
    same = p32 == p64
    assert same.mean() >= 0.9999
    hit = same & (p64 != M.PRIM_NONE)
    assert np.max(np.abs(t32[hit] - t64[hit]) / t64[hit]) < 1e-5


def test_occlusion_limits():
    hs = scene_with(lambda h, m: h.add_sphere((0, 0, 3), 1.0, m))
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    o = np.zeros((3, 3), np.float32)
    d = np.tile(np.array([0, 0, 1], np.float32), (3, 1))
    # the sphere's near side is at t = 2: a light at distance 1.5 is visible,
    # one at 2.5 is blocked, and the far limit (escape test) is blocked
    occ, _, _ = O.occluded_f32(view, bvh, o, d,
                               np.array([1.5, 2.5, 2000.0], np.float32))
    assert list(occ) == [0, 1, 1]


def test_viewport_closed_form():
    """Centre ray along the camera direction; corner rays subtend
    +-fov/2 horizontally (init_viewport_values, viewport.c:38-60)."""
    hs = M.HostScene.empty()
    hs.set_camera((1.0, 2.0, 3.0), (0.0, 0.0, 1.0), fov=90.0)
    view = hs.view()
    vp = O.viewport(M.C.addressof(view.contents.camera), 200, 100)
    origin, p00, du, dv = vp[0:3], vp[3:6], vp[6:9], vp[9:12]
    centre = p00 + du * 99.5 + dv * 49.5 - origin
    np.testing.assert_allclose(centre / np.linalg.norm(centre), [0, 0, 1],
                               atol=1e-12)
    # viewport width = 2 tan(45 deg) * focus = 2, u points to -x (right)
    np.testing.assert_allclose(du * 200, [-2, 0, 0], atol=1e-12)
    np.testing.assert_allclose(dv * 100, [0, -1, 0], atol=1e-12)
