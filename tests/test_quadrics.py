"""Cylinders, cones and hyperboloids (hit_cylinder.c:107-134,
hit_cone.c:111-129, hit_hyperboloid.c:79-104) in the oracle, pinned by
closed forms worked out by hand, by the quadric's own implicit equation and
by brute force; the reference's quirks are pinned on the cases that show
them.  Also the .rt elements cy / co / hb."""
import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O

MODES = [(m, f) for m in (O.TRAV_BRUTE, O.TRAV_REFERENCE, O.TRAV_ORDERED)
         for f in (False, True)]
Q = M.PRIM_QUADRIC


def mats(hs):
    return (hs.add_material(color=(0.5, 0.5, 0.5)),
            hs.add_material(color=(1.0, 0.0, 0.0)),
            hs.add_material(color=(0.0, 1.0, 0.0)))


def query(hs, rays, mode, f64):
    view = hs.view()
    bvh = None if mode == O.TRAV_BRUTE else O.make_bvh(M.BuiltBvh(view))
    o = np.array([r[0] for r in rays], np.float32)
    d = np.array([r[1] for r in rays], np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return O.hit_info(view, bvh, mode, o, d, f64=f64)


@pytest.mark.parametrize("mode,f64", MODES)
def test_cylinder_closed_forms(mode, f64):
    hs = M.HostScene.empty()
    m, mt, mb = mats(hs)
    # centre 0, axis +y, radius 1, |y| <= 1
    hs.add_cylinder((0, 0, 0), (0, 1, 0), 2.0, 2.0, m, mt, mb)
    rays = [((-5, 0.5, 0), (1, 0, 0)),      # body from outside
            ((0, 5, 0.3), (0, -1, 0)),      # top cap
            ((0.2, -5, 0), (0, 1, 0)),      # bottom cap
            ((0, 0, 0), (1, 0, 0)),         # body from inside
            ((-5, 1.5, 0), (1, 0, 0)),      # above: body root out of range
            ((-5, 0, 3), (1, 0, 0)),        # beside: no real root
            ((0.5, 5, 0), (0, -1, 0))]      # along the axis: top cap
    pid, t, n, front, mat = query(hs, rays, mode, f64)
    assert list(pid) == [Q, Q, Q, Q, M.PRIM_NONE, M.PRIM_NONE, Q]
    np.testing.assert_allclose(t[[0, 1, 2, 3, 6]], [4, 4, 4, 1, 4], rtol=1e-6)
    np.testing.assert_allclose(
        n[[0, 1, 2, 3, 6]],
        [(-1, 0, 0), (0, 1, 0), (0, -1, 0), (-1, 0, 0), (0, 1, 0)], atol=1e-6)
    assert list(mat[[0, 1, 2, 3, 6]]) == [m, mt, mb, m, mt]
    # the body is always the front side, also from inside (the reference
    # recomputes front_face from the normal it has already flipped)
    assert front[3]
    assert front[[0, 1, 2, 6]].all()


@pytest.mark.parametrize("mode,f64", MODES)
def test_cylinder_keeps_valid_t_root_only(mode, f64):
    """valid_t picks the near root; when its projection falls outside
    [0, h] the far root is not tried.  An open view into the tube through
    the top is still closed by the top cap."""
    hs = M.HostScene.empty()
    m, mt, mb = mats(hs)
    hs.add_cylinder((0, 0, 0), (0, 1, 0), 2.0, 2.0, m, mt, mb)
    # enters the infinite cylinder above the top (near root out of range),
    # crosses the top cap at y = 1
    d = np.array([1.0, -1.0, 0.0]) / np.sqrt(2)
    o = np.array([-1.5, 2.5, 0.0])
    pid, t, n, front, mat = query(hs, [(o, d)], mode, f64)
    assert pid[0] == Q and mat[0] == mt
    # top cap plane y = 1: t = (2.5 - 1) * sqrt(2)
    assert abs(t[0] - 1.5 * np.sqrt(2)) < 1e-5


@pytest.mark.parametrize("mode,f64", MODES)
def test_cone_closed_forms(mode, f64):
    hs = M.HostScene.empty()
    m, mt, _ = mats(hs)
    # apex 0, axis +y, base radius 1 at y = 2: radius y / 2
    hs.add_cone((0, 0, 0), (0, 1, 0), 2.0, 2.0, m, mt)
    rays = [((-5, 1, 0), (1, 0, 0)),        # body at radius 0.5
            ((0.3, 5, 0), (0, -1, 0)),      # base cap at y = 2
            ((-5, -1, 0), (1, 0, 0)),       # the mirror nappe: no hit
            ((-5, 3, 0), (1, 0, 0))]        # above the base
    pid, t, n, front, mat = query(hs, rays, mode, f64)
    assert list(pid) == [Q, Q, M.PRIM_NONE, M.PRIM_NONE]
    np.testing.assert_allclose(t[:2], [4.5, 3.0], rtol=1e-6)
    # gradient of x^2 + z^2 - y^2/4 at (-0.5, 1, 0): (-1, -0.5, 0)
    g = np.array([-1.0, -0.5, 0.0]) / np.sqrt(1.25)
    np.testing.assert_allclose(n[0], g, atol=1e-6)
    np.testing.assert_allclose(n[1], (0, 1, 0), atol=1e-6)
    assert list(mat[:2]) == [m, mt] and front[:2].all()


@pytest.mark.parametrize("mode,f64", MODES)
def test_cone_steep_ray_from_inside_skips_the_body(mode, f64):
    """With a < 0 (a ray steeper than the cone's side) t0 is the larger
    root; from inside, valid_t takes the hit on the mirror nappe, which is
    out of [0, h], and only the base cap is tried next: a ray heading for
    the apex leaves the cone unseen."""
    hs = M.HostScene.empty()
    m, mt, _ = mats(hs)
    hs.add_cone((0, 0, 0), (0, 1, 0), 2.0, 2.0, m, mt)
    pid, _, _, _, _ = query(hs, [((0.1, 1.5, 0), (0.05, -1, 0))], mode, f64)
    assert pid[0] == M.PRIM_NONE
    # the same ray from outside, above the base, meets the base cap
    pid, t, _, _, mat = query(hs, [((0.0, 3.0, 0), (0.05, -1, 0))], mode, f64)
    assert pid[0] == Q and mat[0] == mt
    assert abs(t[0] - np.sqrt(1 + 0.05 ** 2)) < 1e-5


@pytest.mark.parametrize("mode,f64", MODES)
def test_hyperboloid_closed_forms(mode, f64):
    hs = M.HostScene.empty()
    m, _, _ = mats(hs)
    # x^2 + y^2 - z^2 = 1 about +z, |z| <= 2
    hs.add_hyperboloid((0, 0, 0), (0, 0, 1), 4.0, 1.0, 1.0, 1.0, 1.0, m)
    rays = [((-5, 0, 1), (1, 0, 0)),        # radius sqrt(2) at z = 1
            ((-5, 0, 3), (1, 0, 0)),        # beyond the height
            ((0, 0, 0), (1, 0, 0))]         # from inside the waist
    pid, t, n, front, mat = query(hs, rays, mode, f64)
    assert list(pid) == [Q, M.PRIM_NONE, Q]
    np.testing.assert_allclose(t[[0, 2]], [5 - np.sqrt(2), 1.0], rtol=1e-6)
    # gradient (2x, 2y, -2z) at (-sqrt 2, 0, 1)
    g = np.array([-np.sqrt(2), 0, -1]) / np.sqrt(3)
    np.testing.assert_allclose(n[0], g, atol=1e-6)
    np.testing.assert_allclose(n[2], (-1, 0, 0), atol=1e-6)
    assert front[0] and not front[2]
    assert list(mat[[0, 2]]) == [m, m]


@pytest.mark.parametrize("f64", [False, True])
def test_two_sheet_hyperboloid(f64):
    hs = M.HostScene.empty()
    m, _, _ = mats(hs)
    # x^2 + y^2 - z^2 = -1: sheets z = +-sqrt(1 + r^2), |z| <= 2
    hs.add_hyperboloid((0, 0, 0), (0, 0, 1), 4.0, 1.0, 1.0, 1.0, -1.0, m)
    pid, t, n, front, _ = query(hs, [((0, 0, -5), (0, 0, 1)),
                                     ((0.5, 0, -5), (0, 0, 1))],
                                O.TRAV_BRUTE, f64)
    assert list(pid) == [Q, Q]
    np.testing.assert_allclose(t, [4.0, 5 - np.sqrt(1.25)], rtol=1e-6)
    np.testing.assert_allclose(n[0], (0, 0, -1), atol=1e-6)


def soup(seed, n_each=60):
    rng = np.random.default_rng(seed)
    hs = M.HostScene.empty()
    m, mt, mb = mats(hs)
    for _ in range(n_each):
        ax = rng.normal(size=3)
        hs.add_cylinder(rng.uniform(-4, 4, 3), ax, rng.uniform(0.2, 1.0),
                        rng.uniform(0.2, 1.5), m, mt, mb)
        hs.add_cone(rng.uniform(-4, 4, 3), rng.normal(size=3),
                    rng.uniform(0.2, 1.2), rng.uniform(0.2, 1.5), m, mt)
        hs.add_hyperboloid(rng.uniform(-4, 4, 3), rng.normal(size=3),
                           rng.uniform(0.3, 1.5), *rng.uniform(0.15, 0.6, 3),
                           rng.choice([1.0, -1.0, 0.5]), m)
    for _ in range(20):
        hs.add_sphere(rng.uniform(-4, 4, 3), rng.uniform(0.1, 0.4), m)
    return hs


def random_rays(seed, n):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-6, 6, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    return o, (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


@pytest.mark.parametrize("f64", [False, True])
def test_bvh_traversals_equal_brute_force(f64):
    hs = soup(3)
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    o, d = random_rays(8, 6000)
    pb, tb, nb, fb, mb = O.hit_info(view, None, O.TRAV_BRUTE, o, d, f64=f64)
    assert (pb != M.PRIM_NONE).mean() > 0.3
    for mode in (O.TRAV_REFERENCE, O.TRAV_ORDERED):
        p, t, n, f, mat = O.hit_info(view, bvh, mode, o, d, f64=f64)
        assert np.array_equal(p, pb) and np.array_equal(mat, mb)
        np.testing.assert_array_equal(t, tb)
        np.testing.assert_array_equal(n, nb)
        np.testing.assert_array_equal(f, fb)


def test_hits_lie_on_the_surfaces_and_fp32_tracks_fp64():
    hs = soup(5)
    view = hs.view()
    q = hs.arrays()["quadrics"]
    o, d = random_rays(9, 6000)
    p64, t64, n64, f64_, m64 = O.hit_info(view, None, O.TRAV_BRUTE, o, d,
                                          f64=True)
    p32, t32, n32, f32_, m32 = O.hit_info(view, None, O.TRAV_BRUTE, o, d)
    same = p32 == p64
    assert same.mean() >= 0.999
    hq = same & (p64 != M.PRIM_NONE) & ((p64 & 0xC0000000) == Q)
    assert hq.sum() > 500
    np.testing.assert_allclose(t32[hq], t64[hq], rtol=2e-4, atol=1e-5)
    assert np.mean(np.abs(n32[hq] - n64[hq]).max(1) < 1e-3) >= 0.999
    # each fp64 hit satisfies its quadric's equation (body) or lies on a
    # cap disk, within the height range
    checked = 0
    for i in np.nonzero(hq)[0]:
        Qd = q[p64[i] & 0x3FFFFFFF]
        x = o[i].astype(np.float64) + d[i].astype(np.float64) * t64[i]
        P, ax = np.array(Qd["pos"]), np.array(Qd["axis"], np.float64)
        rel = x - P
        z = rel @ ax
        r2 = rel @ rel - z * z
        h = Qd["height"]
        if Qd["kind"] == M.QUADRIC_CYLINDER:
            on_body = abs(np.sqrt(r2) - Qd["radius"]) < 1e-6 and -h / 2 - 1e-6 <= z <= h / 2 + 1e-6
            on_cap = abs(abs(z) - h / 2) < 1e-6 and np.sqrt(r2) <= Qd["radius"] + 1e-6
            assert on_body or on_cap
        elif Qd["kind"] == M.QUADRIC_CONE:
            rr = Qd["radius"] * z / h
            on_body = abs(np.sqrt(r2) - rr) < 1e-5 and -1e-6 <= z <= h + 1e-6
            on_cap = abs(z - h) < 1e-6 and np.sqrt(r2) <= Qd["radius"] + 1e-6
            assert on_body or on_cap
        else:
            zz = ax
            tmp = np.array([0, 1.0, 0]) if abs(zz[0]) > 0.9 else np.array([1.0, 0, 0])
            xx = np.cross(tmp, zz)
            xx /= np.linalg.norm(xx)
            yy = np.cross(zz, xx)
            lhs = ((rel @ xx) / Qd["a"]) ** 2 + ((rel @ yy) / Qd["b"]) ** 2 \
                - ((rel @ zz) / Qd["c"]) ** 2
            assert abs(lhs - Qd["shape"]) < 1e-5 * (1 + abs(lhs))
            assert abs(z) <= h / 2 + 1e-9
        checked += 1
    assert checked == hq.sum()


def test_box_bounds_hold_every_hit():
    """The builder's boxes contain every hit point (a ray that misses a box
    never reaches the primitive): compare the BVH with the brute force on
    rays aimed at the quadrics' edges."""
    hs = soup(12, n_each=30)
    view = hs.view()
    q = hs.arrays()["quadrics"]
    rng = np.random.default_rng(1)
    o, d = [], []
    for Qd in q:
        P, ax = np.array(Qd["pos"]), np.array(Qd["axis"])
        for _ in range(40):
            s = rng.normal(size=3)
            s -= (s @ ax) * ax
            s /= np.linalg.norm(s)
            r = Qd["radius"] if Qd["kind"] != M.QUADRIC_HYPERBOLOID else \
                max(Qd["a"], Qd["b"]) * np.sqrt(max(0.0, Qd["shape"] + (0.5 * Qd["height"] / Qd["c"]) ** 2))
            zoff = rng.choice([-0.5, 0.5, 1.0]) * Qd["height"]
            tgt = P + ax * zoff + s * r * rng.uniform(0.95, 1.0)
            src = tgt + rng.normal(size=3) * 3
            o.append(src)
            dd = tgt - src
            d.append(dd / np.linalg.norm(dd))
    o = np.array(o, np.float32)
    d = np.array(d, np.float32)
    bvh = O.make_bvh(M.BuiltBvh(view))
    for f64 in (False, True):
        pb, tb, _, _, _ = O.hit_info(view, None, O.TRAV_BRUTE, o, d, f64=f64)
        p, t, _, _, _ = O.hit_info(view, bvh, O.TRAV_ORDERED, o, d, f64=f64)
        assert np.array_equal(p, pb)
        np.testing.assert_array_equal(t, tb)


def test_rt_quadric_elements(tmp_path):
    (tmp_path / "q.rt").write_text(
        "A 0.2 255,255,255\nC 0,1,-6 0,0,1 60\n"
        "mat top 255,0,0 0 0 0 0 1 0 0,0,0\n"
        "mat bot 0,255,0 0 0 0 0 1 0 0,0,0\n"
        "cy 0,0,0 0,1,0 2 3 10,20,30 top bot\n"
        "cy 1,0,0 0,0,-1 1 1 top\n"
        "co 0,2,0 0,-1,0 2 2 bot top\n"
        "hb 0,0,5 0,1,0 4 1 2 3 -0.5 10,10,10\n")
    hs = M.HostScene.parse_file(str(tmp_path / "q.rt"))
    q = hs.arrays()["quadrics"]
    assert [x["kind"] for x in q] == [M.QUADRIC_CYLINDER, M.QUADRIC_CYLINDER,
                                      M.QUADRIC_CONE, M.QUADRIC_HYPERBOLOID]
    top = 0
    bot = 1
    assert q[0]["radius"] == 1.0 and q[0]["height"] == 3.0
    assert (q[0]["mat_top"], q[0]["mat_bot"]) == (top, bot)
    assert q[1]["mat"] == top and q[1]["mat_top"] == top and q[1]["mat_bot"] == top
    assert q[1]["axis"] == [0, 0, -1]
    assert (q[2]["mat"], q[2]["mat_top"]) == (bot, top)
    assert (q[3]["height"], q[3]["a"], q[3]["b"], q[3]["c"], q[3]["shape"]) == \
        (4.0, 1.0, 2.0, 3.0, -0.5)
    mats_ = hs.view().contents.materials
    assert tuple(mats_[q[0]["mat"]].color) == pytest.approx((10 / 255, 20 / 255, 30 / 255))


@pytest.mark.parametrize("line", [
    "cy 0,0,0 0,1,0 2 3",                       # too few fields
    "cy 0,0,0 0,1,0 0 3 10,10,10",              # zero diameter
    "cy 0,0,0 0,2,0 1 3 10,10,10",              # orientation out of range
    "cy 0,0,0 0,1,0 1 3 10,10,10 nosuch",       # unknown cap material
    "co 0,0,0 0,1,0 1 2 10,10,10 a b",          # too many fields
    "hb 0,0,0 0,1,0 1 1 1 10,10,10",            # too few fields
    "hb 0,0,0 0,1,0 1 1 0 1 1 10,10,10"])       # zero semi-axis
def test_rt_quadric_errors(tmp_path, line):
    (tmp_path / "e.rt").write_text("A 0.2 255,255,255\nC 0,1,-6 0,0,1 60\n"
                                   + line + "\n")
    with pytest.raises(M.MrtError) as e:
        M.HostScene.parse_file(str(tmp_path / "e.rt"))
    assert e.value.code == M.MRT_E_INVALID
