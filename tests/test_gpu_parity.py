"""HIP path vs the oracle on the same counter-RNG stream (needs a GPU)."""
import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O

pytestmark = pytest.mark.gpu

SEED = 0x5EED1234


def _setup(hs):
    view = hs.view()
    built = M.BuiltBvh(view)
    return view, O.make_bvh(built)


@pytest.fixture(scope="module")
def plain():
    hs = M.HostScene.plain(7, 160, 90)
    view, bvh = _setup(hs)
    r = M.Renderer(160, 90)
    r.upload(hs)
    r.set_seed(SEED)
    yield hs, view, bvh, r
    r.close()


@pytest.fixture(scope="module")
def mesh():
    hs = M.HostScene.mesh(3, 192, 108, subdiv=6, floor_cells=48)
    view, bvh = _setup(hs)
    r = M.Renderer(192, 108)
    r.upload(hs)
    r.set_seed(SEED)
    yield hs, view, bvh, r
    r.close()


def quadric_scene(w, h):
    """Cylinders, cones and hyperboloids of every material kind over a
    plane, lit by a point light and a soft light, with a glass sphere."""
    hs = M.HostScene.empty()
    hs.set_resolution(w, h)
    hs.set_camera((0.0, 2.2, -7.0), (0.0, -0.22, 1.0), 60.0)
    hs.set_ambient(0.15, (255, 255, 255))
    lam = hs.add_material(color=(0.8, 0.7, 0.6))
    red = hs.add_material(color=(0.9, 0.2, 0.2))
    metal = hs.add_material(color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.2)
    glass = hs.add_material(color=(1.0, 1.0, 1.0), ior=1.5, transmission=1.0)
    mixed = hs.add_material(color=(0.3, 0.6, 0.9), metallic=0.5, roughness=0.4)
    hs.add_plane((0, 0, 0), (0, 1, 0), lam)
    hs.add_cylinder((-2.2, 1.0, 0.5), (0, 1, 0), 1.2, 2.0, red, metal, lam)
    hs.add_cylinder((0.4, 0.6, -1.5), (1, 0.3, 0.2), 0.8, 1.4, glass)
    hs.add_cone((2.0, 2.2, 0.8), (0, -1, 0), 1.6, 2.2, metal, red)
    hs.add_cone((-0.6, 0.0, 1.8), (0.2, 1, -0.1), 1.0, 1.5, mixed)
    hs.add_hyperboloid((0.5, 1.2, 2.5), (0, 1, 0), 2.4, 0.5, 0.5, 0.8, 1.0, lam)
    hs.add_hyperboloid((-2.5, 2.8, 2.5), (0.3, 1, 0), 1.5, 0.4, 0.6, 0.5, -0.3,
                       mixed)
    hs.add_sphere((1.5, 0.5, -2.0), 0.5, glass)
    hs.add_light((1.0, 5.0, -3.0), 0.8, (255, 255, 255))
    hs.add_light((-3.0, 4.0, -1.0), 0.5, (255, 230, 200), radius=0.4,
                 samples=4)
    return hs


def textured_scene(w, h):
    """Every texture kind in every slot: checker floor, image-mapped
    spheres (one normal-mapped, one emissive), a cylinder with checker caps,
    a metal cone with a roughness texture and an alpha-cut quad with
    texture coordinates."""
    rng = np.random.default_rng(21)
    hs = M.HostScene.empty()
    hs.set_resolution(w, h)
    hs.set_camera((0.0, 2.0, -6.5), (0.0, -0.2, 1.0), 60.0)
    hs.set_ambient(0.2, (255, 255, 255))
    img = rng.integers(0, 256, (16, 24, 4), dtype=np.uint8)
    img[..., 3] = 255
    cut = img.copy()
    cut[:, :8, 3] = 0
    cut[:, 8:16, 3] = 100
    nmap = np.zeros((8, 8, 4), np.uint8)
    nmap[..., 0] = rng.integers(60, 200, (8, 8))
    nmap[..., 1] = rng.integers(60, 200, (8, 8))
    nmap[..., 2] = 230
    nmap[..., 3] = 255
    t_img = hs.add_image_texture(img)
    t_cut = hs.add_image_texture(cut)
    t_nrm = hs.add_image_texture(nmap)
    t_cg = hs.add_checker_texture(M.TEX_CHECKER_GLOBAL, 1.5, (230, 230, 230),
                                  (40, 60, 90))
    t_cl = hs.add_checker_texture(M.TEX_CHECKER_LOCAL, 6.0, (255, 200, 0),
                                  (20, 20, 20))
    floor = hs.add_material(color=(0.5, 0.5, 0.5))
    hs.set_material_tex(floor, color=t_cg)
    sph = hs.add_material(color=(0.5, 0.5, 0.5))
    hs.set_material_tex(sph, color=t_img, normal=t_nrm, scale=2.0)
    glow = hs.add_material(color=(0.5, 0.5, 0.5), emission=2.0)
    hs.set_material_tex(glow, emission_color=t_img, emission_strength=t_cl)
    body = hs.add_material(color=(0.7, 0.2, 0.2))
    cap = hs.add_material(color=(0.5, 0.5, 0.5))
    hs.set_material_tex(cap, color=t_cl, normal_intensity=0.5, normal=t_nrm)
    metal = hs.add_material(color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.3)
    hs.set_material_tex(metal, roughness=t_cl, ao=t_cg)
    quad = hs.add_material(color=(1.0, 1.0, 1.0))
    hs.set_material_tex(quad, color=t_cut, metallic=t_cg)
    hs.add_plane((0, 0, 0), (0, 1, 0), floor)
    hs.add_sphere((-1.8, 1.0, 0.5), 1.0, sph)
    hs.add_sphere((1.6, 0.5, -1.5), 0.5, glow)
    hs.add_cylinder((0.3, 0.8, 1.8), (0.2, 1, 0.1), 1.4, 1.6, body, cap, cap)
    hs.add_cone((2.2, 2.0, 1.0), (0, -1, 0), 1.6, 2.0, metal, cap)
    obj = ("v -1 0 0\nv 1 0 0\nv 1 2 0\nv -1 2 0\n"
           "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nf 1/1 2/2 3/3 4/4\n")
    hs.add_obj_text(obj, (-0.2, 0.0, -1.0), (0.3, 0.0, -1.0), 0.9, quad)
    hs.add_light((1.0, 5.0, -3.0), 0.8, (255, 255, 255))
    return hs


def physics_scene(w, h):
    """A thin-lens camera, a directional light beside a point light, and
    glass: a smooth sphere, a rough one, a half-transmissive one and a glass
    slab (a box of twelve triangles) inside which steep rays reflect
    totally."""
    hs = M.HostScene.empty()
    hs.set_resolution(w, h)
    hs.set_camera((0.0, 1.8, -6.0), (0.0, -0.2, 1.0), 55.0, defocus=3.0,
                  focus=6.0)
    hs.set_ambient(0.2, (200, 220, 255))
    lam = hs.add_material(color=(0.8, 0.8, 0.8))
    red = hs.add_material(color=(0.85, 0.25, 0.2))
    glass = hs.add_material(color=(1.0, 1.0, 1.0), ior=1.5, transmission=1.0)
    rough = hs.add_material(color=(0.9, 0.95, 1.0), ior=1.5, transmission=1.0,
                            roughness=0.25)
    part = hs.add_material(color=(0.9, 0.9, 0.7), ior=1.33, transmission=0.5)
    hs.add_plane((0, 0, 0), (0, 1, 0), lam)
    hs.add_sphere((-1.6, 1.0, 0.0), 1.0, glass)
    hs.add_sphere((1.4, 0.7, -0.8), 0.7, rough)
    hs.add_sphere((0.3, 0.5, 1.8), 0.5, red)
    hs.add_sphere((2.4, 0.6, 1.5), 0.6, part)
    # glass slab [-0.6, 0.6] x [0.05, 0.65] x [-2.6, -1.4]
    lo, hi = np.array([-0.6, 0.05, -2.6]), np.array([0.6, 0.65, -1.4])
    c = [[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1],
          lo[2] if i & 4 == 0 else hi[2]] for i in range(8)]
    for a, b, cc, d in ((0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1),
                        (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)):
        hs.add_triangle(c[a] + c[b] + c[cc], glass)
        hs.add_triangle(c[a] + c[cc] + c[d], glass)
    d = np.array([-0.5, -1.0, 0.4])
    hs.add_light(tuple(d / np.linalg.norm(d)), 0.6, (255, 240, 220),
                 kind=M.LIGHT_DIRECTIONAL)
    hs.add_light((2.0, 5.0, -3.0), 0.5, (255, 255, 255))
    return hs


@pytest.fixture(scope="module")
def physics():
    hs = physics_scene(160, 96)
    view, bvh = _setup(hs)
    r = M.Renderer(160, 96)
    r.upload(hs)
    r.set_seed(SEED)
    yield hs, view, bvh, r
    r.close()


@pytest.fixture(scope="module")
def textured():
    hs = textured_scene(160, 96)
    view, bvh = _setup(hs)
    r = M.Renderer(160, 96)
    r.upload(hs)
    r.set_seed(SEED)
    yield hs, view, bvh, r
    r.close()


def test_textured_normals_match_oracle(textured):
    hs, view, bvh, r = textured
    w, h = r.width, r.height
    ids = r.aov(M.AOV_PRIM_ID, 1)
    nrm = r.aov(M.AOV_NORMAL, 1).reshape(-1, 3)
    o, d = O.raygen_f32(view, w, h, 1, SEED)
    rid, _, rn, _, _ = O.hit_info(view, bvh, O.TRAV_ORDERED, o, d)
    same = ids == rid
    assert same.mean() >= 0.999
    hit = same & (rid != M.PRIM_NONE)
    err = np.abs(nrm[hit] - rn[hit]).max(1)
    assert np.mean(err <= 1e-5) >= 0.999


@pytest.fixture(scope="module")
def quadrics():
    hs = quadric_scene(160, 96)
    view, bvh = _setup(hs)
    r = M.Renderer(160, 96)
    r.upload(hs)
    r.set_seed(SEED)
    yield hs, view, bvh, r
    r.close()


def test_quadric_normals_and_sides_match_oracle(quadrics):
    hs, view, bvh, r = quadrics
    w, h = r.width, r.height
    ids = r.aov(M.AOV_PRIM_ID, 2)
    nrm = r.aov(M.AOV_NORMAL, 2).reshape(-1, 3)
    o, d = O.raygen_f32(view, w, h, 2, SEED)
    rid, _, rn, _, _ = O.hit_info(view, bvh, O.TRAV_ORDERED, o, d)
    same = ids == rid
    assert same.mean() >= 0.9999
    q = same & (rid != M.PRIM_NONE) & ((rid & M.PRIM_KIND_MASK) == M.PRIM_QUADRIC)
    assert q.mean() > 0.1
    err = np.abs(nrm[q] - rn[q]).max(1)
    assert np.mean(err <= 1e-5) >= 0.999


@pytest.mark.parametrize("which", ["plain", "mesh", "quadrics", "textured",
                                   "physics"])
@pytest.mark.parametrize("sample", [0, 5])
def test_primary_hits_match_fp32_oracle(request, which, sample):
    hs, view, bvh, r = request.getfixturevalue(which)
    w, h = r.width, r.height
    ids = r.aov(M.AOV_PRIM_ID, sample)
    t = r.aov(M.AOV_T, sample)
    nodes = r.aov(M.AOV_NODES, sample)
    prims = r.aov(M.AOV_PRIMS, sample)
    o, d = O.raygen_f32(view, w, h, sample, SEED)
    rid, rt, rn, rp = O.trace_rays(view, bvh, O.TRAV_ORDERED, o, d)
    print(f"{which} s={sample}: id {np.mean(ids == rid):.6f} "
          f"nodes {np.mean(nodes == rn):.6f} prims {np.mean(prims == rp):.6f}")
    assert np.mean(ids == rid) >= 0.9999
    hit = rid != M.PRIM_NONE
    rel = np.abs(t[hit] - rt[hit]) / rt[hit]
    # a lens point takes cosf / sinf of the concentric map, where the
    # device's and the host's libm may differ by an ulp: 1e-5 there
    tol = 1e-5 if which == "physics" else 1e-6
    print(f"  t: max rel {rel.max():.2e}, within {np.mean(rel <= tol):.6f}")
    assert np.mean(rel <= tol) >= 0.9999
    # same node order on the same fp32 rays: identical traversal counts
    assert np.mean(nodes == rn) >= 0.999
    assert np.mean(prims == rp) >= 0.999


@pytest.mark.parametrize("which", ["plain", "mesh", "quadrics", "textured",
                                   "physics"])
def test_primary_hits_match_fp64_reference_semantics(request, which):
    """IDs match the reference-precision oracle (fp64 vectors, fp32 MT
    scalars, recursive both-children hit_bvh) on >= 99.99% of pixels and t
    within 1e-5 relative (1e-4 on quadrics)."""
    hs, view, bvh, r = request.getfixturevalue(which)
    w, h = r.width, r.height
    ids = r.aov(M.AOV_PRIM_ID, 0)
    t = r.aov(M.AOV_T, 0)
    o, d = O.raygen_f32(view, w, h, 0, SEED)
    rid, rt, _, _ = O.trace_rays(view, bvh, O.TRAV_REFERENCE, o, d, f64=True)
    same = ids == rid
    hit = same & (rid != M.PRIM_NONE)
    rel = np.abs(t[hit].astype(np.float64) - rt[hit]) / rt[hit]
    # quadric roots come from an fp32 quadratic: 1e-4 for those
    tol = np.where((rid[hit] & M.PRIM_KIND_MASK) == M.PRIM_QUADRIC, 1e-4, 1e-5)
    good = same.copy()
    good[hit] &= rel <= tol
    print(f"{which}: fp64 id match {same.mean():.6f}, id+t {good.mean():.6f}")
    assert good.mean() >= 0.9999


@pytest.mark.parametrize("which", ["plain", "mesh", "quadrics", "textured",
                                   "physics"])
def test_image_matches_oracle_same_stream(request, which):
    hs, view, bvh, r = request.getfixturevalue(which)
    w, h = r.width, r.height
    r.clear()
    r.render(3, 4, 8)
    img = r.readback().astype(np.float64)
    ref, _ = O.render(view, bvh, w, h, 3, 4, 8, SEED)
    diff = np.abs(img - ref)
    close = diff <= 1e-4 + 1e-4 * np.abs(ref)
    print(f"{which}: mean abs diff {diff.mean():.3e}, "
          f"close {close.mean():.6f}")
    # paths diverge only where an fp32 branch decision flips
    assert close.mean() >= 0.995
    assert diff.mean() <= 1e-3 * max(ref.mean(), 1e-6)


def test_trace_rays_random(mesh):
    hs, view, bvh, r = mesh
    rng = np.random.default_rng(5)
    n = 20000
    o = rng.uniform(-3, 3, size=(n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(0.2, 3.0, size=n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    pid, t = r.trace_rays(o, d)
    rid, rt, _, _ = O.trace_rays(view, bvh, O.TRAV_BRUTE, o, d)
    assert np.mean(pid == rid) >= 0.9999
    hit = rid != M.PRIM_NONE
    assert np.allclose(t[hit], rt[hit], rtol=1e-6)


def test_lambert_sphere_closed_form():
    """Convex lambert sphere (albedo a) under a uniform background B, no
    lights: every bounce ray escapes, so each sphere pixel is a*B exactly
    (material_default.c:85-91 with add_skybox, path_trace.c:37-48)."""
    hs = M.HostScene.empty()
    hs.set_resolution(64, 64)
    a = 0.6
    m = hs.add_material(color=(a, a, a))
    hs.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    hs.set_ambient(0.25, (255, 255, 255))
    hs.set_camera((0.0, 0.0, -4.0), (0.0, 0.0, 1.0), fov=40.0)
    with M.Renderer(64, 64) as r:
        r.upload(hs)
        r.set_seed(9)
        ids = r.aov(M.AOV_PRIM_ID, 0).reshape(64, 64)
        r.render(0, 1, 8)
        img = r.readback()
    B = np.float32(0.25)
    on = ids == M.PRIM_SPHERE
    assert on.mean() > 0.2
    np.testing.assert_allclose(img[on], np.float32(a) * B, rtol=1e-6)
    np.testing.assert_allclose(img[~on], B, rtol=1e-6)


def test_stats_match_oracle_counts(plain):
    hs, view, bvh, r = plain
    w, h = r.width, r.height
    r.set_option(M.OPT_STATS, M.STATS_COUNT)
    r.reset_stats()
    r.render(0, 2, 8)
    st = r.stats()
    r.set_option(M.OPT_STATS, 0)
    _, cnt = O.render(view, bvh, w, h, 0, 2, 8, SEED)
    for k in ("rays_closest", "rays_shadow", "nodes_closest", "prims_closest",
              "nodes_shadow", "prims_shadow"):
        rel = abs(st[k] - cnt[k]) / max(cnt[k], 1)
        print(k, st[k], cnt[k], rel)
        assert rel <= 2e-3, k


def test_physics_scene_uses_the_lens(physics):
    """Guard on the fixture: its primary rays leave from the lens disk
    (radius focus * tan(angle / 2)), so the parity tests above cover
    defocus rays."""
    hs, view, bvh, r = physics
    o, d = O.raygen_f32(view, r.width, r.height, 0, SEED)
    off = np.linalg.norm(o - np.array([0.0, 1.8, -6.0], np.float32), axis=1)
    R = 6.0 * np.tan(np.radians(3.0) / 2)
    assert off.max() <= R * (1 + 1e-5) and off.mean() > 0.4 * R


def test_directional_light_closed_form_on_gpu():
    """A lambert sphere (albedo a) under one directional light and a black
    sky: pixel = a (c b n.l + c b max(r.v, 0)^64) / 255 on the lit side, 0
    elsewhere (add_dlight, d_light.c:35-57), with exact light sums."""
    w = h = 64
    a, b = 0.5, 0.7
    col = np.array([255.0, 180.0, 90.0])
    ldir = np.array([-0.4, -1.0, 0.6])
    ldir /= np.linalg.norm(ldir)
    hs = M.HostScene.empty()
    hs.set_resolution(w, h)
    m = hs.add_material(color=(a, a, a))
    hs.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    hs.add_light(tuple(ldir), b, tuple(col), kind=M.LIGHT_DIRECTIONAL)
    hs.set_ambient(0.0, (0, 0, 0))
    hs.set_camera((0.0, 0.0, -4.0), (0.0, 0.0, 1.0), fov=40.0)
    hs.set_flags(M.SCENE_LIGHT_EXACT)
    with M.Renderer(w, h) as r:
        r.upload(hs)
        r.set_seed(9)
        ids = r.aov(M.AOV_PRIM_ID, 0)
        t = r.aov(M.AOV_T, 0)
        r.render(0, 1, 8)
        img = r.readback().reshape(-1, 3).astype(np.float64)
    o, d = O.raygen_f32(hs.view(), w, h, 0, 9)
    p = o.astype(np.float64) + d.astype(np.float64) * t[:, None]
    on = ids == M.PRIM_SPHERE
    n = p / np.linalg.norm(p, axis=1, keepdims=True)
    l = -ldir
    ndl = n @ l
    v = np.array([0.0, 0.0, -4.0]) - p
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rr = -l + 2 * ndl[:, None] * n
    spec = np.maximum((rr * v).sum(1), 0) ** 64
    expect = a * (col * b * (ndl + spec)[:, None]) / 255.0
    lit = on & (ndl > 1e-3)
    assert lit.mean() > 0.1
    assert np.max(np.abs(img[lit] - expect[lit])) < 5e-4
    np.testing.assert_array_equal(img[on & (ndl < -1e-3)], 0.0)
    np.testing.assert_array_equal(img[~on], 0.0)


def test_primary_hits_at_full_size_on_the_1080p_mesh():
    """The benchmarked workload itself: 1920x1080 primary rays on the
    1,441,792-triangle tree (icosphere subdivided 8 times over a 256^2
    floor), whose deep 4-wide tree uses the trace stack's overflow columns.
    Ids match the fp32 ordered oracle and the fp64 reference traversal
    (hit_bvh, bvh_hit.c:16-44) on >= 99.99 % of pixels, t within 1e-6 and
    1e-5 relative."""
    w, h = 1920, 1080
    hs = M.HostScene.mesh(2, w, h)
    view = hs.view()
    assert view.contents.n_triangles == 1441792
    bvh = O.make_bvh(M.BuiltBvh(view))
    with M.Renderer(w, h) as r:
        r.upload(hs)
        r.set_seed(SEED)
        assert r.scene_bvh().depth4 >= 10
        ids = r.aov(M.AOV_PRIM_ID, 1)
        t = r.aov(M.AOV_T, 1)
    o, d = O.raygen_f32(view, w, h, 1, SEED)
    rid, rt, _, _ = O.trace_rays(view, bvh, O.TRAV_ORDERED, o, d)
    hit = rid != M.PRIM_NONE
    assert 0.3 < hit.mean() < 0.95
    rel = np.abs(t[hit] - rt[hit]) / rt[hit]
    print(f"1080p fp32: id {np.mean(ids == rid):.6f}, "
          f"t<=1e-6 {np.mean(rel <= 1e-6):.6f}")
    assert np.mean(ids == rid) >= 0.9999
    assert np.mean(rel <= 1e-6) >= 0.9999
    fid, ft, _, _ = O.trace_rays(view, bvh, O.TRAV_REFERENCE, o, d, f64=True)
    same = ids == fid
    fh = same & (fid != M.PRIM_NONE)
    frel = np.abs(t[fh].astype(np.float64) - ft[fh]) / ft[fh]
    good = same.copy()
    good[fh] &= frel <= 1e-5
    print(f"1080p fp64 reference: id {same.mean():.6f}, id+t {good.mean():.6f}")
    assert good.mean() >= 0.9999
