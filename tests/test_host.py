"""Host front-end: .rt reader, PPM / .rts output, HDRI importance tables."""
from pathlib import Path

import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O

GOLDEN = Path(__file__).resolve().parent / "golden"

PLAIN_RT = """
# plain scene in the reference grammar
W 640 360
A 0.1 255,255,255
C 0,1.5,-6 0,0,1 60
L 0,5,-2 1.0 255,255,255
mat lam 204,204,204 0 0 0 0 1 0 0,0,0
mat metal 204,204,204 1 0.1 0 0 1 0 0,0,0
mat glass 255,255,255 0 0 1.5 1 1 0 0,0,0
pl 0,0,0 0,1,0 lam
sp -2,1,0 2 lam
sp 0,1,0 2 metal
sp 2,1,0 2 glass
sp 0,3,4 1 10,20,30
"""


def test_parse_plain_scene():
    hs = M.HostScene.parse(PLAIN_RT)
    s = hs.view().contents
    assert hs.resolution == (640, 360)
    assert (s.n_spheres, s.n_planes, s.n_lights, s.n_triangles) == (4, 1, 1, 0)
    a = hs.arrays()
    np.testing.assert_allclose(a["sph"][0], [-2, 1, 0, 1.0])   # diameter 2
    mats = [s.materials[i] for i in range(s.n_materials)]
    lam, metal, glass = mats[0], mats[1], mats[2]
    assert abs(lam.color[0] - 0.8) < 1e-7 and lam.ao == 1.0
    assert metal.metallic == 1.0 and abs(metal.roughness - 0.1) < 1e-7
    assert glass.ior == 1.5 and glass.transmission == 1.0
    bare = mats[a["sph_mat"][3]]
    assert bare.metallic == 0 and bare.ior == 0 and bare.ao == 1.0
    np.testing.assert_allclose(list(bare.color), [10 / 255, 20 / 255, 30 / 255],
                               rtol=1e-6)
    assert s.camera.fov_deg == 60 and s.camera.focus_dist == 1.0
    assert abs(s.ambient_ratio - 0.1) < 1e-7
    L = s.lights[0]
    assert L.type == M.LIGHT_POINT and L.brightness == 1.0 and L.radius == 0


def test_parse_reference_subject_scene():
    hs = M.HostScene.parse_file(GOLDEN / "subject_sphere.rt")
    s = hs.view().contents
    assert hs.resolution == (800, 800)
    assert s.n_spheres == 1 and s.n_lights == 1
    np.testing.assert_allclose(hs.arrays()["sph"][0], [0, 0, 0, 0.75])
    assert list(s.camera.dir) == [0.0, 0.0, -1.0]


def test_light_radius_defaults_to_five_samples():
    hs = M.HostScene.parse(
        "A 0.2 255,255,255\nC 0,0,0 0,0,1 70\nl 1,2,3 0.5 255,0,0 0.4\n")
    L = hs.view().contents.lights[0]
    assert abs(L.radius - 0.4) < 1e-7 and L.shadow_samples == 5


@pytest.mark.parametrize("text,code", [
    ("C 0,0,0 0,0,1 70\n", M.MRT_E_INVALID),                      # no A
    ("A 0.2 255,255,255\nA 0.2 255,255,255\nC 0,0,0 0,0,1 70\n",
     M.MRT_E_INVALID),
    ("A 0.2 255,255,256\nC 0,0,0 0,0,1 70\n", M.MRT_E_INVALID),   # colour
    ("A 0.2 255,255,255\nC 0,0,0 0,0,2 70\n", M.MRT_E_INVALID),   # dir
    ("A 0.2 255,255,255\nC 0,0,0 0,0,1 181\n", M.MRT_E_INVALID),  # fov
    ("A 0.2 255,255,255\nC 0,0,0 0,0,1 70\nsp 0,0,0 1 nomat\n",
     M.MRT_E_INVALID),
    ("A 0.2 255,255,255\nC 0,0,0 0,0,1 70\ncy 0,0,0 0,1,0 1 2 255,0,0 top\n",
     M.MRT_E_INVALID),                                  # unknown cap material
    ("A 0.2 255,255,255\nC 0,0,0 0,0,1 70\nxx 1\n", M.MRT_E_INVALID),
])
def test_parse_errors(text, code):
    with pytest.raises(M.MrtError) as e:
        M.HostScene.parse(text)
    assert e.value.code == code


def test_material_may_follow_its_use():
    hs = M.HostScene.parse("A 0.2 255,255,255\nC 0,0,0 0,0,1 70\n"
                           "sp 0,0,5 1 red\n"
                           "mat red 255,0,0 0 0 0 0 1 0 0,0,0\n")
    s = hs.view().contents
    assert s.materials[hs.arrays()["sph_mat"][0]].color[0] == 1.0


def test_ppm_export(tmp_path):
    acc = np.array([[[0.5, 2.0, -1.0], [4.0, 1.0, 0.25]]], np.float64) * 4
    p = tmp_path / "x.ppm"
    M.export_ppm(p, acc, samples=4, gamma=1.0)
    data = p.read_bytes()
    assert data.startswith(b"P6\n2 1\n255\n")
    px = np.frombuffer(data[len(b"P6\n2 1\n255\n"):], np.uint8)
    # clamp01(acc / samples) * 255, truncated
    assert list(px) == [127, 255, 0, 255, 255, 63]
    M.export_ppm(p, acc, samples=4, gamma=2.0)
    px = np.frombuffer(p.read_bytes()[len(b"P6\n2 1\n255\n"):], np.uint8)
    assert px[0] == int(np.sqrt(0.5) * 255) and px[5] == 127


def test_heat_map_ramp():
    """get_heatmap_fcolor's four ramps at hand-picked ratios, misses
    untouched, values added to what the buffer holds."""
    depth = np.array([[0, 10, 30, 50], [70, 90, 100, -1]], np.int32)
    acc = np.full((2, 4, 3), 0.5)
    M.heat_map_add(depth, acc)
    want = np.array([[[0, 0, 0], [0, 0, 0.4], [0, 1, 0.8], [0, 1, 0]],
                     [[0.8, 1, 0], [1, 0.4, 0], [1, 0, 0], [0, 0, 0]]])
    np.testing.assert_allclose(acc - 0.5, want, atol=1e-6)
    # a buffer whose largest depth is 0 is blue where hit
    acc = np.zeros((1, 2, 3))
    M.heat_map_add(np.array([0, -1], np.int32), acc)
    assert acc.tolist() == [[[0, 0, 1], [0, 0, 0]]]


def test_rts_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    acc = rng.random((5, 7, 3))
    p = tmp_path / "render_state.rts"
    M.export_rts(p, acc, 17)
    raw = p.read_bytes()
    assert np.frombuffer(raw[:12], np.int32).tolist() == [7, 5, 17]
    assert len(raw) == 12 + 5 * 7 * 3 * 8
    back, s = M.load_rts(p, 7, 5)
    assert s == 17 and np.array_equal(back, acc)
    with pytest.raises(M.MrtError):
        M.load_rts(p, 5, 7)


def brute_tables(rgb):
    """Direct O(W^2 H) re-derivation of calc_pdf.c / calc_cdf.c."""
    h, w = rgb.shape[:2]
    c = rgb.astype(np.float64)
    gray = 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]
    raw = np.cumsum(gray, axis=1)[:, -1] / w      # sequential sums, as C
    col = np.cumsum(raw)[-1] / h
    pdf = (gray / col).astype(np.float32)
    cdf_h = np.cumsum(raw / (col / 255))
    cdf_h = cdf_h / cdf_h[-1]
    marg = np.zeros(h, np.float32)
    for i in range(h):
        u = i / (h - 1)
        for y in range(h):
            if cdf_h[y] >= u:
                marg[i] = y / (h - 1)
                break
    cond = np.zeros((h, w), np.float32)
    for y in range(h):
        c = np.cumsum(gray[y] / (raw[y] / 255))
        c = c / c[-1]
        for i in range(w):
            u = i / (w - 1)
            for x in range(w):
                if c[x] >= u:
                    cond[y, i] = x / (w - 1)
                    break
    return pdf, marg, cond


def test_hdri_tables_match_brute_force_and_oracle():
    rng = np.random.default_rng(4)
    rgb = rng.random((6, 9, 3)).astype(np.float32) ** 3
    rgb[2, 4] = 40.0                     # a bright texel
    pdf, marg, cond, col = M.hdri_tables(rgb)
    bp, bm, bc = brute_tables(rgb)
    np.testing.assert_array_equal(pdf, bp)
    np.testing.assert_array_equal(marg, bm)
    np.testing.assert_array_equal(cond, bc)
    op, om, oc, ocol = O.hdri_tables(rgb)
    np.testing.assert_array_equal(pdf, op)
    np.testing.assert_array_equal(marg, om)
    np.testing.assert_array_equal(cond, oc)
    assert col == ocol


def test_hdri_constant_map_gives_identity_tables():
    rgb = np.full((8, 16, 3), 0.5, np.float32)
    pdf, marg, cond, _ = M.hdri_tables(rgb)
    np.testing.assert_allclose(pdf, 1.0, rtol=1e-12)
    # a uniform CDF reaches i/(n-1) at index ceil(i (n-1)/(n-1) - ...) = i
    # up to rounding one step up; every entry is y/(n-1) and non-decreasing
    assert np.all(np.diff(marg) >= 0) and np.all(np.diff(cond, axis=1) >= 0)
    steps = np.round(marg * 7)
    assert np.all(np.abs(steps - np.arange(8)) <= 1)
    assert marg[-1] == 1.0 and cond[0, -1] == 1.0


def test_hdri_sun_map_concentrates_samples():
    sky = M.hdri_sky(5, 128, 64, sun_radius_deg=3.0, sun_radiance=5e4)
    pdf, marg, cond, _ = M.hdri_tables(sky)
    sun = sky[..., 0] >= 5e4
    assert sun.any()
    rows = np.nonzero(sun.any(1))[0]
    # nearly all inverse-marginal entries land in the sun's rows
    hit_rows = np.round(marg * 63).astype(int)
    assert np.isin(hit_rows, rows).mean() > 0.95


def _triangles(hs):
    v = hs.view().contents
    return np.ctypeslib.as_array(v.tri_pos, (v.n_triangles * 9,)).reshape(-1, 3, 3).copy()


@pytest.mark.parametrize("freq", [1, 2, 3, 7, 100])
def test_geodesic_icosphere_count_and_closed(freq):
    """20 freq^2 triangles; every edge is shared by exactly two triangles
    with the same vertex bits (no cracks where icosahedron faces meet), and
    each of the V - E + F = 2 vertices is used."""
    hs = M.HostScene.empty()
    hs.add_material(color=(0.5, 0.5, 0.5))
    n = hs.add_icosphere((0.0, 1.0, 0.0), 1.0, 0, 0.15, 7, mix=False, mat=0,
                         freq=freq)
    assert n == 20 * freq * freq
    tri = _triangles(hs)
    assert tri.shape[0] == n
    if freq > 7:
        return
    key = lambda p: tuple(np.asarray(p, np.float32).view(np.uint32))
    edges = {}
    verts = set()
    for t in tri:
        for a in range(3):
            p, q = key(t[a]), key(t[(a + 1) % 3])
            verts.add(p)
            edges[frozenset((p, q))] = edges.get(frozenset((p, q)), 0) + 1
    assert set(edges.values()) == {2}
    assert len(verts) - len(edges) + n == 2


def test_geodesic_icosphere_on_the_sphere_and_outward():
    """Without displacement every vertex lies on the sphere, and every
    triangle faces away from the centre; freq 1 is the icosahedron."""
    hs = M.HostScene.empty()
    hs.add_material(color=(0.5, 0.5, 0.5))
    c = np.array([0.5, 1.0, -0.25])
    hs.add_icosphere(tuple(c), 2.0, 0, 0.0, 3, mix=False, mat=0, freq=5)
    tri = _triangles(hs).astype(np.float64)
    r = np.linalg.norm(tri - c, axis=2)
    np.testing.assert_allclose(r, 2.0, rtol=1e-6)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert (np.einsum("ij,ij->i", nrm, tri.mean(axis=1) - c) > 0).all()
    ico = M.HostScene.empty()
    ico.add_material(color=(0.5, 0.5, 0.5))
    ico.add_icosphere((0.0, 0.0, 0.0), 1.0, 0, 0.0, 3, mix=False, mat=0, freq=1)
    rec = M.HostScene.empty()
    rec.add_material(color=(0.5, 0.5, 0.5))
    rec.add_icosphere((0.0, 0.0, 0.0), 1.0, 0, 0.0, 3, mix=False, mat=0)
    np.testing.assert_allclose(_triangles(ico), _triangles(rec), atol=1e-7)
