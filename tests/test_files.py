"""Scene files: Radiance .hdr decoding (src/utils/hdr/parse_hdr*.c and
get_hdr_pixel, get_hdr_value.c:15-40), OBJ meshes placed as the .rt "cu"
element places them (obj_finilizer.c:18-46, obj_ear_clipping.c:57-95), and
.rt files that use an HDR skybox texture and a mesh."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import mrt_hip as M


def rle_channel(vals):
    """Radiance run-length coding of one channel of a scanline: runs of
    equal bytes as (128 + n, byte), other bytes as (n, bytes...)."""
    out, i, n = bytearray(), 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), vals[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 128 and not (
                j + 2 < n and vals[j] == vals[j + 1] == vals[j + 2]):
            j += 1
        j = max(j, i + 1)
        out += bytes([j - i]) + bytes(vals[i:j])
        i = j
    return bytes(out)


def write_hdr(rgbe, exposure=None, gamma=None, rle=True, ysign="-"):
    """rgbe: uint8 [h, w, 4], file row order (first row = first scanline)."""
    h, w, _ = rgbe.shape
    head = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n"
    if exposure is not None:
        head += f"EXPOSURE={exposure}\n".encode()
    if gamma is not None:
        head += f"GAMMA={gamma}\n".encode()
    head += f"\n{ysign}Y {h} +X {w}\n".encode()
    body = bytearray()
    for y in range(h):
        if rle:
            body += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                body += rle_channel(list(rgbe[y, :, c]))
        else:
            body += rgbe[y].tobytes()
    return head + bytes(body)


def expected(rgbe):
    c = rgbe[..., :3].astype(np.float64)
    e = rgbe[..., 3].astype(np.int64)
    return c * np.ldexp(1.0, e - 128)[..., None] / 255.0


@pytest.mark.parametrize("rle", [True, False])
def test_hdr_decode_matches_the_rgbe_formula(rle):
    rng = np.random.default_rng(4)
    h, w = 7, 40
    rgbe = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rgbe[..., 3] = rng.integers(120, 136, (h, w))
    rgbe[2, 5:30] = rgbe[2, 5]                      # long runs
    rgbe[4, :, 3] = 128
    data = write_hdr(rgbe, exposure=0.5, gamma=2.2, rle=rle)
    got, ex, ga = M.decode_hdr(data)
    assert got.shape == (h, w, 3)
    assert ex == pytest.approx(0.5) and ga == pytest.approx(2.2)
    # "-Y": the first scanline is the top row; rows are stored bottom first
    np.testing.assert_allclose(got, expected(rgbe)[::-1], rtol=1e-6)


def test_hdr_positive_y_keeps_the_file_order():
    rgbe = np.full((3, 8, 4), 128, np.uint8)
    rgbe[0, :, 0] = 255
    got, _, _ = M.decode_hdr(write_hdr(rgbe, ysign="+"))
    np.testing.assert_allclose(got, expected(rgbe), rtol=1e-6)


@pytest.mark.parametrize("bad", [
    b"P6\n", b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 2 +X 2\n",
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n\x00"])
def test_hdr_rejects_bad_files(bad):
    with pytest.raises(M.MrtError):
        M.decode_hdr(bad)


OBJ = """# a unit square and a pentagon
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 2 0 0
v 3 0 0
v 3.5 1 0
v 2.5 2 0
v 1.5 1 0
vn 0 0 1
f 1//1 2//1 3//1 4//1
f -5 -4 -3 -2 -1
"""


@pytest.mark.parametrize("direction", [(0, 0, 1), (1, 0, 0), (0.3, -0.5, 0.8),
                                       (0, 0, -1)])
def test_obj_placement_and_ear_clipping(direction):
    hs = M.HostScene.empty()
    m = hs.add_material()
    pos = np.array([1.0, -2.0, 0.5])
    scale = 1.5
    n = hs.add_obj_text(OBJ, pos, direction, scale, m)
    assert n == 2 + 3
    arr = hs.arrays()
    tri = arr["tri_pos"].reshape(-1, 3, 3)
    d = np.asarray(direction, float)
    d /= np.linalg.norm(d)
    ax = np.cross([0, 0, 1], d)
    if np.linalg.norm(ax) < 1e-4:
        rot = Rotation.identity() if d[2] > 0 else Rotation.from_rotvec([np.pi, 0, 0])
    else:
        ang = np.arccos(np.clip(d[2], -1, 1))
        rot = Rotation.from_rotvec(ax / np.linalg.norm(ax) * ang)
    V = np.array([l.split()[1:] for l in OBJ.splitlines() if l.startswith("v ")],
                 float)
    P = rot.apply(V * scale) + pos
    # the reference's order: (v[n-1], v0, v1), drop v0, ...; last three
    faces = [[3, 0, 1], [1, 2, 3], [8, 4, 5], [8, 5, 6], [6, 7, 8]]
    np.testing.assert_allclose(tri, P[faces], atol=1e-5)
    # the square carries its (rotated) normals
    nrm = arr["tri_nrm"].reshape(-1, 3, 3)
    np.testing.assert_allclose(nrm[0], np.tile(rot.apply([0, 0, 1]), (3, 1)),
                               atol=1e-6)


def test_rt_file_with_hdr_skybox_and_mesh(tmp_path):
    rgbe = np.full((4, 16, 4), 128, np.uint8)
    rgbe[..., 0] = 200
    (tmp_path / "sky.hdr").write_bytes(write_hdr(rgbe, exposure=1.0))
    (tmp_path / "mesh.obj").write_text(OBJ)
    (tmp_path / "scene.rt").write_text(
        "tex sky image sky.hdr\n"
        "A 0.5 sky\n"
        "C 0,1,-5 0,0,1 60\n"
        "mat m 200,200,200 0 0 0 0 1 0 0,0,0\n"
        "cu 0,0,0 0,0,1 1 mesh.obj m\n"
        "sp 0,0,3 1 255,0,0\n")
    hs = M.HostScene.parse_file(str(tmp_path / "scene.rt"))
    v = hs.view().contents
    assert v.n_triangles == 5 and v.n_spheres == 1
    assert v.ambient_ratio == pytest.approx(0.5)
    rgb, ex = hs.hdri()
    assert rgb.shape == (4, 16, 3) and ex == pytest.approx(1.0)
    np.testing.assert_allclose(rgb[..., 0], 200 / 255.0, rtol=1e-6)


def test_rt_skybox_gamma_is_baked(tmp_path):
    rgbe = np.full((2, 8, 4), 128, np.uint8)
    rgbe[..., :3] = 100
    (tmp_path / "g.hdr").write_bytes(write_hdr(rgbe, exposure=0.0, gamma=2.0))
    (tmp_path / "s.rt").write_text("tex g image g.hdr\nA 0.8 g\n"
                                   "C 0,0,0 0,0,1 60\n")
    hs = M.HostScene.parse_file(str(tmp_path / "s.rt"))
    rgb, ex = hs.hdri()
    # pow(c / 255 * 2^(0 + 0.8 - 1), 1 / 2), then a lookup scale of 1
    np.testing.assert_allclose(rgb, np.sqrt(100 / 255 * 2 ** -0.2), rtol=1e-6)
    assert 2 ** (ex + 0.8 - 1) == pytest.approx(1.0)


def test_rt_texture_errors(tmp_path):
    (tmp_path / "a.rt").write_text("tex t checker_local 1 0,0,0 255,255,255\n"
                                   "A 0.5 t\nC 0,0,0 0,0,1 60\n")
    with pytest.raises(M.MrtError) as e:
        M.HostScene.parse_file(str(tmp_path / "a.rt"))
    assert e.value.code == M.MRT_E_UNSUPPORTED
    (tmp_path / "b.rt").write_text("A 0.5 nosuch\nC 0,0,0 0,0,1 60\n")
    with pytest.raises(M.MrtError):
        M.HostScene.parse_file(str(tmp_path / "b.rt"))
