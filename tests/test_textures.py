"""Textures: BMP / PNG decoding as the reference's readers lay texels out
(src/utils/bmp, src/utils/png), the .rt "tex" element and textured
materials (parse_texture, parse_material), (u, v) of every primitive, the
checker and image lookups (get_tex_color.c, get_solid_texture.c), the alpha
test (hit_register.c:57-69) and normal maps (apply_map.c) in the oracle,
pinned by hand-worked values; the HIP path against the oracle is in
test_gpu_parity.py."""
import struct
import zlib

import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O


# ---- encoders for fixtures ---------------------------------------------------

def png_bytes(img, ctype, filters=(0,)):
    """8-bit PNG of img (h, w, channels), rows top first; scanline y uses
    filter filters[y % len(filters)]."""
    h, w = img.shape[:2]
    ch = img.shape[2] if img.ndim == 3 else 1
    raw = img.reshape(h, w * ch).astype(np.int32)
    out = bytearray()
    for y in range(h):
        f = filters[y % len(filters)]
        line = raw[y]
        prev = raw[y - 1] if y else np.zeros_like(line)
        enc = np.zeros_like(line)
        for x in range(w * ch):
            a = line[x - ch] if x >= ch else 0
            b = prev[x]
            c = prev[x - ch] if x >= ch else 0
            if f == 0:
                pr = 0
            elif f == 1:
                pr = a
            elif f == 2:
                pr = b
            elif f == 3:
                pr = (a + b) >> 1
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pr = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            enc[x] = (line[x] - pr) & 255
        out += bytes([f]) + bytes(enc.astype(np.uint8))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + \
            struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    hdr = struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", hdr)
            + chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b""))


def bmp_bytes(pixels, bpp, palette=None):
    """BMP with rows as given (file order); pixels (h, w, 4) RGBA, or
    (h, w) palette indices for bpp 8."""
    h, w = pixels.shape[:2]
    rows = []
    for y in range(h):
        if bpp == 8:
            r = bytes(pixels[y].astype(np.uint8))
        elif bpp == 16:
            p = pixels[y].astype(np.int32)
            v = ((p[:, 0] >> 3) << 11) | ((p[:, 1] >> 2) << 5) | (p[:, 2] >> 3)
            r = v.astype("<u2").tobytes()
        elif bpp == 24:
            r = pixels[y][:, [2, 1, 0]].astype(np.uint8).tobytes()
        else:
            r = pixels[y][:, [2, 1, 0, 3]].astype(np.uint8).tobytes()
        rows.append(r + b"\0" * ((4 - len(r) % 4) % 4))
    pal = b""
    if bpp == 8:
        pal = b"".join(bytes([c[2], c[1], c[0], 0]) for c in palette)
    off = 14 + 40 + len(pal)
    body = b"".join(rows)
    info = struct.pack("<IiiHHIIiiII", 40, w, h, 1, bpp, 0, len(body), 2835,
                       2835, len(palette) if bpp == 8 else 0, 0)
    return b"BM" + struct.pack("<IHHI", off + len(body), 0, 0, off) + info + \
        pal + body


# ---- decoders ----------------------------------------------------------------

@pytest.mark.parametrize("filters", [(0,), (1,), (2,), (3,), (4,), (0, 1, 2, 3, 4)])
def test_png_rgba_all_filters_last_scanline_first(filters):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (7, 5, 4), dtype=np.uint8)
    got = M.decode_image(png_bytes(img, 6, filters))
    np.testing.assert_array_equal(got, img[::-1])


def test_png_colour_types_follow_get_pixel_rgva():
    rng = np.random.default_rng(2)
    g = rng.integers(0, 256, (3, 4), dtype=np.uint8)
    got = M.decode_image(png_bytes(g[..., None], 0))
    np.testing.assert_array_equal(got[..., 0], g[::-1])
    assert (got[..., 3] == 255).all()
    rgb = rng.integers(0, 256, (3, 4, 3), dtype=np.uint8)
    got = M.decode_image(png_bytes(rgb, 2, (4,)))
    np.testing.assert_array_equal(got[..., :3], rgb[::-1])
    # palette images: the index is read as grey
    got = M.decode_image(png_bytes(g[..., None], 3))
    np.testing.assert_array_equal(got[..., 1], g[::-1])
    # grey + alpha: grey from byte x of the row, alpha from byte 2x + 1
    ga = rng.integers(0, 256, (3, 4, 2), dtype=np.uint8)
    got = M.decode_image(png_bytes(ga, 4))
    row = ga.reshape(3, 8)[::-1]
    np.testing.assert_array_equal(got[..., 0], row[:, :4])
    np.testing.assert_array_equal(got[..., 3], row[:, 1::2])


@pytest.mark.parametrize("bpp", [8, 16, 24, 32])
def test_bmp_rows_in_file_order(bpp):
    rng = np.random.default_rng(bpp)
    h, w = 3, 5                      # odd width: row padding
    if bpp == 8:
        pal = rng.integers(0, 256, (16, 3))
        idx = rng.integers(0, 16, (h, w))
        got = M.decode_image(bmp_bytes(idx, 8, pal))
        np.testing.assert_array_equal(got[..., :3], pal[idx])
        assert (got[..., 3] == 255).all()
        return
    px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    got = M.decode_image(bmp_bytes(px, bpp))
    if bpp == 16:
        want = px[..., :3] & np.array([0xF8, 0xFC, 0xF8], np.uint8)
        np.testing.assert_array_equal(got[..., :3], want)
    else:
        np.testing.assert_array_equal(got[..., :3], px[..., :3])
    np.testing.assert_array_equal(got[..., 3], px[..., 3] if bpp == 32 else 255)


@pytest.mark.parametrize("data", [b"GIF89a", b"BM" + b"\0" * 60,
                                  b"\x89PNG\r\n\x1a\n" + b"\0" * 30])
def test_bad_images_rejected(data):
    with pytest.raises(M.MrtError):
        M.decode_image(data)


# ---- oracle lookups ------------------------------------------------------------

def textured_plane(kind, scale, c1, c2, mscale=1.0):
    hs = M.HostScene.empty()
    m = hs.add_material(color=(0.5, 0.5, 0.5))
    t = hs.add_checker_texture(kind, scale, c1, c2)
    hs.set_material_tex(m, color=t, scale=mscale)
    hs.add_plane((0, 0, 0), (0, 1, 0), m)
    return hs


def albedo_at(hs, pts, f64=False):
    """Textured colour where vertical rays from above meet the scene."""
    o = np.array([[x, 5.0, z] for x, z in pts], np.float32)
    d = np.tile(np.array([0, -1, 0], np.float32), (len(pts), 1))
    return O.hit_info(hs.view(), None, O.TRAV_BRUTE, o, d, f64=f64,
                      albedo=True)[5]


@pytest.mark.parametrize("f64", [False, True])
def test_checker_global_cells(f64):
    hs = textured_plane(M.TEX_CHECKER_GLOBAL, 2.0, (255, 0, 0), (0, 0, 255))
    # y = 0: floor(2 * 0 + 1/2) = 0; parity of floor(2x) + floor(2z)
    pts = [(0.1, 0.1), (0.6, 0.1), (0.6, 0.6), (-0.1, 0.1), (-0.6, -0.6)]
    want = [(1, 0, 0), (0, 0, 1), (1, 0, 0), (0, 0, 1), (1, 0, 0)]
    np.testing.assert_allclose(albedo_at(hs, pts, f64), want, atol=1e-6)


@pytest.mark.parametrize("f64", [False, True])
def test_checker_local_on_plane_uv(f64):
    """get_uv_plane: normal +Y, so tangent = unit(cross(+X, +Y)) = +Z and
    the second axis cross(+Y, +Z) = +X; u = z * 0.1 * scale, v = x * 0.1 *
    scale, each minus its floor."""
    hs = textured_plane(M.TEX_CHECKER_LOCAL, 4.0, (255, 255, 255), (0, 0, 0),
                        mscale=10.0)
    # u = frac(z), v = frac(x); cell parity floor(4u) + floor(4v)
    pts = [(0.1, 0.1), (0.3, 0.1), (0.3, 0.3), (1.1, 0.1), (-0.1, 0.1)]
    want = [1, 0, 1, 1, 0]      # -0.1 -> v = 0.9: floor(3.6) = 3 (odd)
    got = albedo_at(hs, pts, f64)[:, 0]
    np.testing.assert_allclose(got, want, atol=1e-6)


@pytest.mark.parametrize("f64", [False, True])
def test_sphere_uv_and_image_lookup(f64):
    """get_sphere_uv: u = (atan2((c.z - p.z) / r, (p.x - c.x) / r) + pi) /
    2 pi, v = acos((c.y - p.y) / r) / pi; a 4 x 2 image is read at
    ((int)(3u), (int)(v))."""
    img = np.zeros((2, 4, 4), np.uint8)
    for y in range(2):
        for x in range(4):
            img[y, x] = (40 * x + 10, 100 * y + 5, 7, 255)
    hs = M.HostScene.empty()
    m = hs.add_material()
    hs.set_material_tex(m, color=hs.add_image_texture(img))
    hs.add_sphere((0, 0, 0), 1.0, m)
    # rays toward the centre hit p = -dir
    dirs = np.array([(-1, 0, 0), (1, 0, 0), (0, 0, -1), (0, -1, 0)], np.float32)
    o = -3 * dirs
    _, _, _, _, _, alb = O.hit_info(hs.view(), None, O.TRAV_BRUTE, o, dirs,
                                    f64=f64, albedo=True)
    # p = (+1,0,0): u = 1/2, v = 1/2 -> texel (1, 0)
    # p = (-1,0,0): u = 1 -> (3, 0); p = (0,0,1): atan2(-1, 0) = -pi/2 ->
    # u = 1/4 -> (0, 0); p = (0,1,0): v = acos(-1)/pi = 1 -> row 1
    want = img[[0, 0, 0, 1], [1, 3, 0, 1], :3] / 255.0
    np.testing.assert_allclose(alb, want, atol=1e-6)


def quad_scene(alpha_img, uv=((0, 0), (1, 0), (1, 1), (0, 1))):
    """A textured unit quad at z = 0 (two triangles with (u, v)) in front of
    a backing plane at z = 1."""
    hs = M.HostScene.empty()
    m = hs.add_material(color=(1, 1, 1))
    back = hs.add_material(color=(0.2, 0.2, 0.2))
    hs.set_material_tex(m, color=hs.add_image_texture(alpha_img))
    v = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
    for tri in ((0, 1, 2), (0, 2, 3)):
        k = hs.add_triangle([v[i] for i in tri], m)
        hs.set_triangle_uv(k, [c for i in tri for c in uv[i]])
    hs.add_plane((0, 0, 1), (0, 0, -1), back)
    return hs


def test_alpha_test_and_triangle_uv():
    """texel ((int)(2u), (int)(2v)) of a 3 x 3 image with u = x, v = y on
    the quad: column 0 transparent, column 1 opaque in row 0 and of alpha
    128 in row 1."""
    img = np.full((3, 3, 4), 255, np.uint8)
    img[:, 0, 3] = 0
    img[1, 1, 3] = 128
    hs = quad_scene(img)
    view = hs.view()
    n = 6000
    rng = np.random.default_rng(3)
    xy = rng.uniform(0.01, 0.99, (n, 2))
    xy = xy[(np.abs(xy - 0.5) > 1e-3).all(1)]
    o = np.c_[xy, -np.ones(len(xy))].astype(np.float32)
    d = np.tile(np.array([0, 0, 1], np.float32), (len(xy), 1))
    pid, *_ = O.hit_info(view, None, O.TRAV_BRUTE, o, d)
    bvh = O.make_bvh(M.BuiltBvh(view))
    pid2, *_ = O.hit_info(view, bvh, O.TRAV_ORDERED, o, d)
    assert np.array_equal(pid, pid2)          # the draw is order free
    on_quad = pid != M.PRIM_PLANE
    left, low = xy[:, 0] < 0.5, xy[:, 1] < 0.5
    assert not on_quad[left].any()            # alpha 0: always skipped
    assert on_quad[~left & low].all()         # alpha 255: never
    frac = on_quad[~left & ~low].mean()       # alpha 128: kept 128/255
    assert abs(frac - 128 / 255) < 0.05


def test_normal_map_tbn():
    """apply_normal_map: n = +Y gives tangent cross(n, -Y) = 0, so cross(n,
    +Z) = +X, bitangent unit(cross(n, tangent)) = -Z; a map texel (r, g, b)
    decodes to (r, g, b) / 127.5 - 1."""
    texel = np.array([[[200, 90, 180, 255]]], np.uint8)
    hs = M.HostScene.empty()
    m = hs.add_material()
    nt = hs.add_image_texture(texel)
    hs.set_material_tex(m, normal=nt)
    hs.add_plane((0, 0, 0), (0, 1, 0), m)
    o = np.array([[0.3, 2, 0.2]], np.float32)
    d = np.array([[0, -1, 0]], np.float32)
    nm = texel[0, 0, :3] / 127.5 - 1
    want = nm[0] * np.array([1, 0, 0]) + nm[1] * np.array([0, 0, -1]) \
        + nm[2] * np.array([0, 1, 0])
    want /= np.linalg.norm(want)
    for f64 in (False, True):
        n = O.hit_info(hs.view(), None, O.TRAV_BRUTE, o, d, f64=f64)[2]
        np.testing.assert_allclose(n[0], want, atol=1e-6)
    # intensity 0.5: unit(n/2 + mapped/2)
    hs.set_material_tex(m, normal=nt, normal_intensity=0.5)
    n = O.hit_info(hs.view(), None, O.TRAV_BRUTE, o, d)[2]
    half = (np.array([0, 1, 0]) + want) / 2
    np.testing.assert_allclose(n[0], half / np.linalg.norm(half), atol=1e-6)


# ---- .rt front end -------------------------------------------------------------

def test_rt_textured_materials(tmp_path):
    img = np.zeros((2, 3, 3), np.uint8)
    img[..., 0] = [[10, 20, 30], [40, 50, 60]]
    (tmp_path / "c.png").write_bytes(png_bytes(img, 2))
    px = np.full((2, 2, 4), 255, np.uint8)
    (tmp_path / "r.bmp").write_bytes(bmp_bytes(px, 24))
    (tmp_path / "s.rt").write_text(
        "A 0.2 255,255,255\nC 0,1,-5 0,0,1 60\n"
        "tex col image c.png\n"
        "tex rough image r.bmp\n"
        "tex chk checker_global 3 255,0,0 0,255,0\n"
        "tex missing image nosuch.png\n"
        "mat a col 0 rough 1.5 0 1 0 chk 2.5\n"
        "mat b 100,100,100 0.5 0.1 0 0 missing 0 0,0,0 1 col 0.7\n"
        "sp 0,0,3 2 a\npl 0,-1,0 0,1,0 b\n")
    hs = M.HostScene.parse_file(str(tmp_path / "s.rt"))
    v = hs.view().contents
    assert v.n_textures == 4
    tex = [v.textures[i] for i in range(4)]
    kinds = [t.type for t in tex]
    # textures enter the scene in order of first use
    assert kinds == [M.TEX_IMAGE, M.TEX_IMAGE, M.TEX_CHECKER_GLOBAL,
                     M.TEX_CHECKER_GLOBAL]
    assert (tex[0].width, tex[0].height) == (3, 2)
    # an image that does not load is get_solid_texture_default's pattern
    assert tex[3].scale == 2.0 and list(tex[3].c2) == [243, 0, 243]
    assert list(tex[2].c1) == [255, 0, 0] and tex[2].scale == 3.0
    mt = [v.material_tex[i] for i in range(v.n_materials)]
    assert (mt[0].color, mt[0].roughness, mt[0].emission_color) == (0, 1, 2)
    assert mt[0].metallic == M.NO_TEX and mt[0].scale == 2.5
    assert (mt[1].ao, mt[1].normal) == (3, 0)
    assert abs(mt[1].normal_intensity - 0.7) < 1e-7
    texels = np.ctypeslib.as_array(v.texels, shape=(v.n_texels * 4,))
    assert list(texels[:12:4]) == [40, 50, 60]        # last scanline first


def test_obj_texture_coordinates():
    obj = ("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
           "vt 0.1 0.2\nvt 0.9 0.2\nvt 0.9 0.8\nvt 0.1 0.8\n"
           "f 1/1 2/2 3/3 4/4\n")
    hs = M.HostScene.empty()
    m = hs.add_material()
    assert hs.add_obj_text(obj, (0, 0, 0), (0, 0, 1), 1.0, m) == 2
    uv = np.ctypeslib.as_array(hs.view().contents.tri_uv, shape=(12,))
    # (v3, v0, v1) then (v1, v2, v3)
    np.testing.assert_allclose(uv, [0.1, 0.8, 0.1, 0.2, 0.9, 0.2,
                                    0.9, 0.2, 0.9, 0.8, 0.1, 0.8], rtol=1e-6)
