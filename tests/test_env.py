"""HDRI environment: reference importance-sampling term and the MIS
estimator, on the oracle (CPU) and the HIP path (GPU)."""
import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O

W, H = 48, 48
SEED = 31


def sphere_in_env(albedo, env, ratio=1.0):
    hs = M.HostScene.empty()
    hs.set_resolution(W, H)
    m = hs.add_material(color=(albedo, albedo, albedo))
    hs.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    hs.set_ambient(ratio, (255, 255, 255))
    hs.set_camera((0.0, 0.0, -4.0), (0.0, 0.0, 1.0), fov=40.0)
    return hs


def oracle_render(hs, env, spp, est, f64=True, depth=8, first=0):
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    hd = O.make_hdri(env)
    img, _ = O.render(view, bvh, W, H, first, spp, depth, SEED, f64=f64,
                      traversal=O.TRAV_REFERENCE if f64 else O.TRAV_ORDERED,
                      estimator=est, hdri=hd)
    return img / spp


def sphere_mask(hs):
    view = hs.view()
    o, d = O.raygen_f32(view, W, H, 0, SEED)
    p, _, _, _ = O.trace_rays(view, None, O.TRAV_BRUTE, o, d)
    on = (p == M.PRIM_SPHERE).reshape(H, W)
    inner = on & np.roll(on, 1, 0) & np.roll(on, -1, 0) & \
        np.roll(on, 1, 1) & np.roll(on, -1, 1)
    return on, inner


def const_env(value, w=32, h=16):
    return np.full((h, w, 3), value, np.float32)


def test_mis_tables_match_oracle():
    sky = M.hdri_sky(3, 64, 32, sun_radius_deg=4.0, sun_radiance=500.0)
    hs = M.HostScene.empty()
    mm, mc, mp, tot = O.hdri_mis_tables(sky)
    import ctypes as C
    t = M.HdriTables  # noqa: F841 (keep import of the harness symbols)
    mis = _product_mis(sky)
    np.testing.assert_array_equal(mis[0], mm)
    np.testing.assert_array_equal(mis[1], mc)
    np.testing.assert_array_equal(mis[2], mp)
    assert abs(mis[3] - tot) <= 1e-12 * tot
    # a proper density: cells sum to 1, CDFs end at 1
    assert abs(mp.sum(dtype=np.float64) - 1.0) < 1e-5
    assert mm[-1] == 1.0 and np.all(mc[:, -1] == 1.0)
    del hs, C


def _product_mis(rgb):
    import ctypes as C

    class Mis(C.Structure):
        _fields_ = [("w", C.c_uint32), ("h", C.c_uint32),
                    ("marg", C.POINTER(C.c_float)),
                    ("cond", C.POINTER(C.c_float)),
                    ("pcell", C.POINTER(C.c_float)), ("total", C.c_double)]
    lib = M._lib
    lib.mrt_hdri_build_mis.argtypes = [C.POINTER(C.c_float), C.c_uint32,
                                       C.c_uint32, C.c_float, C.POINTER(Mis)]
    lib.mrt_hdri_free_mis.argtypes = [C.POINTER(Mis)]
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    m = Mis()
    assert lib.mrt_hdri_build_mis(rgb.ctypes.data_as(C.POINTER(C.c_float)),
                                  w, h, 0.0, C.byref(m)) == 0
    out = (np.ctypeslib.as_array(m.marg, shape=(h,)).copy(),
           np.ctypeslib.as_array(m.cond, shape=((h - 1) * w,)).copy()
           .reshape(h - 1, w),
           np.ctypeslib.as_array(m.pcell, shape=((h - 1) * (w - 1),)).copy()
           .reshape(h - 1, w - 1), m.total)
    lib.mrt_hdri_free_mis(C.byref(m))
    return out


@pytest.mark.parametrize("f64", [True, False])
def test_reference_term_constant_env_closed_form(f64):
    """With pdf_joint = 1 (constant map) the reference's term adds
    a L cos and scales the throughput by (1 - cos); the escape then adds
    a (1 - cos) L: every sphere pixel is a L exactly, visible or not."""
    a, L = 0.7, 0.8
    hs = sphere_in_env(a, None)
    img = oracle_render(hs, const_env(L), 3, M.EST_REFERENCE, f64=f64)
    on, inner = sphere_mask(hs)
    np.testing.assert_allclose(img[inner], a * L, rtol=2e-6)
    near = on | np.roll(on, 1, 0) | np.roll(on, -1, 0) | np.roll(on, 1, 1) \
        | np.roll(on, -1, 1)
    np.testing.assert_allclose(img[~near], L, rtol=1e-6)


def test_mis_white_furnace_oracle():
    """A white lambert sphere in a uniform environment of radiance 1
    reflects exactly 1 (white furnace); the MIS estimator is unbiased."""
    hs = sphere_in_env(1.0, None)
    img = oracle_render(hs, const_env(1.0), 64, M.EST_MIS)
    _, inner = sphere_mask(hs)
    assert abs(img[inner].mean() - 1.0) < 0.01


def env_radiance_integral(env, n, sub=4):
    """(1/pi) * integral of L(w) max(n.w, 0) dw over the sphere, by midpoint
    quadrature on sub x sub points per texel cell, with the texel lookup of
    background.c:34-57 (cell (y, x) spans v in [y, y+1)/(h-1), u likewise)."""
    h, w = env.shape[:2]
    nh, nw = h - 1, w - 1
    f = (np.arange(sub) + 0.5) / sub
    v = ((np.arange(nh)[:, None] + f[None, :]) / nh).ravel()
    u = ((np.arange(nw)[:, None] + f[None, :]) / nw).ravel()
    el = (v - 0.5) * np.pi
    ph = (0.5 - u) * 2 * np.pi
    dv = 1.0 / (nh * sub)
    du = 1.0 / (nw * sub)
    E, P = np.meshgrid(el, ph, indexing="ij")
    d = np.stack([np.cos(E) * np.cos(P), np.sin(E), np.cos(E) * np.sin(P)], -1)
    dw = np.cos(E) * (np.pi * dv) * (2 * np.pi * du)
    lum = env[np.repeat(np.arange(nh), sub)][:, np.repeat(np.arange(nw), sub)]
    out = []
    for nn in n:
        c = np.maximum(d @ nn, 0.0)
        out.append((lum * (c * dw)[..., None]).sum((0, 1)) / np.pi)
    return np.array(out)


def test_mis_converges_to_the_irradiance_integral():
    """Convex lambert sphere under a sky with a sun: pixel = a/pi * integral
    of L cos over the hemisphere (one bounce; every bounce ray escapes).
    The MIS estimate over many samples matches direct quadrature."""
    a = 0.5
    sky = M.hdri_sky(3, 48, 24, sun_radius_deg=6.0, sun_radiance=40.0)
    hs = sphere_in_env(a, None)
    spp = 2048
    rows = (20, 28)
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    img, _ = O.render(view, bvh, W, H, 0, spp, 2, SEED, f64=True,
                      traversal=O.TRAV_REFERENCE, estimator=M.EST_MIS,
                      hdri=O.make_hdri(sky), rows=rows)
    img /= spp
    _, inner = sphere_mask(hs)
    ys, xs = np.nonzero(inner[rows[0]:rows[1]])
    ys = ys + rows[0]
    pick = np.linspace(0, len(ys) - 1, 12).astype(int)
    # normals at the pixel centres: the sphere seen from (0, 0, -4)
    o, d = O.raygen_f32(view, W, H, 0, SEED)
    idx = ys[pick] * W + xs[pick]
    oo, dd = o[idx].astype(np.float64), d[idx].astype(np.float64)
    b = (oo * dd).sum(1)
    t = -b - np.sqrt(b * b - ((oo * oo).sum(1) - 1.0))
    nrm = oo + dd * t[:, None]
    expect = a * env_radiance_integral(sky, nrm)
    got = img[ys[pick], xs[pick]]
    # per-pixel jitter moves the normal within the pixel: compare the mean
    rel = np.abs(got.mean(0) - expect.mean(0)) / expect.mean(0)
    assert np.all(rel < 0.03), (got.mean(0), expect.mean(0))


@pytest.mark.gpu
@pytest.mark.parametrize("est", [M.EST_REFERENCE, M.EST_MIS])
def test_gpu_env_matches_oracle_same_stream(est):
    sky = M.hdri_sky(3, 128, 64, sun_radius_deg=3.0, sun_radiance=200.0)
    hs = M.HostScene.plain(7, W, H)
    view = hs.view()
    with M.Renderer(W, H) as r:
        r.upload(hs)
        M.mrt_upload_hdri(r._ctx, np.ascontiguousarray(sky).ctypes.data_as(
            M._fp), 128, 64, 0.0)
        r.set_seed(SEED)
        r.set_option(M.OPT_ESTIMATOR, est)
        r.render(0, 4, 8)
        img = r.readback().astype(np.float64) / 4
    ref = oracle_render(hs, sky, 4, est, f64=False)
    diff = np.abs(img - ref)
    close = diff <= 1e-4 + 1e-4 * np.abs(ref)
    print(f"est {est}: close {close.mean():.6f} mean diff {diff.mean():.3e}")
    assert close.mean() >= 0.99
    del view


@pytest.mark.gpu
def test_gpu_reference_term_constant_env():
    a, L = 0.7, 0.8
    hs = sphere_in_env(a, None)
    env = const_env(L)
    with M.Renderer(W, H) as r:
        r.upload(hs)
        r.set_seed(SEED)
        assert M.mrt_upload_hdri(r._ctx, env.ctypes.data_as(M._fp), 32, 16,
                                 0.0) == 0
        r.render(0, 3, 8)
        img = r.readback() / 3
    _, inner = sphere_mask(hs)
    np.testing.assert_allclose(img[inner], a * L, rtol=1e-5)


@pytest.mark.gpu
def test_gpu_mis_white_furnace():
    hs = sphere_in_env(1.0, None)
    env = const_env(1.0)
    with M.Renderer(W, H) as r:
        r.upload(hs)
        r.set_seed(SEED)
        r.set_option(M.OPT_ESTIMATOR, M.EST_MIS)
        assert M.mrt_upload_hdri(r._ctx, env.ctypes.data_as(M._fp), 32, 16,
                                 0.0) == 0
        r.render(0, 256, 8)
        img = r.readback() / 256
    _, inner = sphere_mask(hs)
    assert abs(img[inner].mean() - 1.0) < 0.005


def _env_maps():
    rng = np.random.default_rng(9)
    rand = (rng.random((37, 53, 3)) ** 4 * 30).astype(np.float32)
    rand[5] = 0.0                                   # a black row: sum 0
    rand[10:12, 7] = 4e4                            # a bright spot
    return [(M.hdri_sky(5, 2048, 1024), 0.0),       # the HDRI workload's size
            (rand, 0.0), (rand, 1.5), (M.hdri_sky(2, 64, 2), 0.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(4))
def test_gpu_hdri_tables_equal_host_tables_bitwise(case):
    """The device build of both table sets is bit for bit the host's
    (mrt_hdri_build_tables) and the oracle's (reference and MIS tables)."""
    rgb, ex = _env_maps()[case]
    h, w = rgb.shape[:2]
    with M.Renderer(8, 8) as r:
        r.upload_hdri(rgb, ex)
        pdf, marg, cond, mm, mc, mp = r.hdri_tables()
    hp, hm, hc, _ = M.hdri_tables(rgb, ex)
    np.testing.assert_array_equal(pdf, hp)
    np.testing.assert_array_equal(marg, hm)
    np.testing.assert_array_equal(cond, hc)
    op, om, oc, _ = O.hdri_tables(rgb, ex)
    np.testing.assert_array_equal(cond, oc)
    om_, oc_, opc, _ = O.hdri_mis_tables(rgb, ex)
    np.testing.assert_array_equal(mm, om_)
    np.testing.assert_array_equal(mc, oc_)
    np.testing.assert_array_equal(mp, opc)


def _cdf_search(cdf, n, u):
    """The kernels' cdf_search: the last index in [0, n-1] whose cdf entry
    is <= u (0 if none), by the same halving."""
    lo, hi = 0, n - 1
    while lo < hi:
        mid = (lo + hi + 1) >> 1
        if cdf[mid] <= u:
            lo = mid
        else:
            hi = mid - 1
    return lo


def test_guided_cdf_search_equals_the_binary_search():
    """cdf_search_guided's scheme (a guide g[k] = cdf_search(k / G), G a power
    of two, then the search within [g[k], g[k + 1]]) returns cdf_search's
    index for every u: random CDFs with flat runs (zero-pdf cells), u on the
    guide boundaries k / G, just below them, and uniform draws on the
    kernels' 2^-24 grid."""
    rng = np.random.default_rng(7)
    for n, G in ((1023, 1024), (2047, 2048), (5, 8), (1, 1), (300, 512)):
        w = rng.exponential(size=n) * (rng.random(n) > 0.3)
        w[0] = 0.0
        cdf = np.concatenate([[0.0], np.cumsum(w[1:])]).astype(np.float32)
        cdf = np.concatenate([cdf / max(cdf[-1], 1e-30), [1.0]]).astype(np.float32)
        guide = [_cdf_search(cdf, n, np.float32(k) / np.float32(G))
                 for k in range(G)] + [n - 1]
        us = list(np.arange(G, dtype=np.float32) / np.float32(G))
        us += [np.nextafter(u, np.float32(0)) for u in us[1:]]
        us += list((rng.integers(0, 1 << 24, 4000) * 2.0 ** -24).astype(np.float32))
        for u in us:
            u = np.float32(u)
            k = min(int(u * np.float32(G)), G - 1)
            lo, hi = guide[k], guide[k + 1]
            while lo < hi:
                mid = (lo + hi + 1) >> 1
                if cdf[mid] <= u:
                    lo = mid
                else:
                    hi = mid - 1
            assert lo == _cdf_search(cdf, n, u), (n, G, float(u))
