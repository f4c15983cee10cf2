"""Caustic photon map: emission, storage rule, gather (k-NN and fixed
radius), and their use in the path tracer.

The oracle is pinned by an independent k-d tree (scipy) for the gather rule
of src/caustic/get_caustic.c:37-88, by the closed-form geometry and power of
a point light above a mirror (src/caustic/trace_photon_path.c:18-75,
caustic_manager.c:20-41), and by an independent Monte-Carlo integral of the
flux an environment delivers through a mirror sphere.  The HIP path is then
compared with the oracle through the C ABI."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import mrt_hip as M
import oracle_ffi as O

SEED = 0x00C0FFEE


# ---- gather rule ----------------------------------------------------------

def kdtree_caustic(pos, pw, pts, k, radius, norm):
    """get_caustic.c:37-88 with scipy's k-d tree: the k nearest photons; no
    estimate unless there are k of them and the farthest lies within
    SQUARED distance `radius`; k == 0: all photons within distance radius."""
    tree = cKDTree(pos.astype(np.float64))
    out = np.zeros((len(pts), 3))
    if k == 0:
        area = np.pi * radius * radius
        for i, idx in enumerate(tree.query_ball_point(pts, radius)):
            out[i] = pw[idx].astype(np.float64).sum(0) / area / norm
        return out
    d, idx = tree.query(pts, k=k)
    d = d.reshape(len(pts), k)
    idx = idx.reshape(len(pts), k)
    for i in range(len(pts)):
        r2 = d[i, -1] ** 2
        if not np.isfinite(r2) or r2 > radius or np.pi * r2 < 1e-9:
            continue
        out[i] = pw[idx[i]].astype(np.float64).sum(0) / (np.pi * r2) / norm
    return out


def random_photons(rng, n):
    pos = rng.random((n, 3)).astype(np.float32)
    pos[:, 1] *= 0.05                       # a thin slab, like a floor
    pw = (rng.random((n, 3)) * 0.01).astype(np.float32)
    return pos, pw


@pytest.mark.parametrize("k,radius", [(15, 0.004), (1, 0.01), (8, 1.2e-3),
                                      (0, 0.05), (0, 0.2)])
def test_oracle_gather_matches_kdtree(k, radius):
    rng = np.random.default_rng(11 + k)
    pos, pw = random_photons(rng, 3000)
    pts = rng.random((300, 3)).astype(np.float32)
    pts[:, 1] *= 0.05
    pm = O.photon_map(pos, pw, k=k, radius=radius, norm=2.5)
    got = O.caustic(pm, pts)
    ref = kdtree_caustic(pos, pw, pts, k, radius, 2.5)
    nz = (ref != 0).any(1)
    assert 0.2 < nz.mean() <= 1.0, nz.mean()
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-9)


def test_oracle_gather_needs_k_photons():
    pos = np.array([[0, 0, 0], [0.01, 0, 0]], np.float32)
    pw = np.ones((2, 3), np.float32)
    pm = O.photon_map(pos, pw, k=3, radius=1.0)
    assert not O.caustic(pm, np.zeros((1, 3), np.float32)).any()
    pm = O.photon_map(pos, pw, k=2, radius=1.0)
    got = O.caustic(pm, np.zeros((1, 3), np.float32))
    np.testing.assert_allclose(got[0], 2.0 / (np.pi * 1e-4), rtol=1e-4)


# ---- hash grid ------------------------------------------------------------

def hash_cell_np(ix, iy, iz, table):
    """hash_cell (src/gpu/convert_caustic_utils.c:15-24): int cell indices
    converted to uint32, multiplied with wrap-around, xor-ed, mod table."""
    m = np.uint64(0xFFFFFFFF)
    u = lambda a, p: (np.asarray(a, np.int64).astype(np.uint64) * np.uint64(p)) & m
    n = u(ix, 73856093) ^ u(iy, 19349663) ^ u(iz, 83492791)
    return (n % np.uint64(table)).astype(np.uint32)


@pytest.mark.parametrize("cell,table,expect", [
    ((0, 0, 0), 1000, 0), ((1, 0, 0), 1000, 93), ((0, 1, 0), 1000, 663),
    ((0, 0, 1), 1000, 791), ((1, 1, 1), 1000, 157),
    ((-1, 0, 0), 1 << 31, 4221111203 - (1 << 31)),
    ((1, 0, 0), 1 << 31, 73856093), ((2, 0, 0), 7, (2 * 73856093) % 7)])
def test_hash_cell_known_answers(cell, table, expect):
    assert int(hash_cell_np(*cell, table)) == expect


# ---- emission: point light above a mirror ---------------------------------

LIGHT_Y, MIRROR_Y = 1.0, 2.0
FLOOR = (0.5, 0.6, 0.7)
LCOL, LBRIGHT = (255.0, 128.0, 64.0), 0.5


def mirror_scene():
    """Diffuse floor y=0, white mirror plane y=2 facing down, point light at
    (0,1,0): a photon is stored iff it leaves upwards (one specular event,
    then the floor); the stored points are hit by lines from the light's
    mirror image (0, 2*MIRROR_Y - LIGHT_Y, 0)."""
    hs = M.HostScene.empty()
    hs.set_resolution(32, 32)
    floor = hs.add_material(color=FLOOR)
    mirror = hs.add_material(color=(1.0, 1.0, 1.0), metallic=1.0,
                             roughness=0.0)
    hs.add_plane((0, 0, 0), (0, 1, 0), floor)
    hs.add_plane((0, MIRROR_Y, 0), (0, -1, 0), mirror)
    hs.add_light((0, LIGHT_Y, 0), LBRIGHT, LCOL)
    hs.set_camera((0, 1, -4), (0, 0, 1), fov=60.0)
    return hs


def check_mirror_photons(pos, pw, gid, n):
    # about half the photons go up; those within ~1e-3 of the horizon reach
    # the floor beyond t = 1000 and are lost
    frac = len(pos) / n
    sig = 0.5 / np.sqrt(n)
    assert abs(frac - 0.5 * (1 - 1.5e-3)) < 5 * sig + 1e-3, frac
    assert np.all(np.diff(gid.astype(np.int64)) > 0)
    assert np.all(np.abs(pos[:, 1]) < 1e-4)
    expect = np.array(LCOL) * LBRIGHT / n * np.array(FLOOR)
    np.testing.assert_allclose(pw, np.broadcast_to(expect, pw.shape),
                               rtol=2e-6)
    # radial law from the image point at height hgt: P(rho < R | up) =
    # 1 - hgt / sqrt(hgt^2 + R^2) (uniform directions); KS distance
    hgt = 2 * MIRROR_Y - LIGHT_Y
    rho = np.sort(np.hypot(pos[:, 0], pos[:, 2]).astype(np.float64))
    cdf = 1 - hgt / np.sqrt(hgt * hgt + rho * rho)
    emp = np.arange(1, len(rho) + 1) / len(rho)
    ks = np.max(np.abs(emp - cdf))
    assert ks < 1.8 / np.sqrt(len(rho)) + 2e-3, ks
    # azimuth uniform
    az = np.arctan2(pos[:, 2], pos[:, 0])
    hist, _ = np.histogram(az, bins=8, range=(-np.pi, np.pi))
    e = len(az) / 8
    assert np.all(np.abs(hist - e) < 5 * np.sqrt(e)), hist


def _bvh(hs):
    view = hs.view()
    return view, O.make_bvh(M.BuiltBvh(view))


@pytest.mark.parametrize("f64", [False, True])
def test_oracle_point_light_mirror_closed_form(f64):
    hs = mirror_scene()
    view, bvh = _bvh(hs)
    n = 40000
    pos, pw, gid = O.emit_photons(view, bvh, n, SEED, f64=f64)
    check_mirror_photons(pos, pw, gid, n)


# ---- emission: environment through a mirror sphere ------------------------

ENV_L = 0.75
SPH_C, SPH_R = np.array([0.0, 1.5, 0.0]), 1.0


def env_mirror_scene():
    hs = M.HostScene.empty()
    hs.set_resolution(32, 32)
    floor = hs.add_material(color=FLOOR)
    mirror = hs.add_material(color=(1.0, 1.0, 1.0), metallic=1.0,
                             roughness=0.0)
    hs.add_plane((0, 0, 0), (0, 1, 0), floor)
    hs.add_sphere(tuple(SPH_C), SPH_R, mirror)
    hs.set_ambient(1.0, (255, 255, 255))
    hs.set_camera((0, 1, -5), (0, 0, 1), fov=60.0)
    return hs


def env_flux_integral(n=1 << 22, seed=5):
    """Flux a constant environment L delivers to the floor through the
    mirror sphere: L * albedo * int_A int_{n.w>0} cos [w_y > 0][r_y < 0]
    dw dA, with x uniform on the sphere and w cosine-distributed (an
    estimator independent of the photon sampler)."""
    rng = np.random.default_rng(seed)
    z = rng.normal(size=(n, 3))
    nrm = z / np.linalg.norm(z, axis=1, keepdims=True)
    u1, u2 = rng.random(n), rng.random(n)
    r, ph = np.sqrt(u1), 2 * np.pi * u2
    loc = np.stack([r * np.cos(ph), r * np.sin(ph), np.sqrt(1 - u1)], 1)
    a = np.where(np.abs(nrm[:, :1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]])
    t1 = np.cross(a, nrm)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(nrm, t1)
    w = loc[:, :1] * t1 + loc[:, 1:2] * t2 + loc[:, 2:] * nrm
    refl = 2 * np.sum(w * nrm, 1, keepdims=True) * nrm - w
    ok = (w[:, 1] > 0) & (refl[:, 1] < 0)
    p = ok.mean()
    scale = 4 * np.pi * SPH_R ** 2 * np.pi
    return ENV_L * scale * p, ENV_L * scale * np.sqrt(p * (1 - p) / n)


def check_env_photons(pos, pw, n):
    """Sum of stored power (one channel) against the integral, to 5 sigma."""
    tot = pw[:, 0].astype(np.float64)
    est = tot.sum() / FLOOR[0]
    # per-photon contributions: zero for the unstored ones
    contrib = np.zeros(n)
    contrib[:len(tot)] = tot / FLOOR[0]
    sig = np.sqrt(n) * contrib.std()
    ref, ref_sig = env_flux_integral()
    assert np.all(np.abs(pos[:, 1]) < 1e-4)
    np.testing.assert_allclose(pw[:, 1] / pw[:, 0], FLOOR[1] / FLOOR[0],
                               rtol=1e-5)
    assert abs(est - ref) < 5 * np.hypot(sig, ref_sig), (est, ref, sig)
    return est, ref, sig


def test_oracle_env_photons_carry_the_env_flux():
    hs = env_mirror_scene()
    view, bvh = _bvh(hs)
    env = np.full((16, 32, 3), ENV_L, np.float32)
    hd = O.make_hdri(env)
    n = 200000
    pos, pw, gid = O.emit_photons(view, bvh, n, SEED, hdri=hd)
    assert len(pos) > 0.05 * n
    check_env_photons(pos, pw, n)


# ---- HIP path -------------------------------------------------------------

def _gpu_photons(hs, n, radius, env=None, k=15, norm=1.0):
    r = M.Renderer(*hs.resolution)
    r.upload(hs)
    if env is not None:
        r.upload_hdri(env)
    r.set_seed(SEED)
    r.build_photons(n, radius, k=k, norm=norm)
    return r


def _by_gid(pos, pw, ids):
    o = np.argsort(ids, kind="stable")
    return pos[o], pw[o], ids[o]


def _match(a, b):
    """Photons of the HIP path against the oracle's, matched by id."""
    pa, wa, ia = a
    pb, wb, ib = b
    common, ja, jb = np.intersect1d(ia, ib, return_indices=True)
    frac = len(common) / max(len(ia), len(ib))
    dp = np.abs(pa[ja] - pb[jb]).max(1)
    same = dp <= 1e-4 * (1 + np.abs(pb[jb]).max(1))
    return frac, same.mean(), wa[ja][same], wb[jb][same]


@pytest.mark.gpu
def test_gpu_point_light_mirror_closed_form():
    hs = mirror_scene()
    n = 40000
    r = _gpu_photons(hs, n, 0.01)
    pos, pw, ids = _by_gid(*r.photons())
    r.close()
    check_mirror_photons(pos, pw, ids, n)


@pytest.mark.gpu
def test_gpu_env_photons_carry_the_env_flux():
    hs = env_mirror_scene()
    env = np.full((16, 32, 3), ENV_L, np.float32)
    n = 200000
    r = _gpu_photons(hs, n, 0.01, env=env)
    pos, pw, ids = r.photons()
    r.close()
    check_env_photons(pos, pw, n)


@pytest.mark.gpu
@pytest.mark.parametrize("with_env", [False, True])
def test_gpu_photons_match_oracle(with_env):
    hs = M.HostScene.plain(7, 64, 36)
    view, bvh = _bvh(hs)
    env = M.hdri_sky(3, 128, 64, sun_radius_deg=3.0, sun_radiance=200.0) \
        if with_env else None
    n = 30000
    r = _gpu_photons(hs, n, 0.01, env=env)
    g = _by_gid(*r.photons())
    r.close()
    hd = O.make_hdri(env) if with_env else None
    o = O.emit_photons(view, bvh, n, SEED, hdri=hd,
                       n_emitters=2 if with_env else 1)
    frac, same, wa, wb = _match(g, o)
    print(f"env={with_env}: {len(g[0])} gpu / {len(o[0])} oracle photons, "
          f"matched {frac:.5f}, same position {same:.5f}")
    assert len(o[0]) > 100
    assert frac >= 0.995 and same >= 0.995
    np.testing.assert_allclose(wa, wb, rtol=1e-4, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("k,radius", [(15, 0.01), (4, 0.002), (0, 0.1)])
def test_gpu_gather_matches_oracle(k, radius):
    hs = M.HostScene.plain(7, 64, 36)
    r = _gpu_photons(hs, 50000, radius, k=k, norm=3.0)
    pos, pw, ids = r.photons()
    rng = np.random.default_rng(k)
    pick = rng.integers(0, len(pos), 2000)
    pts = pos[pick] + rng.normal(scale=0.03, size=(2000, 3)).astype(
        np.float32)
    pts = np.concatenate([pts, rng.random((200, 3)).astype(np.float32)])
    got = r.caustic_at(pts)
    r.close()
    pm = O.photon_map(pos, pw, k=k, radius=radius, norm=3.0)
    ref = O.caustic(pm, pts)
    nz = (ref != 0).any(1)
    assert nz.mean() > 0.3
    close = np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref) + 1e-7, 1)
    print(f"k={k}: nonzero {nz.mean():.3f}, close {close.mean():.5f}")
    # a tie at the k-th distance may pick another photon of equal distance
    assert close.mean() >= 0.999


@pytest.mark.gpu
def test_gpu_image_with_caustics_matches_oracle():
    W, H = 96, 54
    hs = M.HostScene.plain(7, W, H)
    view, bvh = _bvh(hs)
    r = _gpu_photons(hs, 60000, 0.02, k=15, norm=1.0)
    pos, pw, _ = r.photons()
    r.clear()
    r.render(0, 4, 8)
    img = r.readback().astype(np.float64)
    r.close()
    pm = O.photon_map(pos, pw, k=15, radius=0.02, norm=1.0)
    ref, _ = O.render(view, bvh, W, H, 0, 4, 8, SEED, photons=pm)
    base, _ = O.render(view, bvh, W, H, 0, 4, 8, SEED)
    # the caustic term is visible in the image
    assert np.abs(ref - base).max() > 1e-3
    diff = np.abs(img - ref)
    close = diff <= 1e-4 + 1e-4 * np.abs(ref)
    print(f"close {close.mean():.6f} mean diff {diff.mean():.3e}")
    assert close.mean() >= 0.995
    assert diff.mean() <= 1e-3 * max(ref.mean(), 1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("k,radius", [(15, 0.01), (0, 0.05)])
def test_gpu_map_order_is_hash_bucket_then_id(k, radius):
    """The map is sorted by (hash_cell of the photon's cell mod 2n, id)."""
    hs = M.HostScene.plain(7, 64, 36)
    r = _gpu_photons(hs, 100000, radius, k=k)
    pos, _, ids = r.photons()
    r.close()
    n = len(pos)
    cell = np.sqrt(np.float32(radius)) if k else np.float32(radius)
    c = np.floor(pos / np.float32(cell)).astype(np.int64)
    b = hash_cell_np(c[:, 0], c[:, 1], c[:, 2], 2 * n).astype(np.int64)
    key = (b << 32) | ids.astype(np.int64)
    assert n > 1000 and np.all(np.diff(key) > 0)


@pytest.mark.gpu
def test_gpu_photon_map_rebuild_and_disable():
    hs = M.HostScene.plain(7, 64, 36)
    r = _gpu_photons(hs, 5000, 0.01)
    n1 = len(r.photons()[0])
    r.build_photons(2000, 0.01)
    n2 = len(r.photons()[0])
    assert 0 < n2 < n1
    r.build_photons(0, 0.01)            # 0 photons: caustics off
    assert len(r.photons()[0]) == 0
    assert not r.caustic_at(np.zeros((4, 3), np.float32)).any()
    with pytest.raises(M.MrtError):
        r.build_photons(100, -1.0)
    r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("W,H,sub,floor,water,n,radius", [
    (64, 36, 3, 8, 24, 40000, 0.05),
    (128, 72, 4, 16, 64, 200000, 0.03)])
def test_gpu_hdri_caustic_scene_matches_oracle(W, H, sub, floor, water, n,
                                               radius):
    """Water heightfield, glass sphere, HDRI with a sun, MIS estimator and
    environment photons gathered by fixed radius: the full image on the same
    RNG stream against the oracle fed the GPU's photon map (the larger case:
    a 64^2-cell water grid and 200 k photons)."""
    hs = M.HostScene.caustic(4, W, H, subdiv=sub, floor_cells=floor,
                             water_cells=water)
    view, bvh = _bvh(hs)
    sky = M.hdri_sky(5, 128, 64, sun_radius_deg=3.0, sun_radiance=300.0)
    r = M.Renderer(W, H)
    r.upload(hs)
    r.upload_hdri(sky)
    r.set_option(M.OPT_ESTIMATOR, M.EST_MIS)
    r.set_seed(SEED)
    r.build_photons(n, radius, k=0, norm=1.0)
    pos, pw, _ = r.photons()
    r.clear()
    r.render(0, 4, 8)
    img = r.readback().astype(np.float64)
    r.close()
    assert len(pos) > 100
    pm = O.photon_map(pos, pw, k=0, radius=radius, norm=1.0)
    hd = O.make_hdri(sky)
    ref, _ = O.render(view, bvh, W, H, 0, 4, 8, SEED, estimator=M.EST_MIS,
                      hdri=hd, photons=pm)
    base, _ = O.render(view, bvh, W, H, 0, 4, 8, SEED, estimator=M.EST_MIS,
                       hdri=hd)
    assert np.abs(ref - base).max() > 1e-3          # caustics are visible
    diff = np.abs(img - ref)
    close = diff <= 1e-4 + 1e-4 * np.abs(ref)
    print(f"close {close.mean():.6f} mean diff {diff.mean():.3e}")
    # a path whose refraction through the water differs by an ulp may leave
    # on another branch: a few pixels, a small share of the image's energy
    assert close.mean() >= 0.995
    assert diff.mean() <= 1e-3 * ref.mean()
