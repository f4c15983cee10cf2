"""Batched rendering: many scenes of different sizes through one set of
queues.  A batched scene must equal its plain render (same RNG keys, same
estimator) up to the grouping of the fp32 sums."""
import numpy as np
import pytest

import mrt_hip as M

DEPTH = 8


def soft_light_scene():
    hs = M.HostScene.empty()
    lam = hs.add_material(color=(0.7, 0.6, 0.5))
    glass = hs.add_material(color=(1.0, 1.0, 1.0), ior=1.5, transmission=1.0)
    hs.add_plane((0, 0, 0), (0, 1, 0), lam)
    hs.add_sphere((0.0, 1.0, 0.0), 1.0, glass)
    hs.add_sphere((1.5, 0.5, -1.0), 0.5, lam)
    hs.add_light((0, 4, -2), 0.8, (255, 240, 200), radius=0.4, samples=5)
    hs.set_ambient(0.2, (120, 160, 255))
    hs.set_camera((0, 1.5, -6), (0, -0.1, 1), fov=50.0)
    return hs


def planes_only_scene():
    hs = M.HostScene.empty()
    a = hs.add_material(color=(0.8, 0.3, 0.3))
    b = hs.add_material(color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.2)
    hs.add_plane((0, 0, 0), (0, 1, 0), a)
    hs.add_plane((0, 0, 4), (0, 0, -1), b)
    hs.add_light((0, 3, 0), 1.0, (255, 255, 255))
    hs.set_ambient(0.1, (255, 255, 255))
    hs.set_camera((0, 1, -3), (0, 0, 1), fov=70.0)
    return hs


def quadric_scene():
    """Cylinder, cone and hyperboloid with cap materials, after scenes with
    materials of their own (their indices are rebased in the batch)."""
    hs = M.HostScene.empty()
    lam = hs.add_material(color=(0.7, 0.7, 0.7))
    red = hs.add_material(color=(0.9, 0.2, 0.2))
    metal = hs.add_material(color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.3)
    glass = hs.add_material(color=(1.0, 1.0, 1.0), ior=1.5, transmission=1.0)
    hs.add_plane((0, 0, 0), (0, 1, 0), lam)
    hs.add_cylinder((-1.2, 0.8, 0), (0, 1, 0.2), 1.0, 1.6, red, metal, glass)
    hs.add_cone((1.2, 1.6, 0), (0, -1, 0), 1.2, 1.6, glass, red)
    hs.add_hyperboloid((0, 1.0, 1.5), (0, 1, 0), 2.0, 0.4, 0.4, 0.6, 1.0, metal)
    hs.add_light((0, 4, -2), 0.9, (255, 255, 255))
    hs.set_ambient(0.2, (200, 220, 255))
    hs.set_camera((0, 1.5, -5), (0, -0.15, 1), fov=60.0)
    return hs


# (scene factory, width, height, spp, seed)
CASES = [
    (lambda: M.HostScene.plain(7, 64, 36), 64, 36, 5, 0x1111),
    (lambda: M.HostScene.plain(11, 40, 30), 40, 30, 3, 0x2222),
    (lambda: M.HostScene.mesh(3, 48, 27, subdiv=4, floor_cells=16), 48, 27, 4,
     0x3333),
    (soft_light_scene, 33, 21, 6, 0x4444),
    (planes_only_scene, 20, 17, 2, 0x5555),
    (quadric_scene, 40, 32, 4, 0x6666),
]


@pytest.fixture(scope="module")
def scenes():
    return [(f(), w, h, spp, seed) for f, w, h, spp, seed in CASES]


@pytest.fixture(scope="module")
def plain_images(scenes):
    out = []
    for hs, w, h, spp, seed in scenes:
        with M.Renderer(w, h) as r:
            r.upload(hs)
            r.set_seed(seed)
            r.render(0, spp, DEPTH)
            out.append(r.readback())
    return out


def make_batch(scenes, **tuning):
    b = M.Batch()
    for hs, w, h, spp, seed in scenes:
        b.add(hs, w, h, spp, seed)
    b.commit()
    if tuning:
        b.set_tuning(**tuning)
    return b


def test_batch_needs_a_device_or_reports_it():
    """Without a gfx950 device the constructor raises, never crashes."""
    try:
        b = M.Batch()
    except M.MrtError as e:
        assert e.code == M.MRT_E_NODEVICE
        return
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("budget", [1 << 22, 4096, 3001 * 3])
def test_batch_matches_plain_renders(scenes, plain_images, budget):
    b = make_batch(scenes, path_budget=max(budget, 1024))
    b.render(DEPTH)
    for k, ref in enumerate(plain_images):
        img = b.readback(k)
        np.testing.assert_allclose(img, ref, rtol=2e-5, atol=2e-6,
                                   err_msg=f"scene {k}")
    st = b.stats()
    assert st["paths"] == sum(w * h * spp for _, w, h, spp, _ in scenes)
    assert st["nonfinite"] == 0
    b.close()


@pytest.mark.gpu
def test_graph_replay_equals_direct_launches(scenes):
    imgs = []
    for graph in (1, 0):
        b = make_batch(scenes, path_budget=4096, use_graph=graph)
        b.render(DEPTH)
        b.render(DEPTH)              # second frame: replays the same graph
        imgs.append([b.readback(k) for k in range(len(scenes))])
        b.close()
    for a, c in zip(*imgs):
        assert np.array_equal(a, c)


@pytest.mark.gpu
def test_autotune_then_render(scenes, plain_images):
    b = make_batch(scenes)
    t = b.autotune(DEPTH)
    print("chosen", t)
    assert t["path_budget"] in (1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24)
    assert t["ms_per_mpath"] > 0
    # the trial renders are cleared
    assert not b.readback(0).any()
    b.render(DEPTH)
    for k, ref in enumerate(plain_images):
        np.testing.assert_allclose(b.readback(k), ref, rtol=2e-5, atol=2e-6)
    b.close()


@pytest.mark.gpu
def test_batch_counting_mode(scenes):
    b = make_batch(scenes, path_budget=4096)
    b.set_option(M.OPT_STATS, M.STATS_COUNT | M.STATS_TIME)
    b.render(DEPTH)
    st = b.stats()
    assert st["rays_closest"] >= st["paths"]
    assert st["nodes_closest"] > 0 and st["ms_total"] > 0
    b.close()


@pytest.mark.gpu
def test_device_built_batch_matches_device_built_plain_renders(scenes):
    """With MRT_BUILDER_DEVICE every scene of the batch gets the tree a plain
    context builds for it on the device (the build is deterministic)."""
    b = M.Batch()
    b.set_option(M.OPT_BUILDER, M.BUILDER_DEVICE)
    for hs, w, h, spp, seed in scenes:
        b.add(hs, w, h, spp, seed)
    b.commit()
    b.render(DEPTH)
    for k, (hs, w, h, spp, seed) in enumerate(scenes):
        with M.Renderer(w, h) as r:
            r.set_option(M.OPT_BUILDER, M.BUILDER_DEVICE)
            r.upload(hs)
            r.set_seed(seed)
            r.render(0, spp, DEPTH)
            ref = r.readback()
        np.testing.assert_allclose(b.readback(k), ref, rtol=2e-5, atol=1e-6)
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [1, 0])
def test_batch_lanes_equal_one_lane(graph):
    """A pass large enough for four render lanes: every lane count gives the
    same bits, replayed graph or direct launches."""
    big = [(M.HostScene.plain(7, 256, 256), 256, 256, 3, 0x77),
           (M.HostScene.mesh(3, 192, 108, subdiv=5, floor_cells=32), 192, 108,
            4, 0x78),
           (soft_light_scene(), 128, 96, 3, 0x79)]
    imgs = {}
    for lanes in (1, 2, 4):
        b = make_batch(big, path_budget=1 << 20, use_graph=graph)
        b.set_option(M.OPT_LANES, lanes)
        b.render(DEPTH)
        b.render(DEPTH)
        imgs[lanes] = [b.readback(k) for k in range(len(big))]
        b.close()
    for lanes in (2, 4):
        for k in range(len(big)):
            assert np.array_equal(imgs[lanes][k], imgs[1][k]), (lanes, k)
