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
    assert t["path_budget"] in [1 << k for k in range(22, 28)]
    assert t["ms_per_mpath"] > 0
    assert t["use_graph"] in (0, 1) and t["trace_refill"] in (0, 1)
    # the trial renders are cleared
    assert not b.readback(0).any()
    b.render(DEPTH)
    for k, ref in enumerate(plain_images):
        np.testing.assert_allclose(b.readback(k), ref, rtol=2e-5, atol=2e-6)
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [1, 0])
def test_batch_refill_kernel_equals_k_trace(scenes, graph):
    """The batched refill trace kernel (each ray starts at its scene's root,
    read from LDS; planes read per scene at the ray's end) gives the bits and
    the node / primitive counts of k_trace, on scenes with planes, soft
    shadows, a planes-only scene and a mesh; one-sided triangles too."""
    sub = scenes[:5]                 # without the quadric scene
    one = M.HostScene.mesh(5, 40, 24, subdiv=3, floor_cells=8)
    one.set_flags(M.SCENE_TRIANGLE_ONE_SIDED)
    res = {}
    for sided in ("two", "one"):
        cases = sub if sided == "two" else [
            (one, 40, 24, 3, 0x99),
            (M.HostScene.mesh(6, 32, 20, subdiv=3, floor_cells=8), 32, 20, 2, 0x9A)]
        if sided == "one":
            cases[1][0].set_flags(M.SCENE_TRIANGLE_ONE_SIDED)
        for refill in (0, 1):
            b = make_batch(cases, path_budget=4096, use_graph=graph,
                           trace_refill=refill)
            b.render(DEPTH)
            imgs = [b.readback(k) for k in range(len(cases))]
            b.close()
            c = make_batch(cases, path_budget=4096, use_graph=0,
                           trace_refill=refill)
            c.set_option(M.OPT_STATS, M.STATS_COUNT)
            c.render(DEPTH)
            st = c.stats()
            c.close()
            res[sided, refill] = (imgs, st)
    for sided in ("two", "one"):
        (a, sa), (b_, sb) = res[sided, 0], res[sided, 1]
        for k, (x, y) in enumerate(zip(a, b_)):
            assert np.array_equal(x, y), (sided, k)
        for key in ("nodes_closest", "prims_closest", "nodes_shadow",
                    "prims_shadow", "rays_closest", "rays_shadow"):
            assert sa[key] == sb[key], (sided, key)


@pytest.mark.gpu
def test_batch_of_mixed_sidedness_falls_back_to_k_trace():
    """Scenes of both sidedness in one batch: the refill kernel takes one
    sidedness for all, so the batch traces with k_trace and still equals
    the plain renders."""
    one = M.HostScene.mesh(5, 40, 24, subdiv=3, floor_cells=8)
    one.set_flags(M.SCENE_TRIANGLE_ONE_SIDED)
    cases = [(one, 40, 24, 3, 0x99),
             (M.HostScene.mesh(6, 32, 20, subdiv=3, floor_cells=8), 32, 20, 2,
              0x9A)]
    b = make_batch(cases, path_budget=4096, trace_refill=1)
    b.render(DEPTH)
    for k, (hs, w, h, spp, seed) in enumerate(cases):
        with M.Renderer(w, h) as r:
            r.upload(hs)
            r.set_seed(seed)
            r.render(0, spp, DEPTH)
            ref = r.readback()
        np.testing.assert_allclose(b.readback(k), ref, rtol=2e-5, atol=1e-6)
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


@pytest.mark.gpu
def test_batch_matches_the_oracle():
    """Uneven scenes in one batch, each compared with the oracle's
    restatement of path_trace (path_trace.c:50-76) on the same RNG keys:
    the plain three-sphere scene, a small mesh scene, and the smallest
    displaced icosphere of the 64-scene batched workload's generator (its
    geometry as generated, rendered at 64x48)."""
    import oracle_ffi as O
    from mrt_hip.workloads import batched_scene, batched_scenes, triangles
    desc = min(batched_scenes(4, 64), key=triangles)
    cases = [(M.HostScene.plain(7, 64, 36), 64, 36, 5, 0x1111),
             (M.HostScene.mesh(3, 48, 27, subdiv=4, floor_cells=16), 48, 27,
              4, 0x3333),
             (batched_scene(desc, 64, 48), 64, 48, 4, desc["rng_seed"])]
    assert cases[2][0].view().contents.n_triangles == triangles(desc)
    b = make_batch(cases, path_budget=4096)
    b.render(DEPTH)
    for k, (hs, w, h, spp, seed) in enumerate(cases):
        img = b.readback(k).astype(np.float64)
        view = hs.view()
        bvh = O.make_bvh(M.BuiltBvh(view))
        ref, _ = O.render(view, bvh, w, h, 0, spp, DEPTH, seed)
        diff = np.abs(img - ref)
        close = diff <= 1e-4 + 1e-4 * np.abs(ref)
        print(f"scene {k}: close {close.mean():.6f}, "
              f"mean abs diff {diff.mean():.3e}")
        assert close.mean() >= 0.995, k
        assert diff.mean() <= 1e-3 * max(ref.mean(), 1e-6), k
    b.close()


@pytest.mark.gpu
def test_batched_workload_properties():
    """The full 64-scene batched workload (24.96 M triangles, 4.74 G
    samples): every path is traced once, no sample is non-finite, every
    scene's image is finite and lit, and two scenes (the smallest and the
    largest by samples) equal their plain renders."""
    from concurrent.futures import ThreadPoolExecutor
    from mrt_hip.workloads import batched_scene, batched_scenes
    desc = batched_scenes(4, 64)
    with ThreadPoolExecutor(8) as ex:
        scenes = list(ex.map(batched_scene, desc))
    b = M.Batch()
    for hs, d in zip(scenes, desc):
        b.add(hs, d["w"], d["h"], d["spp"], d["rng_seed"])
    b.commit()
    b.render(DEPTH)
    st = b.stats()
    assert st["paths"] == sum(d["w"] * d["h"] * d["spp"] for d in desc)
    assert st["nonfinite"] == 0
    imgs = {}
    for k, d in enumerate(desc):
        img = b.readback(k)
        assert np.isfinite(img).all(), k
        assert img.mean() > 0, k
        imgs[k] = img
    b.close()
    cost = [d["w"] * d["h"] * d["spp"] for d in desc]
    for k in (int(np.argmin(cost)), int(np.argmax(cost))):
        d = desc[k]
        with M.Renderer(d["w"], d["h"]) as r:
            r.upload(scenes[k])
            r.set_seed(d["rng_seed"])
            r.render(0, d["spp"], DEPTH)
            ref = r.readback()
        np.testing.assert_allclose(imgs[k], ref, rtol=1e-4, atol=1e-5,
                                   err_msg=f"scene {k}")


@pytest.mark.gpu
def test_batch_of_more_scenes_than_the_lds_root_table():
    """600 small scenes (more than the refill kernel's 512-entry root
    table): the batch traces with k_trace and every sampled scene equals
    its plain render."""
    cases = [(M.HostScene.plain(100 + k, 12, 8), 12, 8, 2, 0x9000 + k)
             for k in range(600)]
    b = make_batch(cases, path_budget=1 << 16)
    b.render(DEPTH)
    st = b.stats()
    assert st["paths"] == 600 * 12 * 8 * 2
    for k in (0, 1, 511, 512, 599):
        hs, w, h, spp, seed = cases[k]
        with M.Renderer(w, h) as r:
            r.upload(hs)
            r.set_seed(seed)
            r.render(0, spp, DEPTH)
            ref = r.readback()
        np.testing.assert_allclose(b.readback(k), ref, rtol=2e-5, atol=2e-6,
                                   err_msg=f"scene {k}")
    b.close()
