"""Render lanes (MRT_OPT_LANES): the passes of a render larger than the
path budget run on their own HIP streams, overlapping, and a render of one
pass splits it into parts traced at once on the lanes.  Every path keeps
its pixel and sample and the sums are added in the same order, so images
and traversal statistics must equal the one-lane render's bit for bit
(needs a GPU)."""
import numpy as np
import pytest

import mrt_hip as M
from test_gpu_parity import quadric_scene, textured_scene

pytestmark = pytest.mark.gpu

STAT_KEYS = ("rays_closest", "rays_shadow", "nodes_closest", "prims_closest",
             "nodes_shadow", "prims_shadow", "nonfinite")
W, H, SPP = 256, 256, 5
ONE_SAMPLE = W * H               # a path budget of one sample: five passes


def render(hs, lanes, env=None, est=M.EST_MIS, photons=False, spp=SPP,
           count=True, budget=ONE_SAMPLE):
    with M.Renderer(W, H) as r:
        r.upload(hs)
        r.set_seed(0x1A7E5)
        if env is not None:
            r.upload_hdri(env)
            r.set_option(M.OPT_ESTIMATOR, est)
        if photons:
            r.build_photons(20000, 0.05)
        if budget:
            r.set_option(M.OPT_PATH_BUDGET, budget)
        r.set_option(M.OPT_LANES, lanes)
        if count:
            r.set_option(M.OPT_STATS, M.STATS_COUNT)
        r.render(0, spp, 8)
        img = r.readback()
        st = r.stats()
    return img, {k: st[k] for k in STAT_KEYS}


def cases():
    sky = M.hdri_sky(3, 64, 32, sun_radius_deg=8.0, sun_radiance=40.0)
    glassy = M.HostScene.plain(7, W, H)
    return {
        "plain": (M.HostScene.plain(7, W, H), None, M.EST_MIS, False),
        "mesh": (M.HostScene.mesh(3, W, H, subdiv=5, floor_cells=32), None,
                 M.EST_MIS, False),
        "quadrics": (quadric_scene(W, H), None, M.EST_MIS, False),
        "textured": (textured_scene(W, H), None, M.EST_MIS, False),
        "env_mis": (M.HostScene.plain(5, W, H), sky, M.EST_MIS, False),
        "env_reference": (M.HostScene.plain(5, W, H), sky, M.EST_REFERENCE,
                          False),
        "photons": (glassy, None, M.EST_MIS, True),
    }


@pytest.fixture(scope="module")
def scenes():
    return cases()


@pytest.mark.parametrize("name", ["plain", "mesh", "quadrics", "textured",
                                  "env_mis", "env_reference", "photons"])
def test_lanes_equal_one_lane(scenes, name):
    hs, env, est, photons = scenes[name]
    ref, ref_st = render(hs, 1, env, est, photons)
    assert ref.any()
    for lanes in (2, 3, 4):
        img, st = render(hs, lanes, env, est, photons)
        assert np.array_equal(img, ref), (name, lanes)
        assert st == ref_st, (name, lanes)
    # one pass of every sample (327,680 paths): the pass splits into
    # contiguous parts traced at once on the lanes
    one, one_st = render(hs, 1, env, est, photons, budget=None)
    for lanes in (2, 3, 4):
        img, st = render(hs, lanes, env, est, photons, budget=None)
        assert np.array_equal(img, one), (name, lanes)
        assert st == one_st, (name, lanes)


def test_lanes_with_several_batches(scenes):
    """Passes of 2 samples, the last one smaller, on three lanes, then a
    second render call into the same sums: equal to one lane."""
    hs = scenes["mesh"][0]
    out = []
    for lanes in (1, 3):
        with M.Renderer(W, H) as r:
            r.upload(hs)
            r.set_seed(0x1A7E5)
            r.set_option(M.OPT_PATH_BUDGET, 2 * W * H)
            r.set_option(M.OPT_LANES, lanes)
            r.render(0, SPP, 8)
            r.render(SPP, 3, 8)
            out.append(r.readback())
    assert out[0].any()
    assert np.array_equal(out[1], out[0])


def test_lanes_option_is_validated():
    with M.Renderer(8, 8) as r:
        for bad in (0, 5, 1.5, -1):
            with pytest.raises(M.MrtError):
                r.set_option(M.OPT_LANES, bad)
        r.set_option(M.OPT_LANES, 4)
