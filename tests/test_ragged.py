"""Frames whose path count is no multiple of a 64-ray chunk or of a block:
1 x 1, 7 x 5, 33 x 17, 1 x 300, 301 x 1 and 513 x 257 (131,841 pixels, so
a pass of two or more samples splits into parts on the render lanes).
Every sample of every pixel must reach the image exactly once: the same
bits on one render lane and on several, in one pass or in passes of one
sample, tick by tick through the drop-in read-back, and the oracle's
values on the same RNG stream (needs a GPU)."""
import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O

pytestmark = pytest.mark.gpu

SEED = 0x7A66ED
SPP = 4
DEPTH = 8
SIZES = [(1, 1), (7, 5), (33, 17), (1, 300), (301, 1), (513, 257)]
STAT_KEYS = ("rays_closest", "rays_shadow", "nodes_closest", "prims_closest",
             "nodes_shadow", "prims_shadow", "nonfinite")


def _render(hs, w, h, lanes, budget=None):
    with M.Renderer(w, h) as r:
        r.upload(hs)
        r.set_seed(SEED)
        r.set_option(M.OPT_LANES, lanes)
        if budget:
            r.set_option(M.OPT_PATH_BUDGET, budget)
        r.set_option(M.OPT_STATS, M.STATS_COUNT)
        r.render(0, SPP, DEPTH)
        img = r.readback()
        st = r.stats()
    return img, {k: st[k] for k in STAT_KEYS}


@pytest.mark.parametrize("w,h", SIZES)
def test_ragged_frame_same_bits_across_lanes_and_passes(w, h):
    hs = M.HostScene.plain(7, w, h)
    ref, st = _render(hs, w, h, 1)
    # every path is traced at bounce 0; a frame of one pixel still shades
    assert st["rays_closest"] >= w * h * SPP
    assert st["nonfinite"] == 0
    for lanes, budget in ((2, None), (4, None), (1, w * h), (3, w * h)):
        img, st2 = _render(hs, w, h, lanes, budget)
        assert np.array_equal(img, ref), (w, h, lanes, budget)
        assert st2 == st, (w, h, lanes, budget)


@pytest.mark.parametrize("w,h", [(1, 1), (33, 17), (513, 257)])
def test_ragged_frame_ticks_hand_over_every_sample_once(w, h):
    hs = M.HostScene.plain(7, w, h)
    with M.Renderer(w, h) as r:
        r.upload(hs)
        r.set_seed(SEED)
        buf = np.zeros((h, w, 3), np.float64)
        for s in range(SPP):
            r.render(s, 1, DEPTH)
            r.readback_add(buf)
        assert not r.readback().any()
        r.render(0, SPP, DEPTH)
        one = r.readback().astype(np.float64)
    np.testing.assert_allclose(buf, one, rtol=2e-6, atol=1e-6)


def test_ragged_frames_match_oracle():
    """All sizes pooled: per-pixel agreement with the oracle's render of the
    same samples, at the tolerance of the same-stream image tests."""
    diffs, refs, closes = [], [], []
    for w, h in SIZES:
        hs = M.HostScene.plain(7, w, h)
        view = hs.view()
        bvh = O.make_bvh(M.BuiltBvh(view))
        img, _ = _render(hs, w, h, 2)
        ref, _ = O.render(view, bvh, w, h, 0, SPP, DEPTH, SEED)
        assert img.shape == ref.shape == (h, w, 3)
        d = np.abs(img.astype(np.float64) - ref)
        diffs.append(d.ravel())
        refs.append(ref.ravel())
        closes.append((d <= 1e-4 + 1e-4 * np.abs(ref)).ravel())
    d, ref, close = (np.concatenate(x) for x in (diffs, refs, closes))
    print(f"ragged frames vs oracle: close {close.mean():.6f}, "
          f"mean abs diff {d.mean():.3e}")
    assert close.mean() >= 0.995
    assert d.mean() <= 1e-3 * ref.mean()
