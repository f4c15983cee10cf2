"""The drop-in boundary as miniRT's render loop drives it (needs a GPU).

compute_frame_gpu (src/gpu/use_shader.c:81-100) renders one sample per tick
and get_result (:53-79) adds it into the host's double float_render; the
replacement binding is mrt_render(sample, 1) -> mrt_readback_add -> sample++.
These tests drive exactly that sequence (and mrt_frame_tick, and a resume
from a render_state.rts checkpoint, export_render_state.c:15-51) on a scene
read from a .rt file with a camera set through mrt_set_camera, and compare
with one multi-sample render and with the oracle on the same RNG stream.
"""
import numpy as np
import pytest

import mrt_hip as M
from mrt_hip.workloads import config1_rt
import oracle_ffi as O

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE
N = 4
DEPTH = 8
# a viewpoint other than the file's, set with mrt_set_camera
CAM_POS = (0.6, 1.9, -5.5)
CAM_DIR = tuple(np.array([-0.1, -0.15, 1.0]) / np.linalg.norm([-0.1, -0.15, 1.0]))
CAM_FOV = 55.0


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    text, sph = config1_rt()
    path = tmp_path_factory.mktemp("rt") / "plain.rt"
    path.write_text(text)
    hs = M.HostScene.parse_file(path)
    np.testing.assert_array_equal(hs.arrays()["sph"], sph)
    assert hs.resolution == (640, 360)
    return hs


def ticks(r, n, buf, first=0):
    """The compute_frame_gpu binding, tick by tick."""
    for s in range(first, first + n):
        r.render(s, 1, DEPTH)
        r.readback_add(buf)


def test_ticks_equal_one_render_and_the_oracle(scene):
    w, h = scene.resolution
    with M.Renderer(w, h) as r:
        r.upload(scene)
        r.set_seed(SEED)
        r.set_camera(CAM_POS, CAM_DIR, CAM_FOV)
        buf = np.zeros((h, w, 3), np.float64)
        ticks(r, N, buf)
        # the device sums were handed over: nothing is left behind
        assert not r.readback().any()
        r.render(0, N, DEPTH)
        one = r.readback().astype(np.float64)
    # k ticks hold k samples (get_result adds one dispatch per call), not
    # the running totals of every tick
    np.testing.assert_allclose(buf, one, rtol=2e-6, atol=1e-6)
    # the oracle on the same viewpoint and RNG stream
    scene.set_camera(CAM_POS, CAM_DIR, CAM_FOV)
    view = scene.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    ref, _ = O.render(view, bvh, w, h, 0, N, DEPTH, SEED)
    diff = np.abs(buf - ref)
    close = diff <= 1e-4 + 1e-4 * np.abs(ref)
    print(f"ticks vs oracle: close {close.mean():.6f}, "
          f"mean abs diff {diff.mean():.3e}")
    assert close.mean() >= 0.995
    assert diff.mean() <= 1e-3 * ref.mean()


def test_frame_tick_and_rts_resume(scene, tmp_path):
    """mrt_frame_tick (render_frame) gives the tick sequence's bits; a run
    saved to render_state.rts after k samples, loaded into a fresh context
    and continued from screen.sample = k ends with the same bits as the
    uninterrupted run."""
    w, h = scene.resolution
    scene.set_camera((0.0, 1.5, -6.0), (0.0, 0.0, 1.0), 60.0)
    with M.Renderer(w, h) as r:
        r.upload(scene)
        r.set_seed(SEED)
        seq = np.zeros((h, w, 3), np.float64)
        ticks(r, N, seq)
        buf = np.zeros((h, w, 3), np.float64)
        f = r.frame(buf, spp=N, max_bounces=DEPTH)
        added = [f.tick() for _ in range(N + 1)]
        assert added == [1] * N + [0] and f.sample == N
        np.testing.assert_array_equal(buf, seq)
        # two samples per tick, never past spp
        buf2 = np.zeros((h, w, 3), np.float64)
        f2 = r.frame(buf2, spp=N + 1, max_bounces=DEPTH, samples_per_tick=2)
        assert [f2.tick() for _ in range(4)] == [2, 2, 1, 0]
        # interrupted after k samples
        k = 2
        part = np.zeros((h, w, 3), np.float64)
        fp = r.frame(part, spp=N, max_bounces=DEPTH)
        for _ in range(k):
            fp.tick()
        rts = tmp_path / "render_state.rts"
        M.export_rts(rts, part, fp.sample)
    back, s = M.load_rts(rts, w, h)
    assert s == k
    np.testing.assert_array_equal(back, part)
    with M.Renderer(w, h) as r2:
        r2.upload(scene)
        r2.set_seed(SEED)
        resumed = np.ascontiguousarray(back.reshape(h, w, 3))
        fr = r2.frame(resumed, spp=N, max_bounces=DEPTH, sample=s)
        while fr.tick():
            pass
        assert fr.sample == N
    np.testing.assert_array_equal(resumed, seq)


def test_set_camera_moves_the_view(scene):
    """mrt_set_camera (convert_scene) changes primary rays like a scene
    upload with that camera would: the primary ids of both agree."""
    w, h = 160, 90
    scene.set_resolution(w, h)
    try:
        scene.set_camera((0.0, 1.5, -6.0), (0.0, 0.0, 1.0), 60.0)
        with M.Renderer(w, h) as r:
            r.upload(scene)
            r.set_seed(SEED)
            r.set_camera(CAM_POS, CAM_DIR, CAM_FOV, defocus=2.0, focus=5.5)
            ids_a = r.aov(M.AOV_PRIM_ID, 3)
            t_a = r.aov(M.AOV_T, 3)
        scene.set_camera(CAM_POS, CAM_DIR, CAM_FOV, defocus=2.0, focus=5.5)
        with M.Renderer(w, h) as r:
            r.upload(scene)
            r.set_seed(SEED)
            ids_b = r.aov(M.AOV_PRIM_ID, 3)
            t_b = r.aov(M.AOV_T, 3)
        np.testing.assert_array_equal(ids_a, ids_b)
        np.testing.assert_array_equal(t_a, t_b)
        view = scene.view()
        bvh = O.make_bvh(M.BuiltBvh(view))
        o, d = O.raygen_f32(view, w, h, 3, SEED)
        rid, rt, _, _ = O.trace_rays(view, bvh, O.TRAV_ORDERED, o, d)
        assert np.mean(ids_a == rid) >= 0.9999
    finally:
        scene.set_resolution(640, 360)


def test_ticks_at_1080p_hand_over_every_sample_once():
    """At 1920x1080 (6.2 M floats) the read-back copies in chunks and adds
    each on the context's host threads as it arrives: each tick's sums land
    in the host buffer exactly (the float sums widened to double and added),
    and nothing is left on the device."""
    w, h = 1920, 1080
    hs = M.HostScene.plain(3, w, h)
    with M.Renderer(w, h) as r:
        r.upload(hs)
        r.set_seed(SEED)
        buf = np.zeros((h, w, 3), np.float64)
        expect = np.zeros_like(buf)
        for s in range(3):
            # the tick's sums, read without handing them over, then the tick
            with M.Renderer(w, h) as q:
                q.upload(hs)
                q.set_seed(SEED)
                q.render(s, 1, DEPTH)
                expect += q.readback().astype(np.float64)
            r.render(s, 1, DEPTH)
            r.readback_add(buf)
            assert np.array_equal(buf, expect), s
        assert not r.readback().any()
        assert buf.any()
