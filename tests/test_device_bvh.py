"""The device BVH builder (PLOC clustering, leaf collapse, 4-wide collapse)
through the C ABI: the tree's invariants, its records against the host
builder's, traversal against brute force, and renders against the oracle
walking the same tree (needs a GPU)."""
import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O
from test_bvh import collect4
from test_quadrics import soup as quadric_soup

pytestmark = pytest.mark.gpu

SEED = 0x5EED1234


def device_tree(hs, w=96, h=64, builder=M.BUILDER_DEVICE):
    r = M.Renderer(w, h)
    r.set_option(M.OPT_BUILDER, builder)
    r.upload(hs)
    r.set_seed(SEED)
    return r, r.scene_bvh()


def orc(b):
    return O.OrcBvh(None, 0, b.prims.ctypes.data, b.n_prims,
                    b.nodes4.ctypes.data, b.n_nodes4)


@pytest.mark.parametrize("subdiv,floor", [(5, 32), (7, 128)])
def test_device_tree_invariants(subdiv, floor):
    hs = M.HostScene.mesh(4, 64, 36, subdiv=subdiv, floor_cells=floor)
    r, b = device_tree(hs)
    with r:
        assert b.on_device
        seen, depth = collect4(b)
        assert np.all(seen == 1) and depth == b.depth4 <= 40
        # leaves of at most 4, the same records as the host builder's
        # (keyed by source id)
        host = M.BuiltBvh(hs.view())
        dw, hw = b.prim_words(), host.prim_words()
        do, ho = np.argsort(dw[:, 3]), np.argsort(hw[:, 3])
        np.testing.assert_array_equal(dw[do], hw[ho])
        np.testing.assert_array_equal(b.shade_nrm.reshape(-1, 9)[do],
                                      host.shade_nrm.reshape(-1, 9)[ho])
        # a clustering tree is within a modest factor of the SAH tree's size
        assert b.n_nodes4 < 1.5 * host.n_nodes4 + 16


def test_device_tree_traversal_equals_brute_force():
    hs = quadric_soup(7)
    rng = np.random.default_rng(3)
    m = hs.add_material()
    for _ in range(400):
        c = rng.uniform(-4, 4, 3)
        hs.add_triangle(c + rng.normal(scale=0.3, size=(3, 3)), m)
    r, b = device_tree(hs)
    with r:
        assert b.on_device
        view = hs.view()
        n = 5000
        o = rng.uniform(-6, 6, size=(n, 3)).astype(np.float32)
        d = rng.normal(size=(n, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        pb, tb, _, _ = O.trace_rays(view, None, O.TRAV_BRUTE, o, d)
        p4, t4, _, _ = O.trace_rays(view, orc(b), O.TRAV_ORDERED, o, d)
        assert np.array_equal(p4, pb)
        np.testing.assert_array_equal(t4, tb)
        # the GPU on its own tree
        gid, gt = r.trace_rays(o, d)
        assert np.mean(gid == pb) >= 0.9999


def test_device_tree_renders_match_oracle_on_it():
    hs = M.HostScene.mesh(3, 192, 108, subdiv=6, floor_cells=48)
    r, b = device_tree(hs, 192, 108)
    with r:
        view = hs.view()
        bvh = orc(b)
        ids = r.aov(M.AOV_PRIM_ID, 2)
        nodes = r.aov(M.AOV_NODES, 2)
        o, d = O.raygen_f32(view, 192, 108, 2, SEED)
        rid, _, rn, _ = O.trace_rays(view, bvh, O.TRAV_ORDERED, o, d)
        assert np.mean(ids == rid) >= 0.9999
        assert np.mean(nodes == rn) >= 0.999
        r.clear()
        r.render(3, 4, 8)
        img = r.readback().astype(np.float64)
        ref, _ = O.render(view, bvh, 192, 108, 3, 4, 8, SEED)
        close = np.abs(img - ref) <= 1e-4 + 1e-4 * np.abs(ref)
        assert close.mean() >= 0.995


def test_small_scenes_fall_back_to_the_host_builder():
    hs = M.HostScene.plain(7, 32, 18)
    r, b = device_tree(hs, 32, 18)
    with r:
        assert not b.on_device
        seen, _ = collect4(b)
        assert np.all(seen == 1)
