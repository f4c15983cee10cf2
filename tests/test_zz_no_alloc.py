"""mrt_trace_rays_noalloc against mrt_trace_rays through the C ABI (needs a
GPU): the same ids and t, bit for bit, on two calls in a row (the second
smaller, so stale staging contents would show), and the size limit."""
import numpy as np
import pytest

import mrt_hip as M

pytestmark = pytest.mark.gpu


def rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-3, 3, size=(n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(0.2, 3.0, size=n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_noalloc_trace_equals_trace_rays():
    hs = M.HostScene.mesh(3, 64, 36, subdiv=5, floor_cells=32)
    with M.Renderer(64, 36) as r:
        r.upload(hs)
        for n, seed in ((M.TRACE_NOALLOC_MAX, 1), (5000, 2), (1, 3)):
            o, d = rays(n, seed)
            pid, t = r.trace_rays(o, d)
            pid2, t2 = r.trace_rays_noalloc(o, d)
            if n > 1000:
                hit = pid != M.PRIM_NONE
                assert 0.05 < hit.mean() < 0.95
            np.testing.assert_array_equal(pid2, pid)
            np.testing.assert_array_equal(t2.view(np.uint32),
                                          t.view(np.uint32))
        o, d = rays(M.TRACE_NOALLOC_MAX + 1, 4)
        with pytest.raises(M.MrtError) as e:
            r.trace_rays_noalloc(o, d)
        assert e.value.code == -6           # MRT_E_LIMIT
        # a second context on the same GPU shares the staging array
        with M.Renderer(32, 18) as r2:
            r2.upload(hs)
            o, d = rays(777, 5)
            a = r2.trace_rays_noalloc(o, d)
            b = r.trace_rays_noalloc(o, d)
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])
