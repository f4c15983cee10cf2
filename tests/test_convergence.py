"""Converged images: the HIP path at 16k spp against the oracle's 16k-spp
render with the reference's precision and traversal (fp64 vectors, the
recursive hit_bvh of src/ray_tracing/bvh/bvh_hit.c:16-44), under the image
error metric mean |I - R| / max(R, 1e-3) over pixels and channels, and the
Monte-Carlo convergence rate of the GPU image toward it."""
import numpy as np
import pytest

import mrt_hip as M
import oracle_ffi as O

pytestmark = pytest.mark.gpu

W, H = 96, 54
SPP = 16384
SEED = 0x51DE


def image_error(img, ref, eps=1e-3):
    return float(np.mean(np.abs(img - ref) / np.maximum(ref, eps)))


@pytest.fixture(scope="module")
def renders():
    hs = M.HostScene.plain(7, W, H)
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    imgs = {}
    with M.Renderer(W, H) as r:
        r.upload(hs)
        r.set_seed(SEED)
        done = 0
        for target in (64, 256, 1024, 4096, SPP):
            r.render(done, target - done, 8)
            done = target
            imgs[target] = r.readback().astype(np.float64) / target
    ref, _ = O.render(view, bvh, W, H, 0, SPP, 8, SEED, f64=True,
                      traversal=O.TRAV_REFERENCE)
    return imgs, ref / SPP


def test_16k_gpu_image_within_one_percent_of_16k_reference(renders):
    imgs, ref = renders
    err = image_error(imgs[SPP], ref)
    print(f"16k-spp GPU vs 16k-spp fp64 reference: error {err:.3e}")
    assert err <= 0.01


def test_error_falls_like_one_over_sqrt_spp(renders):
    imgs, ref = renders
    errs = [image_error(imgs[s], ref) for s in (64, 256, 1024, 4096)]
    print("error at 64/256/1024/4096 spp:", [f"{e:.4f}" for e in errs])
    for a, b in zip(errs, errs[1:]):
        # 4x the samples: the noise halves (within a generous margin; the
        # reference itself carries 16k-spp noise)
        assert b < a * 0.65
