"""Writes mesh1080_band_16k.npz: the converged reference for the image-error
figure of the benchmarked workload.

The 1920x1080 procedural mesh scene (seed 2: icosphere subdivided 8 times
over a 256^2 floor, 1,441,792 triangles, depth 8) rendered by the oracle in
fp64 with the reference's recursive BVH traversal (the CPU path's
semantics) at 16,384 samples per pixel on a band of rows through the
middle of the frame, samples 2^20 .. 2^20 + 16383 (disjoint from the
samples a GPU render starting at 0 draws), seed 0x5EED.  Stored as the
float32 mean per pixel.  Run from the repository root:
    python tests/golden/make_mesh1080_band.py [threads]
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "42-minirt_amd" / "python"))
sys.path.insert(0, str(ROOT / "oracle"))

import mrt_hip as M          # noqa: E402
import oracle_ffi as O       # noqa: E402

W, H, DEPTH, SEED = 1920, 1080, 8, 0x5EED
ROWS = (538, 542)
SPP = 16384
FIRST = 1 << 20
CHUNK = 256


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 0
    hs = M.HostScene.mesh(2, W, H)
    assert hs.view().contents.n_triangles == 1441792
    view = hs.view()
    bvh = O.make_bvh(M.BuiltBvh(view))
    acc = np.zeros((ROWS[1] - ROWS[0], W, 3), np.float64)
    t0 = time.perf_counter()
    for s in range(0, SPP, CHUNK):
        img, _ = O.render(view, bvh, W, H, FIRST + s, CHUNK, DEPTH, SEED,
                          f64=True, traversal=O.TRAV_REFERENCE,
                          threads=threads, rows=ROWS)
        acc += img[ROWS[0]:ROWS[1]]
        print(f"{s + CHUNK}/{SPP} spp, {time.perf_counter() - t0:.0f} s",
              flush=True)
    out = Path(__file__).resolve().parent / "mesh1080_band_16k.npz"
    np.savez_compressed(out, mean=(acc / SPP).astype(np.float32),
                        rows=np.array(ROWS), spp=SPP, first_sample=FIRST,
                        seed=SEED, depth=DEPTH)
    print("wrote", out)


if __name__ == "__main__":
    main()
