"""Seeded synthetic workloads shared by the benchmark and the tests.

The plain .rt scene: a ground plane and three spheres (lambert, metal,
glass) on seeded jitter, one point light, as the text of a .rt file.

The batched workload: many independent scenes, each a displaced icosphere
over a heightfield floor (the procedural mesh scene at a given subdivision
and floor size), with triangle counts log-uniform over a range, resolutions
from {256^2, 512^2, 960x540, 1920x1080} and spp from {16, 64, 256}.
"""
from __future__ import annotations

import numpy as np

RESOLUTIONS = [(256, 256), (512, 512), (960, 540), (1920, 1080)]
SPPS = [16, 64, 256]


def batched_scenes(seed, n, tri_lo=1e4, tri_hi=2e6, spp_div=1):
    """Descriptions of the n scenes: the icosphere subdivision comes from the
    target triangle count, floor cells make up the rest; resolution and spp
    are drawn from the fixed sets (spp divided by spp_div)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        target = float(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        sub = 3
        while 20 * 4 ** (sub + 1) <= target:
            sub += 1
        rest = max(target - 20 * 4 ** sub, 2.0)
        cells = max(1, int(round((rest / 2) ** 0.5)))
        w, h = RESOLUTIONS[int(rng.integers(len(RESOLUTIONS)))]
        spp = max(1, SPPS[int(rng.integers(len(SPPS)))] // spp_div)
        out.append({"seed": 1000 + k, "subdiv": sub, "floor": cells,
                    "w": w, "h": h, "spp": spp,
                    "rng_seed": int(rng.integers(1 << 62))})
    return out


def batched_scene(d, width=None, height=None):
    """The HostScene of one description (optionally at another size)."""
    from . import HostScene
    return HostScene.mesh(d["seed"], width or d["w"], height or d["h"],
                          subdiv=d["subdiv"], floor_cells=d["floor"],
                          amp=0.15)


def triangles(d):
    """Triangle count of a description: 20 * 4^subdiv + 2 * floor^2."""
    return 20 * 4 ** d["subdiv"] + 2 * d["floor"] ** 2


def config1_rt(seed=11, width=640, height=360):
    """The plain 640x360 scene as .rt text: ground plane, lambert / metal
    (fuzz 0.1) / glass (ior 1.5) unit spheres on seeded jitter within +-0.5
    of x in {-2, 0, 2}, one point light (intensity 1), ambient 0.1, a
    60 degree pinhole camera.  Returns the text and the spheres
    (centre, radius) the generator placed."""
    from . import HostScene
    gen = HostScene.plain(seed, width, height)
    sph = gen.arrays()["sph"]
    lines = [f"W {width} {height}", "A 0.1 255,255,255", "C 0,1.5,-6 0,0,1 60",
             "L 0,5,-2 1.0 255,255,255",
             "mat lam 204,204,204 0 0 0 0 1 0 0,0,0",
             "mat metal 204,204,204 1 0.1 0 0 1 0 0,0,0",
             "mat glass 255,255,255 0 0 1.5 1 1 0 0,0,0",
             "pl 0,0,0 0,1,0 lam"]
    for (x, y, z, r), m in zip(sph, ("lam", "metal", "glass")):
        lines.append(f"sp {x:.9g},{y:.9g},{z:.9g} {2 * r:.9g} {m}")
    return "\n".join(lines) + "\n", sph
