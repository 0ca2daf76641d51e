"""Seeded job mixes of the batched and serving workloads.

Batched: 64 render jobs over the three benchmark scenes (the room, the
1M-triangle soup, the 2048^2 terrain) with square resolutions log-uniform in
[128, 1024], samples per pixel log-uniform in [4, 256] rounded to a perfect
square (the reference's stratified super-sampling uses floor(sqrt(spp))^2
strata, core/Viewport.h:29-36), and maximum depth uniform in {4..8}.
Serving draws jobs the same way, arriving as a Poisson process.
"""
import numpy as np

from . import MPT_SCENE_ROOM, MPT_SCENE_SOUP, MPT_SCENE_TERRAIN, make_job

SCENES = (MPT_SCENE_ROOM, MPT_SCENE_SOUP, MPT_SCENE_TERRAIN)
SCENE_COUNTS = {MPT_SCENE_ROOM: 0, MPT_SCENE_SOUP: 1_000_000, MPT_SCENE_TERRAIN: 2048}


def draw_job(rng, cameras, job_id, side_range=(128, 1024), spp_range=(4, 256), depth_range=(4, 8)):
    kind = SCENES[int(rng.integers(0, len(SCENES)))]
    side = int(round(np.exp(rng.uniform(np.log(side_range[0]), np.log(side_range[1])))))
    spp = int(round(np.sqrt(np.exp(rng.uniform(np.log(spp_range[0]), np.log(spp_range[1])))))) ** 2
    depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
    cam = dict(cameras[kind])
    cam["aspect"] = 1.0
    return kind, make_job(cam, side, side, spp, depth, seed=int(rng.integers(1, 1 << 30)), job_id=job_id)


def batched_jobs(cameras, n=64, seed=11, **kw):
    rng = np.random.default_rng(seed)
    return [draw_job(rng, cameras, i, **kw) for i in range(n)]


def poisson_arrivals(rate, duration, seed=13):
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(t)
