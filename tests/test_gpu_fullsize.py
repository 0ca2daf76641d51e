"""The benchmark configurations at their full sizes: the kd-tree stress
scene (1M triangles, 1920x1080, 64 spp, depth 4) and the procedural terrain
(2048^2 grid = 8.4M triangles + 500 glass spheres + a distant light,
1024x1024, 1024 spp, depth 8).

Full-size properties: the whole film is bit-identical across two renders;
its paths and camera rays are counted exactly; and on pixel runs spread over
the film, the device's per-path radiance agrees with the oracle's on the
same seeds and the film equals the per-pixel mean of those paths.
"""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _full_job(ctx, arrays, job, runs, min_frac, label):
    sc = m.Scene(ctx, arrays)
    W, H = job.width, job.height
    f = m.Film(ctx, W, H)
    st = f.render(sc, job)
    first = f.download_raw()
    f.render(sc, job)
    second = f.download_raw()
    f.close()
    print(f"{label}: {st['paths']} paths in {st['render_ms']:.1f} ms", flush=True)
    assert st["paths"] == W * H * job.spp and st["camera_rays"] == st["paths"]
    assert np.array_equal(first, second)
    img = (first[..., :3] / first[..., 3:]).reshape(-1, 3)
    osc = po.OracleScene(arrays)
    for b, e in runs:
        L_gpu = sc.render_paths(job, b, e).astype(np.float64)
        pix = np.repeat(np.arange(b, e, dtype=np.uint32), job.spp)
        smp = np.tile(np.arange(job.spp, dtype=np.uint32), e - b)
        L_orc = osc.render_paths(job, pix, smp)
        err = np.abs(L_gpu - L_orc).max(axis=1) / np.maximum(1.0, np.abs(L_orc).max(axis=1))
        frac = (err <= 1e-3).mean()
        print(f"{label} pixels {b}-{e}: paths {len(pix)} agree {frac:.5f}", flush=True)
        assert frac >= min_frac
        mean = L_gpu.reshape(e - b, job.spp, 3).mean(axis=1)
        np.testing.assert_allclose(img[b:e], mean, rtol=1e-4, atol=1e-6)


def test_soup_full_job(ctx):
    arrays = m.generate(m.MPT_SCENE_SOUP, 1_000_000, 1)
    cam = dict(arrays.camera)
    job = m.make_job(cam, 1920, 1080, 64, 4, seed=1)
    runs = [(r * 1920 + c, r * 1920 + c + 48) for r, c in ((0, 0), (311, 900), (777, 1500), (1079, 1872))]
    _full_job(ctx, arrays, job, runs, 0.99, "soup 1M 1920x1080x64")


def test_terrain_full_job(ctx):
    arrays = m.generate(m.MPT_SCENE_TERRAIN, 2048, 1)
    job = m.make_job(dict(arrays.camera), 1024, 1024, 1024, 8, seed=1)
    runs = [(r * 1024 + c, r * 1024 + c + 12) for r, c in ((100, 40), (600, 700))]
    _full_job(ctx, arrays, job, runs, 0.999, "terrain 2048^2 1024x1024x1024")
