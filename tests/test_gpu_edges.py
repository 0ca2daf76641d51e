"""Edge cases of the device path, checked against the oracle or against what
the case fixes exactly:
  * a scene with no primitives: every camera ray misses, so every path
    returns the background (a constant, or an environment image looked up
    along the camera ray), through a thin lens and a filtered film too;
  * the smallest job (1 x 1 pixel, one sample, depth 1) and a single pixel
    whose sample count exceeds the batch capacity;
  * empty ray batches, and rays that start outside the scene box and point
    away from it.
"""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from test_gpu_envmap import _sky
from test_gpu_render import _per_path

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def _empty(env=(0.25, 0.5, 1.0), env_map=None):
    mat = np.zeros(1, m.MATERIAL_DTYPE)
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [], np.zeros((0, 9)), [], mat,
                      env, env_map=env_map)
    a.camera = {"eye": [0.0, 0.0, 0.0], "look": [0.3, -0.2, 1.0], "up": [0, 1, 0], "height_angle": 60.0,
                "aspect": 1.25}
    return a


def test_empty_scene_is_the_background(ctx):
    a = _empty()
    sc = m.Scene(ctx, a)
    assert sc.stats()["refs"] == 0
    job = m.make_job(a.camera, 10, 8, 3, 5, seed=1)
    print("[empty scene, constant background] 10x8 spp 3", flush=True)
    L = sc.render_paths(job, 0, 80)
    assert (L == np.float32([0.25, 0.5, 1.0])).all()
    img, st = sc.render(job)
    assert np.allclose(img, np.float32([0.25, 0.5, 1.0]), rtol=1e-6)
    # only the camera rays are traced (the extension count includes them)
    assert st["camera_rays"] == st["extension_rays"] == 240 and st["shadow_rays"] == 0
    rng = np.random.default_rng(2)
    rays = np.zeros(1000, m.RAY_DTYPE)
    rays["o"] = rng.normal(size=(1000, 3))
    rays["d"] = rng.normal(size=(1000, 3))
    hits = sc.trace_closest(rays)
    assert (hits["prim"] == MISS).all() and np.isinf(hits["t"]).all()
    assert not sc.trace_any(rays, np.full(1000, 1e30, np.float32)).any()


def test_empty_scene_env_map_lens_and_filter(ctx):
    a = _empty(env=(0, 0, 0), env_map=m.EnvMap(_sky(), filter=m.MPT_FILTER_GAUSSIAN, support=3, params=(0.8, 0)))
    cam = dict(a.camera, aperture=0.05, focal_length=0.3, focal_distance=2.0, autofocus=True)
    job = m.make_job(cam, 24, 20, 4, 3, seed=7)
    L_gpu, L_orc = _per_path(ctx, a, job, 0, 24 * 20, "empty scene, env map, thin lens")
    assert L_orc.std(axis=0).min() > 0.01
    fjob = m.make_job(cam, 24, 20, 4, 3, seed=7, film_filter=(m.MPT_FILTER_MITCHELL, 2))
    sc = m.Scene(ctx, a)
    f = m.Film(ctx, 24, 20)
    print("[empty scene, filtered film] mitchell support 2", flush=True)
    f.render(sc, fjob)
    dev = f.download_raw().astype(np.float64)
    orc = po.OracleScene(a).render_film(fjob)
    den_err = np.abs(dev[..., 3] - orc[..., 3]) / np.maximum(np.abs(orc[..., 3]), 1e-12)
    num_err = np.abs(dev[..., :3] - orc[..., :3]).max(axis=2) / np.maximum(np.abs(orc[..., :3]).max(axis=2), 1e-3)
    assert np.mean(den_err <= 1e-4) == 1.0
    assert np.mean(num_err <= 1e-3) >= 0.98
    f.close()


def test_smallest_job(ctx):
    arrays = m.generate(m.MPT_SCENE_ROOM)
    sc = m.Scene(ctx, arrays)
    osc = po.OracleScene(arrays)
    print("[room 1x1 spp 1 depth 1]", flush=True)
    job = m.make_job(dict(arrays.camera, aspect=1.0), 1, 1, 1, 1, seed=9)
    L = sc.render_paths(job, 0, 1).astype(np.float64)
    ref = osc.render_paths(job, np.zeros(1, np.uint32), np.zeros(1, np.uint32))
    assert np.allclose(L, ref, rtol=1e-3, atol=1e-6)
    img, st = sc.render(job)
    assert img.shape == (1, 1, 3) and np.allclose(img[0, 0], L[0], rtol=1e-6, atol=1e-7)
    assert st["paths"] == 1 and st["camera_rays"] == 1


def test_pixel_with_more_samples_than_a_batch(ctx):
    arrays = m.generate(m.MPT_SCENE_ROOM)
    sc = m.Scene(ctx, arrays)
    job = m.make_job(arrays.camera, 3, 1, 3000, 4, seed=11)
    saved = ctx.get_option(m.MPT_OPT_BATCH_PATHS)
    print("[room 3x1 spp 3000, batch capacity 1024]", flush=True)
    try:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, 1024)
        img, st = sc.render(job)
        L = sc.render_paths(job, 0, 3)
    finally:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved)
    assert st["paths"] == 9000
    np.testing.assert_allclose(img[0], L.reshape(3, 3000, 3).astype(np.float64).mean(axis=1), rtol=1e-5, atol=1e-6)
    full, _ = sc.render(job)
    assert np.array_equal(full, img)


def test_empty_ray_batches_and_rays_pointing_away(ctx):
    arrays = m.generate(m.MPT_SCENE_ROOM)
    sc = m.Scene(ctx, arrays)
    none = np.zeros(0, m.RAY_DTYPE)
    assert len(sc.trace_closest(none)) == 0
    assert len(sc.trace_any(none, np.zeros(0, np.float32))) == 0
    # the room spans [0, 1]^3: rays from outside heading further out miss it
    rng = np.random.default_rng(4)
    n = 4096
    axis = rng.integers(0, 3, n)
    sign = rng.choice([-1.0, 1.0], n)
    o = rng.uniform(0.1, 0.9, (n, 3))
    o[np.arange(n), axis] = np.where(sign > 0, 1.5, -0.5)
    d = rng.normal(size=(n, 3))
    d[np.arange(n), axis] = sign * (np.abs(d[np.arange(n), axis]) + 0.1)
    rays = np.zeros(n, m.RAY_DTYPE)
    rays["o"], rays["d"] = o, d
    print("[room, rays pointing away]", flush=True)
    hits = sc.trace_closest(rays)
    assert (hits["prim"] == MISS).all()
    assert not sc.trace_any(rays, np.full(n, 1e30, np.float32)).any()
