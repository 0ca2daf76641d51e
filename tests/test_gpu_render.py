"""Estimator parity of the HIP wavefront path tracer against the oracle's
restatement of PathTracer::_evaluate (renderers/renderers/PathTracer.cpp:
28-116) with DirectIllumination::evaluate (renderers/DirectIllumination.cpp:
46-136).

Two checks:
  * same seeds, per path: both sides draw identical random numbers, so a path
    returns the same radiance unless FP32 and FP64 take a different branch
    (an edge tie, a roulette threshold); >= 99% of paths must agree to 1e-3;
  * independent seeds, per pixel: the device image and an oracle image agree
    within 3 sigma of the estimator on >= 99% of pixels.
"""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _per_path(ctx, arrays, job, pix_begin, pix_end, label, min_frac=0.99):
    sc = m.Scene(ctx, arrays)
    L_gpu = sc.render_paths(job, pix_begin, pix_end).astype(np.float64)
    osc = po.OracleScene(arrays)
    pix = np.repeat(np.arange(pix_begin, pix_end, dtype=np.uint32), job.spp)
    smp = np.tile(np.arange(job.spp, dtype=np.uint32), pix_end - pix_begin)
    L_orc = osc.render_paths(job, pix, smp)
    err = np.abs(L_gpu - L_orc).max(axis=1) / np.maximum(1.0, np.abs(L_orc).max(axis=1))
    frac = (err <= 1e-3).mean()
    print(f"[{label}] paths={len(pix)} agree={frac:.5f} mean_gpu={L_gpu.mean(axis=0)} "
          f"mean_orc={L_orc.mean(axis=0)} nonzero={np.mean(L_orc.max(axis=1) > 0):.3f}", flush=True)
    assert np.isfinite(L_gpu).all()
    assert frac >= min_frac
    return L_gpu, L_orc


def _three_sigma(L_a, L_b, spp, label, min_frac=0.99):
    """Per-pixel z-test between two independent per-path sets (pixel-major)."""
    a = L_a.reshape(-1, spp, 3)
    b = L_b.reshape(-1, spp, 3)
    ma, mb = a.mean(axis=1), b.mean(axis=1)
    va, vb = a.var(axis=1, ddof=1) / spp, b.var(axis=1, ddof=1) / spp
    sigma = np.sqrt(va + vb)
    diff = np.abs(ma - mb)
    ok = (diff <= 3 * sigma) | (diff <= 1e-6)
    frac = ok.all(axis=1).mean()
    print(f"[{label}] pixels={len(ma)} within_3sigma={frac:.4f} mean_a={ma.mean(axis=0)} "
          f"mean_b={mb.mean(axis=0)}", flush=True)
    assert frac >= min_frac
    # the image means agree within 3 sigma of the whole-image estimate
    n = a.shape[0] * spp
    s = np.sqrt(a.reshape(-1, 3).var(axis=0) / n + b.reshape(-1, 3).var(axis=0) / n)
    assert np.all(np.abs(a.reshape(-1, 3).mean(axis=0) - b.reshape(-1, 3).mean(axis=0)) <= 4 * s + 1e-9)


def test_room_paths_same_seed(ctx):
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(arrays.camera, 64, 64, 16, 5, seed=1)
    _per_path(ctx, arrays, job, 0, 64 * 64, "room-paths")


def test_room_image_three_sigma(ctx):
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job_gpu = m.make_job(arrays.camera, 48, 48, 64, 5, seed=1)
    job_orc = m.make_job(arrays.camera, 48, 48, 64, 5, seed=2)
    sc = m.Scene(ctx, arrays)
    L_gpu = sc.render_paths(job_gpu, 0, 48 * 48).astype(np.float64)
    osc = po.OracleScene(arrays)
    pix = np.repeat(np.arange(48 * 48, dtype=np.uint32), 64)
    smp = np.tile(np.arange(64, dtype=np.uint32), 48 * 48)
    L_orc = osc.render_paths(job_orc, pix, smp)
    _three_sigma(L_gpu, L_orc, 64, "room-3sigma", min_frac=0.98)


def test_soup_paths_same_seed(ctx):
    arrays = m.generate(m.MPT_SCENE_SOUP, 100_000, 1)
    job = m.make_job(arrays.camera, 96, 54, 16, 4, seed=1)
    _per_path(ctx, arrays, job, 0, 96 * 54, "soup-paths")


def test_film_matches_paths_and_is_deterministic(ctx):
    """The film is the per-pixel mean of the per-path radiance
    (ProgressiveFilterValue numerator/denominator), bit-identical across runs
    and independent of the batch size (queue depth)."""
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(arrays.camera, 40, 30, 9, 5, seed=4)
    sc = m.Scene(ctx, arrays)
    film1, st1 = sc.render(job)
    film2, _ = sc.render(job)
    assert np.array_equal(film1, film2)
    L = sc.render_paths(job, 0, 40 * 30).reshape(30, 40, 9, 3)
    np.testing.assert_allclose(film1, L.mean(axis=2), rtol=1e-5, atol=1e-6)
    saved = ctx.get_option(m.MPT_OPT_BATCH_PATHS)
    ctx.set_option(m.MPT_OPT_BATCH_PATHS, 1152)   # 128 pixels per batch
    try:
        film3, st3 = sc.render(job)
    finally:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved)
    assert np.array_equal(film1, film3)
    assert st1["paths"] == 40 * 30 * 9 and st1["camera_rays"] == st1["paths"]
    assert st1["extension_rays"] == st3["extension_rays"]


def test_ray_counts_match_oracle(ctx):
    """Camera / extension ray counts of the device equal the oracle's on the
    same seeds up to the paths whose branches differ."""
    arrays = m.generate(m.MPT_SCENE_SOUP, 100_000, 1)
    job = m.make_job(arrays.camera, 64, 36, 4, 4, seed=1)
    sc = m.Scene(ctx, arrays)
    _, st = sc.render(job)
    osc = po.OracleScene(arrays)
    _, _, c = osc.render(job)
    assert st["camera_rays"] == c["camera"]
    assert abs(st["extension_rays"] - c["extension"]) <= 0.01 * c["extension"]


def test_terrain_paths_same_seed(ctx):
    """Diffuse + modified Phong terrain, dielectric spheres (hero-wavelength
    channel rule), distant area light, depth 8."""
    arrays = m.generate(m.MPT_SCENE_TERRAIN, 256, 1)
    job = m.make_job(arrays.camera, 64, 64, 16, 8, seed=1)
    _per_path(ctx, arrays, job, 0, 64 * 64, "terrain-paths", min_frac=0.999)


def test_terrain_image_three_sigma(ctx):
    arrays = m.generate(m.MPT_SCENE_TERRAIN, 256, 1)
    job_gpu = m.make_job(arrays.camera, 32, 32, 256, 8, seed=1)
    job_orc = m.make_job(arrays.camera, 32, 32, 256, 8, seed=2)
    sc = m.Scene(ctx, arrays)
    L_gpu = sc.render_paths(job_gpu, 0, 32 * 32).astype(np.float64)
    osc = po.OracleScene(arrays)
    pix = np.repeat(np.arange(32 * 32, dtype=np.uint32), 256)
    smp = np.tile(np.arange(256, dtype=np.uint32), 32 * 32)
    L_orc = osc.render_paths(job_orc, pix, smp)
    _three_sigma(L_gpu, L_orc, 256, "terrain-3sigma", min_frac=0.97)
