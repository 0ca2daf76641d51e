"""EmitterSampler draws and next-event estimation with lights drawn by power
on the device, against the oracle on the same Philox streams: the light
index of every draw (up to draws within a float's rounding of a boundary of
the power table), its point, normal and densities, with the reference's NaN
points on spheres in the same places; and the per-path radiance of jobs that
draw their lights by power, which also come out the same alone and in a
served mix."""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from test_emitter_sampler import _lights_scene, _two_lights
from test_gpu_render import _per_path

pytestmark = pytest.mark.gpu


def test_emitter_sample_matches_oracle(ctx):
    a, share, area = _lights_scene()
    n = 1 << 20
    dev = m.Scene(ctx, a).emitter_sample(n, seed=21)
    orc = po.OracleScene(a).emitter_sample(n, seed=21)
    same = dev["light"].astype(np.int64) == orc["light"]
    print(f"[emitter] draws={n} same light={same.mean():.7f} counts={np.bincount(dev['light'])}", flush=True)
    assert same.mean() >= 1 - 1e-5
    p_d, p_o = dev["p"][same].astype(np.float64), orc["p"][same]
    nan_d, nan_o = np.isnan(p_d).any(axis=1), np.isnan(p_o).any(axis=1)
    # NaN exactly where 2 |s_y| > 1, up to draws at float rounding of the edge
    assert (nan_d != nan_o).mean() < 1e-5 and nan_o.mean() > 0.01
    ok = ~nan_d & ~nan_o
    err = np.abs(p_d[ok] - p_o[ok]).max(axis=1)
    print(f"[emitter] max position error {err.max():.3e}", flush=True)
    assert np.quantile(err, 0.99999) < 1e-5 and err.max() < 1e-3
    nerr = np.abs(dev["n"][same][ok].astype(np.float64) - orc["n"][same][ok]).max(axis=1)
    assert np.quantile(nerr, 0.99999) < 1e-4
    np.testing.assert_allclose(dev["pdf_select"], share[dev["light"]], rtol=1e-6)
    np.testing.assert_allclose(dev["pdf_area"], 1.0 / area, rtol=1e-6)
    again = m.Scene(ctx, a).emitter_sample(n, seed=21)
    assert np.array_equal(again.view(np.uint32), dev.view(np.uint32))


def test_light_select_paths(ctx):
    a, _, _ = _lights_scene()
    job = m.make_job(a.camera, 32, 32, 4, 4, seed=12)
    job.direct_samples = 4
    job.light_select = 1
    _per_path(ctx, a, job, 0, 32 * 32, "lights drawn by power, five lights", min_frac=0.999)
    room = m.generate(m.MPT_SCENE_ROOM)
    room.points = np.array([[0.3, 0.7, 0.4, 0.6, 0.5, 0.4]], np.float32)
    job = m.make_job(room.camera, 48, 48, 8, 5, seed=13)
    job.direct_samples = 2
    job.light_select = 1
    L_gpu, L_orc = _per_path(ctx, room, job, 0, 48 * 48, "room + point light drawn by power", min_frac=0.999)
    assert L_orc.mean() > 0.05


def test_light_select_one_light_and_mix(ctx):
    """One light: the film drawing lights by power equals the film sampling
    every light.  Several: jobs drawing by power are bit-identical alone and
    in a mixed batch with jobs that sample every light."""
    room = m.generate(m.MPT_SCENE_ROOM)
    sc = m.Scene(ctx, room)
    job = m.make_job(room.camera, 64, 48, 16, 5, seed=2)
    job.direct_samples = 4
    loop = sc.render(job)[0]
    job.light_select = 1
    assert np.array_equal(sc.render(job)[0], loop)
    two = _two_lights()
    s2 = m.Scene(ctx, two)
    jobs = []
    for i in range(6):
        j = m.make_job(two.camera, 40 + 8 * i, 32, 4 + 4 * (i % 2), 3, seed=30 + i, job_id=i)
        j.direct_samples = 1 + (i % 3)
        j.light_select = i % 2
        jobs.append(j)
    alone = [s2.render(j)[0] for j in jobs]
    saved = ctx.get_option(m.MPT_OPT_BATCH_PATHS)
    ctx.set_option(m.MPT_OPT_BATCH_PATHS, 1 << 13)
    try:
        srv = m.Server(ctx)
        hs = [srv.submit(s2, j) for j in jobs]
        srv.drain()
        for h, j, ref in zip(hs, jobs, alone):
            assert np.array_equal(srv.result(h, j.width, j.height), ref), f"job {j.job_id}"
        srv.close()
    finally:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved)
