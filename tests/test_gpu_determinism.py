"""Films are bit-identical run to run and alone vs mixed when a vertex casts
several shadow rays: four light samples per vertex on the room's area light
plus an omni point light.  Each vertex's shadow rays are summed in (light,
sample) order before they reach the path (DirectIllumination::evaluate sums
per light in a fixed loop, renderers/DirectIllumination.cpp:90-135), so the
order in which the device traces them cannot change a bit of the film.  The
same jobs also agree path by path with the oracle."""
import numpy as np
import pytest

import milton_amd as m
from test_gpu_render import _per_path

pytestmark = pytest.mark.gpu


def room_with_point_light():
    a = m.generate(m.MPT_SCENE_ROOM)
    a.points = np.array([[0.3, 0.7, 0.4, 0.8, 0.7, 0.5]], np.float32)
    return a


def _job(a, side, spp, depth, seed, direct, job_id=0):
    j = m.make_job(a.camera, side, side, spp, depth, seed=seed, job_id=job_id)
    j.direct_samples = direct
    return j


def test_multi_sample_nee_paths_match_oracle(ctx):
    a = room_with_point_light()
    job = _job(a, 48, 4, 5, 9, 4)
    _per_path(ctx, a, job, 0, 48 * 48, "room + point light, 4 light samples", min_frac=0.99)


def test_multi_sample_nee_film_is_deterministic(ctx):
    a = room_with_point_light()
    sc = m.Scene(ctx, a)
    job = _job(a, 256, 16, 6, 3, 4)
    first, st = sc.render(job)
    assert st["shadow_rays"] > st["extension_rays"]   # several shadow rays per vertex
    for lanes, batch in ((3, 1 << 25), (1, 1 << 16), (2, 12345)):
        saved = (ctx.get_option(m.MPT_OPT_LANES), ctx.get_option(m.MPT_OPT_BATCH_PATHS))
        ctx.set_option(m.MPT_OPT_LANES, lanes)
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, batch)
        try:
            again, _ = sc.render(job)
        finally:
            ctx.set_option(m.MPT_OPT_LANES, saved[0])
            ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved[1])
        print(f"[determinism] lanes={lanes} batch={batch} identical={np.array_equal(first, again)}", flush=True)
        assert np.array_equal(first, again)
    sc.close()


def test_multi_sample_nee_alone_equals_mixed(ctx):
    a = room_with_point_light()
    sc = m.Scene(ctx, a)
    soup = m.generate(m.MPT_SCENE_SOUP, 50_000, 2)
    ssoup = m.Scene(ctx, soup)
    rng = np.random.default_rng(4)
    jobs = []
    for i in range(12):
        side = int(rng.integers(24, 120))
        spp = int(rng.choice([1, 4, 9, 16]))
        direct = int(rng.choice([1, 2, 4, 9]))
        jobs.append((sc, _job(a, side, spp, int(rng.integers(3, 8)), int(rng.integers(1, 999)), direct, i)))
    for i in range(4):
        jobs.append((ssoup, m.make_job(soup.camera, 64, 64, 4, 4, seed=20 + i, job_id=100 + i)))
    alone = [s.render(j)[0] for s, j in jobs]
    saved = ctx.get_option(m.MPT_OPT_BATCH_PATHS)
    ctx.set_option(m.MPT_OPT_BATCH_PATHS, 1 << 15)   # many batches, each mixing several jobs
    try:
        srv = m.Server(ctx)
        handles = [srv.submit(s, j) for s, j in jobs]
        srv.drain()
        same = 0
        for h, (s, j), ref in zip(handles, jobs, alone):
            got = srv.result(h, j.width, j.height)
            same += int(np.array_equal(got, ref))
        print(f"[alone vs mixed] identical {same}/{len(jobs)} batches={srv.stats()['batches']}", flush=True)
        assert same == len(jobs)
        srv.close()
    finally:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved)
    sc.close()
    ssoup.close()
