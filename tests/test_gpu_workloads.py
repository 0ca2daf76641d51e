"""The batched and serving workloads at their specified sizes.

Batched: 64 jobs over the three benchmark scenes (the room, the 1M-triangle
soup, the 2048^2 terrain with 500 glass spheres) with sides log-uniform in
[128, 1024], samples per pixel log-uniform in [4, 256] and depth 4-8, seed 11.
Here the room also holds an omni point light and every other room job draws
four light samples per vertex, so vertices cast several shadow rays.  All 64
jobs go through one server; every film must be bit-identical to the same job
rendered alone.

Serving: jobs drawn the same way arrive as a Poisson process against the three
resident scenes while the server's worker thread runs; every result must be
bit-identical to the job rendered alone.
"""
import time

import numpy as np
import pytest

import milton_amd as m
from milton_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def resident(ctx):
    scenes = {}
    for kind in workloads.SCENES:
        a = m.generate(kind, workloads.SCENE_COUNTS[kind], 1)
        if kind == m.MPT_SCENE_ROOM:
            a.points = np.array([[0.3, 0.7, 0.4, 0.8, 0.7, 0.5]], np.float32)
        t0 = time.perf_counter()
        scenes[kind] = (a, m.Scene(ctx, a))
        print(f"[scenes] kind={kind} prims={a.num_prims} build={time.perf_counter() - t0:.1f}s", flush=True)
    yield scenes
    for _, sc in scenes.values():
        sc.close()


def _with_light_samples(jobs):
    n = 0
    for kind, j in jobs:
        if kind == m.MPT_SCENE_ROOM:
            if n % 2 == 0:
                j.direct_samples = 4
            n += 1
    return jobs


def test_batched_mix_matches_jobs_alone(ctx, resident):
    cams = {k: v[0].camera for k, v in resident.items()}
    jobs = _with_light_samples(workloads.batched_jobs(cams, 64, seed=11))
    assert len(jobs) == 64 and len({k for k, _ in jobs}) == 3
    assert any(j.direct_samples == 4 for _, j in jobs)
    paths = sum(j.width * j.height * j.spp for _, j in jobs)
    t0 = time.perf_counter()
    alone = [resident[k][1].render(j)[0] for k, j in jobs]
    t_alone = time.perf_counter() - t0
    srv = m.Server(ctx)
    t0 = time.perf_counter()
    handles = [srv.submit(resident[k][1], j) for k, j in jobs]
    srv.drain()
    t_mix = time.perf_counter() - t0
    st = srv.stats()
    same = 0
    for h, (k, j), ref in zip(handles, jobs, alone):
        assert srv.status(h)["state"] == m.JOB_DONE
        same += int(np.array_equal(srv.result(h, j.width, j.height), ref))
    srv.close()
    print(f"[batched mix] jobs=64 paths={paths} alone={t_alone:.3f}s server={t_mix:.3f}s "
          f"batches={st['batches']} identical={same}/64", flush=True)
    assert same == 64


def test_poisson_serving_matches_jobs_alone(ctx, resident):
    cams = {k: v[0].camera for k, v in resident.items()}
    rng = np.random.default_rng(13)
    arrivals = workloads.poisson_arrivals(12.0, 3.0, seed=13)
    jobs = _with_light_samples([workloads.draw_job(rng, cams, 1000 + i, side_range=(64, 384))
                                for i in range(len(arrivals))])
    assert len(jobs) >= 20
    srv = m.Server(ctx)
    srv.start()
    try:
        t0 = time.perf_counter()
        handles = []
        for t_arr, (k, j) in zip(arrivals, jobs):
            delay = t_arr - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            handles.append(srv.submit(resident[k][1], j))
        for h in handles:
            assert srv.wait(h, 120000)
        lat = np.array([srv.status(h)["done_ms"] - srv.status(h)["submit_ms"] for h in handles])
        films = [srv.result(h, j.width, j.height) for h, (_, j) in zip(handles, jobs)]
    finally:
        srv.stop()
        srv.close()
    same = sum(int(np.array_equal(f, resident[k][1].render(j)[0])) for f, (k, j) in zip(films, jobs))
    print(f"[poisson] jobs={len(jobs)} p50={np.percentile(lat, 50):.1f}ms p99={np.percentile(lat, 99):.1f}ms "
          f"identical={same}/{len(jobs)}", flush=True)
    assert same == len(jobs)
