"""Batching and serving: jobs of uneven size share batches, yet every job's
film is bit-identical to the same job rendered alone (per-path random
streams keyed by the job, whole pixels per batch); admission respects the HBM
budget; cancellation and failures are reported per job."""
import time

import numpy as np
import pytest

import milton_amd as m

pytestmark = pytest.mark.gpu


def _jobs(arrays, n, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(np.exp(rng.uniform(np.log(16), np.log(96))))
        spp = int(round(np.exp(rng.uniform(np.log(1), np.log(49))) ** 0.5)) ** 2
        depth = int(rng.integers(3, 7))
        out.append(m.make_job(arrays.camera, side, side, max(1, spp), depth, seed=int(rng.integers(1, 1000)),
                              job_id=i))
    return out


@pytest.fixture(scope="module")
def scenes(ctx):
    room = m.generate(m.MPT_SCENE_ROOM)
    soup = m.generate(m.MPT_SCENE_SOUP, 100_000, 1)
    return (room, m.Scene(ctx, room)), (soup, m.Scene(ctx, soup))


def test_mixed_batches_match_alone(ctx, scenes):
    (room, sroom), (soup, ssoup) = scenes
    jobs = [(sroom, j) for j in _jobs(room, 10)] + [(ssoup, j) for j in _jobs(soup, 10, seed=12)]
    alone = [sc.render(j)[0] for sc, j in jobs]
    saved = ctx.get_option(m.MPT_OPT_BATCH_PATHS)
    ctx.set_option(m.MPT_OPT_BATCH_PATHS, 1 << 16)   # force many mixed batches
    try:
        srv = m.Server(ctx)
        handles = [srv.submit(sc, j) for sc, j in jobs]
        srv.drain()
        st = srv.stats()
        for h, (sc, j), ref in zip(handles, jobs, alone):
            assert srv.status(h)["state"] == m.JOB_DONE
            film = srv.result(h, j.width, j.height)
            assert np.array_equal(film, ref), f"job {j.job_id} differs when batched"
        alone_batches = sum(-(-(jj.width * jj.height * jj.spp) // (1 << 16)) for _, jj in jobs)
        print(f"[mix] jobs={len(jobs)} batches={st['batches']} (alone: {alone_batches}) paths={st['paths']}",
              flush=True)
        assert st["batches"] < alone_batches
        srv.close()
    finally:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved)


def test_budget_admission_and_cancel(ctx, scenes):
    (room, sroom), _ = scenes
    job = m.make_job(room.camera, 64, 64, 4, 4, seed=3)
    film_bytes = 64 * 64 * 16
    srv = m.Server(ctx, hbm_budget=2 * film_bytes)
    big = m.make_job(room.camera, 128, 128, 4, 4, seed=3)
    with pytest.raises(m.MptError, match="budget"):
        srv.submit(sroom, big)
    handles = [srv.submit(sroom, job) for _ in range(5)]
    srv.cancel(handles[4])
    # one round admits at most two films
    srv.step()
    states = [srv.status(h)["state"] for h in handles]
    assert states.count(m.JOB_RUNNING) + states.count(m.JOB_DONE) <= 2
    srv.drain()
    st = srv.stats()
    assert st["peak_film_bytes"] <= 2 * film_bytes
    ref = sroom.render(job)[0]
    for h in handles[:4]:
        s = srv.status(h)
        assert s["state"] == m.JOB_DONE and s["done_ms"] >= s["start_ms"] >= s["submit_ms"]
        assert np.array_equal(srv.result(h, 64, 64), ref)
    assert srv.status(handles[4])["state"] == m.JOB_CANCELLED
    with pytest.raises(m.MptError, match="cancelled"):
        srv.result(handles[4], 64, 64)
    srv.close()


def test_worker_thread_serving(ctx, scenes):
    """Jobs submitted while others run are admitted and finish; latencies are
    recorded per job."""
    (room, sroom), (soup, ssoup) = scenes
    srv = m.Server(ctx)
    srv.start()
    try:
        handles = []
        jobs = _jobs(soup, 12, seed=5)
        for j in jobs:
            handles.append(srv.submit(ssoup, j))
            time.sleep(0.002)
        for h in handles:
            assert srv.wait(h, 120000)
        lat = [srv.status(h)["done_ms"] - srv.status(h)["submit_ms"] for h in handles]
        assert all(l >= 0 for l in lat)
        for h, j in zip(handles, jobs):
            film = srv.result(h, j.width, j.height)
            assert np.isfinite(film).all()
    finally:
        srv.stop()
        srv.close()
