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
    """The HBM budget covers the lanes' ray queues and the films: the queues
    take at most half of it, and jobs are admitted while their films fit the
    rest.  Films here are sized so that at most two fit at once."""
    (room, sroom), _ = scenes
    budget = 64 << 20
    srv = m.Server(ctx, hbm_budget=budget)
    mem = srv.memory()
    assert 0 < mem["queue_bytes"] <= budget // 2 and mem["batch_paths"] >= 1024
    avail = budget - mem["queue_bytes"]
    side = int((avail / 2.5 / 16) ** 0.5)
    film_bytes = side * side * 16
    assert 2 * film_bytes <= avail < 3 * film_bytes
    job = m.make_job(room.camera, side, side, 1, 3, seed=3)
    big = m.make_job(room.camera, 2 * side, 2 * side, 1, 3, seed=3)
    with pytest.raises(m.MptError, match="budget"):
        srv.submit(sroom, big)
    handles = [srv.submit(sroom, job) for _ in range(5)]
    srv.cancel(handles[4])
    # one round admits at most two films
    srv.step()
    states = [srv.status(h)["state"] for h in handles]
    assert states.count(m.JOB_RUNNING) + states.count(m.JOB_DONE) <= 2
    srv.drain()
    mem = srv.memory()
    print(f"[budget] {mem}", flush=True)
    assert mem["peak_film_bytes"] <= 2 * film_bytes
    assert mem["peak_film_bytes"] + mem["queue_bytes"] <= budget
    results = []
    for h in handles[:4]:
        s = srv.status(h)
        assert s["state"] == m.JOB_DONE and s["done_ms"] >= s["start_ms"] >= s["submit_ms"]
        results.append(srv.result(h, side, side))
    assert srv.status(handles[4])["state"] == m.JOB_CANCELLED
    with pytest.raises(m.MptError, match="cancelled"):
        srv.result(handles[4], side, side)
    srv.close()
    ref = sroom.render(job)[0]
    assert all(np.array_equal(r, ref) for r in results)


def test_queue_memory_stays_in_budget():
    """On a fresh context, the lanes' workspaces after serving stay within
    the bytes the server reserved for them, also for jobs with several light
    samples per vertex (their batches hold fewer paths)."""
    c = m.Context(0)
    try:
        a = m.generate(m.MPT_SCENE_ROOM)
        a.points = np.array([[0.3, 0.7, 0.4, 0.8, 0.7, 0.5]], np.float32)
        sc = m.Scene(c, a)
        srv = m.Server(c, hbm_budget=96 << 20)
        jobs = []
        for i in range(6):
            j = m.make_job(a.camera, 96, 80, 16, 5, seed=40 + i, job_id=i)
            j.direct_samples = 1 + 3 * (i % 2)
            jobs.append(j)
        hs = [srv.submit(sc, j) for j in jobs]
        srv.drain()
        mem = srv.memory()
        print(f"[queues] {mem}", flush=True)
        assert mem["lane_bytes"] <= mem["queue_bytes"]
        films = [srv.result(h, j.width, j.height) for h, j in zip(hs, jobs)]
        srv.close()
        assert all(np.array_equal(f, sc.render(j)[0]) for f, j in zip(films, jobs))
    finally:
        c.close()


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


@pytest.mark.parametrize("worker", [False, True], ids=["inline", "worker"])
def test_injected_batch_failure(ctx, scenes, worker):
    """A batch that fails (MPT_OPT_FAIL_BATCH makes the third batch from now
    on fail before its kernels are enqueued) fails the jobs it carried and
    only those: the others finish with the same films as alone, their films
    are released, and the server takes new jobs afterwards."""
    (room, sroom), _ = scenes
    jobs = [m.make_job(room.camera, 48, 40, 4, 4, seed=60 + i, job_id=i) for i in range(6)]
    alone = [sroom.render(j)[0] for j in jobs]
    saved = ctx.get_option(m.MPT_OPT_BATCH_PATHS)
    ctx.set_option(m.MPT_OPT_BATCH_PATHS, 4096)   # about two jobs per batch
    try:
        srv = m.Server(ctx)
        ctx.set_option(m.MPT_OPT_FAIL_BATCH, 3)
        if worker:
            srv.start()
        handles = [srv.submit(sroom, j) for j in jobs]
        if worker:
            for h in handles:
                assert srv.wait(h, 120000)
        else:
            with pytest.raises(m.MptError, match="injected"):
                srv.drain()
        assert ctx.get_option(m.MPT_OPT_FAIL_BATCH) == 0
        states = [srv.status(h)["state"] for h in handles]
        print(f"[failure {'worker' if worker else 'inline'}] states={states}", flush=True)
        assert states.count(m.JOB_FAILED) >= 1 and states.count(m.JOB_DONE) >= 1
        assert set(states) <= {m.JOB_FAILED, m.JOB_DONE}
        for h, j, ref, s in zip(handles, jobs, alone, states):
            if s == m.JOB_DONE:
                assert np.array_equal(srv.result(h, j.width, j.height), ref)
            else:
                assert srv.status(h)["error"] == 2   # -MPT_E_DEVICE
                with pytest.raises(m.MptError):
                    srv.result(h, j.width, j.height)
        if worker:
            srv.stop()
        assert srv.memory()["film_bytes"] == 0
        # the server keeps serving
        h = srv.submit(sroom, jobs[0])
        srv.drain()
        assert np.array_equal(srv.result(h, jobs[0].width, jobs[0].height), alone[0])
        srv.close()
    finally:
        ctx.set_option(m.MPT_OPT_FAIL_BATCH, 0)
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved)
