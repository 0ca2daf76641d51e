"""efficientDirect = false on the device (emission at every vertex, no
next-event estimation; renderers/renderers/PathTracer.cpp:50-52, 70-73):
path by path against the oracle on the room and on the floor-under-a-light
closed form of tests/test_direct_off.py, and mixed in one server batch with
jobs of the default estimator, bit-identical to each job alone."""
import numpy as np
import pytest

import milton_amd as m
from test_direct_off import KD, LE, _scene
from test_gpu_render import _per_path

pytestmark = pytest.mark.gpu


def test_room_paths_without_nee(ctx):
    a = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(a.camera, 64, 64, 16, 6, seed=5)
    job.direct_off = 1
    L_gpu, _ = _per_path(ctx, a, job, 0, 64 * 64, "room, efficientDirect false", min_frac=0.999)
    st = m.Scene(ctx, a).render(job)[1]
    assert st["shadow_rays"] == 0


def test_floor_closed_form_without_nee(ctx):
    a = _scene()
    job = m.make_job(a.camera, 24, 24, 16, 4, seed=3)
    job.direct_off = 1
    L = m.Scene(ctx, a).render_paths(job, 0, 24 * 24).astype(np.float64)
    want = float(np.float32(KD) * np.float32(LE))
    hit = (L != 0).any(axis=1)
    assert 0.5 < hit.mean() < 0.95
    assert np.allclose(L[hit], want, rtol=1e-6)


def test_mixed_estimators_alone_equals_batched(ctx):
    a = m.generate(m.MPT_SCENE_ROOM)
    sc = m.Scene(ctx, a)
    jobs = []
    for i in range(6):
        j = m.make_job(a.camera, 40 + 8 * i, 40, 4, 5, seed=10 + i, job_id=i)
        j.direct_off = i % 2
        jobs.append(j)
    alone = [sc.render(j)[0] for j in jobs]
    saved = ctx.get_option(m.MPT_OPT_BATCH_PATHS)
    ctx.set_option(m.MPT_OPT_BATCH_PATHS, 1 << 14)
    try:
        srv = m.Server(ctx)
        hs = [srv.submit(sc, j) for j in jobs]
        srv.drain()
        assert all(np.array_equal(srv.result(h, j.width, j.height), ref) for h, j, ref in zip(hs, jobs, alone))
        srv.close()
    finally:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved)
    sc.close()
