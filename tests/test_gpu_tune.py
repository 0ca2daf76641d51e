"""Autotuner (mpt_tune / mpt_tune_scene): the settings it picks are valid,
never slower than the start on its own sample, and change no result: a film
rendered after tuning is bit-identical to the one rendered before (engine
settings), or agrees up to exact closest-hit ties (kd leaf size, which
changes the tree)."""
import numpy as np
import pytest

import milton_amd as m

pytestmark = pytest.mark.gpu


def _film(ctx, scene, job):
    f = m.Film(ctx, job.width, job.height)
    f.render(scene, job)
    out = f.download()
    f.close()
    return out


def _restore(ctx, saved):
    for opt, v in saved.items():
        ctx.set_option(opt, v)


def test_tune_engine_settings_keep_results(ctx):
    opts = (m.MPT_OPT_BATCH_PATHS, m.MPT_OPT_STACK, m.MPT_OPT_PERSISTENT, m.MPT_OPT_TRACE_BLOCK,
            m.MPT_OPT_SHADE_BLOCK)
    saved = {o: ctx.get_option(o) for o in opts}
    try:
        a = m.generate(m.MPT_SCENE_ROOM, 0, 1)
        scene = m.Scene(ctx, a)
        job = m.make_job(a.camera, 96, 96, 16, 5, seed=3)
        before = _film(ctx, scene, job)
        res = scene.tune(job, sample_paths=1 << 17)
        print("tune:", res, flush=True)
        assert res["stack"] in (8, 12, 16)
        assert res["persistent"] in (0, 1, 2)
        assert res["batch_paths"] == saved[m.MPT_OPT_BATCH_PATHS] or res["batch_paths"] in [1 << b for b in range(20, 27)]
        assert res["trials"] >= 2
        assert res["paths_per_s"] >= res["baseline_paths_per_s"] > 0
        assert ctx.get_option(m.MPT_OPT_STACK) == res["stack"]
        assert ctx.get_option(m.MPT_OPT_PERSISTENT) == res["persistent"]
        assert res["trace_block"] in (128, 256) and res["shade_block"] in (256, 512)
        assert ctx.get_option(m.MPT_OPT_TRACE_BLOCK) == res["trace_block"]
        assert ctx.get_option(m.MPT_OPT_SHADE_BLOCK) == res["shade_block"]
        after = _film(ctx, scene, job)
        assert np.array_equal(before, after)
        # every block size renders the same film
        ctx.set_option(m.MPT_OPT_PERSISTENT, 1)
        for tb in (128, 256):
            for sb in (256, 512):
                ctx.set_option(m.MPT_OPT_TRACE_BLOCK, tb)
                ctx.set_option(m.MPT_OPT_SHADE_BLOCK, sb)
                assert np.array_equal(before, _film(ctx, scene, job)), (tb, sb)
        # every persistent/stack form renders the same film
        for p in (0, 1, 2):
            for k in (8, 16):
                ctx.set_option(m.MPT_OPT_PERSISTENT, p)
                ctx.set_option(m.MPT_OPT_STACK, k)
                assert np.array_equal(before, _film(ctx, scene, job)), (p, k)
        scene.close()
    finally:
        _restore(ctx, saved)


def test_tune_scene_leaf_size(ctx):
    opts = (m.MPT_OPT_BATCH_PATHS, m.MPT_OPT_STACK, m.MPT_OPT_PERSISTENT, m.MPT_OPT_TRACE_BLOCK,
            m.MPT_OPT_SHADE_BLOCK)
    saved = {o: ctx.get_option(o) for o in opts}
    try:
        a = m.generate(m.MPT_SCENE_SOUP, 50_000, 1)
        job = m.make_job(a.camera, 160, 90, 4, 4, seed=5)
        ref_scene = m.Scene(ctx, a)
        ref = _film(ctx, ref_scene, job)
        ref_scene.close()
        scene, res = m.Scene.tuned(ctx, a, job, sample_paths=1 << 16)
        print("tune_scene:", res, scene.stats(), flush=True)
        assert 1 <= res["leaf_size"] <= 16
        assert res["paths_per_s"] > 0 and res["trials"] >= 3
        got = _film(ctx, scene, job)
        # another tree may resolve an exact tie the other way in a handful of pixels
        same = np.all(got == ref, axis=2).mean()
        assert same >= 0.995, same
        scene.close()
    finally:
        _restore(ctx, saved)
