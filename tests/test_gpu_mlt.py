"""Metropolis light transport on the device (MPT_RENDERER_MLT) against the
oracle's restatement of MLTRenderer / MLTMarkovProcess and the three
mutations (renderers/mlt/): chain by chain and mutation by mutation on the
same random streams, the films they make, and the closed form of a light
that fills the view."""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from test_mlt import fullscreen_light, LE

pytestmark = pytest.mark.gpu
MLT = m.MPT_RENDERER_MLT


def _job(a, w, h, seed=3, **mlt):
    return m.make_job(a.camera, w, h, 1, 1, seed=seed, renderer=MLT, mlt=mlt)


def _compare_traces(ctx, a, job, steps, what, min_chains, b_rtol=1e-3):
    """Device and oracle records of the first `steps` mutations of every
    chain; returns the share of chains whose records all agree."""
    sc = m.Scene(ctx, a)
    dev, b_dev, n_dev = sc.mlt_trace(job, steps)
    orc = po.OracleScene(a)
    sd = orc.mlt_seeds(job)
    ref = orc.mlt_trace(job, 0, job.mlt_chains, steps)
    assert n_dev == len(sd["w"]) and abs(b_dev - sd["b"]) <= b_rtol * sd["b"], (what, b_dev, sd["b"])
    same = ((dev[..., 0] == ref[..., 0]) & (dev[..., 2] == ref[..., 2]) & (dev[..., 3] == ref[..., 3]) &
            (dev[..., 4] == ref[..., 4]) & (dev[..., 5] == ref[..., 5]) &
            (np.abs(dev[..., 1] - ref[..., 1]) <= 1e-3 * np.maximum(1e-3, np.abs(ref[..., 1]))))
    chains = same.all(axis=1)
    # a chain's first disagreement, by mutation type
    first = np.argmin(same, axis=1)
    kinds = [int(ref[c, first[c], 0]) for c in np.nonzero(~chains)[0]]
    print(f"[{what}] chains agreeing over {steps} mutations: {chains.mean():.4f}; steps agreeing "
          f"{same.mean():.5f}; first disagreement by mutation {np.bincount(kinds, minlength=3).tolist()}; "
          f"acceptance dev {dev[..., 2].mean():.4f} oracle {ref[..., 2].mean():.4f}", flush=True)
    assert chains.mean() >= min_chains, what
    return dev, ref


def test_fullscreen_light_closed_form(ctx):
    """Every seed path is emitter-camera: b = lum(Le) W H / 3 and the film's
    mean is 2/3 Le times mutations / samples, as in the oracle."""
    a = fullscreen_light()
    W, H = 8, 6
    job = _job(a, W, H, chains=64, mutations=W * H * 4, seed_paths=1000, max_rejections=0, filter_proposed=False,
               probs=(1.0, 1.0, 0.0))
    film, st = m.Scene(ctx, a).render(job)
    lum = 0.2126 * LE[0] + 0.7152 * LE[1] + 0.0722 * LE[2]
    assert abs(st["mlt_b"] - lum * W * H / 3) <= 1e-6 * st["mlt_b"]
    want = 2.0 / 3.0 * np.asarray(LE) * (st["paths"] / st["mlt_samples"])
    assert np.allclose(film.reshape(-1, 3).mean(axis=0), want, rtol=1e-5)


def test_room_traces_match_oracle(ctx):
    a = m.generate(m.MPT_SCENE_ROOM)
    job = _job(a, 32, 24, chains=256, mutations=16, seed_paths=2000, max_depth=10)
    _compare_traces(ctx, a, job, 16, "room", 0.9)


@pytest.mark.parametrize("what,kw", [
    ("bidirectional mutations only", dict(probs=(1.0, 0.0, 0.0))),
    ("lens mutations only", dict(probs=(0.0, 1.0, 0.0))),
    ("perturbations only", dict(probs=(0.0, 0.0, 1.0))),
])
def test_room_traces_per_setting(ctx, what, kw):
    """Each mutation of the aggregate alone."""
    a = m.generate(m.MPT_SCENE_ROOM)
    job = _job(a, 32, 24, chains=256, mutations=16, seed_paths=2000, max_depth=10, **kw)
    _compare_traces(ctx, a, job, 16, f"room, {what}", 0.9)


@pytest.mark.parametrize("what,kw", [("no proposal filter", dict(filter_proposed=False)),
                                     ("3 rejections", dict(max_rejections=3))])
def test_room_film_per_setting(ctx, what, kw):
    """mltFilterProposed off (the film is the numerators times 2 / samples)
    and mltMaxConsequtiveRejections 3 (MLTMarkovProcess.cpp:97-118: a chain
    that has stayed on its path that many times in a row adds no samples
    until it moves), against the oracle's film and sample count."""
    a = m.generate(m.MPT_SCENE_ROOM)
    job = _job(a, 24, 16, chains=128, mutations=64, seed_paths=2000, **kw)
    film, st = m.Scene(ctx, a).render(job)
    orc, _, _, ost = po.OracleScene(a).mlt_render(job)
    print(f"[mlt room film, {what}] dev {film.mean(axis=(0, 1))} oracle {orc.mean(axis=(0, 1))} "
          f"samples {st['mlt_samples']} / {ost['samples']}", flush=True)
    if "max_rejections" in kw:
        # the limit withholds most samples of the default job (limit 500)
        base = po.OracleScene(a).mlt_render(_job(a, 24, 16, chains=128, mutations=64, seed_paths=2000))[3]
        assert ost["samples"] < 0.5 * base["samples"]
    assert abs(st["mlt_samples"] - ost["samples"]) <= 0.02 * ost["samples"]
    assert np.allclose(film.mean(axis=(0, 1)), orc.mean(axis=(0, 1)), rtol=0.05)


def test_room_film_matches_oracle_and_is_deterministic(ctx):
    a = m.generate(m.MPT_SCENE_ROOM)
    job = _job(a, 24, 16, chains=128, mutations=64, seed_paths=2000)
    sc = m.Scene(ctx, a)
    film, st = sc.render(job)
    again, _ = sc.render(job)
    assert np.array_equal(film, again)
    orc, _, _, ost = po.OracleScene(a).mlt_render(job)
    print(f"[mlt room film] dev {film.mean(axis=(0, 1))} oracle {orc.mean(axis=(0, 1))} "
          f"samples {st['mlt_samples']} / {ost['samples']} accepted {st['mlt_accepted']} / {ost['accepted']}",
          flush=True)
    assert abs(st["mlt_samples"] - ost["samples"]) <= 0.02 * ost["samples"]
    assert np.allclose(film.mean(axis=(0, 1)), orc.mean(axis=(0, 1)), rtol=0.05)


def test_mixed_room_traces_match_oracle(ctx):
    """Phong lobes, a glass sphere (caustic and multi-chain perturbations,
    dielectric events that keep their mode) and a point light."""
    from test_gpu_bdpt import _room_mixed
    a = _room_mixed()
    job = _job(a, 32, 24, chains=256, mutations=24, seed_paths=2000, max_depth=10)
    _compare_traces(ctx, a, job, 24, "room, Phong, glass, point light", 0.9)


def test_terrain_traces_match_oracle(ctx):
    """The terrain's Phong (exponent 50) and glass.  A few seed splits there
    weigh about 1e143: the pdfs of vertices whose event lobe is the specular
    one (ModifiedPhongBSDF::getPd by event index) are cosines near 1e-5 to the
    50th power, which FP32 positions change by tens of percent, and b is the
    weights' mean; so b agrees to 50% while the chains agree step by step."""
    a = m.generate(m.MPT_SCENE_TERRAIN, 256, 1)
    job = _job(a, 32, 24, chains=256, mutations=16, seed_paths=1000, max_depth=8)
    _compare_traces(ctx, a, job, 16, "terrain 256^2", 0.9, b_rtol=0.5)


def test_chains_are_independent_of_each_other(ctx):
    """A chain's records depend only on its index: the first 64 chains of a
    256-chain job step exactly as a 64-chain job's, and runs repeat."""
    a = m.generate(m.MPT_SCENE_ROOM)
    sc = m.Scene(ctx, a)
    big, _, _ = sc.mlt_trace(_job(a, 32, 24, chains=256, mutations=12, seed_paths=1000), 12)
    small, _, _ = sc.mlt_trace(_job(a, 32, 24, chains=64, mutations=12, seed_paths=1000), 12)
    again, _, _ = sc.mlt_trace(_job(a, 32, 24, chains=64, mutations=12, seed_paths=1000), 12)
    assert np.array_equal(big[:64], small) and np.array_equal(small, again)


def test_refused_jobs(ctx):
    room = m.generate(m.MPT_SCENE_ROOM)
    sr = m.Scene(ctx, room)
    soup = m.generate(m.MPT_SCENE_SOUP, 1000, 1)   # lit by the environment only
    with pytest.raises(m.MptError, match="light"):
        m.Scene(ctx, soup).render(_job(soup, 8, 8, chains=4, mutations=4, seed_paths=16))
    lens = dict(room.camera, aperture=0.05, focal_length=0.05, focal_distance=2.0)
    with pytest.raises(m.MptError, match="pinhole"):
        sr.render(m.make_job(lens, 8, 8, 1, 1, renderer=MLT, mlt=dict(chains=4, mutations=4)))
    for bad, what in ((dict(max_depth=17), "mlt_max_depth"), (dict(max_depth=1), "mlt_max_depth"),
                      (dict(chains=0), "mlt_chains"), (dict(mutations=40000), "32767"),
                      (dict(chains=1 << 18, mutations=1 << 14), "2\\^32")):
        p = dict(chains=4, mutations=4, seed_paths=16)
        p.update(bad)
        with pytest.raises(m.MptError, match=what):
            sr.render(_job(room, 8, 8, **p))
    with pytest.raises(m.MptError, match="whole films"):
        sr.render_paths(_job(room, 8, 8, chains=4, mutations=4, seed_paths=16), 0, 8)


def test_server_refuses_metropolis_jobs(ctx):
    """A Metropolis job's chains splat over the whole film: the server,
    which batches pixel ranges of many jobs, refuses it."""
    room = m.generate(m.MPT_SCENE_ROOM)
    job = _job(room, 16, 12, chains=64, mutations=32, seed_paths=500)
    srv = m.Server(ctx)
    try:
        with pytest.raises(m.MptError, match="server"):
            srv.submit(m.Scene(ctx, room), job)
    finally:
        srv.close()
