"""The oracle's Metropolis light transport (renderers/mlt/) pinned without
the device: closed forms of a light that fills the view (the seed weights'
mean b, the film's mean), the proposal filter against scipy's median
filter, the lens stratification's even coverage and the aggregate's
mutation frequencies."""
import types

import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po

LE = (0.5, 1.0, 2.0)


def fullscreen_light():
    """A 30 x 30 emitting quad (normal +y) under a camera looking down: every
    camera ray hits it."""
    mats = np.zeros(1, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIFFUSE
    mats[0]["Le"] = LE
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [],
                      np.asarray([[0, 0, 0, 30, 0, 0, 0, 0, 30]], np.float32), [0], mats, [0, 0, 0])
    a.camera = {"eye": [0.1, 2.0, 0.05], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 40.0, "aspect": 1.0}
    return a


def _job(a, w, h, **kw):
    p = dict(chains=16, mutations=64, seed_paths=1000, max_depth=10, max_rejections=500, filter_proposed=True,
             probs=(1.0, 1.0, 1.0))
    p.update(kw)
    base = m.make_job(a.camera, w, h, 1, 1, seed=kw.get("seed", 3))
    return types.SimpleNamespace(camera=base.camera, width=w, height=h, spp=1, max_depth=1, seed=base.seed,
                                 job_id=0, renderer=m.MPT_RENDERER_MLT, mlt_chains=p["chains"],
                                 mlt_mutations=p["mutations"], mlt_seed_paths=p["seed_paths"],
                                 mlt_max_depth=p["max_depth"], mlt_max_rejections=p["max_rejections"],
                                 mlt_flags=0 if p["filter_proposed"] else 1, mlt_probs=p["probs"])


LUM = 0.2126 * LE[0] + 0.7152 * LE[1] + 0.0722 * LE[2]


def test_fullscreen_light_seed_weights():
    """Each generated path is (emitter, camera): its splits are (0, 2), a
    direct view with contribution Le / pA (pA = 1 / (W H), the film bin's
    area), and (1, 1) and (2, 0), which a pinhole cannot take: three entries,
    one of weight lum(Le) W H.  So b = lum(Le) W H / 3."""
    a = fullscreen_light()
    W, H = 8, 6
    sd = po.OracleScene(a).mlt_seeds(_job(a, W, H))
    assert len(sd["w"]) == 3 * 1000
    assert (sd["len"] == 2).all() or ((sd["len"] == 2) | (sd["len"] == 0)).all()
    assert abs(sd["b"] - LUM * W * H / 3) <= 1e-9 * sd["b"]
    assert (np.count_nonzero(sd["w"]) == 1000) and np.allclose(sd["w"][sd["w"] > 0], 1.0 / 1000)


@pytest.mark.parametrize("probs", [(1.0, 1.0, 0.0), (1.0, 1.0, 1.0), (0.0, 1.0, 0.0)])
def test_fullscreen_light_film(probs):
    """Every step adds samples of total weight 1 whose value has luminance
    b, so the film's sum is 2 b (mutations / samples) W H / ... : its mean is
    2/3 Le (mutations / samples) per channel, whatever the mutations."""
    a = fullscreen_light()
    W, H = 8, 6
    job = _job(a, W, H, chains=16, mutations=W * H * 10, max_rejections=0, filter_proposed=False, probs=probs)
    film, num, prop, st = po.OracleScene(a).mlt_render(job)
    want = 2.0 / 3.0 * np.asarray(LE) * (st["mutations"] / st["samples"])
    assert np.allclose(film.reshape(-1, 3).mean(axis=0), want, rtol=1e-9)
    # the lens and perturbation mutations always land on the light and are
    # accepted; the bidirectional one keeps the path or fails
    if probs == (0.0, 1.0, 0.0):
        assert st["accepted"] == st["mutations"]


def test_lens_stratification_is_even():
    """One chain's lens mutations on the full-view light (every proposal
    accepted): the stratification (MLTLensSubpathMutation.cpp:65-103) keeps
    each bin's proposals within a few of the mean, where independent film
    points would spread about +-10 around 20."""
    a = fullscreen_light()
    W, H = 8, 6
    job = _job(a, W, H, chains=1, mutations=W * H * 20, max_rejections=0, filter_proposed=False,
               probs=(0.0, 1.0, 0.0))
    film, num, prop, st = po.OracleScene(a).mlt_render(job)
    assert prop.sum() == W * H * 20
    assert prop.max() - prop.min() <= 8 and prop.min() >= 14


def test_proposal_median_filter_is_scipys():
    """RenderOutput's MLT film (RenderOutput.cpp:153-200): the numerator
    times median(3 x 3 proposals, edges clamped) / proposals times
    2 / samples."""
    from scipy.ndimage import median_filter
    a = m.generate(m.MPT_SCENE_ROOM)
    job = _job(a, 12, 10, chains=32, mutations=200, seed_paths=500)
    film, num, prop, st = po.OracleScene(a).mlt_render(job)
    med = median_filter(prop.astype(np.float64), size=3, mode="nearest")
    scale = np.where((med > 0) & (prop > 0), med / np.maximum(prop, 1), 1.0)
    want = num * (scale * 2.0 / st["samples"])[..., None]
    assert np.allclose(film, want, rtol=1e-12, atol=1e-300)
    job2 = _job(a, 12, 10, chains=32, mutations=200, seed_paths=500, filter_proposed=False)
    film2, num2, _, st2 = po.OracleScene(a).mlt_render(job2)
    assert np.allclose(film2, num2 * 2.0 / st2["samples"], rtol=1e-12)


def test_mutation_frequencies():
    """MLTAggregatePathMutation picks each enabled mutation with its
    normalised weight."""
    a = m.generate(m.MPT_SCENE_ROOM)
    for probs, want in (((1.0, 1.0, 1.0), (1 / 3, 1 / 3, 1 / 3)), ((2.0, 0.0, -1.0), (2 / 3, 0, 1 / 3)),
                        ((0.0, 0.0, 0.0), (1 / 3, 1 / 3, 1 / 3))):
        job = _job(a, 16, 12, chains=64, mutations=100, seed_paths=500, probs=probs)
        rec = po.OracleScene(a).mlt_trace(job, 0, 64, 100)
        f = np.bincount(rec[..., 0].astype(int).ravel(), minlength=3) / rec[..., 0].size
        assert np.allclose(f, want, atol=0.02), (probs, f)
