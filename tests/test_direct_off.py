"""efficientDirect = false (renderers/renderers/PathTracer.cpp:23, 50-52,
70-73) in the oracle: emitted radiance counts at every vertex and
next-event estimation is off.

Closed form: a diffuse floor (kd) under a downward-facing square light (Le)
and a black background.  Without next-event estimation every camera path
hits the floor and bounces once; its radiance is exactly kd * Le (f / pdf =
kd for the cosine-sampled Lambertian) when the bounce sees the light, else
0.  With next-event estimation (the default) the expected radiance is the
same kd * Le * (the light's cosine-weighted solid angle), so the two
estimators' means agree."""
import math

import numpy as np

import milton_amd as m
from oracle import pyoracle as po

KD, LE = 0.6, 2.0


def _scene():
    mats = np.zeros(2, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIFFUSE
    mats[0]["kd"] = KD
    mats[1]["type"] = m.MPT_BSDF_DIFFUSE   # black emitter
    mats[1]["Le"] = LE
    quads = np.array([[0, 0, 0, 40, 0, 0, 0, 0, 40],         # floor, normal +y
                      [0, 1, 0, 4, 0, 0, 0, 0, -4]],         # light, normal -y
                     np.float32)
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [], quads, [0, 1], mats, [0, 0, 0])
    a.camera = {"eye": [0, 0.5, 0], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 40.0, "aspect": 1.0}
    return a


def _paths(a, direct_off, spp=16, side=24):
    job = m.make_job(a.camera, side, side, spp, 4, seed=3)
    job.direct_off = direct_off
    pix = np.repeat(np.arange(side * side, dtype=np.uint32), spp)
    smp = np.tile(np.arange(spp, dtype=np.uint32), side * side)
    return po.OracleScene(a).render_paths(job, pix, smp)


def test_no_nee_counts_emission_at_every_vertex():
    a = _scene()
    L = _paths(a, 1)
    want = float(np.float32(KD)) * float(np.float32(LE))
    hit = (L != 0).any(axis=1)
    assert 0.5 < hit.mean() < 0.95
    assert np.allclose(L[hit], want, rtol=1e-12, atol=0)


def test_both_estimators_have_the_same_mean():
    a = _scene()
    off, on = _paths(a, 1), _paths(a, 0)
    mo, mn = off.mean(axis=0), on.mean(axis=0)
    so, sn = off.std(axis=0) / math.sqrt(len(off)), on.std(axis=0) / math.sqrt(len(on))
    print(f"efficientDirect false {mo} +- {so}, true {mn} +- {sn}", flush=True)
    assert np.all(np.abs(mo - mn) <= 4 * np.sqrt(so ** 2 + sn ** 2) + 1e-12)
    # with next-event estimation the light's own emission after the bounce is
    # not counted again: no path sees kd * Le from a bounce alone
    assert not np.allclose(on, float(np.float32(KD)) * float(np.float32(LE)))


def test_scenefile_efficient_direct_key():
    base = ('{ "scenefile" : { "renderer" : { "type" : "pathTracer" %s }, '
            '"scene" : { "sphere" : { } } } }')
    _, job = m.load_scenefile(text=base % "")
    assert job.direct_off == 0
    _, job = m.load_scenefile(text=base % ', "efficientDirect" : false')
    assert job.direct_off == 1
    _, job = m.load_scenefile(text=base % ', "efficientDirect" : true')
    assert job.direct_off == 0
