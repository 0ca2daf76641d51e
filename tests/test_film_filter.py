"""Film reconstruction filters on the CPU: the oracle's restatement of
ReconstructionRenderOutput::addSample (renderers/outputs/
ReconstructionRenderOutput.cpp:30-83) and the scene file's output filter.

Pinned by closed forms of the splat geometry.  A sample at (X, Y) inside
pixel (c, r) reaches the integer corners of the 3 x 3 pixels around (c, r);
with a constant radiance every filtered pixel equals that constant (the
denominator normalises).  For samples uniform in their pixel:
  * box, support 2 (a unit disc around each corner): only the four corners
    of the sample's own pixel can be within distance 1, each with
    probability pi/4, so the weights per sample average pi;
  * triangle, support 2 (1 - dx)(1 - dy) on the signed offsets: the three
    column weights are 2 + fx, 1 + fx and fx, so the weights per sample
    average (3 + 3/2)^2 = 20.25;
  * no filter: each pixel's denominator is its sample count.
"""
import math

import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po


def _far_scene(env=(0.7, 1.3, 2.0)):
    """Every camera ray misses: L = the environment on every path."""
    mat = np.zeros(1, m.MATERIAL_DTYPE)
    mat["kd"] = 0.5
    quad = np.array([[100, 100, 100, 1, 0, 0, 0, 0, 1]], np.float32)
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [], quad, [0], mat, env)
    a.camera = {"eye": [0, 0, 0], "look": [0, 0, -1], "up": [0, 1, 0], "height_angle": 40.0, "aspect": 1.0}
    return a


def _interior_weight_mean(film, spp):
    """Total splat weight per sample over the interior pixels' own samples:
    sum of denominators of the interior divided by the interior's samples."""
    den = film[..., 3]
    return den[2:-2, 2:-2].sum() / (den[2:-2, 2:-2].size * spp)


@pytest.mark.parametrize("kind, support", [(None, 0), (m.MPT_FILTER_BOX, 2), (m.MPT_FILTER_TRIANGLE, 2),
                                           (m.MPT_FILTER_GAUSSIAN, 2), (m.MPT_FILTER_MITCHELL, 2),
                                           (m.MPT_FILTER_LANCZOS, 3)])
def test_constant_radiance_is_preserved(kind, support):
    a = _far_scene()
    ff = None if kind is None else (kind, support, (0.0, 0.0))
    job = m.make_job(a.camera, 12, 10, 9, 3, seed=4, film_filter=ff)
    film = po.OracleScene(a).render_film(job)
    assert (film[..., 3] > 0).all()
    val = film[..., :3] / film[..., 3:]
    assert np.allclose(val, np.float32([0.7, 1.3, 2.0]).astype(np.float64), rtol=1e-12)   # the scene holds FP32
    if kind is None:
        assert (film[..., 3] == 9).all()


def test_box_weights_average_pi():
    a = _far_scene()
    job = m.make_job(a.camera, 40, 40, 64, 3, seed=7, film_filter=(m.MPT_FILTER_BOX, 2))
    w = _interior_weight_mean(po.OracleScene(a).render_film(job), 64)
    assert w == pytest.approx(math.pi, rel=0.01)


def test_triangle_weights_average():
    a = _far_scene()
    job = m.make_job(a.camera, 40, 40, 64, 3, seed=8, film_filter=(m.MPT_FILTER_TRIANGLE, 2))
    w = _interior_weight_mean(po.OracleScene(a).render_film(job), 64)
    assert w == pytest.approx(20.25, rel=0.01)


def test_unfiltered_film_matches_pixel_sums():
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(arrays.camera, 16, 16, 4, 4, seed=5)
    osc = po.OracleScene(arrays)
    film = osc.render_film(job)
    s, _, _ = osc.render(job)
    assert np.allclose(film[..., :3].reshape(-1, 3), s, rtol=1e-12)


def test_scenefile_output_filter():
    base = '{ "scenefile" : { "output" : { "width" : 8, "height" : 8 %s }, "scene" : { "sphere" : { } } } }'
    _, j = m.load_scenefile(text=base % ', "filter" : { "type" : "gaussian", "support" : 3, "sigma" : 0.5 }')
    assert (j.filter, j.filter_support) == (1 + m.MPT_FILTER_GAUSSIAN, 3) and j.filter_params[0] == 0.5
    _, j = m.load_scenefile(text=base % ', "filter" : { "type" : "mitchell" }')
    assert j.filter == 1 + m.MPT_FILTER_MITCHELL and j.filter_params[0] == pytest.approx(1 / 3)
    _, j = m.load_scenefile(text=base % ', "type" : "naive", "filter" : { "type" : "box" }')
    assert j.filter == 0
    _, j = m.load_scenefile(text=base % '')
    assert j.filter == 0
    for bad in (', "filter" : { "type" : "dynamic" }', ', "filter" : { }', ', "type" : "dynamic"'):
        with pytest.raises(m.MptError):
            m.load_scenefile(text=base % bad)
