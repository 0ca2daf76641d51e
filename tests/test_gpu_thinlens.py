"""The thin-lens camera through the device path: the room rendered with a
lens, per path against the oracle with the same seeds (the lens draws its
own random stream on both sides), with an explicit focal distance and with
autofocus (the device traces the focus ray at the start of the render)."""
import numpy as np
import pytest

import milton_amd as m
from test_gpu_render import _per_path

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lens", [dict(aperture=0.05, focal_length=0.2, focal_distance=2.5),
                                  dict(aperture=0.08, focal_length=0.3, focal_distance=7.0, autofocus=True),
                                  dict(aperture=0.03, focal_length=0.4, focal_distance=1.0, autofocus=True,
                                       focal_point=(0.2, 0.8))])
def test_thin_lens_paths_same_seed(ctx, lens):
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(dict(arrays.camera, **lens), 48, 48, 4, 5, seed=13)
    _per_path(ctx, arrays, job, 0, 48 * 48, f"thin lens {lens}")


def test_thin_lens_film_matches_paths(ctx):
    """The film of a thin-lens job is the per-pixel mean of its paths (the
    render and the path hook resolve the same focus)."""
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(dict(arrays.camera, aperture=0.08, focal_length=0.3, autofocus=True), 32, 32, 4, 5, seed=2)
    sc = m.Scene(ctx, arrays)
    L = sc.render_paths(job, 0, 32 * 32).astype(np.float64).reshape(32 * 32, 4, 3).mean(axis=1)
    f = m.Film(ctx, 32, 32)
    f.render(sc, job)
    img = f.download().reshape(-1, 3)
    f.close()
    assert np.allclose(img, L, rtol=1e-5, atol=1e-6)
