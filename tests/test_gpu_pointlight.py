"""Point lights through the device path: the floor of tests/test_pointlight.py
(every path equals the closed form) and the scene-file room with an omni
point light added to its area light, per path against the oracle with the
same seeds."""
import numpy as np
import pytest

import milton_amd as m
from test_gpu_render import _per_path
from test_pointlight import _floor
from test_scenefile import ROOM

pytestmark = pytest.mark.gpu


def test_point_light_floor_paths(ctx):
    a = _floor()
    job = m.make_job(a.camera, 32, 32, 4, 3, seed=5)
    L_gpu, L_orc = _per_path(ctx, a, job, 0, 32 * 32, "point light over a floor")
    assert np.allclose(L_gpu, L_orc, rtol=1e-4)


def test_room_with_point_light_paths(ctx):
    light = ('"material" : { "emitter" : { "type" : "omni", "power" : [ 0.8, 0.7, 0.5 ] },'
             ' "transform" : { "translate" : [ 0.3, 0.7, 0.4 ], "point" : { } } },\n         "material" : {\n'
             '            "bsdf" : { "type" : "diffuse", "kd" : 0.75 },')
    text = ROOM.replace('"material" : {\n            "bsdf" : { "type" : "diffuse", "kd" : 0.75 },', light, 1)
    assert text != ROOM
    arrays, job = m.load_scenefile(text=text)
    assert len(arrays.points) == 1
    _per_path(ctx, arrays, job, 0, job.width * job.height, "scenefile room + point light", min_frac=0.999)


def test_occluder_next_to_point_light(ctx):
    """An occluder 5e-6 below the point light shadows the floor on the
    device too (shadow rays end 2 EPSILON before the light, as in the
    reference), path by path equal to the oracle."""
    from test_pointlight import occluded_near_light
    a = occluded_near_light()
    job = m.make_job(a.camera, 32, 32, 2, 3, seed=8)
    L_gpu, L_orc = _per_path(ctx, a, job, 0, 32 * 32, "occluder next to a point light", min_frac=1.0)
    assert (L_orc == 0).all() and (L_gpu == 0).all()
