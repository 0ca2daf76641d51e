"""Point lights on the CPU: the oracle's restatement of DirectIllumination's
point-light branch (renderers/DirectIllumination.cpp:96-121 with isPoint: one
sample, no area weight, cos at the light 1 since a PointShape has no normal,
core/SurfacePoint.h:143) with OmniEmitter's radiance P / (4 pi)
(emitters/OmniEmitter.cpp:15-17), and the scene file's "point" shape.

Pinned by the closed form of a diffuse floor under one point light and a
black background: every path's radiance is exactly
kd / pi * P / (4 pi) * cos / r^2, cos = h / r, at the point its camera ray
hits (the bounce after it escapes to the black background).
"""
import math

import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po

KD, H = 0.6, 1.5
P = np.array([4.0, 2.0, 1.0])
LIGHT = np.array([0.2, H, -0.1])


def _floor():
    mat = np.zeros(1, m.MATERIAL_DTYPE)
    mat["type"] = m.MPT_BSDF_DIFFUSE
    mat["kd"] = KD
    quad = np.array([[0, 0, 0, 20, 0, 0, 0, 0, 20]], np.float32)   # normal ez x ex = +y
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [], quad, [0], mat, [0, 0, 0],
                      points=[[*LIGHT, *P]])
    a.camera = {"eye": [0.3, 3.0, -0.2], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 50.0, "aspect": 1.0}
    return a


def test_point_light_closed_form():
    a = _floor()
    job = m.make_job(a.camera, 16, 16, 4, 3, seed=5)
    pix = np.repeat(np.arange(256, dtype=np.uint32), 4)
    smp = np.tile(np.arange(4, dtype=np.uint32), 256)
    L = po.OracleScene(a).render_paths(job, pix, smp)
    want = np.zeros_like(L)
    lp = LIGHT.astype(np.float32).astype(np.float64)
    pw = P.astype(np.float32).astype(np.float64)
    for i, (px, s) in enumerate(zip(pix, smp)):
        r = po.camera_ray(job, int(px), int(s))
        t = -r[1] / r[4]
        x = r[:3] + t * r[3:]
        w = lp - x
        dist = np.linalg.norm(w)
        want[i] = float(np.float32(KD)) / math.pi * pw / (4 * math.pi) * (w[1] / dist) / dist ** 2   # kd is FP32
    assert np.allclose(L, want, rtol=1e-9, atol=0)
    assert L.min() > 0


def test_point_light_is_shadowed():
    """A large quad between the light and the floor leaves the floor black."""
    a = _floor()
    a.quads = np.concatenate([a.quads, np.float32([[0, 1.0, 0, 20, 0, 0, 0, 0, 20]])])
    a.quad_material = np.concatenate([a.quad_material, np.uint32([0])])
    a.camera = dict(a.camera, eye=[0.3, 0.5, -0.2])
    job = m.make_job(a.camera, 8, 8, 2, 3, seed=6)
    pix = np.repeat(np.arange(64, dtype=np.uint32), 2)
    smp = np.tile(np.arange(2, dtype=np.uint32), 64)
    L = po.OracleScene(a).render_paths(job, pix, smp)
    assert (L == 0).all()


def test_scenefile_point_light():
    base = '{ "scenefile" : { "scene" : { "sphere" : { }, %s } } }'
    a, _ = m.load_scenefile(text=base % '"material" : { "emitter" : { "type" : "omni", "power" : [3, 2, 1] },'
                                        ' "transform" : { "translate" : [1, 2, 3], "point" : { } } }')
    assert np.allclose(a.points, [[1, 2, 3, 3, 2, 1]])
    a, _ = m.load_scenefile(text=base % '"point" : { }')
    assert len(a.points) == 0
    for bad, word in (('"material" : { "emitter" : { "type" : "oriented", "power" : 1 }, "point" : { } }', "omni"),
                      ('"material" : { "emitter" : { "type" : "omni", "power" : 1 }, "plane" : { } }', "point")):
        with pytest.raises(m.MptError) as e:
            m.load_scenefile(text=base % bad)
        assert word in str(e.value)


def occluded_near_light():
    """The floor with a small quad 5e-6 below the point light, facing down:
    it shadows the whole floor in the reference's FP64 window
    t < dist - 2 EPSILON (DirectIllumination.cpp:123, Scene.cpp:113)."""
    a = _floor()
    gap = 5e-6
    a.quads = np.concatenate([a.quads, np.float32([[LIGHT[0], LIGHT[1] - gap, LIGHT[2], 0.01, 0, 0, 0, 0, -0.01]])])
    a.quad_material = np.concatenate([a.quad_material, np.uint32([0])])
    assert np.float32(LIGHT[1] - gap) < np.float32(LIGHT[1])
    return a


def test_occluder_next_to_point_light():
    a = occluded_near_light()
    job = m.make_job(a.camera, 16, 16, 2, 3, seed=8)
    pix = np.repeat(np.arange(256, dtype=np.uint32), 2)
    smp = np.tile(np.arange(2, dtype=np.uint32), 256)
    L = po.OracleScene(a).render_paths(job, pix, smp)
    assert (L == 0).all()
