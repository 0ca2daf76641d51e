"""Emitting meshes and spheres as next-event-estimation lights in the oracle
(DirectIllumination::evaluate over every emitting shape,
renderers/DirectIllumination.cpp:90-135), pinned by closed forms in the
small-light limit on a diffuse floor under a black background, with max
depth 1 so a path's radiance is the floor point's direct light alone:

  * a tiny square mesh light of area A facing down gives
    kd / pi * Le * A * cos_wo * cos_wi / r^2 (Mesh::getPosition lands on the
    mesh, Mesh::_getSurfaceArea weights it, the interpolated normal is the
    face normal);
  * a mesh light of two tiny triangles of different areas, four light
    samples per vertex: Mesh::getPosition picks a triangle by the sample's u
    (MESH_GET_INDEX, shapes/Mesh.cpp:26), so the 2 x 2 strata send two
    samples to each triangle whatever its area, and the estimate is
    (A_total / 2) (c_0 + c_1), not the area-weighted mean;
  * a tiny sphere light turned by its transform: Sphere::getPosition puts
    every sample at object-space distance 1 near +z (shapes/Sphere.cpp:
    117-124), the normal is that point's direction, the area is pi
    (Sphere.cpp:138-141); the estimate matches the integral of the
    contribution over the sample square, computed here by quadrature.
"""
import math

import numpy as np

import milton_amd as m
from oracle import pyoracle as po

KD, LE = 0.6, 5.0


def _materials(n_extra=0):
    mats = np.zeros(2, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIFFUSE
    mats[0]["kd"] = KD
    mats[1]["type"] = m.MPT_BSDF_DIFFUSE   # black emitter
    mats[1]["Le"] = LE
    return mats


CAMERA = {"eye": [0, 0.5, 0], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 60.0, "aspect": 1.0}
FLOOR = [[0, 0, 0, 40, 0, 0, 0, 0, 40]]   # y = 0, normal +y


def _paths(a, direct=1, side=16, spp=4):
    job = m.make_job(a.camera, side, side, spp, 1, seed=2)
    job.direct_samples = direct
    pix = np.repeat(np.arange(side * side, dtype=np.uint32), spp)
    smp = np.tile(np.arange(spp, dtype=np.uint32), side * side)
    return job, pix, smp, po.OracleScene(a).render_paths(job, pix, smp)


def _floor_points(job, pix, smp):
    pts = []
    for p, s in zip(pix, smp):
        r = po.camera_ray(job, int(p), int(s))
        t = -r[1] / r[4]
        pts.append(r[:3] + t * r[3:])
    return np.array(pts)


def _direct(p, c, n_light, area):
    """kd / pi * Le * area * cos_wo * cos_wi / r^2 for a point light-like patch."""
    w = c - p
    r = np.linalg.norm(w, axis=1)
    w = w / r[:, None]
    cos_wo = np.abs(w[:, 1])
    cos_wi = np.abs(w @ n_light)
    lit = (w @ n_light) < 0
    return np.where(lit, KD / math.pi * LE * area * cos_wo * cos_wi / r ** 2, 0.0)


def _square_mesh(c, a, down=True):
    x, y, z = c
    v = np.array([[x - a / 2, y, z - a / 2], [x + a / 2, y, z - a / 2], [x + a / 2, y, z + a / 2],
                  [x - a / 2, y, z + a / 2]], np.float64)
    idx = [[0, 1, 2], [0, 2, 3]] if down else [[0, 2, 1], [0, 3, 2]]
    return v, np.array(idx)


def test_tiny_mesh_light():
    a_side = 2e-3
    c = np.array([0.1, 1.0, 0.2])
    v, idx = _square_mesh(c, a_side)
    a = m.SceneArrays(v, idx, [1, 1], np.zeros((0, 4)), [], FLOOR, [0], _materials(), [0, 0, 0])
    a.camera = CAMERA
    job, pix, smp, L = _paths(a)
    want = _direct(_floor_points(job, pix, smp), c, np.array([0, -1.0, 0]), a_side ** 2)
    assert (want > 0).all()
    assert np.allclose(L[:, 0], want, rtol=1e-2) and np.allclose(L[:, 0], L[:, 2])
    # the same mesh facing up lights nothing below it (one-sided emitter)
    v2, idx2 = _square_mesh(c, a_side, down=False)
    a2 = m.SceneArrays(v2, idx2, [1, 1], np.zeros((0, 4)), [], FLOOR, [0], _materials(), [0, 0, 0])
    a2.camera = CAMERA
    assert (_paths(a2)[3] == 0).all()


def test_mesh_light_picks_triangles_uniformly():
    e = 1e-3
    c0, c1 = np.array([-0.4, 1.0, 0.0]), np.array([0.5, 1.0, 0.3])
    # a small and a four times larger right triangle, both facing down
    v = np.array([c0, c0 + [e, 0, 0], c0 + [0, 0, e], c1, c1 + [2 * e, 0, 0], c1 + [0, 0, 2 * e]], np.float64)
    idx = np.array([[0, 1, 2], [3, 4, 5]])
    a = m.SceneArrays(v, idx, [1, 1], np.zeros((0, 4)), [], FLOOR, [0], _materials(), [0, 0, 0],
                      tri_shape=[7, 7])
    a.camera = CAMERA
    job, pix, smp, L = _paths(a, direct=4)
    p = _floor_points(job, pix, smp)
    down = np.array([0, -1.0, 0])
    area = 0.5 * e * e + 0.5 * 4 * e * e
    k0 = _direct(p, c0 + [e / 3, 0, e / 3], down, 1.0)
    k1 = _direct(p, c1 + [2 * e / 3, 0, 2 * e / 3], down, 1.0)
    want = area / 2 * (k0 + k1)                       # two samples on each triangle
    area_weighted = area * (0.2 * k0 + 0.8 * k1)      # what uniform-area sampling would give
    assert np.allclose(L[:, 1], want, rtol=1e-2)
    assert np.mean(np.abs(want - area_weighted) / want) > 0.1


def test_tiny_sphere_light_turned_by_its_transform():
    r = 1e-3
    c = np.array([0.1, 1.0, 0.2])
    # a quarter turn about x: object +z goes to world -y, towards the floor
    frame = 2 * r * np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], np.float64)   # rows: images of x, y, z
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], [[*c, r]], [1], FLOOR, [0], _materials(),
                      [0, 0, 0], sphere_frames=frame.reshape(1, 9))
    a.camera = CAMERA
    job, pix, smp, L = _paths(a, direct=64, side=8, spp=2)
    p = _floor_points(job, pix, smp)
    # quadrature of Sphere::getPosition's normals over the sample square
    g = (np.arange(400) + 0.5) / 400
    U, V = np.meshgrid(g, g, indexing="ij")
    x = np.clip(np.tan((1 - U) / (2 * math.pi)), -0.5, 0.5)
    y = np.clip(np.sin((0.5 - V) / math.pi) * 0.5, -0.5, 0.5)
    z = np.sqrt(1 - x * x - y * y)
    nrm = np.stack([x, y, z], -1).reshape(-1, 3) @ (frame / (2 * r))   # world normals
    want = np.zeros(len(p))
    for i, q in enumerate(p):
        w = c - q
        dist = np.linalg.norm(w)
        w /= dist
        cw = nrm @ w
        want[i] = KD / math.pi * LE * math.pi * abs(w[1]) * np.mean(np.where(cw < 0, -cw, 0.0)) / dist ** 2
    assert (want > 0).all()
    ratio = L[:, 0] / want
    print(f"sphere light: ratio mean {ratio.mean():.4f} min {ratio.min():.4f} max {ratio.max():.4f}", flush=True)
    assert abs(ratio.mean() - 1) < 5e-3 and np.all(np.abs(ratio - 1) < 3e-2)
    # without the turn the samples sit on the +z side and face sideways: the
    # floor gets only the few whose normals lean down
    a.sphere_frames = None
    assert _paths(a, direct=64, side=8, spp=2)[3][:, 0].mean() < 0.2 * L[:, 0].mean()
