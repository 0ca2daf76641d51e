"""Emitting meshes and spheres on the device, path by path against the
oracle: the small-light scenes of tests/test_area_lights.py, the room with
its ceiling light replaced by a two-triangle mesh light and an emitting
sphere turned by its transform (several light samples per vertex, depth 6),
and the arrangement of Milton's plates.js (sphere lights over Phong plates)."""
import numpy as np
import pytest

import milton_amd as m
from test_area_lights import CAMERA, FLOOR, _materials
from test_gpu_render import _per_path

pytestmark = pytest.mark.gpu


def _room_with_area_lights():
    a = m.generate(m.MPT_SCENE_ROOM)
    q = a.quads[0].astype(np.float64)                  # the ceiling light: centre, ex, ez
    c, ex, ez = q[:3], q[3:6], q[6:]
    v0 = len(a.vertices)
    corners = [c - ex / 2 - ez / 2, c + ex / 2 - ez / 2, c + ex / 2 + ez / 2, c - ex / 2 + ez / 2]
    light_mat = a.quad_material[0]
    verts = np.concatenate([a.vertices, np.array(corners, np.float32)])
    # the winding gives the quad's own normal, ez x ex (facing down)
    tris = np.concatenate([a.indices, np.array([[v0, v0 + 2, v0 + 1], [v0, v0 + 3, v0 + 2]], np.uint32)])
    mats = np.concatenate([a.materials, a.materials[light_mat:light_mat + 1]])
    mats[-1]["Le"] = [6.0, 5.0, 3.0]
    tri_mat = np.concatenate([a.tri_material, [light_mat, light_mat]]).astype(np.uint32)
    sph_frame = 0.12 * np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], np.float32)   # +z turned towards -y
    b = m.SceneArrays(verts, tris, tri_mat, [[0.7, 0.75, 0.3, 0.06]], [len(mats) - 1], np.zeros((0, 9)), [],
                      mats, a.env, tri_shape=np.concatenate([np.zeros(len(a.indices)), [1, 1]]),
                      sphere_frames=sph_frame.reshape(1, 9))
    b.camera = a.camera
    return b


def test_small_area_lights_paths(ctx):
    from test_area_lights import _square_mesh
    v, idx = _square_mesh(np.array([0.1, 1.0, 0.2]), 0.05)
    a = m.SceneArrays(v, idx, [1, 1], [[-0.3, 0.8, 0.1, 0.04]], [1], FLOOR, [0], _materials(), [0, 0, 0],
                      sphere_frames=(0.08 * np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]])).reshape(1, 9))
    a.camera = CAMERA
    job = m.make_job(a.camera, 32, 32, 4, 3, seed=6)
    job.direct_samples = 4
    _per_path(ctx, a, job, 0, 32 * 32, "small mesh and sphere lights", min_frac=0.999)


def test_room_with_mesh_and_sphere_lights_paths(ctx):
    a = _room_with_area_lights()
    job = m.make_job(a.camera, 48, 48, 16, 6, seed=4)
    job.direct_samples = 4
    L_gpu, L_orc = _per_path(ctx, a, job, 0, 48 * 48, "room, mesh + sphere lights", min_frac=0.999)
    assert L_orc.mean() > 0.05
    sc = m.Scene(ctx, a)
    film, st = sc.render(job)
    assert st["shadow_rays"] > 0
    again, _ = sc.render(job)
    assert np.array_equal(film, again)


def test_plates_like_scene_paths(ctx):
    """The arrangement of Milton's plates.js: Phong plates of rising
    exponent lit by emitting spheres of falling size and rising power, under
    a large emitting plane."""
    mats = np.zeros(10, m.MATERIAL_DTYPE)
    for i, n in enumerate((10, 100, 1000, 10000)):
        mats[i]["type"] = m.MPT_BSDF_PHONG
        mats[i]["kd"] = 0.1
        mats[i]["ks"] = 0.8
        mats[i]["exponent"] = n
    for i, (le, r) in enumerate(((2.0, 0.2), (8.0, 0.1), (30.0, 0.05), (120.0, 0.025))):
        mats[4 + i]["type"] = m.MPT_BSDF_DIFFUSE
        mats[4 + i]["Le"] = le
    mats[8]["type"] = m.MPT_BSDF_DIFFUSE
    mats[8]["Le"] = 0.05
    mats[9]["type"] = m.MPT_BSDF_DIFFUSE
    mats[9]["kd"] = 0.5
    quads, qmat = [], []
    for i in range(4):   # plates tilted towards the camera
        z = -0.6 + 0.4 * i
        quads.append([0, 0.2 * i, z, 2.0, 0, 0, 0, 0.2, 0.35])
        qmat.append(i)
    quads.append([0, 4.0, 0, 20, 0, 0, 0, 0, -20])   # the overarching emitter, facing down
    qmat.append(8)
    quads.append([0, -0.5, 0, 20, 0, 0, 0, 0, 20])   # floor
    qmat.append(9)
    spheres = [[-0.6 + 0.4 * i, 1.2, 0.0, r] for i, r in enumerate((0.2, 0.1, 0.05, 0.025))]
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], spheres, [4, 5, 6, 7], quads, qmat, mats, [0, 0, 0])
    a.camera = {"eye": [0, 1.5, -3.0], "look": [0, -0.3, 1], "up": [0, 1, 0], "height_angle": 40.0, "aspect": 1.5}
    job = m.make_job(a.camera, 48, 32, 4, 5, seed=3)
    _per_path(ctx, a, job, 0, 48 * 32, "plates-like scene: sphere lights over Phong plates", min_frac=0.999)
