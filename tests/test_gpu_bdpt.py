"""The bidirectional path tracer on the device (MPT_RENDERER_BIDIR), path by
path against the oracle's restatement of BidirectionalPathTracer::sample
(renderers/renderers/BidirectionalPathTracer.cpp:34-174) on the same random
streams: the closed-form scenes of tests/test_bdpt.py, the room (diffuse
boxes, a ceiling quad light), the room with a Phong box, a glass sphere and a
point light, and a 256^2 terrain (Phong, glass spheres, a distant light).
Films are the same on every run and alone or served beside path-tracing
jobs; scenes and cameras the renderer does not take are refused."""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from test_bdpt import KD, H, _floor_mats, _scene
from test_gpu_render import _per_path

pytestmark = pytest.mark.gpu
BIDIR = m.MPT_RENDERER_BIDIR


def _job(arrays, w, h, spp, depth, seed, job_id=0):
    return m.make_job(arrays.camera, w, h, spp, depth, seed=seed, job_id=job_id, renderer=BIDIR)


def test_closed_form_scenes_paths(ctx):
    mats = np.zeros(1, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIFFUSE
    mats[0]["kd"] = 0.5
    mats[0]["Le"] = 0.6
    a = _scene([[0, 0, 0, 1, 0, 0, 0, 0, 1]], [0], mats)
    a.camera = {"eye": [0.1, 2.0, 0.05], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 40.0, "aspect": 1.0}
    _per_path(ctx, a, _job(a, 16, 16, 2, 4, 3), 0, 256, "bidir: emitter seen directly", min_frac=1.0)

    lm = _floor_mats(1e7)
    a = _scene([[0, 0, 0, 30, 0, 0, 0, 0, 30], [0.2, H, -0.1, 1e-4, 0, 0, 0, 0, -1e-4]], [0, 1], lm)
    a.camera = {"eye": [0.3, 3.0, -0.2], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 50.0, "aspect": 1.0}
    _per_path(ctx, a, _job(a, 16, 16, 4, 4, 7), 0, 256, "bidir: tiny light over a floor", min_frac=1.0)

    a = _scene([[0, 0, 0, 20, 0, 0, 0, 0, 20]], [0], _floor_mats(0.0)[:1], points=[[0.2, H, -0.1, 4.0, 2.0, 1.0]])
    a.camera = {"eye": [0.3, 3.0, -0.2], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 50.0, "aspect": 1.0}
    _per_path(ctx, a, _job(a, 16, 16, 4, 3, 5), 0, 256, "bidir: point light over a floor", min_frac=1.0)

    mats = np.zeros(2, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIELECTRIC
    mats[0]["ks"] = 1.0
    mats[0]["ior"] = 1.5
    mats[0]["transparency"] = 0.0
    mats[1]["type"] = m.MPT_BSDF_DIFFUSE
    mats[1]["Le"] = 0.5
    a = _scene([[0, 0, 0, 4, 0, 0, 0, 0, 4], [0, 1, 1, 1, 0, 0, 0, 0, -1]], [0, 1], mats)
    a.camera = {"eye": [0.0, 1.0, -1.0], "look": [0, -1, 1], "up": [0, 1, 1], "height_angle": 20.0, "aspect": 1.0}
    _per_path(ctx, a, _job(a, 16, 16, 4, 6, 2), 0, 256, "bidir: mirror into a light", min_frac=1.0)


def test_room_paths(ctx):
    a = m.generate(m.MPT_SCENE_ROOM)
    job = _job(a, 48, 48, 8, 5, 1)
    L_gpu, L_orc = _per_path(ctx, a, job, 0, 48 * 48, "bidir: room", min_frac=0.995)
    assert L_orc.mean() > 0.02 and L_orc.max() <= 1.0


def _room_mixed():
    """The room with its second box Phong, a glass sphere and a point light."""
    a = m.generate(m.MPT_SCENE_ROOM)
    mats = np.concatenate([a.materials, np.zeros(2, m.MATERIAL_DTYPE)])
    ph, gl = len(a.materials), len(a.materials) + 1
    mats[ph]["type"] = m.MPT_BSDF_PHONG
    mats[ph]["kd"] = 0.4
    mats[ph]["ks"] = 0.4
    mats[ph]["exponent"] = 30.0
    mats[gl]["type"] = m.MPT_BSDF_DIELECTRIC
    mats[gl]["ks"] = 1.0
    mats[gl]["ior"] = 1.5
    mats[gl]["transparency"] = 1.0
    tri_mat = a.tri_material.copy()
    tri_mat[-12:] = ph   # the last box's twelve triangles
    b = m.SceneArrays(a.vertices, a.indices, tri_mat, [[0.3, 0.15, 0.7, 0.12]], [gl], a.quads, a.quad_material,
                      mats, a.env, points=[[0.7, 0.8, 0.3, 0.6, 0.6, 0.5]])
    b.camera = a.camera
    return b


def test_room_phong_glass_point_light_paths(ctx):
    a = _room_mixed()
    job = _job(a, 40, 40, 8, 6, 9)
    L_gpu, L_orc = _per_path(ctx, a, job, 0, 40 * 40, "bidir: room, Phong, glass, point light", min_frac=0.99)
    assert L_orc.mean() > 0.02


def test_terrain_paths(ctx):
    a = m.generate(m.MPT_SCENE_TERRAIN, 256, 1)
    job = _job(a, 32, 32, 4, 5, 3)
    _per_path(ctx, a, job, 0, 32 * 32, "bidir: terrain 256^2", min_frac=0.99)


def test_films_are_deterministic_and_batch_independent(ctx):
    a = m.generate(m.MPT_SCENE_ROOM)
    sc = m.Scene(ctx, a)
    job = _job(a, 64, 48, 16, 5, 4)
    film, st = sc.render(job)
    assert st["camera_rays"] == 64 * 48 * 16 and st["shadow_rays"] > 0
    again, _ = sc.render(job)
    assert np.array_equal(film, again)
    saved = ctx.get_option(m.MPT_OPT_BATCH_PATHS)
    try:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, 4096)   # many small batches over the lanes
        small, _ = sc.render(job)
    finally:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, saved)
    assert np.array_equal(film, small)
    # the film is the per-pixel mean of the paths
    L = sc.render_paths(job, 0, 64 * 48).reshape(64 * 48, 16, 3).astype(np.float64)
    assert np.allclose(film.reshape(-1, 3), L.mean(axis=1), rtol=1e-5, atol=1e-6)


def test_served_beside_path_tracing_jobs():
    """On a fresh context (the lanes hold only what the server asked for),
    bidirectional and path-tracing jobs through one server: each film equals
    the job rendered alone, and the lanes stay within the bytes the server
    reserved for ray queues and subpath records."""
    c = m.Context(0)
    try:
        a = m.generate(m.MPT_SCENE_ROOM)
        sc = m.Scene(c, a)
        jobs = [_job(a, 40, 32, 8, 5, 20 + i, job_id=i) for i in range(3)]
        pt = [m.make_job(a.camera, 48, 40, 8, 5, seed=30 + i, job_id=10 + i) for i in range(3)]
        srv = m.Server(c)
        hs = [srv.submit(sc, j) for j in jobs + pt]
        srv.drain()
        films = [srv.result(h, j.width, j.height) for h, j in zip(hs, jobs + pt)]
        mem = srv.memory()
        print(f"[bidir served] {mem}", flush=True)
        srv.close()
        assert mem["lane_bytes"] <= mem["queue_bytes"]
        for f, j in zip(films, jobs + pt):
            assert np.array_equal(f, sc.render(j)[0])
    finally:
        c.close()


def test_refused_scenes_and_cameras(ctx):
    soup = m.generate(m.MPT_SCENE_SOUP, 1000, 1)   # lit by the environment only
    sc = m.Scene(ctx, soup)
    with pytest.raises(m.MptError, match="light"):
        sc.render(_job(soup, 8, 8, 1, 3, 1))
    room = m.generate(m.MPT_SCENE_ROOM)
    sr = m.Scene(ctx, room)
    with pytest.raises(m.MptError, match="max_depth"):
        sr.render(_job(room, 8, 8, 1, 13, 1))
    lens = dict(room.camera, aperture=0.05, focal_length=0.05, focal_distance=2.0)
    with pytest.raises(m.MptError, match="pinhole"):
        sr.render(m.make_job(lens, 8, 8, 1, 3, renderer=BIDIR))
    a = _room_mixed()
    mats = a.materials.copy()
    mats[a.sphere_material[0]]["Le"] = 2.0
    b = m.SceneArrays(a.vertices, a.indices, a.tri_material, a.spheres, a.sphere_material, a.quads,
                      a.quad_material, mats, a.env)
    b.camera = a.camera
    with pytest.raises(m.MptError, match="sphere"):
        m.Scene(ctx, b).render(_job(b, 8, 8, 1, 3, 1))


@pytest.mark.parametrize("kind, support", [(m.MPT_FILTER_MITCHELL, 2), (m.MPT_FILTER_GAUSSIAN, 2)])
def test_filtered_film_matches_oracle(ctx, kind, support):
    """A bidirectional job with a reconstruction filter: each sample splats
    its clamped value onto the 3 x 3 pixels around it (the film kernels the
    path tracer uses), as the oracle's ReconstructionRenderOutput does."""
    a = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(a.camera, 40, 36, 9, 5, seed=17, film_filter=(kind, support), renderer=BIDIR)
    f = m.Film(ctx, 40, 36)
    f.render(m.Scene(ctx, a), job)
    dev = f.download_raw().astype(np.float64)
    again = m.Film(ctx, 40, 36)
    again.render(m.Scene(ctx, a), job)
    assert (again.download_raw() == dev).all()
    f.close()
    again.close()
    orc = po.OracleScene(a).render_film(job)
    dv = dev[..., :3] / np.maximum(dev[..., 3:], 1e-30)
    ov = orc[..., :3] / np.maximum(orc[..., 3:], 1e-30)
    close = np.abs(dv - ov).max(axis=2) <= 1e-3 * np.maximum(1, np.abs(ov).max(axis=2))
    dden = np.abs(dev[..., 3] - orc[..., 3]) / np.maximum(orc[..., 3], 1e-30)
    print(f"[bidir filtered {kind}] pixels close {close.mean():.5f}, denominators {np.mean(dden <= 1e-4):.5f}",
          flush=True)
    assert close.mean() >= 0.97 and np.mean(dden <= 1e-4) == 1.0
