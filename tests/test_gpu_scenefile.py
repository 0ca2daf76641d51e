"""A Milton scene file through the device path: the room of
tests/test_scenefile.py (one-sided walls, a transformed light plane, a glass
sphere, a Phong triangle) loaded by mpt_scenefile_*, rendered on the GPU and
compared with the oracle on the same arrays, per path (same seeds) and per
pixel (independent seeds); several light samples per vertex
(DirectIllumination's noDirectSamples); and a rotated, scaled cube."""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from test_gpu_render import _per_path, _three_sigma
from test_scenefile import ROOM

pytestmark = pytest.mark.gpu


def test_scenefile_room_paths_same_seed(ctx):
    arrays, job = m.load_scenefile(text=ROOM)
    print("scenefile room:", len(arrays.indices), "one-sided triangles,", len(arrays.quads), "quad,",
          len(arrays.spheres), "sphere", flush=True)
    _per_path(ctx, arrays, job, 0, job.width * job.height, "scenefile-room-paths", min_frac=0.999)


def test_scenefile_room_image_three_sigma(ctx):
    arrays, job = m.load_scenefile(text=ROOM)
    job.spp = 64
    job_orc = m.Job.from_buffer_copy(job)
    job_orc.seed = job.seed + 1
    sc = m.Scene(ctx, arrays)
    n = job.width * job.height
    L_gpu = sc.render_paths(job, 0, n).astype(np.float64)
    osc = po.OracleScene(arrays)
    pix = np.repeat(np.arange(n, dtype=np.uint32), job.spp)
    smp = np.tile(np.arange(job.spp, dtype=np.uint32), n)
    L_orc = osc.render_paths(job_orc, pix, smp)
    _three_sigma(L_gpu, L_orc, job.spp, "scenefile-room-3sigma", min_frac=0.97)


@pytest.mark.parametrize("direct", [4, 5])
def test_direct_samples_paths_same_seed(ctx, direct):
    """noDirectSamples > 1: per light, that many shadow rays jittered over
    floor(sqrt(n))^2 strata (5 cycles over 2x2), each weighted area / n."""
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(arrays.camera, 48, 48, 16, 5, seed=4, direct_samples=direct)
    L_gpu, _ = _per_path(ctx, arrays, job, 0, 48 * 48, f"room-direct{direct}-paths")
    sc = m.Scene(ctx, arrays)
    f = m.Film(ctx, 48, 48)
    st = f.render(sc, job)
    f.close()
    one = m.make_job(arrays.camera, 48, 48, 16, 5, seed=4)
    f1 = m.Film(ctx, 48, 48)
    st1 = f1.render(sc, one)
    f1.close()
    # the same vertices draw `direct` times the candidate shadow rays
    assert st["shadow_rays"] > 2 * st1["shadow_rays"]


def test_dispersive_glass_paths_same_seed(ctx):
    """A glass sphere whose IOR differs per channel: a refracting path keeps
    one channel, refracted with that channel's index."""
    arrays, job = m.load_scenefile(text=ROOM.replace('"medium" : { "ior" : 1.5 }',
                                                     '"medium" : { "ior" : [1.4, 1.55, 1.8] }'))
    assert arrays.materials[arrays.sphere_material[0]]["ior_b"] > 0
    _per_path(ctx, arrays, job, 0, job.width * job.height, "scenefile-room-dispersive", min_frac=0.999)


def test_cube_paths_same_seed(ctx):
    """The room with the Phong triangle replaced by a rotated, non-uniformly
    scaled Phong cube (six quads)."""
    tri = '"triangle" : { "vertices" : [ -0.5,0,-0.5,  -0.5,0,0.5,  0.5,0,0.5 ] },'
    assert tri in ROOM
    # lifted off the floor so that no face is coplanar with a wall
    text = ROOM.replace('"translate" : [ 0.7, 0.15, 0.35 ],\n               "scale" : [ 0.2, 0.3, 0.2 ],',
                        '"translate" : [ 0.7, 0.17, 0.35 ], "rotate" : [ 0, 0, 0, 0, 1, 0, 25 ],'
                        ' "scale" : [ 0.2, 0.3, 0.15 ],')
    assert text != ROOM
    text = text.replace(tri, '"cube" : { },')
    arrays, job = m.load_scenefile(text=text)
    assert len(arrays.quads) == 7
    _per_path(ctx, arrays, job, 0, job.width * job.height, "scenefile-room-cube", min_frac=0.98)
