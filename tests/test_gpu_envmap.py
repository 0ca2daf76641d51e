"""Environment images through the device path: a diffuse and a glass sphere
under a lat-long environment image, rendered on the GPU and by the oracle
with the same seeds, per path, for every reconstruction filter; and a scene
file whose background is an image file.

The device computes the lookup in double precision from the FP32 ray
direction (the oracle from its FP64 one), so a path differs only where a
lookup position lands within rounding of a texel boundary or an earlier
FP32/FP64 branch differs; >= 99 % of paths agree to 1e-3.
"""
import numpy as np
import pytest

import milton_amd as m
from test_gpu_render import _per_path

pytestmark = pytest.mark.gpu


def _sky(h=24, w=48, seed=5):
    """Smooth colour gradients with a few bright texels (a 'sun')."""
    rng = np.random.default_rng(seed)
    v = np.linspace(0, 1, h)[:, None]
    u = np.linspace(0, 1, w)[None, :]
    img = np.stack([0.2 + 0.8 * v + 0 * u, 0.3 + 0.5 * u + 0 * v, 1.0 - 0.6 * v * u], axis=2)
    img *= rng.uniform(0.8, 1.2, (h, w, 1))
    img[3:5, 30:33] = (40.0, 35.0, 20.0)
    return img.astype(np.float32)


def _spheres(env_map):
    mats = np.zeros(2, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIFFUSE
    mats[0]["kd"] = (0.7, 0.6, 0.5)
    mats[1]["type"] = m.MPT_BSDF_DIELECTRIC
    mats[1]["ks"] = 1.0
    mats[1]["ior"] = 1.5
    mats[1]["transparency"] = 1.0
    spheres = np.array([[-0.55, 0, 0, 0.5], [0.55, 0, 0, 0.5]], np.float32)
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], spheres, [0, 1], np.zeros((0, 9)), [], mats,
                      [0, 0, 0], env_map=env_map)
    a.camera = {"eye": [0.0, 0.4, -3.0], "look": [0.0, -0.1, 1.0], "up": [0, 1, 0], "height_angle": 50.0,
                "aspect": 1.0}
    return a


@pytest.mark.parametrize("filt, support, params", [
    (m.MPT_FILTER_TRIANGLE, 2, (0, 0)), (m.MPT_FILTER_BOX, 2, (0, 0)), (m.MPT_FILTER_GAUSSIAN, 3, (0.75, 0)),
    (m.MPT_FILTER_MITCHELL, 4, (0, 0)), (m.MPT_FILTER_LANCZOS, 3, (2, 0)), (m.MPT_FILTER_TRIANGLE, 1, (0, 0))])
def test_env_map_paths_same_seed(ctx, filt, support, params):
    env = m.EnvMap(_sky(), repeat_u=1.0, filter=filt, support=support, params=params)
    arrays = _spheres(env)
    job = m.make_job(arrays.camera, 40, 40, 4, 5, seed=21)
    L_gpu, L_orc = _per_path(ctx, arrays, job, 0, 40 * 40, f"envmap filter {filt} support {support}")
    # misses at depth 0 see the image itself: the mean is far from any constant
    assert L_orc.std(axis=0).min() > 0.05


def test_env_map_repeat_and_large_image(ctx):
    rng = np.random.default_rng(8)
    img = rng.lognormal(-0.5, 1.0, (256, 512, 3)).astype(np.float32)
    env = m.EnvMap(img, repeat_u=3.0, repeat_v=2.0)
    arrays = _spheres(env)
    job = m.make_job(arrays.camera, 32, 32, 4, 4, seed=3)
    _per_path(ctx, arrays, job, 0, 32 * 32, "envmap 512x256 repeat 3x2")


def test_scenefile_environment(ctx, tmp_path):
    m.write_image(tmp_path / "sky.hdr", _sky(16, 32))
    text = """{ "scenefile" : {
      "renderer" : { "noSuperSamples" : 4, "maxDepth" : 4, "seed" : 9 },
      "camera" : { "type" : "pinhole", "eye" : [0, 0.3, -2.5], "look" : [0, -0.1, 1], "heightAngle" : 50 },
      "output" : { "width" : 32, "height" : 32 },
      "scene" : {
        "background" : { "emitter" : { "type" : "environment", "path" : "sky.hdr", "filter" : "box" } },
        "material" : { "bsdf" : { "type" : "diffuse", "kd" : [0.6, 0.6, 0.6] }, "sphere" : { } } } } }"""
    arrays, job = m.load_scenefile(text=text, base_dir=str(tmp_path))
    assert arrays.env_map is not None and arrays.env_map.filter == m.MPT_FILTER_BOX
    _per_path(ctx, arrays, job, 0, job.width * job.height, "scenefile envmap")
