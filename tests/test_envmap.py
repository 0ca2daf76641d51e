"""Environment images on the CPU: the oracle's background lookup
(materials/emitters/EnvironmentMap.cpp:44-95 through
filters/KernelFilter.cpp:124-176), the image reader and the scene-file
"environment" background.

The oracle is pinned by closed forms: a constant image is returned
unchanged by every filter (the kernel is normalised); the lat-long mapping
sends the axis directions to known texels (a ray travelling along d sees the
image at the UV of -d); the triangle filter of support 2 weighs the 3 x 3
neighbourhood by (2.5, 1.5, 0.5) x (2.5, 1.5, 0.5) / 20.25 (its taps sit at
offsets -1.5, -0.5, 0.5 and the reference evaluates 1 - x / r without an
absolute value); support 1 is a nearest-texel lookup; the box filter of
support 2 averages the 2 x 2 texels at and after the position; edges clamp;
repeatU = 2 tiles the image twice around; straight down (v = 1) wraps to
the top row, as (t - floor(t)) does in the reference.
"""
import math
import os

import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from test_tonemap import read_rgbe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GRACE = "/root/reference/mlt/src/tests/image/grace_probe.hdr"


def _scene(env_map):
    """A tiny scene far from the origin: the background is all a ray sees."""
    mat = np.zeros(1, m.MATERIAL_DTYPE)
    mat["kd"] = 0.5
    quad = np.array([[100, 100, 100, 1, 0, 0, 0, 0, 1]], np.float32)
    return m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [], quad, [0], mat,
                         [0, 0, 0], env_map=env_map)


def _bg(env_map, dirs):
    return po.OracleScene(_scene(env_map)).background(np.asarray(dirs, np.float64))


def _index_image(h, w):
    img = np.zeros((h, w, 3), np.float32)
    img[..., 0] = np.arange(h)[:, None]
    img[..., 1] = np.arange(w)[None, :]
    img[..., 2] = 1
    return img


@pytest.mark.parametrize("filt, support", [(m.MPT_FILTER_BOX, 2), (m.MPT_FILTER_TRIANGLE, 2),
                                           (m.MPT_FILTER_GAUSSIAN, 3), (m.MPT_FILTER_MITCHELL, 4),
                                           (m.MPT_FILTER_LANCZOS, 2)])
def test_constant_image_any_filter(filt, support):
    img = np.broadcast_to(np.array([0.25, 1.5, 3.0], np.float32), (5, 9, 3))
    rng = np.random.default_rng(1)
    d = rng.normal(size=(50, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    got = _bg(m.EnvMap(img, filter=filt, support=support), d)
    assert np.allclose(got, [0.25, 1.5, 3.0], rtol=1e-12)


def test_lat_long_mapping_nearest():
    """Support 1: a single tap.  8 x 4 image whose texels hold (row, col)."""
    env = m.EnvMap(_index_image(4, 8), support=1)
    s = 1 / math.sqrt(1.25)
    dirs = [(0, 0, -1),    # w = +z: u = 0.5, v = 0.5
            (-1, 0, 0),    # w = +x: phi = pi/2, u = 0.25
            (1, 0, 0),     # w = -x: u = 0.75
            (0, 0, 1),     # w = -z: phi = pi, u = 0
            (0, -1, 0),    # w = +y: theta = pi/2, v = 0 (top row)
            (0, 1, 0),     # w = -y: v = 1, which t - floor(t) wraps to row 0
            (0, -0.5 * s, -s)]   # atan(0.5) up: v = 0.5 - 0.1476, y = 1.41
    want_rc = [(2, 4), (2, 2), (2, 6), (2, 0), (0, 4), (0, 4), (1, 4)]
    got = _bg(env, dirs)
    assert [(int(r), int(c)) for r, c, _ in got] == want_rc


def test_repeat_u_tiles_twice():
    env = m.EnvMap(_index_image(4, 8), repeat_u=2.0, support=1)
    got = _bg(env, [(-1, 0, 0), (0, 0, -1)])   # u = 0.25 -> s = 0.5; u = 0.5 -> s = 1.0 -> 0
    assert got[0, 1] == 4 and got[1, 1] == 0


def _one_hot(h, w, r, c):
    img = np.zeros((h, w, 3), np.float32)
    img[r, c] = 1
    return img


def test_triangle_kernel_weights():
    """Position (row 2, col 4) of an 8 x 4 image; the taps cover rows 1..3
    and cols 3..5 with weights (2.5, 1.5, 0.5) each way."""
    w1 = np.array([2.5, 1.5, 0.5])
    for r, c in [(1, 3), (1, 5), (2, 4), (3, 3), (3, 5), (0, 4)]:
        got = _bg(m.EnvMap(_one_hot(4, 8, r, c)), [(0, 0, -1)])[0, 0]
        want = w1[r - 1] * w1[c - 3] / 20.25 if 1 <= r <= 3 and 3 <= c <= 5 else 0.0
        assert got == pytest.approx(want, rel=1e-12), (r, c)


def test_box_kernel_and_edge_clamp():
    """Box support 2: the taps at offsets (+-0.5, +-0.5) are the ones inside
    the unit disc, i.e. rows y, y+1 and cols x, x+1, each 1/4."""
    for r, c, want in [(2, 4, 0.25), (3, 5, 0.25), (1, 4, 0.0), (2, 3, 0.0)]:
        got = _bg(m.EnvMap(_one_hot(4, 8, r, c), filter=m.MPT_FILTER_BOX), [(0, 0, -1)])[0, 0]
        assert got == pytest.approx(want, rel=1e-12), (r, c)
    # top row (w = +y: row 0, col 4): the triangle taps above it clamp to row 0,
    # which gets (2.5 + 1.5) of the row weights
    got = _bg(m.EnvMap(_one_hot(4, 8, 0, 4)), [(0, -1, 0)])[0, 0]
    assert got == pytest.approx(4.0 * 1.5 / 20.25, rel=1e-12)


def test_gaussian_weights():
    sig = 0.75   # exact in the float the descriptor carries
    g = np.exp(-0.5 * np.array([-1.5, -0.5, 0.5]) ** 2 / sig ** 2)
    got = _bg(m.EnvMap(_one_hot(4, 8, 1, 5), filter=m.MPT_FILTER_GAUSSIAN, params=(sig, 0)), [(0, 0, -1)])[0, 0]
    assert got == pytest.approx(g[0] * g[2] / g.sum() ** 2, rel=1e-12)


# ---------------------------------------------------------------------------
# Image reader.
# ---------------------------------------------------------------------------
def test_read_back_written_images(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.lognormal(0, 1.5, (6, 11, 3)).astype(np.float32)
    m.write_image(tmp_path / "a.pfm", img)
    assert (m.read_image(tmp_path / "a.pfm") == img).all()
    for w in (5, 40):   # flat and run-length encoded scanlines
        im = rng.lognormal(0, 1.5, (3, w, 3)).astype(np.float32)
        im[1] = 0.5
        m.write_image(tmp_path / "b.hdr", im)
        got = m.read_image(tmp_path / "b.hdr")
        b = read_rgbe(tmp_path / "b.hdr").astype(np.float64)
        dec = b[..., :3] * np.where(b[..., 3:] > 0, np.ldexp(1.0, (b[..., 3:] - 136).astype(int)), 0)
        assert (got == dec).all()   # byte * 2^(e - 136), as rgbe2float decodes
        assert (got <= im).all() and (got[1] == 0.5).all()
    ldr = rng.integers(0, 256, (4, 7, 4), dtype=np.uint8)
    m.write_image(tmp_path / "c.ppm", ldr)
    assert np.allclose(m.read_image(tmp_path / "c.ppm"), ldr[..., :3] / 255.0, rtol=1e-7)


def test_read_flat_scanlines_and_errors(tmp_path):
    """Scanlines that are not run-length encoded (first bytes other than
    2, 2) are read flat, as the reference's RGBE_ReadPixels_RLE falls back."""
    w, h = 9, 2
    px = np.zeros((h, w, 4), np.uint8)
    px[..., 0] = np.arange(w)[None] + 10
    px[..., 3] = 129
    head = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y 2 +X 9\n"
    (tmp_path / "f.hdr").write_bytes(head + px.tobytes())
    got = m.read_image(tmp_path / "f.hdr")
    assert np.allclose(got[..., 0], (np.arange(w)[None] + 10) * 2.0 ** -7)
    (tmp_path / "t.hdr").write_bytes(head + px.tobytes()[:-5])
    (tmp_path / "n.hdr").write_bytes(b"#?RADIANCE\n\n-Y 2 +X 9\n" + px.tobytes())
    for bad in ("t.hdr", "n.hdr", "missing.hdr", "x.png"):
        with pytest.raises(m.MptError):
            m.read_image(tmp_path / bad)


@pytest.mark.skipif(not os.path.exists(GRACE), reason="reference checkout not present")
def test_read_reference_light_probe():
    """The reference's own RGBE test image (HDR Shop output, run-length
    encoded, an EXPOSURE line in the header) decodes to the same values with
    the library's reader and the independent reader of the tests."""
    got = m.read_image(GRACE)
    b = read_rgbe(GRACE).astype(np.float64)
    dec = b[..., :3] * np.where(b[..., 3:] > 0, np.ldexp(1.0, (b[..., 3:] - 136).astype(int)), 0)
    assert got.shape == dec.shape and got.shape[0] > 100
    assert (got == dec).all()


# ---------------------------------------------------------------------------
# Scene files.
# ---------------------------------------------------------------------------
def test_scenefile_environment_background(tmp_path):
    img = _index_image(4, 8) + 0.5
    m.write_image(tmp_path / "sky.pfm", img)
    text = """{ "scenefile" : { "scene" : {
        "background" : { "emitter" : { "type" : "environment", "coords" : "spherical", "path" : "sky.pfm",
                                       "filter" : "gaussian", "support" : 3, "sigma" : 0.7, "repeatU" : 2 } },
        "sphere" : { } } } }"""
    a, _ = m.load_scenefile(text=text, base_dir=str(tmp_path))
    e = a.env_map
    assert e is not None and (e.image == img).all()
    assert (e.filter, e.support, e.repeat_u, e.repeat_v) == (m.MPT_FILTER_GAUSSIAN, 3, 2.0, 1.0)
    assert e.params[0] == pytest.approx(0.7)
    plain, _ = m.load_scenefile(text=text.replace('"filter" : "gaussian", "support" : 3, "sigma" : 0.7, ', ""),
                                base_dir=str(tmp_path))
    assert (plain.env_map.filter, plain.env_map.support) == (m.MPT_FILTER_TRIANGLE, 2)
    for snippet, word in [('"path" : "none.hdr"', "cannot be read"), ('"path" : "sky.pfm", "filter" : "cubic"',
                                                                       "unexpected filter")]:
        with pytest.raises(m.MptError) as err:
            m.load_scenefile(text='{ "scenefile" : { "scene" : { "background" : { "emitter" : '
                                  '{ "type" : "environment", %s } } } } }' % snippet, base_dir=str(tmp_path))
        assert word in str(err.value)
