"""Tone maps and film files through the device path (mpt_film_tonemap,
mpt_film_save, raw film download/upload), compared with the oracle's tone
maps on the same film contents.

Bar: the device computes in double precision like the oracle, so bytes are
identical except where a value falls within rounding of a truncation
boundary (device and host libm pow/log/exp may differ in the last bit):
at least 99.99 % of bytes identical and none more than 1 apart.
"""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from test_tonemap import read_pfm, read_png, read_ppm, read_rgbe, rgbe_to_float, rgbe_within_step

pytestmark = pytest.mark.gpu


def _random_film(h, w, seed):
    """Wide dynamic range, some black pixels and some never-sampled ones."""
    rng = np.random.default_rng(seed)
    n = rng.integers(1, 65, (h, w, 1)).astype(np.float32)
    rgb = rng.lognormal(-1.0, 2.0, (h, w, 3)).astype(np.float32) * n
    rgb[rng.random((h, w)) < 0.02] = 0
    f = np.concatenate([rgb, n], axis=2)
    f[rng.random((h, w)) < 0.01] = 0
    return f


def _compare(dev, orc, label):
    d = dev.astype(np.int16) - orc.astype(np.int16)
    same = (d == 0).mean()
    print(f"{label}: identical bytes {same:.6f}, max |diff| {np.abs(d).max()}", flush=True)
    assert np.abs(d).max() <= 1
    assert same >= 0.9999


@pytest.mark.parametrize("shape", [(1, 1), (37, 53), (1080, 1920)])
@pytest.mark.parametrize("mode", [m.MPT_TONEMAP_REINHARD, m.MPT_TONEMAP_GAMMA])
def test_tonemap_matches_oracle(ctx, shape, mode):
    h, w = shape
    raw = _random_film(h, w, h * 7 + w + mode)
    f = m.Film(ctx, w, h)
    f.upload_raw(raw)
    assert (f.download_raw() == raw).all()
    dev = f.tonemap(mode)
    again = f.tonemap(mode)
    f.close()
    assert (dev == again).all()   # fixed-order reduction: the same bytes every run
    _compare(dev, po.tonemap(raw, mode), f"tonemap mode {mode} {w}x{h}")


def test_tonemap_rendered_room(ctx):
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(arrays.camera, 64, 48, 16, 5, seed=9)
    sc = m.Scene(ctx, arrays)
    f = m.Film(ctx, 64, 48)
    f.render(sc, job)
    raw = f.download_raw()
    for mode in (m.MPT_TONEMAP_REINHARD, m.MPT_TONEMAP_GAMMA):
        _compare(f.tonemap(mode), po.tonemap(raw, mode), f"room tonemap mode {mode}")
    f.close()


def test_film_save_formats(ctx, tmp_path):
    arrays = m.generate(m.MPT_SCENE_ROOM)
    job = m.make_job(arrays.camera, 40, 30, 8, 5, seed=2)
    sc = m.Scene(ctx, arrays)
    f = m.Film(ctx, 40, 30)
    f.render(sc, job)
    ldr = f.tonemap("default")
    lin = f.download()
    f.save(tmp_path / "o.png")
    f.save(tmp_path / "o.ppm", "reinhard")
    f.save(tmp_path / "o.hdr")
    f.save(tmp_path / "r.hdr", radiance=True)
    f.save(tmp_path / "r.pfm", radiance=True)
    f.save(tmp_path / "g.png", "null")
    assert (read_png(tmp_path / "o.png") == ldr[..., :3]).all()
    assert (read_ppm(tmp_path / "o.ppm") == ldr[..., :3]).all()
    assert (read_png(tmp_path / "g.png") == f.tonemap(m.MPT_TONEMAP_GAMMA)[..., :3]).all()
    # an HDR file of the tone-mapped image holds bytes / 255, as the
    # reference's RgbaImage output does
    assert rgbe_within_step(rgbe_to_float(read_rgbe(tmp_path / "o.hdr")), ldr[..., :3] / 255.0)
    assert rgbe_within_step(rgbe_to_float(read_rgbe(tmp_path / "r.hdr")), lin)
    assert (read_pfm(tmp_path / "r.pfm") == lin).all()
    with pytest.raises(m.MptError):
        f.save(tmp_path / "o.jpg")
    f.close()


def test_tiles_merge_by_summing_raw_films(ctx):
    """Two halves of a job rendered into separate films and summed equal the
    whole job's film: the merge step of a job split across devices."""
    arrays = m.generate(m.MPT_SCENE_ROOM)
    W, H = 48, 40
    job = m.make_job(arrays.camera, W, H, 8, 5, seed=11)
    sc = m.Scene(ctx, arrays)
    whole, a, b = m.Film(ctx, W, H), m.Film(ctx, W, H), m.Film(ctx, W, H)
    whole.render(sc, job)
    cut = W * H // 2 + 7
    a.render_range(sc, job, 0, cut)
    b.render_range(sc, job, cut, W * H)
    merged = a.download_raw() + b.download_raw()
    assert (merged == whole.download_raw()).all()
    a.upload_raw(merged)
    assert (a.tonemap() == whole.tonemap()).all()
    for f in (whole, a, b):
        f.close()
