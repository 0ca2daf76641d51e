"""The allocation-free tone map (mpt_film_tonemap_noalloc) against
mpt_film_tonemap, through the C ABI: the same bytes on two calls in a row,
the second on a smaller film, for both tone maps; films above the scratch
limit are refused."""
import ctypes

import numpy as np
import pytest

import milton_amd as m

pytestmark = pytest.mark.gpu


def _film(ctx, w, h, seed):
    rng = np.random.default_rng(seed)
    raw = np.zeros((h, w, 4), np.float32)
    raw[..., :3] = rng.lognormal(-1.0, 1.5, (h, w, 3)).astype(np.float32)
    raw[..., 3] = rng.integers(0, 5, (h, w)).astype(np.float32)   # some pixels never sampled
    f = m.Film(ctx, w, h)
    f.upload_raw(raw)
    return f


def _noalloc(ctx, film, tonemap):
    out = np.zeros((film.height, film.width, 4), np.uint8)
    rc = ctx.L.mpt_film_tonemap_noalloc(film.h, tonemap, out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, ctx.L.mpt_last_error(ctx.h)
    return out


@pytest.mark.parametrize("tonemap", [m.MPT_TONEMAP_REINHARD, m.MPT_TONEMAP_GAMMA])
def test_noalloc_tonemap_matches_allocating_entry_point(tonemap):
    ctx = m.Context(0)
    try:
        # two calls in a row, the second on a smaller film: the scratch array
        # keeps the first call's contents, which must not leak into the second
        for (w, h), seed in (((1920, 1080), 1), ((53, 37), 2)):
            print(f"noalloc tonemap {tonemap} film {w}x{h}", flush=True)
            f = _film(ctx, w, h, seed)
            got = _noalloc(ctx, f, tonemap)
            want = f.tonemap(tonemap)
            assert np.array_equal(got, want)
            f.close()
    finally:
        ctx.close()


def test_noalloc_tonemap_of_a_rendered_film():
    ctx = m.Context(0)
    try:
        arrays = m.generate(m.MPT_SCENE_ROOM)
        scene = m.Scene(ctx, arrays)
        job = m.make_job(arrays.camera, 96, 64, 4, 5, seed=3)
        f = m.Film(ctx, 96, 64)
        print("noalloc tonemap: rendered room 96x64", flush=True)
        f.render(scene, job)
        for tm in (m.MPT_TONEMAP_REINHARD, m.MPT_TONEMAP_GAMMA):
            assert np.array_equal(_noalloc(ctx, f, tm), f.tonemap(tm))
        f.close()
        scene.close()
    finally:
        ctx.close()


def test_noalloc_tonemap_refuses_films_above_the_limit():
    ctx = m.Context(0)
    try:
        w = h = 2049   # 4,198,401 pixels > 2^22
        assert w * h > 4194304
        f = m.Film(ctx, w, h)
        out = np.zeros((h, w, 4), np.uint8)
        rc = ctx.L.mpt_film_tonemap_noalloc(f.h, m.MPT_TONEMAP_GAMMA, out.ctypes.data_as(ctypes.c_void_p))
        assert rc == -1
        assert b"scratch limit" in ctx.L.mpt_last_error(ctx.h)
        f.close()
    finally:
        ctx.close()
