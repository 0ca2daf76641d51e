"""Tone maps and image writers on the CPU.

The oracle's tone maps (renderers/tonemaps/ReinhardToneMap.cpp:25-98,
NullToneMap.cpp:21-46) are pinned by closed forms: a grey image keeps its
chromaticity, so Reinhard reduces to 255 (Ld)^(1/2.22) with Ld = L (1 + L /
Ymax^2) / (1 + L) and L = 0.18 Y / key; the key is the geometric mean of the
luminances; gamma-only is 255 c^(1/2.22).  Expected bytes are chosen away
from truncation boundaries, so the matrices' 1e-6 round-trip error cannot
move them.  The writers (PNG, PPM, PFM, Radiance RGBE) are checked by
decoding their files with independent readers written here from the
published formats (zlib for PNG's deflate stream).
"""
import math
import re
import struct
import zlib

import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po

GAMMA = 1 / 2.22
# luminance of RGB (1, 1, 1) under the sRGB primaries (row 2 of the matrix)
Y_GREY = 0.21267112134122 + 0.71515920531078 + 0.07216877677613


def _film(rgb, weight=1.0):
    rgb = np.asarray(rgb, np.float64)
    w = np.full(rgb.shape[:-1] + (1,), weight)
    return np.concatenate([rgb * weight, w], axis=-1).astype(np.float32)


def _byte(v):
    """Expected byte, asserting it is not within 0.02 of a truncation boundary."""
    x = 255 * v ** GAMMA
    assert 0.02 < x - math.floor(x) < 0.98 or x >= 255, f"ambiguous expectation {x}"
    return min(255, int(x))


@pytest.mark.parametrize("c", [0.05, 0.3, 2.0, 7.5, 40.0])
def test_reinhard_grey_closed_form(c):
    Y = c * Y_GREY
    key = 1e-7 + Y
    L = 0.18 * Y / key
    Ld = L * (1 + L / (Y * Y)) / (1 + L)
    want = _byte(Ld / Y_GREY)
    out = po.tonemap(_film(np.full((3, 5, 3), c)))
    assert (out[..., :3] == want).all(), (out[0, 0], want)
    assert (out[..., 3] == 255).all()


def test_reinhard_key_is_geometric_mean():
    """Two grey pixels: key = sqrt(Y1 Y2) (1e-7 is negligible here), white =
    1 / Y2^2 with Y2 the brighter."""
    c1, c2 = 0.2, 3.0
    Y1, Y2 = c1 * Y_GREY, c2 * Y_GREY
    key = math.sqrt((Y1 + 1e-7) * (Y2 + 1e-7))
    out = po.tonemap(_film([[[c1] * 3, [c2] * 3]]))
    for i, Y in enumerate((Y1, Y2)):
        L = 0.18 * Y / key
        Ld = L * (1 + L / (Y2 * Y2)) / (1 + L)
        assert (out[0, i, :3] == _byte(Ld / Y_GREY)).all(), (i, out[0, i], Ld)


def test_reinhard_keeps_chromaticity():
    """A saturated pixel among grey ones: the output, back through the
    gamma, keeps the input's RGB ratios (up to byte quantisation)."""
    img = np.full((4, 4, 3), 0.5)
    img[1, 2] = (0.9, 0.45, 0.15)
    out = po.tonemap(_film(img)).astype(np.float64)
    lin = (out[1, 2, :3] / 255) ** 2.22
    assert lin[1] / lin[0] == pytest.approx(0.5, rel=0.03)
    assert lin[2] / lin[0] == pytest.approx(1 / 6, rel=0.06)


def test_black_and_empty_pixels():
    """Y = 0 makes the chromaticity 0/0; the reference's clamp passes the NaN
    and the byte conversion yields 0.  A pixel with denominator 0 reads as
    black (ProgressiveFilterValue::getValue)."""
    f = _film(np.full((2, 3, 3), 1.5))
    f[0, 0] = 0
    f[1, 1, :3] = 5.0
    f[1, 1, 3] = 0
    for mode in (0, 1):
        out = po.tonemap(f, mode)
        assert (out[0, 0, :3] == 0).all() and (out[1, 1, :3] == 0).all()
        assert (out[0, 1, :3] > 0).all()


def test_denominator_divides():
    a = po.tonemap(_film(np.full((2, 2, 3), 0.7), weight=1.0))
    b = po.tonemap(_film(np.full((2, 2, 3), 0.7), weight=37.0))
    assert (a == b).all()


def test_gamma_closed_form():
    vals = np.array([0.0, 0.001, 0.2, 0.5, 0.999, 1.0, 3.0])
    out = po.tonemap(_film(np.repeat(vals[None, :, None], 3, axis=2)), 1)
    want = [0 if v == 0 else _byte(v) for v in vals]
    assert out[0, :, 0].tolist() == want
    assert want[3] == 186 and want[5] == 255 and want[6] == 255


def test_tonemap_names():
    assert m.tonemap_from_name("reinhard") == m.MPT_TONEMAP_REINHARD
    assert m.tonemap_from_name("default") == m.MPT_TONEMAP_REINHARD
    assert m.tonemap_from_name("null") == m.MPT_TONEMAP_GAMMA
    assert m.tonemap_from_name("") == m.MPT_TONEMAP_GAMMA


# ---------------------------------------------------------------------------
# Readers for the written files (from the published format descriptions).
# ---------------------------------------------------------------------------
def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body), typ
        chunks.append((typ, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def read_ppm(path):
    """Header tokens separated by whitespace; exactly one whitespace byte
    after maxval, then the raster (which may itself start with a byte that
    reads as whitespace)."""
    data = open(path, "rb").read()
    hdr = re.match(rb"P6\s+(\d+)\s+(\d+)\s+(\d+)\s", data)
    assert hdr and hdr.group(3) == b"255"
    w, h = int(hdr.group(1)), int(hdr.group(2))
    rest = data[hdr.end():]
    assert len(rest) == 3 * w * h
    return np.frombuffer(rest, np.uint8).reshape(h, w, 3)


def read_pfm(path):
    data = open(path, "rb").read()
    magic, dims, scale, rest = data.split(b"\n", 3)
    assert magic == b"PF" and float(scale) < 0
    w, h = map(int, dims.split())
    img = np.frombuffer(rest, "<f4").reshape(h, w, 3)
    return img[::-1]


def read_rgbe(path):
    """Returns the (H, W, 4) stored bytes of a Radiance RGBE file."""
    data = open(path, "rb").read()
    head, _, body = data.partition(b"\n\n")
    assert head.startswith(b"#?") and b"FORMAT=32-bit_rle_rgbe" in head
    res, _, body = body.partition(b"\n")
    _, h, _, w = res.split()
    h, w = int(h), int(w)
    out = np.zeros((h, w, 4), np.uint8)
    pos = 0
    for y in range(h):
        if 8 <= w <= 0x7fff:
            assert body[pos:pos + 2] == b"\x02\x02" and (body[pos + 2] << 8 | body[pos + 3]) == w
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = body[pos]
                    pos += 1
                    assert n != 0
                    if n > 128:
                        out[y, x:x + n - 128, c] = body[pos]
                        pos += 1
                        x += n - 128
                    else:
                        out[y, x:x + n, c] = np.frombuffer(body[pos:pos + n], np.uint8)
                        pos += n
                        x += n
                assert x == w
        else:
            out[y] = np.frombuffer(body[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
    assert pos == len(body)
    return out


def rgbe_to_float(b):
    """Decoded value (byte + 0.5) 2^(e - 136), zero where e = 0."""
    b = b.astype(np.float64)
    f = np.where(b[..., 3:4] > 0, np.ldexp(1.0, (b[..., 3:4] - 136).astype(np.int64)), 0.0)
    return (b[..., :3] + 0.5) * f


def rgbe_within_step(dec, ref):
    """Each decoded value lies within half a step of the pixel's shared
    exponent grid, 2^(e-8) with e the exponent of its largest component."""
    ref = np.asarray(ref, np.float64)
    mx = ref.max(axis=-1, keepdims=True)
    step = np.ldexp(1.0, np.frexp(mx)[1] - 8)
    step = np.where(mx < 1e-32, 0.0, step)
    return bool((np.abs(dec - np.where(mx < 1e-32, 0.0, ref)) <= 0.5 * step).all())


def test_png_and_ppm_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (13, 21, 4), dtype=np.uint8)
    img[0, 0, :3] = (10, 32, 9)   # raster starting with bytes that read as whitespace
    m.write_image(tmp_path / "a.png", img)
    m.write_image(tmp_path / "a.PPM", img)
    assert (read_png(tmp_path / "a.png") == img[..., :3]).all()
    assert (read_ppm(tmp_path / "a.PPM") == img[..., :3]).all()


def test_pfm_round_trip(tmp_path):
    rng = np.random.default_rng(6)
    img = rng.gamma(0.7, 2.0, (9, 14, 3)).astype(np.float32)
    m.write_image(tmp_path / "a.pfm", img)
    assert (read_pfm(tmp_path / "a.pfm") == img).all()


def test_rgbe_known_values(tmp_path):
    img = np.array([[[1.0, 1.0, 1.0], [0.5, 0.5, 0.5], [1.0, 0.5, 0.25], [0, 0, 0],
                     [1e-40, 0, 0], [3.0, 0.0, 1.5]]], np.float32)
    m.write_image(tmp_path / "k.hdr", img)
    b = read_rgbe(tmp_path / "k.hdr")[0]
    assert b.tolist() == [[128, 128, 128, 129], [128, 128, 128, 128], [128, 64, 32, 129], [0, 0, 0, 0],
                          [0, 0, 0, 0], [192, 0, 96, 130]]


@pytest.mark.parametrize("w", [5, 8, 300])
def test_rgbe_round_trip(tmp_path, w):
    """Flat rows below 8 pixels, run-length rows from 8; runs and literals
    both occur (constant stretches and noise).  Every decoded value is within
    one step of its shared-exponent grid: |decoded - c| <= 2^(e-8)."""
    rng = np.random.default_rng(w)
    img = rng.lognormal(0.0, 2.0, (7, w, 3)).astype(np.float32)
    img[2, :] = (0.25, 4.0, 1.0)
    img[3, : w // 2] = 0
    img[4, ::3] = (1e6, 2e-3, 7.0)
    path = tmp_path / "r.rgbe"
    m.write_image(path, img)
    b = read_rgbe(path)
    assert rgbe_within_step(rgbe_to_float(b), img)
    mx = img.max(axis=2)
    assert ((b[..., 3] == 0) == (mx < 1e-32)).all()
    assert (b[..., :3].max(axis=2)[mx > 1e-32] >= 128).all()


def test_ldr_to_hdr_is_bytes_over_255(tmp_path):
    img = np.zeros((2, 3, 4), np.uint8)
    img[..., 0] = [[0, 51, 102], [153, 204, 255]]
    m.write_image(tmp_path / "l.pfm", img)
    got = read_pfm(tmp_path / "l.pfm")
    assert np.allclose(got[..., 0], img[..., 0] / 255.0, rtol=1e-7)


def test_unsupported_formats(tmp_path):
    img = np.zeros((2, 2, 4), np.uint8)
    for name in ("a.jpg", "a.gif", "a.tiff", "a.exr", "noext"):
        with pytest.raises(m.MptError):
            m.write_image(tmp_path / name, img)
    with pytest.raises(m.MptError):
        m.write_image(tmp_path / "a.png", np.zeros((2, 2, 3), np.float32))
