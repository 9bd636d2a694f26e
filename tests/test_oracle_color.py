"""Pins of the oracle's colour pipeline and TIFF writer (no GPU): the TIFF restatement read back by an
independent reader written from the TIFF 6.0 field layout; the tone curve's cubic reproducing cubic
polynomials and its levels; round-to-nearest-even conversion; and the whole pipeline on inputs for which every
step has a closed form."""
import struct

import numpy as np
import pytest

from oracle import oracle as O


def _read_tiff(b: bytes):
    """Minimal little-endian TIFF reader: header, first IFD (tag -> (type, count, values)), strip data."""
    assert b[:4] == b"II*\x00"
    (ifd,) = struct.unpack_from("<I", b, 4)
    (n,) = struct.unpack_from("<H", b, ifd)
    size = {2: 1, 3: 2, 4: 4, 5: 8}
    fmt = {2: "c", 3: "H", 4: "I", 5: "II"}
    tags, order, value_at = {}, [], {}
    for i in range(n):
        tag, typ, cnt = struct.unpack_from("<HHI", b, ifd + 2 + 12 * i)
        field = ifd + 2 + 12 * i + 8
        at = field if size[typ] * cnt <= 4 else struct.unpack_from("<I", b, field)[0]
        vals = struct.unpack_from("<" + fmt[typ] * cnt, b, at)
        tags[tag] = (typ, cnt, vals)
        order.append(tag)
        value_at[tag] = at
    (nxt,) = struct.unpack_from("<I", b, ifd + 2 + 12 * n)
    return tags, order, value_at, nxt


def test_tiff_layout_reads_back():
    rng = np.random.default_rng(71)
    h, w = 5, 7
    rgb = rng.integers(0, 65536, (h, w, 3)).astype(np.uint16)
    b = O.tiff16_bytes(rgb)
    tags, order, at, nxt = _read_tiff(b)
    assert nxt == 0
    # the reference's construction order (R:PefFile/ImageFileDirectory.cs:76-92), Make last
    assert order == [256, 257, 258, 259, 262, 273, 277, 278, 279, 282, 283, 296, 271]
    assert tags[256] == (4, 1, (w,)) and tags[257] == (4, 1, (h,))
    assert tags[258] == (3, 3, (16, 16, 16))
    assert tags[259][2] == (1,) and tags[262][2] == (2,) and tags[277][2] == (3,) and tags[296][2] == (1,)
    assert tags[278] == (4, 1, (h,)) and tags[279] == (4, 1, (w * h * 6,))
    assert tags[282] == (5, 1, (720, 10)) and tags[283] == (5, 1, (720, 10))
    assert b"".join(tags[271][2]) == b"MK TIFF\x00"
    # values appended in entry order after the directory, then the pixels
    assert (at[258], at[282], at[283], at[271]) == (170, 176, 184, 192)
    off = tags[273][2][0]
    assert off == 200 and len(b) == off + w * h * 6
    np.testing.assert_array_equal(np.frombuffer(b[off:], "<u2").reshape(h, w, 3), rgb)


def test_lut_cubic_reproduces_cubics_and_levels():
    """The cubic through four levels reproduces any cubic polynomial (non-uniform levels, both ends), returns
    the level's value exactly at a level, and clamps outside the levels."""
    rng = np.random.default_rng(72)
    lx = np.sort(rng.uniform(0, 1, 12)).astype(np.float32)
    lx[0], lx[-1] = 0.0, 1.0
    p = lambda x: 0.3 - 1.2 * x + 2.5 * x ** 2 - 0.9 * x ** 3
    ly = p(lx.astype(np.float64)).astype(np.float32)
    for v in np.concatenate([rng.uniform(0, 1, 200), [0.0, 1.0, lx[1], lx[-2]]]).astype(np.float32):
        assert abs(O.lut_cubic(v, lx, ly) - p(float(v))) < 2e-5
    for i in range(len(lx)):
        assert O.lut_cubic(lx[i], lx, ly) == ly[i]
    assert O.lut_cubic(-0.5, lx, ly) == ly[0] and O.lut_cubic(1.7, lx, ly) == ly[-1]


def test_convert_rounds_to_nearest_even_and_saturates():
    cases = [(0.5, 0), (1.5, 2), (2.5, 2), (2.4999998, 2), (65534.5, 65534), (-3.0, 0), (70000.0, 65535), (254.5, 254)]
    for v, want in cases:
        assert O.convert_near(v, 65535.0) == want
    assert O.convert_near(300.0, 255.0) == 255.0


def _srgb(v):
    v = np.clip(np.nan_to_num(v.astype(np.float32), nan=0.0), 0, 1).astype(np.float32)
    hi = (np.float32(1.055) * np.power(v.astype(np.float64), 1 / 2.4).astype(np.float32) - np.float32(0.055))
    return np.where(v <= np.float32(0.0031308), np.float32(12.92) * v, hi).astype(np.float32)


def test_pipeline_closed_forms():
    """Inputs on the tone curve's levels with an identity curve, permutation / power-of-two twists: every step is
    exact, so the output is rint(sRGB(2^e s v_perm) scale) with the first threshold's clip, in both formats."""
    rng = np.random.default_rng(73)
    lx = (np.arange(256) / 255.0).astype(np.float32)  # the reference's LUTx
    h, w = 6, 9
    img = lx[rng.integers(0, 256, (h, w, 3))].copy()
    img[0, 0] = [-0.2, 1.5, np.nan]  # below 0, above the white level, NaN
    wb = np.array([lx[200], lx[255], lx[128]], np.float32)
    perm = np.array([[0, 0, 1, 0], [0, 1, 0, 0], [1, 0, 0, 0]], np.float32)  # camera -> ProPhoto: swap R and B
    half = np.array([[0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 0.5, 0]], np.float32)
    crop = (2, 1, 5, 4)
    cx, cy, cw, ch = crop
    wants = {}
    for bgra8 in (False, True):
        out = O.color_pipeline(img, wb, perm, lx, lx, half, 1.0, crop=crop, bgra8=bgra8)
        v = np.minimum(np.maximum(np.nan_to_num(img, nan=np.nan), 0), wb)  # NaN passes the thresholds
        v = v[..., ::-1]
        v = np.where(np.isnan(v), 0.0, v).astype(np.float32)  # NaN reads the curve's first level
        want = _srgb(np.float32(2.0) * (np.float32(0.5) * v)) * np.float32(255.0 if bgra8 else 65535.0)
        want = np.clip(np.rint(want), 0, 255 if bgra8 else 65535)
        want = wants[bgra8] = want[cy:cy + ch, cx:cx + cw]
        if bgra8:
            np.testing.assert_array_equal(out[..., :3], want[..., ::-1])
            assert np.all(out[..., 3] == 255)
        else:
            np.testing.assert_array_equal(out, want)
    # the pixel with NaN (outside the crop): 0 x NaN is NaN, so the twist spreads it to every channel, the tone
    # curve reads its first level for each, and the pixel is black
    full = O.color_pipeline(img, wb, perm, lx, lx, half, 1.0)
    np.testing.assert_array_equal(full[0, 0], [0, 0, 0])
    np.testing.assert_array_equal(full[cy:cy + ch, cx:cx + cw], wants[False])
