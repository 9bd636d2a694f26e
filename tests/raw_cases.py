"""Raw-file cases shared by the CPU (oracle) and GPU ingest tests: each builds file bytes with the
independent writer (tests/raw_writer.py) and states the image a correct reader must decode."""
from __future__ import annotations

import numpy as np

import raw_writer as W


def scene(h, w, bits, seed=0, noise=40.0):
    """Smooth structure + noise + a few steps, clipped to the bit depth."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    top = (1 << bits) - 1
    v = (np.sin(x / 9.0 + 0.3 * seed) + np.cos(y / 6.0) + 2.0) / 4.0 * 0.8 * top
    v += (x > w / 3) * 0.1 * top - (y > h / 2) * 0.05 * top
    v += rng.normal(0, noise * top / 16383.0, (h, w))
    return np.clip(np.rint(v), 0, top).astype(np.uint16)


def dng_expected(img, lin=None, bh=None, bv=None):
    v = img.astype(np.int64)
    if lin is not None:
        lin = np.asarray(lin, np.int64)
        v = lin[np.minimum(v, len(lin) - 1)]
    if bh is not None:
        v = v - np.asarray(bh, np.int64)[None, :]
    if bv is not None:
        v = v - np.asarray(bv, np.int64)[:, None]
    return np.maximum(v, 0).astype(np.uint16)


def cases():
    """name -> (file bytes, expected RawImage or None when only the oracle states it)."""
    out = {}
    a = scene(37, 54, 14, 1)
    for order in "<>":
        o = "le" if order == "<" else "be"
        out[f"pef_huff_{o}"] = (W.write_pef(a, order=order), a)
        out[f"pef_huff_pentaxhdr_{o}"] = (W.write_pef(a, order=order, mn_header="PENTAX"), a)
        out[f"pef_unc14_{o}"] = (W.write_pef(a, order=order, compressed=False, bits=14), a)
    b = scene(64, 80, 12, 2)
    out["pef_huff12_fixedtable"] = (W.write_pef(b, bits=12, lengths={0: 3, 1: 2, 2: 2, 3: 3, 4: 3, 5: 4, 6: 5, 7: 6, 8: 7,
                                                                       9: 8, 10: 9, 11: 10, 12: 11, 13: 12, 14: 12}), b)
    out["pef_unc12_strips"] = (W.write_pef(b, bits=12, compressed=False, rows_per_strip=16), b)
    out["pef_unc16"] = (W.write_pef(b, bits=16, compressed=False), b)
    out["pef_huff_odd"] = (W.write_pef(scene(29, 33, 14, 3)), scene(29, 33, 14, 3))
    c = scene(48, 64, 14, 4)
    for comps in (1, 2, 4):
        for psv in range(1, 8):
            out[f"dng_tiles_c{comps}_p{psv}"] = (W.write_dng(c, tile=(32, 16), comps=comps, psv=psv), dng_expected(c, bh=None))
    # black level 256 is metadata only: the pixels keep their values; deltas are subtracted at decode time
    d = scene(45, 70, 14, 5)
    out["dng_tiles_edge_be"] = (W.write_dng(d, order=">", tile=(32, 16)), d)
    out["dng_tiles_shared_table"] = (W.write_dng(d, tile=(32, 16), tables="shared"), d)
    out["dng_tiles_restart"] = (W.write_dng(c, tile=(32, 16), restart_rows=3), c)
    out["dng_tiles_restart_each_row"] = (W.write_dng(c, tile=(32, 16), restart_rows=1, comps=1), c)
    out["dng_strips_ljpeg"] = (W.write_dng(d, layout="strips", rows_per_strip=15), d)
    out["dng_ifd0_raw"] = (W.write_dng(d, tile=(32, 16), sub_ifd=False), d)
    e = scene(40, 48, 16, 6, noise=2000)
    out["dng_16bit_large"] = (W.write_dng(e, bits=16, tile=(24, 20), white=65535), e)
    out["dng_unc16"] = (W.write_dng(d, layout="strips", compression=1, bits=16, rows_per_strip=9), d)
    out["dng_unc12"] = (W.write_dng((d >> 2).astype(np.uint16), layout="strips", compression=1, bits=12,
                                    rows_per_strip=5), (d >> 2).astype(np.uint16))
    out["dng_unc14_be"] = (W.write_dng(d, order=">", layout="strips", compression=1, bits=14), d)
    lin = (np.arange(16384) * 3 // 4 + 10).astype(np.uint16)[: 12000]
    rng = np.random.default_rng(7)
    bh = rng.integers(0, 40, 70)
    bv = rng.integers(0, 30, 45)
    out["dng_lin_deltas"] = (W.write_dng(d, tile=(32, 16), linearization=lin, delta_h=bh, delta_v=bv),
                             dng_expected(d, lin, bh, bv))
    aa = (4, 6, 41, 66)  # top, left, bottom, right
    bh2 = np.zeros(70, np.int64)
    bv2 = np.zeros(45, np.int64)
    bh2[6:66] = rng.integers(0, 50, 60)
    bv2[4:41] = rng.integers(0, 20, 37)
    out["dng_active_area_deltas"] = (W.write_dng(d, tile=(32, 16), active_area=aa, delta_h=bh2[6:66],
                                                 delta_v=bv2[4:41]), dng_expected(d, None, bh2, bv2))
    return out
