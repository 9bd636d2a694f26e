"""GPU parity: every C-ABI kernel entry point of libisa_hip.so against the CPU
oracle (reference semantics, FP32) on the same seeded inputs.

Tolerances: integer / index / copy work is bit-exact; arithmetic built only
from +,-,*,/,sqrt (IEEE-correct on both sides, FMA contraction off) is
bit-exact; expressions that call expf/powf/atan2f/sinf/cosf are compared at
a few ULP (device and host libm differ in the last bits)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.gpu_util import dev, empty, host, p, stream, ulp_diff, within_reorder_bound, workspace  # noqa: E402

RGGB = [0, 1, 1, 2]
GRBG = [1, 0, 2, 1]
GBRG = [1, 2, 0, 1]
BGGR = [2, 1, 1, 0]


def _raw(rng, h, w, lo=200, hi=16000):
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.45 + 0.25 * np.sin(xx / 6.3 + 0.4) * np.cos(yy / 4.1) + 0.1 * np.sin((xx + yy) / 2.3)
    img = base * hi + rng.normal(0, 60, (h, w))
    return np.clip(img, lo, 16383).astype(np.uint16)


@pytest.mark.parametrize("cfa", [RGGB, GRBG, GBRG, BGGR])
@pytest.mark.parametrize("h,w", [(70, 94), (37, 301)])  # one block; odd sizes over 3 x 3 blocks
def test_debayer_full(isa, oracle, gpu, cfa, h, w):
    rng = np.random.default_rng(1)
    raw = _raw(rng, h, w)
    black, white, cw = [64.0, 60.0, 62.0], [16383.0, 16000.0, 15800.0], [1.9, 1.0, 1.4]
    out = empty((h, w, 3), fill=7.0)
    isa.call("isa_debayer_full", w, h, p(dev(raw)), w * 2, p(out), w * 12, isa.f32s(*black), isa.f32s(*white),
             isa.f32s(*cw), isa.i32s(*cfa), stream())
    ref = oracle.debayer_full(raw, black, white, cw, cfa)
    np.testing.assert_array_equal(host(out), ref)


@pytest.mark.parametrize("cfa", [GBRG, BGGR])
def test_merger_grey_front_end(isa, oracle, gpu, cfa):
    """The merger's fused debayer + sRGB + grey + Gaussian (its Lucas-Kanade images), bit-exact
    against the oracle for CFA layouts other than the pipeline tests' RGGB."""
    w, h, n = 290, 166, 2
    cfg = isa.default_synth(seed=5, width=w, height=h, frames=n)
    raw_t, _ = isa.synth_burst(cfg)
    params = isa.default_params(levels=[(1, 32, 2)], cfa=cfa)
    m = isa.BurstMerger(w, h, n, params)
    m.run(raw_t)
    raw = host(raw_t).view(np.uint16)
    for f in range(n):
        ref = oracle.tracking_image(raw[f], params, 0, params.sigma_debayer_accumulation)
        np.testing.assert_array_equal(host(m.buffer_tensor(1, f, 1)), ref)


def test_debayer_green_redblue(isa, oracle, gpu):
    rng = np.random.default_rng(2)
    h, w = 40, 52
    raw = _raw(rng, h, w).astype(np.float32)
    black, scale = [10.0, 12.0, 8.0], [0.5, 1.0, 0.7]
    out = empty((h, w, 3), fill=0.0)
    d_in = dev(raw)
    isa.call("isa_debayer_green", w, h, p(d_in), w * 4, p(out), w * 12, isa.f32s(*black), isa.f32s(*scale),
             isa.i32s(*RGGB), stream())
    isa.call("isa_debayer_redblue", w, h, p(d_in), w * 4, p(out), w * 12, isa.f32s(*black), isa.f32s(*scale),
             isa.i32s(*RGGB), stream())
    ref = np.zeros((h, w, 3), np.float32)
    fg, frb = oracle.fn("debayer_green"), oracle.fn("debayer_redblue")
    b, s, c = oracle.arr(black, np.float32), oracle.arr(scale, np.float32), oracle.arr(RGGB, np.int32)
    fg(w, h, oracle.ptr(raw), oracle.ptr(ref), oracle.ptr(b), oracle.ptr(s), oracle.ptr(c))
    frb(w, h, oracle.ptr(raw), oracle.ptr(ref), oracle.ptr(b), oracle.ptr(s), oracle.ptr(c))
    np.testing.assert_array_equal(host(out), ref)


def test_gamma_gray(isa, oracle, gpu):
    rng = np.random.default_rng(3)
    h, w = 33, 47
    img = rng.uniform(-0.2, 1.2, (h, w, 3)).astype(np.float32)
    img[0, 0, 0] = np.nan
    d = dev(img)
    isa.call("isa_gamma_srgb", p(d), w, h, w * 12, stream())
    ref = img.copy()
    oracle.fn("gamma_srgb")(w, h, oracle.ptr(ref))
    np.testing.assert_array_equal(host(d), ref)
    g = empty((h, w))
    coef = [0.2989, 0.5870, 0.1141]
    isa.call("isa_color_to_gray", p(d), w * 12, p(g), w * 4, w, h, isa.f32s(*coef), stream())
    gref = np.zeros((h, w), np.float32)
    oracle.fn("color_to_gray")(w, h, oracle.ptr(host(d)), oracle.ptr(gref), oracle.ptr(oracle.arr(coef, np.float32)))
    np.testing.assert_array_equal(host(g), gref)


# one channel with w % 4 == 0 takes the register form (float4 lanes, neighbours by shuffles): widths
# of one vector, of a partial wave, of several waves (wave-edge lanes load their neighbours) and
# ragged heights; other widths the LDS form
@pytest.mark.parametrize("sigma,ch,h,w", [(0.5, 1, 29, 61), (1.0, 3, 29, 61), (2.5, 1, 29, 61), (1.0, 4, 29, 61),
                                          (0.5, 1, 29, 8), (0.5, 1, 70, 260), (0.5, 1, 33, 1028),
                                          (1.0, 1, 70, 516), (0.7, 1, 65, 772), (1.5, 1, 40, 516)])
def test_gauss_filter(isa, oracle, gpu, sigma, ch, h, w):
    rng = np.random.default_rng(4)
    img = rng.uniform(0, 1, (h, w, ch) if ch > 1 else (h, w)).astype(np.float32)
    taps = (C.c_float * 99)()
    n = isa.lib().isa_gaussian_taps(sigma, taps, 99)
    np.testing.assert_array_equal(np.array(taps[:n], np.float32), oracle.gaussian_taps(sigma))
    out = empty(img.shape)
    tmp = empty(img.shape)
    isa.call("isa_gauss_filter_border", p(dev(img)), w * 4 * ch, p(out), w * 4 * ch, p(tmp), w, h, ch, taps, n,
             stream())
    ref = oracle.gauss_filter(img, np.array(taps[:n], np.float32))
    np.testing.assert_array_equal(host(out), ref)


# Rows of 4 w bytes: 16-byte aligned for w = 96 and 700 (matrix-core GEMMs; at 700 x 300 with
# hp + hps = 0.12 the band has 11 column tiles in 2 groups and the K sum 11 chunks, the last ragged),
# 8-byte aligned for w = 130 (SIMT GEMMs).
@pytest.mark.parametrize("w,h,hp,hps", [(96, 64, 0.01, 0.0025), (130, 90, 0.01, 0.0025), (700, 300, 0.1, 0.02)])
def test_fourier_highpass(isa, oracle, gpu, w, h, hp, hps):
    rng = np.random.default_rng(5)
    yy, xx = np.mgrid[0:h, 0:w]
    img = (0.5 + 0.3 * np.sin(xx / 17.0) * np.cos(yy / 23.0) + 0.05 * rng.standard_normal((h, w))).astype(np.float32)
    ws_bytes = isa.lib().isa_fourier_highpass_workspace(w, h, 0, C.c_float(hp), C.c_float(hps))
    ws = empty((ws_bytes // 4 + 64,))
    d = dev(img)
    isa.call("isa_fourier_highpass", p(d), w * 4, w, h, 0, C.c_float(hp), C.c_float(hps), p(ws), stream())
    ref = oracle.fourier_highpass(img, hp, hps, dt=np.float64)
    err = np.abs(host(d).astype(np.float64) - ref).max()
    assert err < 2e-5, err


# At 1024 x 512 the row DFT sums 8 K chunks of 128 columns; at 700 x 300 the band has 4 column tiles.
@pytest.mark.parametrize("w,h,hp,hps", [(96, 64, 0.01, 0.0025), (700, 300, 0.1, 0.02), (1024, 512, 0.01, 0.0025)])
def test_fourier_highpass_bounded(isa, oracle, gpu, w, h, hp, hps):
    """The fp16-split matrix-core form for an image in [0, 1]: within 2e-5 of the FP64 oracle (the bar of the
    FP32-product form) and within 2e-6 of the FP32-product form; a bound too large for fp16 operands gives the
    FP32-product form bit for bit."""
    rng = np.random.default_rng(6)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 0.5 + 0.3 * np.sin(xx / 17.0) * np.cos(yy / 23.0) + 0.05 * rng.standard_normal((h, w))
    img = np.clip(img, 0.0, 1.0).astype(np.float32)
    ws = empty((isa.lib().isa_fourier_highpass_workspace(w, h, 0, C.c_float(hp), C.c_float(hps)) // 4 + 64,))
    d32, d16, dbig = dev(img), dev(img), dev(img)
    args = (w * 4, w, h, 0, C.c_float(hp), C.c_float(hps))
    isa.call("isa_fourier_highpass", p(d32), *args, p(ws), stream())
    isa.call("isa_fourier_highpass_bounded", p(d16), *args, C.c_float(1.0), p(ws), stream())
    isa.call("isa_fourier_highpass_bounded", p(dbig), *args, C.c_float(1e9), p(ws), stream())
    ref = oracle.fourier_highpass(img, hp, hps, dt=np.float64)
    e64 = np.abs(host(d16).astype(np.float64) - ref).max()
    e32 = np.abs(host(d16).astype(np.float64) - host(d32)).max()
    print(f"highpass bounded {w}x{h}: max |fp16 split - FP64| {e64:.3g}, max |fp16 split - FP32 products| {e32:.3g}")
    assert e64 < 2e-5 and e32 < 2e-6
    np.testing.assert_array_equal(host(dbig), host(d32))
    with pytest.raises(isa.IsaError):
        isa.call("isa_fourier_highpass_bounded", p(d16), *args, C.c_float(float("nan")), p(ws), stream())


# clearAxis: h = 41 with c = 1 clears |ky| = 1 by the reference's float test; c = 40 on 96 x 64 clears every bin
# (output 0); 130 x 90 runs the SIMT GEMMs, 700 x 300 the matrix cores with a ragged band.
@pytest.mark.parametrize("w,h,c,hp,hps", [(64, 41, 1, 0.01, 0.0025), (96, 64, 3, 0.01, 0.0025), (130, 90, 2, 0.01, 0.0025),
                                          (700, 300, 5, 0.1, 0.02), (96, 64, 40, 0.01, 0.0025)])
def test_fourier_highpass_clear_axis(isa, oracle, gpu, w, h, c, hp, hps):
    """isa_fourier_highpass and its bounded form with clear_axis > 0 against the FP64 oracle (a full DFT with the
    reference's cleared bins, itself pinned against numpy's FFT), at the bar of the plain filter."""
    rng = np.random.default_rng(50 + c)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 0.5 + 0.2 * np.sin(xx / 7.0) + 0.1 * np.cos(yy / 5.0) + 0.1 * np.sin(xx / 11.0 + yy / 9.0)
    img = np.clip(img + 0.05 * rng.standard_normal((h, w)), 0.0, 1.0).astype(np.float32)
    nbytes = isa.lib().isa_fourier_highpass_workspace(w, h, c, C.c_float(hp), C.c_float(hps))
    assert nbytes > isa.lib().isa_fourier_highpass_workspace(w, h, 0, C.c_float(hp), C.c_float(hps))
    ws = empty((nbytes // 4 + 64,))
    d, db = dev(img), dev(img)
    args = (w * 4, w, h, c, C.c_float(hp), C.c_float(hps))
    print(f"highpass clear_axis {c} {w}x{h}", flush=True)
    isa.call("isa_fourier_highpass", p(d), *args, p(ws), stream())
    isa.call("isa_fourier_highpass_bounded", p(db), *args, C.c_float(1.0), p(ws), stream())
    ref = oracle.fourier_highpass(img, hp, hps, clear_axis=c, dt=np.float64)
    e = np.abs(host(d).astype(np.float64) - ref).max()
    eb = np.abs(host(db).astype(np.float64) - ref).max()
    print(f"  max |GPU - FP64| {e:.3g}, bounded {eb:.3g}")
    assert e < 2e-5 and eb < 2e-5
    # the cleared axes matter: the result differs from the plain filter's
    plain = oracle.fourier_highpass(img, hp, hps, dt=np.float64)
    assert np.abs(plain - ref).max() > 1e-3
    with pytest.raises(isa.IsaError):
        isa.call("isa_fourier_highpass", p(d), w * 4, w, h, -1, C.c_float(hp), C.c_float(hps), p(ws), stream())


@pytest.mark.parametrize("h,w,which", [(64, 96, (2, 4, 8)), (40, 520, (2, 4, 8)), (16, 8, (2, 8)), (48, 1032, (4,))])
def test_resize_box_levels(isa, oracle, gpu, h, w, which):
    """The one-pass pyramid equals isa_resize_box (the oracle's box means) bit for bit, for any subset of levels."""
    rng = np.random.default_rng(16)
    img = rng.uniform(0, 1, (h, w)).astype(np.float32)
    outs = {l: (empty((h // l, w // l), fill=-1.0) if l in which else None) for l in (2, 4, 8)}
    ptr = {l: (p(outs[l]) if outs[l] is not None else None) for l in (2, 4, 8)}
    isa.call("isa_resize_box_levels", p(dev(img)), w * 4, w, h, ptr[2], (w // 2) * 4, ptr[4], (w // 4) * 4, ptr[8],
             (w // 8) * 4, stream())
    for l in which:
        np.testing.assert_array_equal(host(outs[l]), oracle.downscale_box(img, l))


def test_resize_box(isa, oracle, gpu):
    rng = np.random.default_rng(6)
    img = rng.uniform(0, 1, (64, 96)).astype(np.float32)
    for l in (2, 4, 8):
        out = empty((64 // l, 96 // l))
        isa.call("isa_resize_box", p(dev(img)), 96 * 4, p(out), (96 // l) * 4, 96, 64, l, stream())
        np.testing.assert_array_equal(host(out), oracle.downscale_box(img, l))


def _track_inputs(rng, h=80, w=112):
    yy, xx = np.mgrid[0:h, 0:w]
    a = (0.5 + 0.2 * np.sin(xx / 3.1) * np.cos(yy / 2.7) + 0.05 * rng.standard_normal((h, w))).astype(np.float32)
    b = np.roll(a, (1, -2), axis=(0, 1)) + (0.01 * rng.standard_normal((h, w))).astype(np.float32)
    return a, b.astype(np.float32)


# (tile, M, base shift, rotation, image h, w). Tiles dividing 64 with M in {1, 2, 4} take the
# row-parallel tile-search kernel, the others the block-parallel one; both must give the same bits.
TRACK_CASES = [
    (16, 2, (1.25, -0.5), 0.0, 80, 112),
    (16, 4, (1.25, -0.5), 0.0, 80, 112),
    (32, 3, (1.25, -0.5), 0.0, 80, 112),
    (64, 2, (6.5, -3.0), 0.01, 150, 212),   # tiles cut by the image edge (clamped reference rows)
    (32, 4, (-5.0, 4.5), -0.02, 120, 170),
    (8, 1, (0.0, 0.0), 0.0, 60, 70),
    (48, 2, (2.0, 1.0), 0.0, 110, 150),     # 64 % 48 != 0: block-parallel form
]


@pytest.mark.parametrize("tile,M,bs,rot,h,w", TRACK_CASES)
def test_tiles_ssd_findmin(isa, oracle, gpu, tile, M, bs, rot, h, w):
    rng = np.random.default_rng(7 + M + tile)
    a, b = _track_inputs(rng, h, w)
    cx, cy = (w - 2 * M) // tile, (h - 2 * M) // tile
    B, S = tile + 2 * M, 2 * M + 1
    pre = rng.uniform(-3, 3, (cy, cx, 2)).astype(np.float32)
    print(f"track case tile={tile} M={M} tiles={cx}x{cy}", flush=True)
    rt, tt = empty((cx * cy, B, B)), empty((cx * cy, B, B))
    isa.call("isa_convert_to_tiles_border", p(dev(a)), w * 4, p(rt), w, h, M, tile, cx, cy, isa.f32s(*bs),
             C.c_float(rot), stream())
    isa.call("isa_convert_to_tiles_preshift", p(dev(b)), w * 4, p(tt), p(dev(pre)), cx * 8, w, h, M, tile, cx, cy,
             isa.f32s(*bs), C.c_float(rot), stream())
    rt_ref = oracle.tiles_border(a, M, tile, cx, cy, bs, rot)
    tt_ref = oracle.tiles_preshift(b, pre, M, tile, cx, cy, bs, rot)
    np.testing.assert_array_equal(host(rt), rt_ref)
    np.testing.assert_array_equal(host(tt), tt_ref)
    simg = empty((cx * cy, S, S))
    isa.call("isa_tile_ssd", p(rt), p(tt), p(simg), M, tile, cx * cy, stream())
    s_ref = oracle.tile_ssd(rt_ref, tt_ref, M, tile)
    np.testing.assert_array_equal(host(simg), s_ref)  # FP64 in the same order: bit-exact
    coords = empty((cy, cx, 2))
    isa.call("isa_find_minimum", p(simg), p(coords), cx * 8, M, cx, cy, C.c_float(1.0), stream())
    np.testing.assert_array_equal(host(coords).reshape(-1, 2), oracle.find_minimum(s_ref, M, 1.0))
    # fused PatchTracker.Track
    d_pre = dev(pre)
    ws = workspace(isa.lib().isa_track_workspace(w, h, w * 4, tile, M, 0))
    isa.call("isa_track", p(dev(b)), p(dev(a)), w * 4, p(d_pre), cx * 8, w, h, tile, M, isa.f32s(*bs),
             C.c_float(rot), isa.f32s(*bs), C.c_float(rot), C.c_float(1.0), p(ws), stream())
    np.testing.assert_array_equal(host(d_pre).reshape(-1, 2), pre.reshape(-1, 2) + oracle.find_minimum(s_ref, M, 1.0))


@pytest.mark.parametrize("tile,M", [(16, 1), (16, 2), (32, 4)])
def test_track_ties_take_the_first_minimum(isa, oracle, gpu, tile, M):
    """On an image of period 2 in x, shifts 2 px apart give exactly equal FP64 SSD values, so the
    shift images hold tied minima; the fused search must take the first one in scan order, as
    findMinimum does."""
    rng = np.random.default_rng(31 + M)
    h, w = 96, 128
    col = rng.uniform(0.2, 0.8, (h, 2)).astype(np.float32)
    a = np.ascontiguousarray(np.tile(col, (1, w // 2)))
    b = np.ascontiguousarray(np.roll(a, 1, axis=0))
    cx, cy = (w - 2 * M) // tile, (h - 2 * M) // tile
    bs, rot = (0.0, 0.0), 0.0
    pre = rng.integers(-1, 2, (cy, cx, 2)).astype(np.float32)
    print(f"track ties tile={tile} M={M}", flush=True)
    rt_ref = oracle.tiles_border(a, M, tile, cx, cy, bs, rot)
    tt_ref = oracle.tiles_preshift(b, pre, M, tile, cx, cy, bs, rot)
    s_ref = oracle.tile_ssd(rt_ref, tt_ref, M, tile)
    mins = s_ref.reshape(len(s_ref), -1)
    assert np.mean((mins == mins.min(axis=1, keepdims=True)).sum(axis=1) > 1) > 0.5  # ties are common
    d_pre = dev(pre)
    ws = workspace(isa.lib().isa_track_workspace(w, h, w * 4, tile, M, 0))
    isa.call("isa_track", p(dev(b)), p(dev(a)), w * 4, p(d_pre), cx * 8, w, h, tile, M, isa.f32s(*bs),
             C.c_float(rot), isa.f32s(*bs), C.c_float(rot), C.c_float(1.0), p(ws), stream())
    np.testing.assert_array_equal(host(d_pre).reshape(-1, 2), pre.reshape(-1, 2) + oracle.find_minimum(s_ref, M, 1.0))


def test_upsample_shifts(isa, oracle, gpu):
    rng = np.random.default_rng(8)
    ocx, ocy, ncx, ncy = 7, 5, 15, 11
    inp = rng.uniform(-4, 4, (ocy, ocx, 2)).astype(np.float32)
    out = empty((ncy, ncx, 2))
    isa.call("isa_upsample_shifts", p(dev(inp)), p(out), ocx * 8, ncx * 8, 4, 2, ocx, ocy, ncx, ncy, 32, 32, stream())
    np.testing.assert_array_equal(host(out), oracle.upsample_shifts(inp, 4, 2, ocx, ocy, ncx, ncy, 32, 32))


# frame counts up to 8, 16, 24 and 32 take the register forms of the minimizer (lane per frame), 40 the
# LDS form; the larger cases carry extra outliers so that tiles iterate several times
@pytest.mark.parametrize("n,strategy,block", [(5, 0, 5), (8, 0, 5), (6, 1, 5), (9, 2, 4), (7, 3, 3), (20, 0, 5),
                                              (17, 2, 6), (30, 3, 4), (40, 0, 5)])
def test_minimize_shifts(isa, oracle, gpu, n, strategy, block):
    rng = np.random.default_rng(9 + n)
    ref = 0 if strategy != 2 else 3
    A, pairs, index, rows = oracle.shift_tables(n, ref, strategy, block)
    m = A.shape[0]
    tiles = 37
    x = np.concatenate([np.zeros((tiles, 1, 2)), rng.uniform(-6, 6, (tiles, n - 1, 2))], axis=1)
    d = x[:, rows[:, 1]] - x[:, rows[:, 0]] + rng.normal(0, 0.05, (tiles, m, 2))
    # outliers in some tiles
    for t in range(0, tiles, 3):
        d[t, rng.integers(0, m)] += rng.uniform(3, 8, 2)
    if n > 9:
        for t in range(1, tiles, 4):
            for k in rng.choice(m, 4, replace=False):
                d[t, k] += rng.uniform(2, 9, 2)
    d = d.astype(np.float32).astype(np.float64)
    dh_ref, r_ref, st_ref, it_ref = oracle.minimize_shifts(d, n, ref, strategy, block)
    meas = dev(np.ascontiguousarray(np.transpose(d, (1, 0, 2))).astype(np.float32))  # [m][tile][2]
    r = empty((tiles, n - 1, 2))
    st = empty((tiles,), dtype=__import__("torch").int32)
    it = empty((tiles,), dtype=__import__("torch").int32)
    ws = workspace(isa.lib().isa_minimize_shifts_workspace(n, ref, strategy, block))
    isa.call("isa_minimize_shifts", p(meas), tiles * 8, p(r), p(st), p(it), tiles, n, ref, strategy, block, p(ws),
             stream())
    np.testing.assert_array_equal(host(st), st_ref)
    np.testing.assert_array_equal(host(it), it_ref)
    np.testing.assert_allclose(host(r), r_ref, atol=2e-5)
    np.testing.assert_allclose(np.transpose(host(meas), (1, 0, 2)), dh_ref, atol=2e-5)


def test_get_optimal_shifts(isa, oracle, gpu):
    rng = np.random.default_rng(10)
    n, cx, cy = 6, 5, 4
    r = rng.uniform(-3, 3, (cx * cy, n - 1, 2)).astype(np.float32)
    for ref, trk in [(0, 4), (3, 1), (2, 2)]:
        out = empty((cy, cx, 2))
        isa.call("isa_get_optimal_shifts", p(out), cx * 8, p(dev(r)), n, cx, cy, ref, trk, stream())
        np.testing.assert_array_equal(host(out).reshape(-1, 2), oracle.get_optimal_shifts(r, n, ref, trk))


@pytest.mark.parametrize("quant", [0, 1])
def test_flow_warp_derivatives(isa, oracle, gpu, quant):
    rng = np.random.default_rng(11)
    h, w = 48, 64
    tiles = rng.uniform(-2, 2, (3, 4, 2)).astype(np.float32)
    flow = empty((h, w, 2))
    isa.call("isa_create_flow_field_from_tiles", p(flow), w * 8, p(dev(tiles)), 4 * 8, 4, 3, w, h,
             isa.f32s(0.5, -0.25), C.c_float(0.0), quant, stream())
    fref = np.zeros((h, w, 2), np.float32)
    oracle.fn("flow_from_tiles")(oracle.ptr(tiles), 4, 3, oracle.ptr(fref), w, h, 0.5, -0.25, 0.0, quant)
    np.testing.assert_array_equal(host(flow), fref)
    src, tgt = _track_inputs(rng, h, w)
    out = empty((h, w))
    isa.call("isa_warp", p(dev(src)), w * 4, p(out), w * 4, p(flow), w * 8, w, h, quant, stream())
    wref = np.zeros((h, w), np.float32)
    oracle.fn("warp")(oracle.ptr(src), oracle.ptr(fref), oracle.ptr(wref), w, h, quant)
    np.testing.assert_array_equal(host(out), wref)
    ix, iy, iz = empty((h, w)), empty((h, w)), empty((h, w))
    isa.call("isa_compute_derivatives", p(dev(wref)), p(dev(tgt)), w * 4, p(ix), p(iy), p(iz), w * 4, w, h, quant,
             stream())
    rx, ry, rz = (np.zeros((h, w), np.float32) for _ in range(3))
    oracle.fn("derivatives")(oracle.ptr(wref), oracle.ptr(tgt), oracle.ptr(rx), oracle.ptr(ry), oracle.ptr(rz), w, h,
                             quant)
    for a, b in ((ix, rx), (iy, ry), (iz, rz)):
        np.testing.assert_array_equal(host(a), b)
    isa.call("isa_compute_derivatives2", p(dev(tgt)), w * 4, p(ix), p(iy), w * 4, w, h, quant, stream())
    oracle.fn("derivatives2")(oracle.ptr(tgt), oracle.ptr(rx), oracle.ptr(ry), w, h, quant)
    np.testing.assert_array_equal(host(ix), rx)
    np.testing.assert_array_equal(host(iy), ry)


def test_lucas_kanade_step(isa, oracle, gpu):
    rng = np.random.default_rng(12)
    h, w = 48, 64
    src, tgt = _track_inputs(rng, h, w)
    rx, ry, rz = (np.zeros((h, w), np.float32) for _ in range(3))
    oracle.fn("derivatives")(oracle.ptr(src), oracle.ptr(tgt), oracle.ptr(rx), oracle.ptr(ry), oracle.ptr(rz), w, h, 0)
    flow0 = rng.uniform(-0.5, 0.5, (h, w, 2)).astype(np.float32)
    f = dev(flow0)
    isa.call("isa_lucas_kanade_optim", p(f), w * 8, p(dev(rx)), p(dev(ry)), p(dev(rz)), w * 4, w, h, 11,
             C.c_float(0.01), stream())
    ref = flow0.copy()
    oracle.fn("lucas_kanade_step")(oracle.ptr(ref), oracle.ptr(rx), oracle.ptr(ry), oracle.ptr(rz), w, h, 5, 0.01,
                                   0.0)
    np.testing.assert_array_equal(host(f), ref)


@pytest.mark.parametrize("quant,iters,window,rot", [(1, 3, 11, 0.0), (0, 2, 7, 0.004), (1, 0, 11, 0.0)])
def test_lucas_kanade_drop_in(isa, oracle, gpu, quant, iters, window, rot):
    """OpticalFlow.LucasKanade (R:PEFStudioDX/OpticalFlow.cs:147-164) as one call: flow from the tile
    shifts, `iters` x {warp, derivatives, LK step} with the reference's per-pixel window loops, then the
    final warp written over the source. Bit for bit against the oracle's lucas_kanade."""
    rng = np.random.default_rng(40 + iters + window)
    h, w = 56, 72
    src, tgt = _track_inputs(rng, h, w)
    tw, th = 5, 4
    tiles = rng.uniform(-1.5, 1.5, (th, tw, 2)).astype(np.float32)
    bs = (0.75, -0.5)
    print(f"lucas_kanade quant={quant} iters={iters} window={window}", flush=True)
    d_src, flow, work = dev(src), empty((h, w, 2)), empty((4, h, w))
    isa.call("isa_lucas_kanade", p(d_src), p(dev(tgt)), w * 4, p(dev(tiles)), tw * 8, tw, th, p(flow), w * 8, p(work),
             w, h, iters, isa.f32s(*bs), C.c_float(rot), C.c_float(0.01), window, quant, stream())
    s_ref, f_ref = src.copy(), np.zeros((h, w, 2), np.float32)
    oracle.fn("lucas_kanade")(oracle.ptr(s_ref), oracle.ptr(tgt), oracle.ptr(tiles), tw, th, oracle.ptr(f_ref), w, h,
                              iters, bs[0], bs[1], rot, 0.01, window, quant, 0.0, None)
    np.testing.assert_array_equal(host(flow), f_ref)
    np.testing.assert_array_equal(host(d_src), s_ref)


def test_add_shifts(isa, gpu):
    """NPP Add of the float view of the shift images (R:PEFStudioDX/PatchTracker.cs:355-358), pitched."""
    rng = np.random.default_rng(44)
    cx, cy = 13, 7
    pre = rng.uniform(-5, 5, (cy, cx + 3, 2)).astype(np.float32)  # 3 columns of row padding
    co = rng.uniform(-1, 1, (cy, cx + 1, 2)).astype(np.float32)
    d_pre = dev(pre)
    isa.call("isa_add_shifts", p(d_pre), (cx + 3) * 8, p(dev(co)), (cx + 1) * 8, cx, cy, stream())
    want = pre.copy()
    want[:, :cx] += co[:, :cx]
    np.testing.assert_array_equal(host(d_pre), want)


def test_structure_tensor_kernel_param(isa, oracle, gpu):
    rng = np.random.default_rng(13)
    h, w = 20, 30
    dx, dy = (rng.normal(0, 0.05, (h, w)).astype(np.float32) for _ in range(2))
    dx[0, :5] = 0
    dy[0, :5] = 0  # zero gradient: NaN kernel (cross fallback later)
    t = empty((h, w, 3))
    isa.call("isa_compute_structure_tensor", p(dev(dx)), p(dev(dy)), w * 4, p(t), w * 12, w, h, stream())
    tref = np.zeros((h, w, 3), np.float32)
    oracle.fn("structure_tensor")(oracle.ptr(dx), oracle.ptr(dy), oracle.ptr(tref), w, h)
    np.testing.assert_array_equal(host(t), tref)
    args = (0.001, 0.006, 0.25, 3.0, 4.0, 2.0)
    isa.call("isa_compute_kernel_param", p(t), w, h, w * 12, *[C.c_float(a) for a in args], stream())
    oracle.fn("kernel_param")(oracle.ptr(tref), w, h, *args)
    np.testing.assert_array_equal(host(t), tref)


@pytest.mark.parametrize("cfa", [RGGB, GRBG])
@pytest.mark.parametrize("hw,hh", [(36, 10), (38, 9), (260, 7)])
def test_subsample3_forms(isa, oracle, gpu, cfa, hw, hh):
    """deBayersSubSample3 bit for bit: four output pixels per thread (half width a multiple of 4, aligned rows)
    and the per-pixel form (38)."""
    rng = np.random.default_rng(hw + hh)
    raw = _raw(rng, 2 * hh, 2 * hw)
    out = empty((hh, hw, 3), fill=-1.0)
    isa.call("isa_debayer_subsample3", p(dev(raw)), p(out), C.c_float(16383.0), hw, hh, hw * 12, isa.i32s(*cfa),
             stream())
    ref = np.zeros((hh, hw, 3), np.float32)
    oracle.fn("subsample3")(oracle.ptr(raw), oracle.ptr(ref), 16383.0, hw, hh, oracle.ptr(oracle.arr(cfa, np.int32)))
    np.testing.assert_array_equal(host(out), ref)


def test_subsample_robustness_erode(isa, oracle, gpu):
    rng = np.random.default_rng(14)
    h, w = 64, 80
    hw, hh = w // 2, h // 2
    ra, rb = _raw(rng, h, w), _raw(rng, h, w)
    rb[20:40, 30:50] = np.clip(rb[20:40, 30:50].astype(int) + 3000, 0, 16383)
    ha, hb = empty((hh, hw, 3)), empty((hh, hw, 3))
    isa.call("isa_debayer_subsample3", p(dev(ra)), p(ha), C.c_float(16383.0), hw, hh, hw * 12, isa.i32s(*RGGB),
             stream())
    isa.call("isa_debayer_subsample3", p(dev(rb)), p(hb), C.c_float(16383.0), hw, hh, hw * 12, isa.i32s(*RGGB),
             stream())
    sa, sb = np.zeros((hh, hw, 3), np.float32), np.zeros((hh, hw, 3), np.float32)
    c = oracle.arr(RGGB, np.int32)
    oracle.fn("subsample3")(oracle.ptr(ra), oracle.ptr(sa), 16383.0, hw, hh, oracle.ptr(c))
    oracle.fn("subsample3")(oracle.ptr(rb), oracle.ptr(sb), 16383.0, hw, hh, oracle.ptr(c))
    np.testing.assert_array_equal(host(ha), sa)
    np.testing.assert_array_equal(host(hb), sb)
    flow = rng.uniform(-3, 3, (h, w, 2)).astype(np.float32)
    for quant in (0, 1):
        mask = empty((hh, hw, 4), fill=1.0)
        alpha, beta = 0.5 / 16383, 9 / 16383 ** 2
        isa.call("isa_compute_robustness_mask", p(ha), p(hb), hw * 12, p(mask), hw * 16, p(dev(flow)), w * 8, w, h,
                 hw, hh, C.c_float(alpha), C.c_float(beta), C.c_float(1e5), quant, stream())
        mref = np.ones((hh, hw, 4), np.float32)
        oracle.fn("robustness_mask")(oracle.ptr(sa), oracle.ptr(sb), oracle.ptr(mref), oracle.ptr(flow), w, h, hw, hh,
                                     alpha, beta, 1e5, quant)
        np.testing.assert_array_equal(host(mask), mref)
        er = empty((hh, hw, 4))
        isa.call("isa_erode_border4", p(dev(mref)), hw * 16, p(er), hw * 16, hw, hh, 5, stream())
        eref = np.zeros_like(mref)
        oracle.fn("erode")(oracle.ptr(mref), oracle.ptr(eref), hw, hh, 5)
        np.testing.assert_array_equal(host(er), eref)
        # window sizes of the staged (separable) form and of the direct form
        for size in (3, 7, 17):
            er = empty((hh, hw, 4))
            isa.call("isa_erode_border4", p(dev(mref)), hw * 16, p(er), hw * 16, hw, hh, size, stream())
            eref = np.zeros_like(mref)
            oracle.fn("erode")(oracle.ptr(mref), oracle.ptr(eref), hw, hh, size)
            np.testing.assert_array_equal(host(er), eref)


@pytest.mark.parametrize("cfa", [RGGB, GRBG])
@pytest.mark.parametrize("layout", ["aligned", "offset"])
@pytest.mark.parametrize("h,w", [(72, 160), (38, 130)])
def test_accumulate_frames(isa, oracle, gpu, cfa, layout, h, w):
    """Several frames in one pass against the oracle's accumulateImages called once per frame in the
    same order: the parity-grouped wide-load form (aligned frames) and the general form (one raw
    frame at an odd 2-byte offset), with NaN kernel and mask values, wild flows and windows that
    leave the image."""
    rng = np.random.default_rng(41 + cfa[0])
    n = 6
    hw, hh = w // 2, h // 2
    raws = [_raw(rng, h, w) for _ in range(n)]
    kern = np.zeros((h, w, 3), np.float32)
    kern[..., 0] = rng.uniform(0.2, 2.0, (h, w))
    kern[..., 1] = rng.uniform(0.2, 2.0, (h, w))
    kern[..., 2] = rng.uniform(-0.1, 0.1, (h, w))
    kern[5, 9] = np.nan
    masks = [rng.uniform(0, 1, (hh, hw, 4)).astype(np.float32) for _ in range(n)]
    masks[2][3, 4, 1] = np.nan
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    flows = [None]
    for f in range(1, n):
        g = rng.uniform(-6, 6, 2)
        fl = np.stack([g[0] + 1.5 * np.sin(xx / 17.0 + f), g[1] + 1.5 * np.cos(yy / 13.0 - f)], -1)
        fl += rng.uniform(-0.6, 0.6, fl.shape)
        flows.append(fl.astype(np.float32))
    flows[3][8:16, 40:72] = rng.uniform(-40, 40, (8, 32, 2))  # one block's windows spread past the LDS box
    flows[4][:, :, 0] += 30.0                                  # windows leaving the image on the right
    s0 = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    w0 = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    white, black = [16383.0, 16000.0, 15000.0], [10.0, 20.0, 30.0]
    print(f"accumulate_frames {n} x {w}x{h} cfa={cfa}", flush=True)
    d_raw = [dev(r) for r in raws]
    if layout == "offset":  # frame 1 one sample into a larger buffer: not 4-byte aligned
        buf = dev(np.concatenate([np.zeros(1, np.uint16), raws[1].ravel()]))
        d_raw[1] = buf[1:]
        assert d_raw[1].data_ptr() % 4 == 2
    d_mask = [dev(m) for m in masks]
    d_flow = [None if fl is None else dev(fl) for fl in flows]
    ptrs = lambda ts: (C.c_void_p * n)(*[None if t is None else t.data_ptr() for t in ts])  # noqa: E731
    ds, dw = dev(s0), dev(w0)
    isa.call("isa_accumulate_frames", ptrs(d_raw), ptrs(d_mask), ptrs(d_flow), n, p(ds), p(dw), p(dev(kern)),
             isa.f32s(*white), isa.f32s(*black), w, h, w * 12, hw * 16, w * 8, isa.i32s(*cfa), stream())
    zero = np.zeros((h, w, 2), np.float32)
    res = {}
    for order in ("accumulate_class_order", "accumulate"):
        sr_, wr_ = s0.copy(), w0.copy()
        for f in range(n):
            oracle.fn(order)(oracle.ptr(raws[f]), oracle.ptr(sr_), oracle.ptr(wr_), oracle.ptr(masks[f]),
                             oracle.ptr(kern), oracle.ptr(zero if flows[f] is None else flows[f]),
                             oracle.ptr(oracle.arr(white, np.float32)), oracle.ptr(oracle.arr(black, np.float32)),
                             w, h, oracle.ptr(oracle.arr(cfa, np.int32)))
        res[order] = (sr_, wr_)
    # bit for bit against the same tap order (class by class); within the reordering bound of the
    # reference's raster order
    got_s, got_w = host(ds), host(dw)
    np.testing.assert_array_equal(got_s, res["accumulate_class_order"][0])
    np.testing.assert_array_equal(got_w, res["accumulate_class_order"][1])
    assert within_reorder_bound(got_s, got_w, *res["accumulate"], s0, w0, 25 * n, vmax=16383 / 15000)


@pytest.mark.parametrize("sr,w", [(0, 64), (1, 64), (0, 62)])
def test_accumulate_apply_weighting(isa, oracle, gpu, sr, w):
    rng = np.random.default_rng(15 + sr)
    h = 48
    hw, hh = w // 2, h // 2
    raw = _raw(rng, h, w)
    kch = 4 if sr else 3
    kern = np.zeros((h, w, kch), np.float32)
    kern[..., 0] = rng.uniform(0.2, 2.0, (h, w))
    kern[..., 1] = rng.uniform(0.2, 2.0, (h, w))
    kern[..., 2] = rng.uniform(-0.1, 0.1, (h, w))
    kern[3, 3, :3] = np.nan
    mask = rng.uniform(0, 1, (hh, hw, 4)).astype(np.float32)
    flow = rng.uniform(-2.6, 2.6, (h, w, 2)).astype(np.float32)
    s0 = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    w0 = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    white, black = [16383.0, 16000.0, 15000.0], [10.0, 20.0, 30.0]
    ds, dw = dev(s0), dev(w0)
    if sr:
        isa.call("isa_accumulate_images_superres", p(dev(raw)), p(ds), p(dw), p(dev(mask)), p(dev(kern)),
                 p(dev(flow)), isa.f32s(*white), isa.f32s(*black), w, h, w * 12, hw * 16, w * 16, w * 8,
                 isa.i32s(*RGGB), 0, stream())
    else:
        isa.call("isa_accumulate_images", p(dev(raw)), p(ds), p(dw), p(dev(mask)), p(dev(kern)), p(dev(flow)),
                 isa.f32s(*white), isa.f32s(*black), w, h, w * 12, hw * 16, w * 8, isa.i32s(*RGGB), stream())
    sr_, wr_ = s0.copy(), w0.copy()
    args = [oracle.ptr(raw), oracle.ptr(sr_), oracle.ptr(wr_), oracle.ptr(mask), oracle.ptr(kern), oracle.ptr(flow),
            oracle.ptr(oracle.arr(white, np.float32)), oracle.ptr(oracle.arr(black, np.float32)), w, h,
            oracle.ptr(oracle.arr(RGGB, np.int32))]
    if sr:
        oracle.fn("accumulate_sr")(*args, 0)  # raster order on both sides
    else:
        oracle.fn("accumulate_class_order")(*args)
        s_r, w_r = s0.copy(), w0.copy()
        oracle.fn("accumulate")(*(args[:1] + [oracle.ptr(s_r), oracle.ptr(w_r)] + args[3:]))
        assert within_reorder_bound(host(ds), host(dw), s_r, w_r, s0, w0, 25, vmax=16383 / 15000)
    np.testing.assert_array_equal(host(ds), sr_)
    np.testing.assert_array_equal(host(dw), wr_)
    io = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    dio = dev(io)
    isa.call("isa_apply_weighting", p(dio), p(dev(sr_)), p(dev(wr_)), w, h, w * 12, C.c_float(1.0), stream())
    ioref = io.copy()
    oracle.fn("apply_weighting")(oracle.ptr(ioref), oracle.ptr(sr_), oracle.ptr(wr_), w, h, 1.0)
    np.testing.assert_array_equal(host(dio), ioref)


def test_srgb_fast_power_is_exact(isa, oracle, gpu):
    """The certified fast power of the sRGB OETF returns the float of the exact one (pow in double,
    rounded once) for every float of the power branch (0.0031308, 1], and matches the oracle."""
    import torch

    lo = int(np.float32(0.0031308).view(np.int32)) - 4
    hi = int(np.float32(1.0).view(np.int32)) + 4
    print(f"srgb exhaustive over {hi - lo + 1} floats", flush=True)
    n_all = hi - lo + 1
    chunk = 1 << 24
    for c0 in range(lo, hi + 1, chunk):
        bits = torch.arange(c0, min(hi + 1, c0 + chunk), dtype=torch.int32, device="cuda")
        x = bits.view(torch.float32)
        fast, exact = torch.empty_like(x), torch.empty_like(x)
        isa.call("isa_srgb_oetf", p(x), p(fast), C.c_size_t(x.numel()), 0, stream())
        isa.call("isa_srgb_oetf", p(x), p(exact), C.c_size_t(x.numel()), 1, stream())
        torch.cuda.synchronize()
        assert torch.equal(fast.view(torch.int32), exact.view(torch.int32)), c0
    assert n_all > 7 * 10**7
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(0, 1, 100000), rng.uniform(-0.5, 1.5, 1000),
                         [np.nan, 0.0, -0.0, 1.0, 0.0031308, np.float32(0.0031308) * 1.0000001, 2.0, -1.0]]).astype(np.float32)
    d_out = empty(xs.shape)
    isa.call("isa_srgb_oetf", p(dev(xs)), p(d_out), C.c_size_t(xs.size), 0, stream())
    img = np.ascontiguousarray(np.repeat(xs[:, None], 3, axis=1))
    oracle.fn("gamma_srgb")(xs.size, 1, oracle.ptr(img))
    np.testing.assert_array_equal(host(d_out), img[:, 0])


@pytest.mark.parametrize("grid", ["reference", "full"])
@pytest.mark.parametrize("quant", [0, 1])
def test_accumulate_frames_superres(isa, oracle, gpu, grid, quant):
    """accumulateImagesSuperRes over four frames in one pass, on the reference's central-half grid (w x h,
    offset (w/2, h/2)) and on the full-frame 2x grid (2w x 2h, offset 0), bit for bit against the oracle's
    per-frame accumulation in the same order (raster taps); frame 0 has the zero (null) flow, with NaN kernel
    and certainty values."""
    rng = np.random.default_rng(81 + quant)
    h, w, n = 40, 56, 4
    hw, hh = w // 2, h // 2
    ow, oh, ox, oy = (w, h, w // 2, h // 2) if grid == "reference" else (2 * w, 2 * h, 0, 0)
    raws = [_raw(rng, h, w) for _ in range(n)]
    kern = np.zeros((h, w, 4), np.float32)
    kern[..., 0] = rng.uniform(0.2, 2.0, (h, w))
    kern[..., 1] = rng.uniform(0.2, 2.0, (h, w))
    kern[..., 2] = rng.uniform(-0.1, 0.1, (h, w))
    kern[7, 11, :3] = np.nan
    masks = [rng.uniform(0, 1, (hh, hw, 4)).astype(np.float32) for _ in range(n)]
    masks[1][5, 6, 0] = np.nan
    flows = [None] + [rng.uniform(-2.6, 2.6, (h, w, 2)).astype(np.float32) for _ in range(n - 1)]
    s0 = rng.uniform(0, 1, (oh, ow, 3)).astype(np.float32)
    w0 = rng.uniform(0, 1, (oh, ow, 3)).astype(np.float32)
    white, black = [16383.0, 16000.0, 15000.0], [10.0, 20.0, 30.0]
    print(f"superres accumulation {grid} {ow}x{oh} quant={quant}", flush=True)
    d_raw, d_mask = [dev(r) for r in raws], [dev(m) for m in masks]
    d_flow = [None if f is None else dev(f) for f in flows]
    ptrs = lambda ts: (C.c_void_p * n)(*[None if t is None else t.data_ptr() for t in ts])  # noqa: E731
    ds, dw = dev(s0), dev(w0)
    isa.call("isa_accumulate_frames_superres", ptrs(d_raw), ptrs(d_mask), ptrs(d_flow), n, p(ds), p(dw), p(dev(kern)),
             isa.f32s(*white), isa.f32s(*black), w, h, ow, oh, ox, oy, ow * 12, hw * 16, w * 16, w * 8,
             isa.i32s(*RGGB), quant, stream())
    sr_, wr_ = s0.copy(), w0.copy()
    zero = np.zeros((h, w, 2), np.float32)
    for f in range(n):
        oracle.fn("accumulate_sr_grid")(oracle.ptr(raws[f]), oracle.ptr(sr_), oracle.ptr(wr_), oracle.ptr(masks[f]),
                                        oracle.ptr(kern), oracle.ptr(zero if flows[f] is None else flows[f]),
                                        oracle.ptr(oracle.arr(white, np.float32)),
                                        oracle.ptr(oracle.arr(black, np.float32)), w, h,
                                        oracle.ptr(oracle.arr(RGGB, np.int32)), quant, ow, oh, ox, oy)
    np.testing.assert_array_equal(host(ds), sr_)
    np.testing.assert_array_equal(host(dw), wr_)
    with pytest.raises(isa.IsaError):  # a grid reaching past the frame is rejected before any launch
        isa.call("isa_accumulate_frames_superres", ptrs(d_raw), ptrs(d_mask), ptrs(d_flow), n, p(ds), p(dw),
                 p(dev(kern)), isa.f32s(*white), isa.f32s(*black), w, h, 2 * w, 2 * h, 2, 0, 2 * w * 12, hw * 16,
                 w * 16, w * 8, isa.i32s(*RGGB), quant, stream())


@pytest.mark.parametrize("fmt,crop", [("rgb16", None), ("rgb16", (3, 2, 41, 27)), ("bgra8", (1, 3, 50, 30))])
def test_color_pipeline(isa, oracle, gpu, fmt, crop):
    """isa_color_pipeline (ThresholdLTGT, ColorTwist, ThresholdLTGT, LUTCubic, ColorTwist, Mul, GammasRGB, Mul,
    ROI, Convert) equals the oracle bit for bit: values outside [0, white] and NaN, a tone curve of 256 levels,
    realistic twists with offsets, a non-zero exposure, both output formats."""
    from imagestackalignator_amd import export

    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(91 + len(fmt))
    h, w = 33, 57
    img = rng.uniform(-0.1, 1.3, (h, w, 3)).astype(np.float32)
    img[5, 7, 1] = np.nan
    wb = [0.92, 1.0, 0.81]
    m1 = [[0.71, 0.21, 0.05, 0.002], [0.12, 0.88, -0.03, 0.0], [0.01, -0.09, 1.07, -0.001]]
    m2 = [[1.82, -0.51, -0.21, 0.0], [-0.19, 1.22, -0.03, 0.0], [0.01, -0.13, 1.12, 0.0]]
    lx = (np.arange(256) / 255.0).astype(np.float32)
    ly = (lx + 0.15 * np.sin(2 * np.pi * lx) * lx * (1 - lx)).astype(np.float32)  # a mild S-curve
    out = export.color_pipeline(dev(img), wb, m1, dev(lx), dev(ly), m2, exposure=0.35, crop=crop, fmt=fmt)
    want = oracle.color_pipeline(img, wb, m1, lx, ly, m2, 0.35, crop=crop, bgra8=fmt == "bgra8")
    got = host(out)
    if fmt == "rgb16":
        got = got.view(np.uint16)
    np.testing.assert_array_equal(got, want)
    assert want.max() > (200 if fmt == "bgra8" else 50000)
    with pytest.raises(isa.IsaError):  # a crop outside the image
        export.color_pipeline(dev(img), wb, m1, dev(lx), dev(ly), m2, crop=(0, 0, w + 1, h), fmt=fmt)
