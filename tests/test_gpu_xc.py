"""GPU: the mixed-precision tile search (cross terms on the matrix cores) against the exact FP64 search.

The MFMA forms change D(s) by the rounding of their operands (bf16 hi+lo: ~2^-16 relative per
product; fp16 hi+lo: ~2^-22; fp32: ~2^-24) and of the fp32 accumulation. Each
lag's value carries a proven error bound; where the bounds cannot decide the integer minimum and the
threshold test, the tile is searched exactly. So:
  - per tile, the integer part of findMinimum (argmin and threshold decision) equals the exact
    search's, bit for bit, on every input: textured, flat (every lag tied) and low-contrast noise,
    at several thresholds;
  - the sub-pixel shifts of well-matched tiles agree to a few hundredths of a pixel;
  - whole merge: the merged image stays within the stated PSNR of the exact path, and the loss
    against the noiseless ground truth is at most a few hundredths of a dB.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from tests.gpu_util import dev, empty, host, p, stream, workspace  # noqa: E402

XC = {"bf16x3": 1, "fp16x3": 2, "fp32": 3}
# fraction of tiles within TOL px of the exact shift, and the tolerance
BOUNDS = {"bf16x3": (0.97, 0.05), "fp16x3": (0.985, 0.01), "fp32": (0.985, 0.01)}


def _pair(rng, h, w, kind="textured"):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    if kind == "flat":
        a = np.full((h, w), 0.25)
        return a.astype(np.float32), a.astype(np.float32)
    if kind == "noise":
        a = 0.5 + 1e-3 * rng.standard_normal((h, w))
        b = np.roll(a, (1, -2), axis=(0, 1)) + 1e-3 * rng.standard_normal((h, w))
        return a.astype(np.float32), b.astype(np.float32)
    a = 0.5 + 0.2 * np.sin(xx / 3.1) * np.cos(yy / 2.7) + 0.1 * np.sin((xx - 2 * yy) / 5.3)
    a = a + 0.03 * rng.standard_normal((h, w))
    b = np.roll(a, (1, -2), axis=(0, 1)) + 0.01 * rng.standard_normal((h, w))
    return a.astype(np.float32), np.clip(b, 0, 1).astype(np.float32)


def _track(isa, a, b, pre, w, h, tile, M, code, thr=1.0):
    cx, cy = (w - 2 * M) // tile, (h - 2 * M) // tile
    d_pre = dev(pre)
    dec = empty((cy * cx,), dtype=torch.int32, fill=-7)
    n_exact = empty((1,), dtype=torch.int32, fill=-1)
    bs = (0.75, -0.25)
    ws = workspace(isa.lib().isa_track_workspace(w, h, w * 4, tile, M, code))
    isa.call("isa_track_xc_ex", p(dev(b)), p(dev(a)), w * 4, p(d_pre), cx * 8, w, h, tile, M, isa.f32s(*bs),
             C.c_float(0.0), isa.f32s(*bs), C.c_float(0.0), C.c_float(thr), code, p(dec), p(n_exact), p(ws),
             stream())
    return host(d_pre).reshape(-1, 2) - pre.reshape(-1, 2), host(dec), int(host(n_exact)[0])


CASES = [(32, 2, 200, 260), (64, 2, 210, 280), (32, 4, 200, 240), (16, 4, 120, 150), (48, 3, 160, 200), (20, 2, 100, 130)]


@pytest.mark.parametrize("prec", list(XC))
@pytest.mark.parametrize("tile,M,h,w", CASES)
def test_track_xc_close_to_exact(isa, gpu, prec, tile, M, h, w):
    rng = np.random.default_rng(tile * 10 + M)
    a, b = _pair(rng, h, w)
    cx, cy = (w - 2 * M) // tile, (h - 2 * M) // tile
    pre = rng.uniform(-1.5, 1.5, (cy, cx, 2)).astype(np.float32)
    print(f"xc case {prec} tile={tile} M={M} tiles={cx}x{cy}", flush=True)
    ex, dec_ex, n_ex = _track(isa, a, b, pre, w, h, tile, M, 0)
    got, dec, n_exact = _track(isa, a, b, pre, w, h, tile, M, XC[prec])
    assert n_ex == cx * cy and 0 <= n_exact <= cx * cy
    np.testing.assert_array_equal(dec, dec_ex)
    d = np.abs(got - ex).max(axis=1)
    frac, tol = BOUNDS[prec]
    print(f"  max {d.max():.3g} px, median {np.median(d):.2e}, within {tol}: {np.mean(d <= tol):.4f}, "
          f"searched exactly {n_exact}/{cx * cy}")
    assert np.mean(d <= tol) >= frac and np.median(d) < tol / 5


@pytest.mark.parametrize("prec", list(XC))
@pytest.mark.parametrize("kind", ["flat", "noise"])
@pytest.mark.parametrize("thr", [0.0, 1e-5, 1e-3, 1.0])
def test_track_xc_decisions_on_near_ties(isa, gpu, prec, kind, thr):
    """Inputs whose lags are tied or nearly tied, and thresholds near the spread of D: the certificate
    must send every undecidable tile to the exact search (flat: all of them), so the decisions match."""
    tile, M, h, w = 32, 2, 168, 232
    rng = np.random.default_rng(77)
    a, b = _pair(rng, h, w, kind)
    cx, cy = (w - 2 * M) // tile, (h - 2 * M) // tile
    pre = rng.uniform(-1.5, 1.5, (cy, cx, 2)).astype(np.float32)
    print(f"xc near-tie case {prec} {kind} thr={thr}", flush=True)
    ex, dec_ex, _ = _track(isa, a, b, pre, w, h, tile, M, 0, thr)
    got, dec, n_exact = _track(isa, a, b, pre, w, h, tile, M, XC[prec], thr)
    np.testing.assert_array_equal(dec, dec_ex)
    print(f"  searched exactly {n_exact}/{cx * cy}")
    if kind == "flat":
        assert n_exact == cx * cy
        np.testing.assert_array_equal(got, ex)


def test_track_xc_rejects_unknown_precision(isa, gpu):
    a = np.zeros((64, 64), np.float32)
    d_pre = dev(np.zeros((1, 1, 2), np.float32))
    ws = workspace(isa.lib().isa_track_workspace(64, 64, 64 * 4, 32, 2, 0))
    with pytest.raises(isa.IsaError):
        isa.call("isa_track_xc", p(dev(a)), p(dev(a)), 64 * 4, p(d_pre), 8, 64, 64, 32, 2, isa.f32s(0, 0),
                 C.c_float(0.0), isa.f32s(0, 0), C.c_float(0.0), C.c_float(1.0), 7, p(ws), stream())
    assert isa.lib().isa_track_workspace(64, 64, 64 * 4, 32, 2, 7) == 0


def _psnr(x, y):
    mse = float(np.mean((x.astype(np.float64) - y.astype(np.float64)) ** 2))
    return 99.0 if mse == 0 else 10 * np.log10(1.0 / mse)


@pytest.mark.parametrize("xc_levels", [-1, 7])
@pytest.mark.parametrize("prec", list(XC))
def test_merge_xc_against_exact_and_ground_truth(isa, gpu, prec, xc_levels):
    """xc_levels -1 (default): the 32-pixel levels on the matrix cores, the 16-pixel level exact; 7: every
    level on the matrix cores."""
    w, h, n = 320, 256, 5
    cfg = isa.default_synth(seed=91, width=w, height=h, frames=n, global_shift_max=6.0, occluded_fraction=0.03)
    raw_t, _ = isa.synth_burst(cfg)
    levels = [(1, 32, 2), (2, 32, 4), (4, 16, 4)]
    on_mfma = [True, True, xc_levels == 7]
    res = {}
    for name, code in [("exact", 0), (prec, XC[prec])]:
        m = isa.BurstMerger(w, h, n, isa.default_params(levels=levels, xc_precision=code, xc_levels=xc_levels))
        res[name] = (host(m.run(raw_t)).copy(), host(m.buffer_tensor(2, 0, 2)).copy(), m.reference_index,
                     m.exact_items())
        for level, (n_exact, items) in enumerate(res[name][3]):
            assert 0 <= n_exact <= items and (code != 0 or n_exact == items), (level, n_exact, items)
            if code != 0:  # a level on the matrix cores certifies most items; the others are all searched exactly
                assert (n_exact < items) == on_mfma[level], (level, n_exact, items)
    gt = host(isa.ground_truth(cfg, res["exact"][2], w, h))
    sl = (slice(4, -4), slice(4, -4))
    out_x, out_e = res[prec][0][sl], res["exact"][0][sl]
    fin = np.isfinite(out_x) & np.isfinite(out_e)
    p_pair = _psnr(out_x[fin], out_e[fin])
    loss = _psnr(out_e[fin], gt[sl][fin]) - _psnr(out_x[fin], gt[sl][fin])
    ds = np.abs(res[prec][1] - res["exact"][1]).max(axis=-1)
    print(f"{prec}: merged PSNR vs exact {p_pair:.1f} dB, PSNR loss vs ground truth {loss:+.4f} dB, "
          f"one-to-one shift |d| median {np.median(ds):.2e} max {ds.max():.3g} px, items searched exactly "
          f"{res[prec][3]}")
    assert p_pair > 40.0 and loss < 0.05
