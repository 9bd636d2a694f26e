"""GPU: the mixed-precision tile search (cross terms on the matrix cores) against the exact FP64 search.

The MFMA forms change D(s) by the rounding of their operands (bf16 hi+lo: ~2^-16 relative per
product; fp16 hi+lo with per-tile scaling: ~2^-22; fp32: ~2^-24) and of the fp32 accumulation, so
they are compared with the exact path by error bounds, not bit for bit:
  - per pair: the sub-pixel shifts of well-matched tiles agree to a few hundredths of a pixel
    (a near-tie of two lags in a flat tile may still pick the other integer minimum);
  - whole merge: the merged image stays within the stated PSNR of the exact path, and the loss
    against the noiseless ground truth is at most a few hundredths of a dB.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.gpu_util import dev, host, p, stream  # noqa: E402

XC = {"bf16x3": 1, "fp16x3": 2, "fp32": 3}
# fraction of tiles within TOL px of the exact shift, and the tolerance
BOUNDS = {"bf16x3": (0.97, 0.05), "fp16x3": (0.985, 0.01), "fp32": (0.985, 0.01)}


def _pair(rng, h, w):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    a = 0.5 + 0.2 * np.sin(xx / 3.1) * np.cos(yy / 2.7) + 0.1 * np.sin((xx - 2 * yy) / 5.3)
    a = a + 0.03 * rng.standard_normal((h, w))
    b = np.roll(a, (1, -2), axis=(0, 1)) + 0.01 * rng.standard_normal((h, w))
    return a.astype(np.float32), np.clip(b, 0, 1).astype(np.float32)


@pytest.mark.parametrize("prec", list(XC))
@pytest.mark.parametrize("tile,M,h,w", [(32, 2, 200, 260), (64, 2, 210, 280), (32, 4, 200, 240), (16, 4, 120, 150),
                                        (48, 3, 160, 200)])
def test_track_xc_close_to_exact(isa, gpu, prec, tile, M, h, w):
    rng = np.random.default_rng(tile * 10 + M)
    a, b = _pair(rng, h, w)
    cx, cy = (w - 2 * M) // tile, (h - 2 * M) // tile
    pre = rng.uniform(-1.5, 1.5, (cy, cx, 2)).astype(np.float32)
    bs = (0.75, -0.25)
    print(f"xc case {prec} tile={tile} M={M} tiles={cx}x{cy}", flush=True)
    out = {}
    for name, code in [("exact", 0), (prec, XC[prec])]:
        d_pre = dev(pre)
        isa.call("isa_track_xc", p(dev(b)), p(dev(a)), w * 4, p(d_pre), cx * 8, w, h, tile, M, isa.f32s(*bs),
                 C.c_float(0.0), isa.f32s(*bs), C.c_float(0.0), C.c_float(1.0), code, stream())
        out[name] = host(d_pre).reshape(-1, 2) - pre.reshape(-1, 2)
    d = np.abs(out[prec] - out["exact"]).max(axis=1)
    frac, tol = BOUNDS[prec]
    print(f"  max {d.max():.3g} px, median {np.median(d):.2e}, within {tol}: {np.mean(d <= tol):.4f}")
    assert np.mean(d <= tol) >= frac and np.median(d) < tol / 5


def test_track_xc_rejects_unknown_precision(isa, gpu):
    a = np.zeros((64, 64), np.float32)
    d_pre = dev(np.zeros((1, 1, 2), np.float32))
    with pytest.raises(isa.IsaError):
        isa.call("isa_track_xc", p(dev(a)), p(dev(a)), 64 * 4, p(d_pre), 8, 64, 64, 32, 2, isa.f32s(0, 0),
                 C.c_float(0.0), isa.f32s(0, 0), C.c_float(0.0), C.c_float(1.0), 7, stream())


def _psnr(x, y):
    mse = float(np.mean((x.astype(np.float64) - y.astype(np.float64)) ** 2))
    return 99.0 if mse == 0 else 10 * np.log10(1.0 / mse)


@pytest.mark.parametrize("prec", list(XC))
def test_merge_xc_against_exact_and_ground_truth(isa, gpu, prec):
    w, h, n = 320, 256, 5
    cfg = isa.default_synth(seed=91, width=w, height=h, frames=n, global_shift_max=6.0, occluded_fraction=0.03)
    raw_t, _ = isa.synth_burst(cfg)
    levels = [(1, 32, 2), (2, 32, 4), (4, 16, 4)]
    res = {}
    for name, code in [("exact", 0), (prec, XC[prec])]:
        m = isa.BurstMerger(w, h, n, isa.default_params(levels=levels, xc_precision=code))
        res[name] = (host(m.run(raw_t)).copy(), host(m.buffer_tensor(2, 0, 2)).copy(), m.reference_index)
    gt = host(isa.ground_truth(cfg, res["exact"][2], w, h))
    sl = (slice(4, -4), slice(4, -4))
    out_x, out_e = res[prec][0][sl], res["exact"][0][sl]
    fin = np.isfinite(out_x) & np.isfinite(out_e)
    p_pair = _psnr(out_x[fin], out_e[fin])
    loss = _psnr(out_e[fin], gt[sl][fin]) - _psnr(out_x[fin], gt[sl][fin])
    ds = np.abs(res[prec][1] - res["exact"][1]).max(axis=-1)
    print(f"{prec}: merged PSNR vs exact {p_pair:.1f} dB, PSNR loss vs ground truth {loss:+.4f} dB, "
          f"one-to-one shift |d| median {np.median(ds):.2e} max {ds.max():.3g} px")
    assert p_pair > 40.0 and loss < 0.05
