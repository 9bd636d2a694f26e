"""GPU global pre-alignment (isa_prealign) against the oracle's full-frame FFT cross-correlation: the same
integer shift on every frame (the device searches the lags a coarse-to-fine pyramid leaves open, the oracle
all of them) and the same correlation value."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.gpu_util import dev, empty, host, p, stream, workspace  # noqa: E402


def _texture(rng, h, w, margin=0):
    """A high-passed-looking texture: band-limited noise plus blobs, clamped to [0, 1] like the tracking images."""
    H, W = h + 2 * margin, w + 2 * margin
    f = np.fft.rfft2(rng.standard_normal((H, W)))
    ky = np.fft.fftfreq(H)[:, None]
    kx = np.fft.rfftfreq(W)[None, :]
    r = np.sqrt(kx ** 2 + ky ** 2)
    band = (r > 0.0125) * np.exp(-(r / 0.12) ** 2)
    img = np.fft.irfft2(f * band, s=(H, W))
    img = img / img.std() * 0.15 + 0.1
    return np.clip(img, 0, 1)


def _run(isa, frames, max_shift):
    n = len(frames)
    h, w = frames[0].shape
    d = [dev(np.ascontiguousarray(f, np.float32)) for f in frames]
    ptrs = (C.c_void_p * n)(*[t.data_ptr() for t in d])
    out = empty((n, 4), fill=-99.0)
    ws = workspace(isa.lib().isa_prealign_workspace(w, h, max_shift))
    isa.call("isa_prealign", ptrs, w * 4, w, h, n, max_shift, p(out), p(ws), stream())
    return host(out)


@pytest.mark.parametrize("h,w,max_shift,shifts", [
    (256, 320, 64, [(0, 0), (37, -21), (-60, 44), (5, 63)]),                   # coarsest level 4
    (768, 1024, 256, [(0, 0), (-130, 97), (201, -188), (3, -2), (-255, 0)]),   # coarsest level 8
    (250, 333, 40, [(0, 0), (17, -39), (-40, 12)]),                           # coarsest level 2, odd width
    (1024, 1536, 128, [(0, 0), (-101, 77), (120, -127), (64, 9)]),            # coarsest level 16
])
def test_prealign_circular_shifts(isa, oracle, gpu, h, w, max_shift, shifts):
    """Frames circularly shifted by known integer offsets (np.roll by d: the content moves by d): every frame's
    result is (-d, angle 0) and equals the oracle's, value within 1e-6 relative."""
    rng = np.random.default_rng(h + w)
    f0 = _texture(rng, h, w)
    frames = [np.roll(f0, (d[1], d[0]), axis=(0, 1)) + 0.01 * rng.standard_normal((h, w)) for d in shifts]
    frames[0] = f0
    print(f"prealign {w}x{h} max_shift {max_shift}", flush=True)
    got = _run(isa, frames, max_shift)
    want = oracle.prealign(frames)
    np.testing.assert_array_equal(got[:, :3], want[:, :3])
    np.testing.assert_allclose(got[1:, 3], want[1:, 3], rtol=1e-6)
    for i, d in enumerate(shifts):
        assert (got[i, 0], got[i, 1]) == (-d[0], -d[1])


def test_prealign_crops_of_one_scene(isa, oracle, gpu):
    """Crops of a larger scene (no wrap in the content: the circular correlation also sees the mismatched
    edges), up to 64 px apart, with sensor noise: the device's result is the oracle's global maximum."""
    rng = np.random.default_rng(61)
    h, w, m = 384, 512, 80
    scene = _texture(rng, h, w, margin=m)
    offs = [(0, 0), (64, -51), (-47, 60), (12, 33), (-64, -64)]
    frames = [scene[m + oy:m + oy + h, m + ox:m + ox + w] + 0.01 * rng.standard_normal((h, w)) for ox, oy in offs]
    got = _run(isa, frames, 72)
    want = oracle.prealign(frames)
    np.testing.assert_array_equal(got[:, :3], want[:, :3])
    for i, (ox, oy) in enumerate(offs):
        # frame_i(y) = scene(y + off_i) = frame_0(y + off_i): the content moved by -off, the result is +off
        assert (got[i, 0], got[i, 1]) == (ox, oy)
