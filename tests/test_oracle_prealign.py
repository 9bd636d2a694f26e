"""Pins of the oracle's global pre-alignment (no GPU): the numpy-FFT cross-correlation against its
definition by brute force, the sign and wrap of the returned shift on circularly shifted images, and the
PreAlignmentStore rule."""
import numpy as np
import pytest


def test_cross_correlation_matches_its_definition(oracle):
    """c[sy, sx] = sum_x ref(x) trk((x + s) mod size): brute force over every lag of a tiny pair; the
    conjugate on the first operand fixes the direction of s."""
    rng = np.random.default_rng(51)
    h, w = 7, 9
    ref, trk = rng.uniform(-1, 1, (h, w)), rng.uniform(-1, 1, (h, w))
    cc = oracle.cross_correlation(ref, trk)
    for sy in range(h):
        for sx in range(w):
            want = sum(ref[y, x] * trk[(y + sy) % h, (x + sx) % w] for y in range(h) for x in range(w))
            assert abs(cc[sy, sx] - want) < 1e-12


@pytest.mark.parametrize("d", [(5, -3), (-17, 11), (0, 0), (31, 23), (-30, -22)])
def test_prealign_returns_minus_the_content_motion(oracle, d):
    """A frame whose content moved by d (frame_i(y) = frame_0(y - d), np.roll) correlates best at s = d,
    and ScanAngles returns -s, wrapped into (-W/2, W/2]."""
    rng = np.random.default_rng(52)
    h, w = 48, 64
    f0 = rng.uniform(0, 1, (h, w))
    f1 = np.roll(f0, (d[1], d[0]), axis=(0, 1))
    out = oracle.prealign([f0, f1])
    np.testing.assert_array_equal(out[0], 0.0)
    assert (out[1, 0], out[1, 1]) == (-d[0], -d[1])
    assert out[1, 2] == 0.0 and abs(out[1, 3] - np.sum(f0 * f0)) < 1e-9  # the peak is the autocorrelation


def test_prealign_wrap_boundary(oracle):
    """The maximum at index W/2 stays +W/2 (only > W/2 wraps), one past it becomes negative."""
    h, w = 16, 20
    rng = np.random.default_rng(53)
    f0 = rng.uniform(0, 1, (h, w))
    out = oracle.prealign([f0, np.roll(f0, (0, w // 2), axis=(0, 1)), np.roll(f0, (h // 2 + 1, 0), axis=(0, 1))])
    assert out[1, 0] == -(w // 2)
    assert out[2, 1] == -(h // 2 + 1 - h)


def test_prealignment_store_rule(oracle):
    """Frames drifting by +10 px each: the middle frame minimises the summed shift length and becomes the
    reference; shifts are relative to it (R:PEFStudioDX/PreAlignmentStore.cs:36-110)."""
    init = np.array([[0, 0], [10, 0], [20, 0], [30, 0], [40, 0]], np.float64)
    ref, per = oracle.prealignment_store(init)
    assert ref == 2
    np.testing.assert_array_equal(per[:, 0], [-20, -10, 0, 10, 20])
    # no motion: frame 0
    ref, per = oracle.prealignment_store(np.zeros((4, 2)))
    assert ref == 0 and np.all(per == 0)
