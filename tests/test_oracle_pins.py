"""Known-answer tests for oracle functions whose values the other pins leave open (no GPU).

Each expected value is derived by hand from the reference's definition (the controller's grey
coefficients, the texture unit's 1.8 fixed-point filter weights, the robustness statistic's
overwritten 5x5 loop, the accumulation's raster tap order), so that a swapped coefficient, an
unquantised weight, a corrected loop or a reordered sum fails one of them."""
import numpy as np
import pytest

F32, F64 = np.float32, np.float64


def _f(v):
    """A float literal of the reference (its kernels and controller write 0.2989f, 2.4f, ...)."""
    return float(np.float32(v))


def _srgb(v):
    """sRGB OETF, IEC 61966-2-1, with the reference's float constants (R:Kernels/kernel.cu:271-313)."""
    return np.where(v <= _f(0.0031308), _f(12.92) * v, (1 + _f(0.055)) * np.power(v, 1 / _f(2.4)) - _f(0.055))


# ---------------------------------------------------------------- a2 ColorToGray
def test_color_to_gray_unit_colours(oracle):
    """Pure red, green and blue return the three coefficients, in that order
    (R:PEFStudioDX/ImageStackAlignatorController.cs:2404)."""
    coef = [0.2989, 0.5870, 0.1141]
    img = np.zeros((1, 4, 3), F64)
    img[0, 0, 0] = img[0, 1, 1] = img[0, 2, 2] = 1.0
    img[0, 3] = (0.5, 0.25, 2.0)
    out = np.zeros((1, 4), F64)
    oracle.fn("color_to_gray", F64)(4, 1, oracle.ptr(img), oracle.ptr(out), oracle.ptr(oracle.arr(coef, F64)))
    np.testing.assert_array_equal(out[0, :3], coef)
    assert abs(out[0, 3] - (0.5 * 0.2989 + 0.25 * 0.5870 + 2.0 * 0.1141)) < 1e-15


@pytest.mark.parametrize("green_only", [0, 1])
def test_tracking_image_grey_of_a_constant_colour(oracle, isa, green_only):
    """DeBayerBWGaussWB without the high-pass on a constant-colour mosaic: the interior is
    0.2989 s(R) + 0.5870 s(G) + 0.1141 s(B) (s the sRGB OETF; green only: s(G)), the Gaussian
    leaving a constant unchanged. R != B, so swapping the red and blue weights fails."""
    h, w = 28, 36
    rgb = {0: 5000, 1: 9000, 2: 1500}
    cfa = [0, 1, 1, 2]
    raw = np.zeros((h, w), np.uint16)
    for y in range(h):
        for x in range(w):
            raw[y, x] = rgb[cfa[(y % 2) * 2 + (x % 2)]]
    params = isa.default_params(levels=[(1, 8, 2)], green_channel_only=green_only)
    out = oracle.tracking_image(raw, params, 0, 0.5, dt=F64)
    s = {c: float(_srgb(rgb[c] / 16383.0)) for c in range(3)}
    want = s[1] if green_only else _f(0.2989) * s[0] + _f(0.5870) * s[1] + _f(0.1141) * s[2]
    np.testing.assert_allclose(out[4:-4, 4:-4], want, rtol=1e-13)
    swapped = _f(0.1141) * s[0] + _f(0.5870) * s[1] + _f(0.2989) * s[2]
    assert green_only or abs(want - swapped) > 1e-3


# ---------------------------------------------------------------- texture weights (tex_quant = 1)
def test_warp_filter_weights_are_8bit_fixed_point(oracle):
    """Linear filtering at fractional texel position 0.3 on a ramp along x: with quant the weight is
    the nearest multiple of 1/256, 77/256 = 0.30078125 (0.3 * 256 = 76.8), without it 0.3; the ramp
    makes the interpolated value a * (i + alpha) + b exactly."""
    h, w = 8, 40
    a, b = 0.125, 2.0
    src = np.ascontiguousarray(np.broadcast_to(a * np.arange(w) + b, (h, w)), F64)
    flow = np.zeros((h, w, 2), F64)
    flow[..., 0] = 3.3
    flow[..., 1] = -1.0  # integer: no vertical interpolation
    got = {}
    for q in (0, 1):
        out = np.zeros((h, w), F64)
        oracle.fn("warp", F64)(oracle.ptr(src), oracle.ptr(flow), oracle.ptr(out), w, h, q)
        got[q] = out
    xs = np.arange(2, w - 5)
    np.testing.assert_allclose(got[0][3, xs], a * (xs + 3 + 0.3) + b, atol=1e-12)
    np.testing.assert_allclose(got[1][3, xs], a * (xs + 3 + 77 / 256) + b, atol=1e-12)
    assert np.all(np.abs(got[1][3, xs] - got[0][3, xs] - a * (77 / 256 - 0.3)) < 1e-12)


def test_flow_from_tiles_filter_weights(oracle):
    """The tile-shift texture (linear, clamp) at every pixel: with quant the weights are multiples of
    1/256 within half a step of the exact fraction, and the two modes differ where the exact fraction
    is not such a multiple. Tile values form a ramp along x, so the sampled value is a * (i + alpha) + b
    with i the lower texel and alpha its weight."""
    tw, th, w, h = 7, 3, 50, 6
    a, b = 0.5, -1.0
    tiles = np.zeros((th, tw, 2), F64)
    tiles[..., 0] = a * np.arange(tw) + b
    res = {}
    for q in (0, 1):
        flow = np.zeros((h, w, 2), F64)
        oracle.fn("flow_from_tiles", F64)(oracle.ptr(tiles), tw, th, oracle.ptr(flow), w, h, 0.0, 0.0, 0.0, q)
        res[q] = flow[2, :, 0]
    X = (np.arange(w) + 0.5) / w * tw - 0.5  # texel coordinate of each pixel
    inside = (X >= 0) & (X <= tw - 1)
    i = np.floor(X)
    alpha_q = (res[1] - b) / a - i
    assert np.all(np.abs(alpha_q[inside] * 256 - np.round(alpha_q[inside] * 256)) < 1e-9)  # k / 256
    assert np.all(np.abs(alpha_q[inside] - (X - i)[inside]) <= 1 / 512 + 1e-12)             # nearest
    np.testing.assert_allclose(res[0][inside], a * X[inside] + b, atol=1e-12)
    assert np.any(np.abs(res[1] - res[0]) > 1e-4)
    np.testing.assert_allclose(res[1][~inside], np.where(X[~inside] < 0, b, a * (tw - 1) + b), atol=1e-12)  # clamp


def test_warp_mirror_addressing(oracle):
    """A sample left of texel 0 reads the mirrored texel -1 -> 0 (half-sample symmetric): at texel
    coordinate -0.7 both taps are texel 0, whatever the weight."""
    h, w = 4, 12
    src = np.ascontiguousarray(np.broadcast_to(np.arange(w, dtype=F64) * 3 + 1, (h, w)))
    flow = np.zeros((h, w, 2), F64)
    flow[:, 0, 0] = -0.7
    out = np.zeros((h, w), F64)
    oracle.fn("warp", F64)(oracle.ptr(src), oracle.ptr(flow), oracle.ptr(out), w, h, 0)
    np.testing.assert_allclose(out[:, 0], src[:, 0], atol=1e-12)


# ---------------------------------------------------------------- two-image warp / derivatives
def test_warp_by_an_integer_flow_shifts_the_image(oracle):
    """WarpingKernel reads src at x + flow(x): a constant integer flow (2, -1) gives src shifted by it."""
    rng = np.random.default_rng(21)
    h, w = 20, 26
    src = rng.uniform(0, 1, (h, w))
    flow = np.zeros((h, w, 2), F64)
    flow[..., 0], flow[..., 1] = 2.0, -1.0
    for q in (0, 1):
        out = np.zeros((h, w), F64)
        oracle.fn("warp", F64)(oracle.ptr(src), oracle.ptr(flow), oracle.ptr(out), w, h, q)
        if q:  # the normalised-coordinate round trip leaves a fraction of ~1e-15, which 8 bits round away
            np.testing.assert_array_equal(out[1:, : w - 2], src[:-1, 2:])
        else:
            np.testing.assert_allclose(out[1:, : w - 2], src[:-1, 2:], atol=1e-13)


def test_two_image_derivatives(oracle):
    """ComputeDerivativesKernel on two different cubics f (source) and g (target): Ix and Iy are the
    negated mean of both images' exact derivatives (the 5-point stencil is exact on cubics), Iz is
    f - g (source minus target)."""
    h, w = 24, 36
    yy, xx = np.mgrid[0:h, 0:w].astype(F64)
    f = 1e-4 * (xx ** 3 - 2 * xx * yy ** 2) + 0.01 * yy
    g = 2e-4 * (yy ** 3 + xx * xx * yy) - 0.02 * xx + 0.3
    fx, fy = 1e-4 * (3 * xx ** 2 - 2 * yy ** 2), 1e-4 * (-4 * xx * yy) + 0.01
    gx, gy = 2e-4 * (2 * xx * yy) - 0.02, 2e-4 * (3 * yy ** 2 + xx * xx)
    ix, iy, iz = np.zeros_like(f), np.zeros_like(f), np.zeros_like(f)
    oracle.fn("derivatives", F64)(oracle.ptr(np.ascontiguousarray(f)), oracle.ptr(np.ascontiguousarray(g)),
                                  oracle.ptr(ix), oracle.ptr(iy), oracle.ptr(iz), w, h, 0)
    s = np.s_[3:-3, 3:-3]
    np.testing.assert_allclose(ix[s], -(fx[s] + gx[s]) / 2, atol=1e-9)
    np.testing.assert_allclose(iy[s], -(fy[s] + gy[s]) / 2, atol=1e-9)
    np.testing.assert_allclose(iz, f - g, atol=1e-12)


def test_derivatives_of_a_warped_integer_translation(oracle):
    """Source = target translated by an integer (3, 2); warping it back by that flow and taking the
    two-image derivatives leaves no temporal difference and the target's own gradients."""
    rng = np.random.default_rng(22)
    h, w = 30, 40
    tgt = rng.uniform(0, 1, (h, w))
    src = np.zeros_like(tgt)
    src[2:, 3:] = tgt[:-2, :-3]  # src(x + (3, 2)) = tgt(x)
    flow = np.zeros((h, w, 2), F64)
    flow[..., 0], flow[..., 1] = 3.0, 2.0
    warped = np.zeros_like(tgt)
    oracle.fn("warp", F64)(oracle.ptr(src), oracle.ptr(flow), oracle.ptr(warped), w, h, 1)
    ix, iy, iz = np.zeros_like(tgt), np.zeros_like(tgt), np.zeros_like(tgt)
    oracle.fn("derivatives", F64)(oracle.ptr(warped), oracle.ptr(tgt), oracle.ptr(ix), oracle.ptr(iy), oracle.ptr(iz),
                                  w, h, 1)
    dx, dy = np.zeros_like(tgt), np.zeros_like(tgt)
    oracle.fn("derivatives2", F64)(oracle.ptr(tgt), oracle.ptr(dx), oracle.ptr(dy), w, h, 1)
    s = np.s_[: h - 2 - 2, : w - 3 - 2]  # away from the bottom/right band the warp read past the image
    np.testing.assert_array_equal(iz[s], 0.0)
    np.testing.assert_allclose(ix[s], dx[s], atol=1e-15)
    np.testing.assert_allclose(iy[s], dy[s], atol=1e-15)


# ---------------------------------------------------------------- a19 robustness M statistic
def test_robustness_motion_statistic_uses_the_last_sample_only(oracle):
    """ComputeRobustnessMask's 5x5 min/max loop overwrites its result each pass
    (R:Kernels/RobustnessModell.cu:62-72), so M compares the centre flow with the (+2, +2) sample only:
    on a linear flow (a x, b y) sampled at half-resolution centres (full-resolution texel 2 p + 0.5)
    the two samples differ by (4 a, 4 b), and M = 0.5 meandist * |(4 a, 4 b)|. With mov = ref + 0.05,
    meandist = 0.05. A min/max over the whole window would see (8 a, 8 b) and double M. Above the
    threshold M the mask is clamp(0 - 0.12) = 0."""
    rng = np.random.default_rng(23)
    hh, hw = 16, 20
    ref = rng.uniform(0.2, 0.6, (hh, hw, 3))
    mov = ref + 0.05
    a, b = 0.01, 0.02
    fy, fx = np.mgrid[0:2 * hh, 0:2 * hw].astype(F64)
    flow = np.ascontiguousarray(np.stack([a * fx - 0.2, b * fy - 0.3], -1))  # |flow| < 1: no moved-sample shift
    alpha, beta = 0.5 / 16383, 9 / 16383 ** 2
    want_m = 0.5 * 0.05 * np.hypot(4 * a, 4 * b)
    inner = np.s_[1:hh - 3, 1:hw - 3]
    for thr, nonzero in ((1e5, True), (want_m * 0.99, False)):
        mask = np.zeros((hh, hw, 4))
        oracle.fn("robustness_mask", F64)(oracle.ptr(ref), oracle.ptr(mov), oracle.ptr(mask), oracle.ptr(flow), 2 * hw,
                                          2 * hh, hw, hh, alpha, beta, thr, 0)
        np.testing.assert_allclose(mask[inner][..., 3], want_m, rtol=1e-9)
        if nonzero:
            assert np.all(mask[inner][..., :3] > 0)
        else:
            np.testing.assert_array_equal(mask[inner][..., :3], 0.0)
    # a window-wide min/max would double M
    assert abs(want_m - 0.5 * 0.05 * np.hypot(8 * a, 8 * b)) > 0.5 * want_m


# ---------------------------------------------------------------- a20 tap order
def _acc_inputs(rng, h, w):
    raw = rng.integers(0, 16384, (h, w)).astype(np.uint16)
    kern = np.zeros((h, w, 3), F32)
    kern[..., 0] = rng.uniform(0.2, 2.0, (h, w))
    kern[..., 1] = rng.uniform(0.2, 2.0, (h, w))
    kern[..., 2] = rng.uniform(-0.1, 0.1, (h, w))
    mask = rng.uniform(0, 1, (h // 2, w // 2, 4)).astype(F32)
    flow = rng.uniform(-4, 4, (h, w, 2)).astype(F32)
    flow[:, :4, 0] -= 10.0  # windows leaving the image on the left: clamped taps of several classes
    return raw, kern, mask, flow


def test_accumulate_tap_orders_agree_within_the_summation_bound(oracle):
    """accumulateImages visits its 25 taps in raster order; the HIP kernels visit them class by class.
    Both add the same terms to each colour's sums, so the results differ by at most the reordering
    bound of a floating-point sum: |raster - classes| <= 2 gamma_k (|s0| + sum |terms|), k = 25 frames
    additions per colour, gamma_k = k u / (1 - k u), u = 2^-24. The terms of the weight sums are
    non-negative, those of the value sums at most max|v| = 1 times them. The orders do differ, so the
    test compares something."""
    rng = np.random.default_rng(31)
    h, w, frames = 40, 52, 4
    white, black = oracle.arr([16383.0] * 3, F32), oracle.arr([0.0] * 3, F32)
    cfa = oracle.arr([0, 1, 1, 2], np.int32)
    s0 = rng.uniform(0, 1, (h, w, 3)).astype(F32)
    w0 = rng.uniform(0, 1, (h, w, 3)).astype(F32)
    res = {}
    for name in ("accumulate", "accumulate_class_order"):
        s, wt = s0.copy(), w0.copy()
        r2 = np.random.default_rng(32)
        for _ in range(frames):
            raw, kern, mask, flow = _acc_inputs(r2, h, w)
            oracle.fn(name)(oracle.ptr(raw), oracle.ptr(s), oracle.ptr(wt), oracle.ptr(mask), oracle.ptr(kern),
                            oracle.ptr(flow), oracle.ptr(white), oracle.ptr(black), w, h, oracle.ptr(cfa))
        res[name] = (s, wt)
    (sr, wr), (sc, wc) = res["accumulate"], res["accumulate_class_order"]
    k = 25 * frames
    u = 2.0 ** -24
    gamma = k * u / (1 - k * u)
    wsum = wr.astype(F64)
    bound_w = 2 * gamma * wsum
    bound_s = 2 * gamma * (np.abs(s0.astype(F64)) + (wsum - w0))
    assert np.all(np.abs(wr.astype(F64) - wc) <= bound_w)
    assert np.all(np.abs(sr.astype(F64) - sc) <= bound_s)
    assert np.any(sr != sc) and np.any(wr != wc)
    # red and blue sit on one tap class each away from the edges: there the orders agree exactly
    inner = np.s_[6:-6, 8:-6]  # no tap clamped: |flow| <= 4 there
    np.testing.assert_array_equal(sr[inner][..., 0], sc[inner][..., 0])
    np.testing.assert_array_equal(sr[inner][..., 2], sc[inner][..., 2])


# ---------------------------------------------------------------- a20 full-frame 2x grid
def test_full_frame_superres_contains_the_reference_crop(oracle):
    """The full-frame 2x grid (2w x 2h, offset 0) is the reference's accumulateImagesSuperRes rule with a zero
    offset: inside the central half, away from where the reference's clamp to that half bites, its output
    pixel (x + w/2, y + h/2) equals the reference's pixel (x, y) bit for bit; the two differ near the crop edge,
    where only the reference clamps."""
    rng = np.random.default_rng(71)
    h, w = 40, 48
    raw = rng.integers(0, 16384, (h, w)).astype(np.uint16)
    kern4 = np.zeros((h, w, 4), F32)
    kern4[..., 0] = rng.uniform(0.2, 2.0, (h, w))
    kern4[..., 1] = rng.uniform(0.2, 2.0, (h, w))
    kern4[..., 2] = rng.uniform(-0.1, 0.1, (h, w))
    mask = rng.uniform(0, 1, (h // 2, w // 2, 4)).astype(F32)
    flow = rng.uniform(-1.5, 1.5, (h, w, 2)).astype(F32)
    white, black = oracle.arr([16383.0] * 3, F32), oracle.arr([0.0] * 3, F32)
    cfa = oracle.arr([0, 1, 1, 2], np.int32)
    s_ref, w_ref = np.zeros((h, w, 3), F32), np.zeros((h, w, 3), F32)
    oracle.fn("accumulate_sr")(oracle.ptr(raw), oracle.ptr(s_ref), oracle.ptr(w_ref), oracle.ptr(mask),
                               oracle.ptr(kern4), oracle.ptr(flow), oracle.ptr(white), oracle.ptr(black), w, h,
                               oracle.ptr(cfa), 1)
    s_full, w_full = np.zeros((2 * h, 2 * w, 3), F32), np.zeros((2 * h, 2 * w, 3), F32)
    oracle.fn("accumulate_sr_grid")(oracle.ptr(raw), oracle.ptr(s_full), oracle.ptr(w_full), oracle.ptr(mask),
                                    oracle.ptr(kern4), oracle.ptr(flow), oracle.ptr(white), oracle.ptr(black), w, h,
                                    oracle.ptr(cfa), 1, 2 * w, 2 * h, 0, 0)
    m = 8  # output pixels: sample shifts up to round(2 * 1.5) = 3 plus the 2-tap window, halved to raw pixels
    crop = np.s_[h // 2 + m:h // 2 + h - m, w // 2 + m:w // 2 + w - m]
    np.testing.assert_array_equal(s_full[crop], s_ref[m:h - m, m:w - m])
    np.testing.assert_array_equal(w_full[crop], w_ref[m:h - m, m:w - m])
    assert np.any(s_full[h // 2 + 1:h // 2 + h - 1, w // 2 + 1] != s_ref[1:h - 1, 1])  # the reference's clamp
    # outside the crop the full grid is written too (the reference's grid has no such pixels)
    assert np.all(w_full[2:6, 2:-2].sum(axis=-1) > 0)


def test_lk_fp64_solve_of_ill_conditioned_windows(oracle):
    """The FP64 Lucas-Kanade solve used for ill-conditioned windows: G^-1 b for det > 0 (numpy's solve), the
    Moore-Penrose pseudo-inverse for a singular G (numpy's pinv: G / trace^2 for rank one), the reference's
    sigma1 gate (lambda1 < min_det skips), NaN -> 0; the selection test is det <= ratio * trace^2."""
    from oracle import oracle as O

    G, rhs = np.array([[4.0, 1.0], [1.0, 2.0]]), np.array([1.0, 2.0])
    np.testing.assert_allclose(O.lk_solve_fp64(4.0, 1.0, 2.0, 1.0, 2.0, 0.1), np.linalg.solve(G, rhs), rtol=1e-15)
    G1, rhs1 = np.array([[1.0, 2.0], [2.0, 4.0]]), np.array([0.5, -3.0])  # det 0
    np.testing.assert_allclose(O.lk_solve_fp64(1.0, 2.0, 4.0, 0.5, -3.0, 0.1), np.linalg.pinv(G1) @ rhs1, rtol=1e-14)
    # lambda1 of [[0.3, 0], [0, 0.2]] is 0.3: below min_det 0.5 the window is skipped, above it solved
    assert O.lk_solve_fp64(0.3, 0.0, 0.2, 1.0, 1.0, 0.5) is None
    a32, d32 = float(np.float32(0.3)), float(np.float32(0.2))  # the sums are floats
    np.testing.assert_allclose(O.lk_solve_fp64(0.3, 0.0, 0.2, 1.0, 1.0, 0.25), (1 / a32, 1 / d32), rtol=1e-15)
    assert O.lk_solve_fp64(np.inf, 0.0, 1.0, 1.0, 1.0, 0.1) == (0.0, 0.0)  # u0 = 1 / inf, u1 = inf / inf -> NaN -> 0
    # det / trace^2 against the ratio: diag(1, 0.01) has det / tr^2 = 0.0098 (ill at 0.01), diag(1, 0.02) 0.0192
    assert O.lk_ill_conditioned(1.0, 0.0, 0.01, 0.01) and not O.lk_ill_conditioned(1.0, 0.0, 0.02, 0.01)
    assert O.lk_ill_conditioned(1.0, 2.0, 4.0, 1e-9)  # singular
    assert not O.lk_ill_conditioned(1.0, 0.0, 1.0, 0.0) and O.lk_ill_conditioned(0.0, 0.0, 0.0, 0.0)
