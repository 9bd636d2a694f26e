"""Pins the FP64 relaxed-merge oracle (merge_diff_forward / merge_diff_backward) on the CPU.

Forward: a constant scene is reproduced whatever the flow; one frame with flow -1/4 and anchor 0
reduces to the reference's accumulateImages (oracle `accumulate`, itself pinned against the GPU
kernel and by known answers) with ComputeKernelParam weights; a second frame that is an even-pixel
translation of the first, aligned by an integer flow, leaves the merge unchanged (this fixes the
sign of the flow, the anchor offset and the moved-mean sampling of the mask).
Gradient: central differences of the FP64 forward, for flow entries and the four kernel parameters.
"""
import numpy as np
import pytest

F64 = np.float64
RGGB = [0, 1, 1, 2]
BLACK, WHITE = [64.0, 64.0, 64.0], [16000.0, 16000.0, 16000.0]
THETA = [0.25, 3.0, 0.001, 0.006]  # k_detail, k_denoise, D_th, D_tr


def _scene(h, w, off=(0.0, 0.0)):
    yy, xx = np.mgrid[0:h, 0:w].astype(F64)
    yy, xx = yy + off[0], xx + off[1]
    return 0.45 + 0.25 * np.sin(xx / 2.3 + 0.4 * yy) * np.cos(yy / 3.1) + 0.1 * np.sin((xx - 2 * yy) / 4.7)


def _raw_from(img):
    return np.clip(np.round(BLACK[0] + WHITE[0] * img), 0, 16383).astype(np.uint16)


def _half(oracle, raw):
    n, h, w = raw.shape
    out = np.zeros((n, h // 2, w // 2, 3), F64)
    cfa = np.array(RGGB, np.int32)
    for f in range(n):
        oracle.fn("subsample3", F64)(oracle.ptr(np.ascontiguousarray(raw[f])), oracle.ptr(out[f]), 16383.0 + 64.0,
                                     w // 2, h // 2, oracle.ptr(cfa))
    return out


def _tensor(rng, h, w, scale=2e-5):
    u = rng.uniform(0.2, 1.0, (h, w)) * np.sqrt(scale)
    v = rng.uniform(0.0, 1.0, (h, w)) * np.sqrt(scale)
    rho = rng.uniform(-0.9, 0.9, (h, w))
    t = np.stack([u * u, v * v, rho * u * v], axis=-1)
    return t


def _prob(raw, half, tensor, anchor, ref=0, alpha=0.5 / 16383, beta=9.0 / 16383**2):
    return dict(raw=raw, half=half, tensor=tensor, anchor=anchor, black=BLACK, white=WHITE, cfa=RGGB, reference=ref,
                alpha=alpha, beta=beta, k_stretch=4.0, k_shrink=2.0)


def test_constant_scene_is_reproduced(oracle):
    rng = np.random.default_rng(1)
    n, h, w = 3, 20, 24
    vals = {0: 3000, 1: 9000, 2: 5000}
    raw = np.zeros((n, h, w), np.uint16)
    for y in range(h):
        for x in range(w):
            raw[:, y, x] = vals[RGGB[(y % 2) * 2 + x % 2]]
    anchor = rng.integers(-2, 3, (n, h, w, 2)).astype(F64)
    flow = anchor + rng.uniform(-0.45, 0.45, anchor.shape)
    prob = _prob(raw, _half(oracle, raw), _tensor(rng, h, w), anchor)
    _, out, wsum, _ = oracle.merge_diff_forward(prob, flow, THETA)
    assert np.all(wsum > 0)
    for c in range(3):
        np.testing.assert_allclose(out[..., c], (vals[c] - BLACK[c]) / WHITE[c], rtol=1e-12)


def test_single_frame_reduces_to_reference_accumulate(oracle):
    rng = np.random.default_rng(2)
    h, w = 22, 26
    img = _scene(h, w) + 0.02 * rng.standard_normal((h, w))
    raw = _raw_from(img)[None]
    tensor = _tensor(rng, h, w)
    tensor[5, 7] = 0.0  # zero gradient: NaN kernel, cross-kernel fallback in both
    anchor = np.zeros((1, h, w, 2))
    flow = np.full((1, h, w, 2), -0.25)
    _, out, _, _ = oracle.merge_diff_forward(_prob(raw, _half(oracle, raw), tensor, anchor), flow, THETA)
    # reference accumulate (1x): taps at integer offsets around b, certainty 1
    kern = np.ascontiguousarray(tensor.copy())
    oracle.fn("kernel_param", F64)(oracle.ptr(kern), w, h, THETA[2], THETA[3], THETA[0], THETA[1], 4.0, 2.0)
    s, wt = np.zeros((h, w, 3)), np.zeros((h, w, 3))
    mask = np.ones((h // 2, w // 2, 4))
    zero = np.zeros((h, w, 2))
    oracle.fn("accumulate", F64)(oracle.ptr(np.ascontiguousarray(raw[0])), oracle.ptr(s), oracle.ptr(wt),
                                 oracle.ptr(mask), oracle.ptr(kern), oracle.ptr(zero),
                                 oracle.ptr(np.array(WHITE)), oracle.ptr(np.array(BLACK)), w, h,
                                 oracle.ptr(np.array(RGGB, np.int32)))
    with np.errstate(invalid="ignore"):
        ref = s[1:-1, 1:-1] / wt[1:-1, 1:-1]
    # output pixel 2b+1 sits at b + 1/4; with flow -1/4 its sample position is b itself
    mine = out[3:-2:2, 3:-2:2]
    # the cross-kernel fallback at the zero-tensor pixel (blue site) leaves red without weight: 0/0 there
    assert np.isnan(ref[4, 6, 0]) and mine[4, 6, 0] == 0.0
    ok = np.isfinite(ref)
    assert np.count_nonzero(~ok) == 1
    np.testing.assert_allclose(mine[ok], ref[ok], rtol=1e-12, atol=1e-15)


def test_aligned_translated_frame_leaves_merge_unchanged(oracle):
    rng = np.random.default_rng(3)
    h, w, sh = 32, 36, (2, -4)  # even shift keeps the Bayer phase
    img = _scene(h, w) + 0.02 * rng.standard_normal((h, w))
    r0 = _raw_from(img)
    r1 = np.roll(r0, sh, axis=(0, 1))  # r1[y + sh] = r0[y]
    tensor = _tensor(rng, h, w)
    one = np.stack([r0])
    two = np.stack([r0, r1])
    a1 = np.zeros((1, h, w, 2))
    a2 = np.zeros((2, h, w, 2))
    a2[1, ..., 0], a2[1, ..., 1] = sh[1], sh[0]
    _, o1, _, _ = oracle.merge_diff_forward(_prob(one, _half(oracle, one), tensor, a1), a1.copy(), THETA)
    _, o2, _, m2 = oracle.merge_diff_forward(_prob(two, _half(oracle, two), tensor, a2), a2.copy(), THETA)
    inner = (slice(16, -16), slice(20, -20))  # taps of frame 1 stay clear of the wrapped rows and columns
    np.testing.assert_allclose(o2[inner], o1[inner], rtol=1e-13)
    # a wrong flow sign reads other samples: the merge changes
    a3 = a2.copy()
    a3[1] = -a3[1]
    _, o3, _, _ = oracle.merge_diff_forward(_prob(two, _half(oracle, two), tensor, a3), a3.copy(), THETA)
    assert np.abs(o3[inner] - o1[inner]).max() > 1e-3
    assert np.all(m2[1, 4:-4, 4:-4] == 1.0)


def _burst(oracle, rng, n=3, h=20, w=24):
    frames = []
    for f in range(n):
        off = (0.0, 0.0) if f == 0 else tuple(rng.uniform(-1.5, 1.5, 2))
        img = _scene(h, w, off) + 0.02 * rng.standard_normal((h, w))
        if f == 2:
            img[6:11, 8:14] += 0.3  # an occluder: masks between 0 and 1
        frames.append(_raw_from(np.clip(img, 0, 1)))
    raw = np.stack(frames)
    tensor = _tensor(rng, h, w)
    tensor[3, 4] = 0.0
    anchor = rng.integers(-1, 2, (n, h, w, 2)).astype(F64)
    anchor[0] = 0
    flow = anchor + rng.uniform(-0.45, 0.45, anchor.shape)
    flow[0] = rng.uniform(-0.3, 0.3, flow[0].shape)
    prob = _prob(raw, _half(oracle, raw), tensor, anchor, alpha=2e-3, beta=1e-5)
    _, out, _, mask = oracle.merge_diff_forward(prob, flow, THETA)
    gt = out + 0.05 * rng.standard_normal(out.shape)
    return prob, flow, gt, mask


def test_gradient_matches_central_differences(oracle):
    rng = np.random.default_rng(4)
    prob, flow, gt, mask = _burst(oracle, rng)
    # the masks are neither all clamped nor all open: the mask chain is exercised
    inner = mask[1:, 1:-1, 1:-1]
    assert np.mean((inner > 0) & (inner < 1)) > 0.1
    dflow, dth = oracle.merge_diff_backward(prob, flow, THETA, gt)
    assert np.abs(dflow).max() > 0
    eps = 1e-6
    n, h, w, _ = flow.shape
    picks = [(f, int(rng.integers(0, h)), int(rng.integers(0, w)), int(rng.integers(0, 2))) for f in range(n)
             for _ in range(25)]
    # also the even raw pixels whose flow moves a mask
    picks += [(2, 2 * y, 2 * x, k) for y in (3, 4, 5) for x in (4, 5, 6) for k in (0, 1)]
    for f, y, x, k in picks:
        fp, fm = flow.copy(), flow.copy()
        fp[f, y, x, k] += eps
        fm[f, y, x, k] -= eps
        fd = (oracle.merge_diff_forward(prob, fp, THETA, gt)[0] - oracle.merge_diff_forward(prob, fm, THETA, gt)[0])
        fd /= 2 * eps
        assert abs(fd - dflow[f, y, x, k]) <= 1e-6 * abs(fd) + 1e-8, (f, y, x, k, fd, dflow[f, y, x, k])
    for i in range(4):
        e = 1e-6 * THETA[i]
        tp, tm = list(THETA), list(THETA)
        tp[i] += e
        tm[i] -= e
        fd = (oracle.merge_diff_forward(prob, flow, tp, gt)[0] - oracle.merge_diff_forward(prob, flow, tm, gt)[0]) / (2 * e)
        assert abs(fd - dth[i]) <= 1e-5 * abs(fd) + 1e-7, (i, fd, dth[i])
    assert np.count_nonzero(dth) == 4


@pytest.mark.parametrize("which", ["flow", "theta"])
def test_gradient_is_zero_at_the_target(oracle, which):
    rng = np.random.default_rng(5)
    prob, flow, _, _ = _burst(oracle, rng)
    _, out, _, _ = oracle.merge_diff_forward(prob, flow, THETA)
    dflow, dth = oracle.merge_diff_backward(prob, flow, THETA, out)
    assert (np.abs(dflow).max() if which == "flow" else np.abs(dth).max()) == 0.0
