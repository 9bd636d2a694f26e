"""GPU: the differentiable merge (isa_merge_forward / isa_merge_backward) against the FP64 oracle.

The device computes in FP32 (weights, sums, gradients; the loss and the theta reduction in FP64);
the oracle in FP64. Bounds:
  - forward: merged values within 2e-5 absolute (values in [0, 1]), loss within 1e-4 relative;
  - backward: d loss / d flow within 1e-3 of the oracle in relative L2 norm, 99.9 % of entries within
    1e-3 of the largest gradient, d loss / d theta within 2e-3 relative.
The oracle's gradient is itself pinned against central differences of its forward (CPU tests).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.gpu_util import dev, host  # noqa: E402
from tests.test_oracle_diff import BLACK, RGGB, THETA, WHITE, _half, _raw_from, _scene, _tensor  # noqa: E402


def _problem(oracle, rng, n=4, h=48, w=64):
    frames = []
    for f in range(n):
        off = (0.0, 0.0) if f == 0 else tuple(rng.uniform(-1.5, 1.5, 2))
        img = _scene(h, w, off) + 0.02 * rng.standard_normal((h, w))
        if f == 2:
            img[10:20, 12:30] += 0.3
        frames.append(_raw_from(np.clip(img, 0, 1)))
    raw = np.stack(frames)
    half = _half(oracle, raw)
    tensor = _tensor(rng, h, w)
    tensor[7, 9] = 0.0
    anchor = rng.integers(-1, 2, (n, h, w, 2)).astype(np.float64)
    anchor[0] = 0
    flow = anchor + rng.uniform(-0.45, 0.45, anchor.shape)
    flow[0] = 0.0
    prob = dict(raw=raw, half=half, tensor=tensor, anchor=anchor, black=BLACK, white=WHITE, cfa=RGGB, reference=0,
                alpha=2e-3, beta=1e-5, k_stretch=4.0, k_shrink=2.0)
    # flow, half, tensor and anchor as the device sees them (FP32)
    for k in ("half", "tensor", "anchor"):
        prob[k] = prob[k].astype(np.float32).astype(np.float64)
    flow = flow.astype(np.float32).astype(np.float64)
    return prob, flow


def _device(isa, prob):
    from imagestackalignator_amd.merge_diff import MergeDiff

    params = isa.default_params(black=BLACK, white=WHITE, cfa=RGGB, noise_alpha=prob["alpha"],
                                noise_beta=prob["beta"], k_stretch=prob["k_stretch"], k_shrink=prob["k_shrink"])
    f32 = lambda a: dev(np.ascontiguousarray(a, np.float32))  # noqa: E731
    raw = dev(prob["raw"])
    return MergeDiff(raw, f32(prob["half"]), f32(prob["tensor"]), f32(prob["anchor"]), prob["reference"], params)


def test_forward_matches_fp64_oracle(isa, gpu, oracle):
    rng = np.random.default_rng(11)
    prob, flow = _problem(oracle, rng)
    print("adjoint forward 4x48x64", flush=True)
    _, out_o, _, _ = oracle.merge_diff_forward(prob, flow, THETA)
    gt = out_o + 0.05 * rng.standard_normal(out_o.shape)
    loss_o = oracle.merge_diff_forward(prob, flow, THETA, gt)[0]
    md = _device(isa, prob)
    out, loss = md.forward(dev(flow.astype(np.float32)), THETA, dev(gt.astype(np.float32)))
    gt32 = gt.astype(np.float32).astype(np.float64)
    loss_o32 = oracle.merge_diff_forward(prob, flow, THETA, gt32)[0]
    o = host(out).astype(np.float64)
    err = np.abs(o - out_o).max()
    print(f"  max |out - oracle| {err:.2e}, loss {loss:.6f} vs {loss_o32:.6f} (fp64 gt {loss_o:.6f})")
    assert err < 2e-5
    assert abs(loss - loss_o32) <= 1e-4 * loss_o32


def test_backward_matches_fp64_oracle(isa, gpu, oracle):
    rng = np.random.default_rng(12)
    prob, flow = _problem(oracle, rng)
    print("adjoint backward 4x48x64", flush=True)
    _, out_o, _, mask = oracle.merge_diff_forward(prob, flow, THETA)
    inner = mask[1:, 1:-1, 1:-1]
    assert np.mean((inner > 0) & (inner < 1)) > 0.05  # the mask chain is exercised
    gt = (out_o + 0.05 * rng.standard_normal(out_o.shape)).astype(np.float32)
    g_o, dth_o = oracle.merge_diff_backward(prob, flow, THETA, gt.astype(np.float64))
    md = _device(isa, prob)
    g, dth = md.backward(dev(flow.astype(np.float32)), THETA, dev(gt))
    g = host(g).astype(np.float64)
    rel = np.linalg.norm(g - g_o) / np.linalg.norm(g_o)
    scale = np.abs(g_o).max()
    frac = np.mean(np.abs(g - g_o) <= 1e-3 * scale)
    rth = np.abs(np.array(dth) - dth_o) / np.abs(dth_o)
    print(f"  dflow rel L2 {rel:.2e}, within 1e-3 of max: {frac:.5f}; dtheta rel {rth}")
    assert rel < 1e-3 and frac >= 0.999
    assert np.all(rth < 2e-3)


def test_rejects_bad_problem(isa, gpu):
    from imagestackalignator_amd.merge_diff import IsaMergeDiff

    d = IsaMergeDiff()
    d.width, d.height, d.frames, d.reference = 15, 32, 2, 0
    assert isa.lib().isa_merge_diff_workspace(C.byref(d)) == 0
    with pytest.raises(isa.IsaError):
        isa.call("isa_merge_forward", C.byref(d), None, None, None, None, None, None, None)


def test_refinement_lowers_loss_on_a_synthetic_burst(isa, gpu):
    from imagestackalignator_amd.merge_diff import MergeDiff

    w, h, n = 320, 256, 5
    cfg = isa.default_synth(seed=31, width=w, height=h, frames=n, global_shift_max=4.0, occluded_fraction=0.02)
    raw, _ = isa.synth_burst(cfg)
    m = isa.BurstMerger(w, h, n, isa.default_params(levels=[(1, 32, 2), (2, 32, 4), (4, 16, 4)]))
    m.run(raw)
    print("adjoint refinement 5x256x320", flush=True)
    md, flow = MergeDiff.from_merger(m, raw)
    gt = isa.ground_truth(cfg, m.reference_index, 2 * w, 2 * h, 0.5, -0.25, -0.25)
    theta = md.theta_of(m.params)
    new_flow, new_theta, losses = md.refine(flow, theta, gt, steps=5, lr_flow=2.0)
    print("  losses", ", ".join(f"{v:.4f}" for v in losses))
    assert all(b <= a for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.995 * losses[0]
    # the step's reported loss is the forward loss at the flow it returned (the forward is deterministic), the
    # flow stays within 0.5 px of its anchor and theta is unchanged without lr_theta
    assert md.forward(new_flow, new_theta, gt)[1] == losses[-1]
    assert float((new_flow - md.anchor).abs().max()) <= 0.5 + 1e-6
    assert new_theta == [float(np.float32(v)) for v in theta]
