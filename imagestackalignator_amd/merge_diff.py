"""Differentiable merge over device tensors: isa_merge_forward / isa_merge_backward.

The reference's merge (robustness R:Kernels/RobustnessModell.cu:28-158, accumulation
R:Kernels/DeBayerKernels.cu:288-468, kernel shape R:Kernels/kernel.cu:627-700) rounds the flow before
it gathers, so its output is piecewise constant in the flow. The relaxed merge keeps the integer tap
centres of the initial flow (the anchors) and lets the flow move the sample position inside them,
which makes the merged image, and a squared loss against a target, differentiable in the flow and
in the kernel parameters theta = (k_detail, k_denoise, D_th, D_tr).
"""
from __future__ import annotations

import ctypes as C

from . import call, lib
from .merge import _stream_ptr, _torch, check_tensor, raw_dtypes


class IsaMergeDiff(C.Structure):
    """Mirror of `isa_merge_diff` (include/isa_hip.h), field for field."""

    _fields_ = [
        ("width", C.c_int),
        ("height", C.c_int),
        ("frames", C.c_int),
        ("reference", C.c_int),
        ("raw_dev", C.c_void_p),
        ("half_dev", C.c_void_p),
        ("tensor_dev", C.c_void_p),
        ("anchor_dev", C.c_void_p),
        ("black", C.c_float * 3),
        ("white", C.c_float * 3),
        ("cfa", C.c_int * 4),
        ("noise_alpha", C.c_float),
        ("noise_beta", C.c_float),
        ("k_stretch", C.c_float),
        ("k_shrink", C.c_float),
    ]


def half_res(raw, params):
    """isa_debayer_subsample3 of every frame: (frames, H/2, W/2, 3) float32 (R:Kernels/DeBayerKernels.cu:242-283)."""
    torch = _torch()
    if raw.dim() != 3:
        raise ValueError("raw: expected (frames, H, W)")
    n, h, w = raw.shape
    check_tensor("raw", raw, (n, h, w), raw_dtypes())
    if h % 2 or w % 2:
        raise ValueError("raw: width and height must be even")
    out = torch.empty((n, h // 2, w // 2, 3), dtype=torch.float32, device=raw.device)
    maxv = max(params.white[c] + params.black[c] for c in range(3))
    for f in range(n):
        call("isa_debayer_subsample3", C.c_void_p(raw[f].data_ptr()), C.c_void_p(out[f].data_ptr()), C.c_float(maxv),
             w // 2, h // 2, (w // 2) * 12, params.cfa, _stream_ptr())
    return out


def smoothed_tensor(image, params):
    """Structure tensor of a float image (H, W), Gaussian-smoothed with sigma_structure_tensor, as the
    merge's PrepareAccumulation computes it before ComputeKernelParam: (H, W, 3) = (a11, a22, a12)."""
    torch = _torch()
    if image.dim() != 2:
        raise ValueError("image: expected (H, W)")
    h, w = image.shape
    check_tensor("image", image, (h, w), (torch.float32,))
    dx = torch.empty((h, w), dtype=torch.float32, device=image.device)
    dy = torch.empty_like(dx)
    t = torch.empty((h, w, 3), dtype=torch.float32, device=image.device)
    out = torch.empty_like(t)
    tmp = torch.empty_like(t)
    s = _stream_ptr()
    call("isa_compute_derivatives2", C.c_void_p(image.data_ptr()), image.stride(0) * 4, C.c_void_p(dx.data_ptr()),
         C.c_void_p(dy.data_ptr()), w * 4, w, h, params.tex_quant, s)
    call("isa_compute_structure_tensor", C.c_void_p(dx.data_ptr()), C.c_void_p(dy.data_ptr()), w * 4,
         C.c_void_p(t.data_ptr()), w * 12, w, h, s)
    taps = (C.c_float * 99)()
    nt = lib().isa_gaussian_taps(C.c_float(params.sigma_structure_tensor), taps, 99)
    if nt <= 0:
        raise ValueError("isa_gaussian_taps failed")
    call("isa_gauss_filter_border", C.c_void_p(t.data_ptr()), w * 12, C.c_void_p(out.data_ptr()), w * 12,
         C.c_void_p(tmp.data_ptr()), w, h, 3, taps, nt, s)
    return out


class MergeDiff:
    """One relaxed-merge problem: raw frames, their half-resolution debayers, the reference's smoothed
    structure tensor and the integer anchors, all device tensors kept alive by this object."""

    def __init__(self, raw, half, tensor, anchor, reference: int, params):
        torch = _torch()
        if raw.dim() != 3:
            raise ValueError("raw: expected (frames, H, W)")
        n, h, w = raw.shape
        check_tensor("raw", raw, (n, h, w), raw_dtypes())
        check_tensor("half", half, (n, h // 2, w // 2, 3), (torch.float32,))
        check_tensor("tensor", tensor, (h, w, 3), (torch.float32,))
        check_tensor("anchor", anchor, (n, h, w, 2), (torch.float32,))
        self.raw, self.half, self.tensor, self.anchor = raw, half, tensor, anchor
        self.width, self.height, self.frames = w, h, n
        d = IsaMergeDiff()
        d.width, d.height, d.frames, d.reference = w, h, n, reference
        d.raw_dev, d.half_dev = raw.data_ptr(), half.data_ptr()
        d.tensor_dev, d.anchor_dev = tensor.data_ptr(), anchor.data_ptr()
        d.black[:] = list(params.black)
        d.white[:] = list(params.white)
        d.cfa[:] = list(params.cfa)
        d.noise_alpha, d.noise_beta = params.noise_alpha, params.noise_beta
        d.k_stretch, d.k_shrink = params.k_stretch, params.k_shrink
        self.problem = d
        nbytes = lib().isa_merge_diff_workspace(C.byref(d))
        if nbytes == 0:
            call("isa_merge_forward", C.byref(d), None, None, None, None, None, None, None)  # raises with the reason
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=raw.device)
        self.refine_workspace = None  # allocated by the first refine()

    @classmethod
    def from_merger(cls, merger, raw):
        """Problem around the last run of a BurstMerger: flows = the merger's LK flows, anchors = their
        rounding, tensor from the reference's LK image. Returns (problem, initial flow)."""
        torch = _torch()
        n, h, w = raw.shape
        ref = merger.reference_index
        flow = torch.stack([merger.buffer_tensor(3, f, 2) for f in range(n)]).contiguous()
        flow[ref].zero_()
        # roundf (half away from zero), as the reference's gather rounds the flow
        anchor = (torch.sign(flow) * torch.floor(flow.abs() + 0.5)).contiguous()
        lk_ref = merger.buffer_tensor(1, ref, 1).contiguous()
        tensor = smoothed_tensor(lk_ref, merger.params)
        half = half_res(raw, merger.params)
        return cls(raw, half, tensor, anchor, ref, merger.params), flow

    def theta_of(self, params):
        return [params.k_detail, params.k_denoise, params.dth, params.dtr]

    def _flow(self, flow, name="flow"):
        check_tensor(name, flow, (self.frames, self.height, self.width, 2), (_torch().float32,))

    def _gt(self, gt):
        if gt is not None:
            check_tensor("gt", gt, (2 * self.height, 2 * self.width, 3), (_torch().float32,))

    def forward(self, flow, theta, gt=None, out=None):
        """Merged (2H, 2W, 3) image and, with a target gt of that shape, the loss (sum of squared
        differences over the interior, 4-pixel border excluded)."""
        torch = _torch()
        self._flow(flow)
        self._gt(gt)
        if out is None:
            out = torch.empty((2 * self.height, 2 * self.width, 3), dtype=torch.float32, device=flow.device)
        check_tensor("out", out, (2 * self.height, 2 * self.width, 3), (torch.float32,))
        loss = C.c_double(0.0)
        th = (C.c_float * 4)(*[float(v) for v in theta])
        call("isa_merge_forward", C.byref(self.problem), C.c_void_p(flow.data_ptr()), th, C.c_void_p(out.data_ptr()),
             None if gt is None else C.c_void_p(gt.data_ptr()), C.byref(loss) if gt is not None else None,
             C.c_void_p(self.workspace.data_ptr()), _stream_ptr())
        return out, (loss.value if gt is not None else None)

    def backward(self, flow, theta, gt, dflow=None):
        """d loss / d flow (frames, H, W, 2) and d loss / d theta (4 floats)."""
        torch = _torch()
        self._flow(flow)
        if gt is None:
            raise ValueError("gt: the loss needs a target")
        self._gt(gt)
        if dflow is None:
            dflow = torch.empty_like(flow)
        self._flow(dflow, "dflow")
        th = (C.c_float * 4)(*[float(v) for v in theta])
        dth = (C.c_float * 4)()
        call("isa_merge_backward", C.byref(self.problem), C.c_void_p(flow.data_ptr()), th, C.c_void_p(gt.data_ptr()),
             C.c_void_p(dflow.data_ptr()), dth, C.c_void_p(self.workspace.data_ptr()), _stream_ptr())
        return dflow, list(dth)

    def refine(self, flow, theta, gt, steps: int, lr_flow: float, lr_theta=None, max_step: float = 0.25):
        """`steps` calls of isa_merge_refine_step: gradient descent with backtracking on the flow (each step clipped
        to max_step px, the flow kept within 0.5 px of its anchor so the tap window stays centred) and, with
        lr_theta (4 step sizes), on theta. A step that raises the loss is retried at half the step size (at most 9
        tries). Returns (flow, theta, losses: the initial loss and the loss after each step)."""
        torch = _torch()
        self._flow(flow)
        self._gt(gt)
        flow = flow.clone()
        if self.refine_workspace is None:
            nbytes = lib().isa_merge_refine_workspace(C.byref(self.problem))
            self.refine_workspace = torch.empty(nbytes, dtype=torch.uint8, device=flow.device)
        th = (C.c_float * 4)(*[float(v) for v in theta])
        lr_th = None if lr_theta is None else (C.c_float * 4)(*[float(v) for v in lr_theta])
        loss = C.c_double(self.forward(flow, list(th), gt)[1])
        losses = [loss.value]
        scale = C.c_float(1.0)
        acc = C.c_int(0)
        for _ in range(steps):
            call("isa_merge_refine_step", C.byref(self.problem), C.c_void_p(flow.data_ptr()), th,
                 C.c_void_p(gt.data_ptr()), C.c_float(lr_flow), lr_th, C.c_float(max_step), C.byref(scale),
                 C.byref(loss), C.byref(acc), C.c_void_p(self.refine_workspace.data_ptr()), _stream_ptr())
            losses.append(loss.value)
        return flow, list(th), losses
