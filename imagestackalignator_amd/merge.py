"""Python handle over the device-resident merge (isa_merger_*) and the
synthetic burst generator. Device memory comes from torch (ROCm) tensors."""
from __future__ import annotations

import ctypes as C

from . import IsaParams, IsaSynthConfig, call, lib

STAGES = ["tracking_images", "track_patches", "minimizer", "prepare_accumulation", "lucas_kanade", "robustness",
          "accumulate", "final_weighting", "prealign"]


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("a ROCm GPU is required (torch.cuda.is_available() is False)")
    return torch


def _stream_ptr():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def check_tensor(name, t, shape, dtypes):
    """Raises ValueError unless t is a contiguous GPU tensor of the given shape and one of the dtypes: the
    library reads and writes through raw pointers, so a wrong shape, a view or a host tensor would make a
    kernel access memory outside the tensor."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name}: must be a GPU tensor")
    if t.dtype not in dtypes:
        raise ValueError(f"{name}: dtype {t.dtype} not in {tuple(str(d) for d in dtypes)}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")


def raw_dtypes():
    torch = _torch()
    return tuple(d for d in (torch.int16, getattr(torch, "uint16", None)) if d is not None)


class BurstMerger:
    """Mirror of the controller's pipeline (TrackPatches, PrepareAccumulation,
    Accumulate) on one GPU; all buffers are allocated once at construction."""

    def __init__(self, width: int, height: int, frames: int, params: IsaParams):
        self.width, self.height, self.frames = width, height, frames
        self.params = params
        self._h = C.c_void_p()
        call("isa_merger_create", C.byref(self._h), width, height, frames, C.byref(params))

    def close(self):
        if self._h:
            lib().isa_merger_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, raw, out=None, base_shift=None, base_rotation=None):
        """raw: uint16-valued torch tensor (frames, H, W) on the GPU (torch.int16/uint16 storage).
        Returns the merged (H, W, 3) float32 image ((2H, 2W, 3) with super_resolution = 2)."""
        torch = _torch()
        check_tensor("raw", raw, (self.frames, self.height, self.width), raw_dtypes())
        oh, ow = self.out_shape
        if out is None:
            out = torch.empty((oh, ow, 3), dtype=torch.float32, device=raw.device)
        check_tensor("out", out, (oh, ow, 3), (torch.float32,))
        if base_shift is not None and len(base_shift) != 2 * self.frames:
            raise ValueError(f"base_shift: expected {2 * self.frames} values (x, y per frame)")
        if base_rotation is not None and len(base_rotation) != self.frames:
            raise ValueError(f"base_rotation: expected {self.frames} values")
        bs = None if base_shift is None else (C.c_float * (2 * self.frames))(*[float(v) for v in base_shift])
        br = None if base_rotation is None else (C.c_float * self.frames)(*[float(v) for v in base_rotation])
        call("isa_merger_run", self._h, C.c_void_p(raw.data_ptr()), bs, br, C.c_void_p(out.data_ptr()), ow * 12,
             _stream_ptr())
        return out

    @property
    def out_shape(self):
        """(rows, columns) of the merged image: the frame size, doubled for the full-frame 2x merge."""
        s = 2 if self.params.super_resolution == 2 else 1
        return s * self.height, s * self.width

    @property
    def reference_index(self) -> int:
        return lib().isa_merger_reference_index(self._h)

    KERNELS = {"lucas_kanade": 0, "track": 1, "exact_queue": 2, "accumulate": 3}

    def kernel_ms(self, kernel: str, level: int = -1):
        """(device ms, launches) of the last run spent in one kernel family (KERNELS); `level` selects one pyramid
        level of the tile search (-1: all)."""
        ms, cnt = C.c_float(), C.c_int()
        call("isa_merger_kernel_ms", self._h, self.KERNELS[kernel], level, C.byref(ms), C.byref(cnt))
        return ms.value, cnt.value

    def stage_ms(self):
        buf = (C.c_float * 16)()
        n = lib().isa_merger_stage_ms(self._h, buf, 16)
        return {STAGES[i]: buf[i] for i in range(n)}

    def exact_items(self):
        """Per pyramid level (finest first) of the last run: ((pair, tile) items the exact FP64 search
        decided, all items). With a matrix-core precision, the exact items are those whose error bounds
        left the integer minimum or the threshold test open."""
        out = []
        for level in range(self.params.n_levels):
            n_exact, items = C.c_int(), C.c_int()
            call("isa_merger_exact_items", self._h, level, C.byref(n_exact), C.byref(items))
            out.append((n_exact.value, items.value))
        return out

    def buffer(self, which: int, index: int = 0):
        """(device pointer, pitch bytes, width, height) of an internal buffer."""
        p, pitch, w, h = C.c_void_p(), C.c_int(), C.c_int(), C.c_int()
        call("isa_merger_buffer", self._h, which, index, C.byref(p), C.byref(pitch), C.byref(w), C.byref(h))
        return p.value, pitch.value, w.value, h.value

    def buffer_tensor(self, which: int, index: int, channels: int, rows: int | None = None):
        """Copies an internal buffer into a new (rows, width[, channels]) float32 tensor."""
        torch = _torch()
        ptr, pitch, w, h = self.buffer(which, index)
        rows = h if rows is None else rows
        out = torch.empty((rows, pitch // 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        copy_device(out.data_ptr(), ptr, rows * pitch)
        img = out[:, : w * channels]
        return img.reshape(rows, w, channels) if channels > 1 else img


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _HIP = C.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise RuntimeError("libamdhip64.so not found")
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipMemcpy.restype = C.c_int
    return _HIP


def copy_device(dst: int, src: int, nbytes: int) -> None:
    """Synchronous device-to-device copy of raw pointers (hipMemcpy, kind DeviceToDevice)."""
    rc = _hip().hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")


def synth_burst(cfg: IsaSynthConfig):
    """Renders cfg.frames RGGB frames on the GPU; returns (raw int16-storage tensor, per-frame global shifts)."""
    torch = _torch()
    raw = torch.empty((cfg.frames, cfg.height, cfg.width), dtype=torch.int16, device="cuda")
    shifts = (C.c_float * (2 * cfg.frames))()
    call("isa_synth_burst", C.byref(cfg), C.c_void_p(raw.data_ptr()), shifts, _stream_ptr())
    return raw, [(shifts[2 * i], shifts[2 * i + 1]) for i in range(cfg.frames)]


def ground_truth(cfg: IsaSynthConfig, frame: int, out_w: int, out_h: int, scale: float = 1.0, ox: float = 0.0,
                 oy: float = 0.0):
    torch = _torch()
    out = torch.empty((out_h, out_w, 3), dtype=torch.float32, device="cuda")
    call("isa_synth_ground_truth", C.byref(cfg), frame, C.c_void_p(out.data_ptr()), out_w * 12, out_w, out_h,
         C.c_float(scale), C.c_float(ox), C.c_float(oy), _stream_ptr())
    return out
