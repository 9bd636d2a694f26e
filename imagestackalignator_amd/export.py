"""The step after the merge: the colour pipeline and the 16-bit TIFF export of ImageStackAlignatorController
(R:PEFStudioDX/ImageStackAlignatorController.cs:2198-2322) over the C ABI (isa_color_pipeline,
isa_tiff16_encode, isa_write_tiff16)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import call, lib
from .merge import _stream_ptr, _torch, check_tensor, raw_dtypes

FORMATS = {"rgb16": 0, "bgra8": 1}


def color_pipeline(img, camera_white, camera_to_prophoto, lut_x, lut_y, prophoto_to_srgb, exposure: float = 0.0,
                   crop=None, fmt: str = "rgb16", out=None):
    """Merged float3 image (H, W, 3) on the GPU -> the crop (left, top, width, height) as (h, w, 3) uint16
    ("rgb16", SaveAs16BitTiff) or (h, w, 4) uint8 B, G, R, 255 ("bgra8", the display image). camera_white: the
    first threshold's upper bounds; camera_to_prophoto / prophoto_to_srgb: 3 x 4 colour twists (row-major);
    lut_x / lut_y: the tone curve's levels (device tensors, float32, 4..1024 of them)."""
    torch = _torch()
    if not isinstance(img, torch.Tensor) or img.dim() != 3 or img.shape[2] != 3:
        raise ValueError("img: expected an (H, W, 3) tensor")
    h, w = int(img.shape[0]), int(img.shape[1])
    check_tensor("img", img, (h, w, 3), (torch.float32,))
    n = int(lut_x.numel())
    check_tensor("lut_x", lut_x, (n,), (torch.float32,))
    check_tensor("lut_y", lut_y, (n,), (torch.float32,))
    cx, cy, cw, ch = crop if crop is not None else (0, 0, w, h)
    if fmt not in FORMATS:
        raise ValueError(f"fmt: one of {sorted(FORMATS)}")
    shape, dt, bpp = ((ch, cw, 3), torch.int16, 6) if fmt == "rgb16" else ((ch, cw, 4), torch.uint8, 4)
    if out is None:
        out = torch.empty(shape, dtype=dt, device=img.device)
    check_tensor("out", out, shape, (dt,) if fmt == "bgra8" else raw_dtypes())
    f3 = (C.c_float * 3)(*[float(v) for v in camera_white])
    m1 = (C.c_float * 12)(*[float(v) for v in np.asarray(camera_to_prophoto, np.float64).reshape(12)])
    m2 = (C.c_float * 12)(*[float(v) for v in np.asarray(prophoto_to_srgb, np.float64).reshape(12)])
    call("isa_color_pipeline", C.c_void_p(img.data_ptr()), w * 12, w, h, f3, m1, C.c_void_p(lut_x.data_ptr()),
         C.c_void_p(lut_y.data_ptr()), n, m2, C.c_float(exposure), cx, cy, cw, ch, FORMATS[fmt],
         C.c_void_p(out.data_ptr()), cw * bpp, _stream_ptr())
    return out


def tiff16_bytes(rgb: np.ndarray) -> bytes:
    """The TIFF file ImageFileDirectory.SaveAsTiff writes for a (h, w, 3) uint16 image."""
    rgb = np.ascontiguousarray(rgb, np.uint16)
    if rgb.ndim != 3 or rgb.shape[2] != 3:
        raise ValueError(f"rgb: expected (h, w, 3), got {rgb.shape}")
    h, w = rgb.shape[:2]
    n = lib().isa_tiff16_encode(None, 0, rgb.ctypes.data, w, h)
    if n == 0:
        call("isa_write_tiff16", None, None, 0, 0)  # raises with the reason
    buf = (C.c_ubyte * n)()
    if lib().isa_tiff16_encode(buf, n, rgb.ctypes.data, w, h) != n:
        raise RuntimeError("isa_tiff16_encode failed")
    return bytes(buf)


def write_tiff16(path: str, rgb: np.ndarray) -> None:
    """SaveAs16BitTiff's file write for a (h, w, 3) uint16 image."""
    rgb = np.ascontiguousarray(rgb, np.uint16)
    if rgb.ndim != 3 or rgb.shape[2] != 3:
        raise ValueError(f"rgb: expected (h, w, 3), got {rgb.shape}")
    call("isa_write_tiff16", str(path).encode(), rgb.ctypes.data, rgb.shape[1], rgb.shape[0])
