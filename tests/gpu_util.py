"""Helpers for GPU parity tests: numpy <-> device tensors, stream handle."""
import ctypes as C

import numpy as np
import torch


# Device tensors handed to the library by raw pointer must outlive the launch that uses
# them; the caching allocator would otherwise recycle their memory. They are kept here
# until the end of the test (cleared by an autouse fixture in conftest).
KEEP = []


def dev(a: np.ndarray) -> torch.Tensor:
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        t = torch.from_numpy(a.view(np.int16)).cuda()
    else:
        t = torch.from_numpy(a).cuda()
    KEEP.append(t)
    return t


def host(t: torch.Tensor, dtype=None) -> np.ndarray:
    torch.cuda.synchronize()
    a = t.detach().cpu().numpy()
    if dtype is not None:
        a = a.view(dtype) if a.dtype.itemsize == np.dtype(dtype).itemsize else a.astype(dtype)
    return a


def empty(shape, dtype=torch.float32, fill=None) -> torch.Tensor:
    t = torch.empty(shape, dtype=dtype, device="cuda")
    if fill is not None:
        t.fill_(fill)
    KEEP.append(t)
    return t


def p(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ulp_diff(a: np.ndarray, b: np.ndarray) -> int:
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, np.int64(-2**31) - a, a)
    b = np.where(b < 0, np.int64(-2**31) - b, b)
    return int(np.abs(a - b).max()) if a.size else 0
