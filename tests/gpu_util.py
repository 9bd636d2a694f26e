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


def within_reorder_bound(s_got, w_got, s_ref, w_ref, s0, w0, additions, vmax=1.0):
    """Whether two accumulations of the same terms in different orders (value sums s, weight sums w,
    started from s0, w0) are within the reordering bound of a floating-point sum: 2 gamma_k times
    |s0| + sum |terms|, k additions, gamma_k = k u / (1 - k u), u = 2^-24. Weight terms are
    non-negative; value terms are at most vmax (the largest |normalised raw|) times their weight term."""
    u = 2.0 ** -24
    g = additions * u / (1 - additions * u)
    wref = np.asarray(w_ref, np.float64)
    bound_w = 2 * g * np.abs(wref)
    bound_s = 2 * g * (np.abs(np.asarray(s0, np.float64)) + vmax * np.abs(wref - np.asarray(w0, np.float64)))
    return bool(np.all(np.abs(np.asarray(w_got, np.float64) - wref) <= bound_w) and
                np.all(np.abs(np.asarray(s_got, np.float64) - np.asarray(s_ref, np.float64)) <= bound_s))


def workspace(nbytes: int) -> torch.Tensor:
    """A device workspace of nbytes (256-byte aligned, as the caching allocator returns)."""
    assert nbytes > 0, "workspace query returned 0 (invalid arguments)"
    t = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
    assert t.data_ptr() % 256 == 0
    KEEP.append(t)
    return t
