"""ctypes bindings of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product (libisa_hip.so) never
uses it. Arrays are numpy, C-contiguous; `dt` selects the instantiation:
np.float32 = reference semantics (FP32, written operation order),
np.float64 = FP64 oracle.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _declare(_lib)
    return _lib


P = C.c_void_p
I = C.c_int


def _declare(L):
    spec = {
        # name: (restype, [argtypes]) with T-typed scalars marked "T"
        "debayer_full": (None, [I, I, P, P, P, P, P, P]),
        "debayer_green": (None, [I, I, P, P, P, P, P]),
        "debayer_redblue": (None, [I, I, P, P, P, P, P]),
        "gamma_srgb": (None, [I, I, P]),
        "color_to_gray": (None, [I, I, P, P, P]),
        "fourier_highpass": (None, [I, I, P, "T", "T", I]),
        "fourier_filter_value": ("T", [I, I, I, I, "T", "T", I]),
        "gaussian_taps": (I, ["T", P, I]),
        "gauss_filter": (None, [I, I, I, P, P, P, I]),
        "tracking_image": (None, [I, I, P, P, P, I, "T"]),
        "downscale_box": (None, [I, I, P, I, P]),
        "tiles_border": (None, [I, I, P, P, I, I, I, I, "T", "T", "T"]),
        "tiles_preshift": (None, [I, I, P, P, P, I, I, I, I, "T", "T", "T"]),
        "tile_ssd": (None, [P, P, P, I, I, I]),
        "find_minimum": (None, [P, P, I, I, "T"]),
        "upsample_shifts": (None, [P, P, I, I, I, I, I, I, I, I]),
        "get_optimal_shifts": (None, [P, P, I, I, I, I]),
        "flow_from_tiles": (None, [P, I, I, P, I, I, "T", "T", "T", I]),
        "warp": (None, [P, P, P, I, I, I]),
        "derivatives": (None, [P, P, P, P, P, I, I, I]),
        "derivatives2": (None, [P, P, P, I, I, I]),
        "lucas_kanade_step": (None, [P, P, P, P, I, I, I, "T", C.c_double]),
        "lucas_kanade": (None, [P, P, P, I, I, P, I, I, I, "T", "T", "T", "T", I, I, C.c_double, P]),
        "structure_tensor": (None, [P, P, P, I, I]),
        "kernel_param": (None, [P, I, I, "T", "T", "T", "T", "T", "T"]),
        "subsample3": (None, [P, P, "T", I, I, P]),
        "robustness_mask": (None, [P, P, P, P, I, I, I, I, "T", "T", "T", I]),
        "erode": (None, [P, P, I, I, I]),
        "accumulate": (None, [P, P, P, P, P, P, P, P, I, I, P]),
        "accumulate_class_order": (None, [P, P, P, P, P, P, P, P, I, I, P]),
        "accumulate_sr": (None, [P, P, P, P, P, P, P, P, I, I, P, I]),
        "accumulate_sr_grid": (None, [P, P, P, P, P, P, P, P, I, I, P, I, I, I, I, I]),
        "apply_weighting": (None, [P, P, P, I, I, "T"]),
        "merge": (I, [P, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    }
    for sfx, T in (("f32", C.c_float), ("f64", C.c_double)):
        for name, (rt, args) in spec.items():
            fn = getattr(L, f"ora_{name}_{sfx}")
            fn.restype = T if rt == "T" else rt
            fn.argtypes = [T if a == "T" else a for a in args]
    L.ora_set_threads.argtypes = [I]
    L.ora_set_threads.restype = None
    L.ora_get_threads.restype = I
    L.ora_default_params.argtypes = [P]
    L.ora_default_params.restype = None
    L.ora_shift_tables.argtypes = [I, I, I, I, P, P, P, P]
    L.ora_shift_tables.restype = I
    L.ora_minimize_shifts.argtypes = [I, I, I, I, I, P, P, P, P, P]
    L.ora_minimize_shifts.restype = I
    L.ora_lk_ill_counts.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.ora_lk_ill_counts.restype = None
    L.ora_lk_solve_fp64.argtypes = [C.c_float] * 5 + [C.c_double, C.POINTER(C.c_double)]
    L.ora_lk_solve_fp64.restype = C.c_int
    L.ora_lk_ill_conditioned.argtypes = [C.c_float] * 3 + [C.c_double]
    L.ora_lk_ill_conditioned.restype = C.c_int
    L.ora_color_pipeline.argtypes = [I, I, P, P, P, P, P, I, P, C.c_float, I, I, I, I, I, P]
    L.ora_color_pipeline.restype = None
    L.ora_lut_cubic.argtypes = [C.c_float, P, P, I]
    L.ora_lut_cubic.restype = C.c_float
    L.ora_convert_near.argtypes = [C.c_float, C.c_float]
    L.ora_convert_near.restype = C.c_float


def lk_solve_fp64(a, b, d, b0, b1, min_det):
    """The FP64 solve of one Lucas-Kanade window (G = [[a, b], [b, d]], rhs (b0, b1)): (u0, u1), or None when the
    sigma1 gate skips the window."""
    u = (C.c_double * 2)()
    ok = lib().ora_lk_solve_fp64(a, b, d, b0, b1, min_det, u)
    return (u[0], u[1]) if ok else None


def lk_ill_conditioned(a, b, d, ratio):
    return bool(lib().ora_lk_ill_conditioned(a, b, d, ratio))


def color_pipeline(img3, camera_white, cam_to_prophoto, lut_x, lut_y, prophoto_to_srgb, exposure, crop=None,
                   bgra8=False):
    """The colour pipeline after the merge (float32): img3 (h, w, 3) -> (ch, cw, 3) u16, or (ch, cw, 4) u8 BGRA.
    crop = (left, top, width, height); the amplification 2^exposure is formed in double and rounded to float."""
    img = np.ascontiguousarray(img3, np.float32)
    h, w = img.shape[:2]
    cx, cy, cw, ch = crop if crop is not None else (0, 0, w, h)
    f = lambda a, n: np.ascontiguousarray(np.asarray(a, np.float32).reshape(n))
    wb, m1, m2 = f(camera_white, 3), f(cam_to_prophoto, 12), f(prophoto_to_srgb, 12)
    lx, ly = f(lut_x, -1), f(lut_y, -1)
    out = np.zeros((ch, cw, 4), np.uint8) if bgra8 else np.zeros((ch, cw, 3), np.uint16)
    lib().ora_color_pipeline(w, h, ptr(img), ptr(wb), ptr(m1), ptr(lx), ptr(ly), len(lx), ptr(m2),
                             float(np.float32(2.0 ** exposure)), cx, cy, cw, ch, int(bgra8), ptr(out))
    return out


def lut_cubic(v, lut_x, lut_y):
    lx, ly = np.ascontiguousarray(lut_x, np.float32), np.ascontiguousarray(lut_y, np.float32)
    return lib().ora_lut_cubic(float(v), ptr(lx), ptr(ly), len(lx))


def convert_near(v, maxv):
    return lib().ora_convert_near(float(v), float(maxv))


def tiff16_bytes(rgb):
    """ImageFileDirectory(width, height).SaveAsTiff (R:PefFile/ImageFileDirectory.cs:76-123) restated: header,
    13 entries in the reference's construction order, the values that do not fit an entry appended in that
    order (each entry's offset field pointing at them), then the pixels, whose position is StripOffsets."""
    import struct

    rgb = np.ascontiguousarray(rgb, np.uint16)
    h, w = rgb.shape[:2]
    SHORT, LONG, RATIONAL, ASCII = 3, 4, 5, 2
    entries = [(256, LONG, 1, w), (257, LONG, 1, h), (258, SHORT, 3, None), (259, SHORT, 1, 1), (262, SHORT, 1, 2),
               (273, LONG, 1, "pixels"), (277, SHORT, 1, 3), (278, LONG, 1, h), (279, LONG, 1, w * h * 6),
               (282, RATIONAL, 1, None), (283, RATIONAL, 1, None), (296, SHORT, 1, 1), (271, ASCII, 8, None)]
    extra = {258: struct.pack("<3H", 16, 16, 16), 282: struct.pack("<2I", 720, 10), 283: struct.pack("<2I", 720, 10),
             271: b"MK TIFF\0"}
    pos = 8 + 2 + 12 * len(entries) + 4
    offsets = {}
    for tag, _, _, v in entries:
        if v is None:
            offsets[tag] = pos
            pos += len(extra[tag])
    pix_at = pos
    out = bytearray(b"II*\0" + struct.pack("<I", 8) + struct.pack("<H", len(entries)))
    for tag, typ, cnt, v in entries:
        if v is None:
            val = offsets[tag]
        elif v == "pixels":
            val = pix_at
        else:
            val = v
        out += struct.pack("<HHI", tag, typ, cnt)
        out += struct.pack("<HH", val, 0) if (typ == SHORT and cnt == 1) else struct.pack("<I", val)
    out += struct.pack("<I", 0)
    for tag, _, _, v in entries:
        if v is None:
            out += extra[tag]
    assert len(out) == pix_at
    out += rgb.astype("<u2").tobytes()
    return bytes(out)


def lk_ill_counts():
    """(windows solved in FP64, all windows) of the FP32 Lucas-Kanade passes with lk_ill_ratio > 0 since the
    last call."""
    a, b = C.c_longlong(), C.c_longlong()
    lib().ora_lk_ill_counts(C.byref(a), C.byref(b))
    return a.value, b.value


def set_threads(n: int) -> None:
    lib().ora_set_threads(int(n))


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(P)


def fn(name: str, dt=np.float32):
    sfx = "f32" if np.dtype(dt) == np.float32 else "f64"
    return getattr(lib(), f"ora_{name}_{sfx}")


def arr(values, dt):
    return np.ascontiguousarray(np.asarray(values, dtype=dt))


# ---------------------------------------------------------------------------
# convenience wrappers (numpy in, numpy out)

def debayer_full(raw: np.ndarray, black, white, camwhite, cfa, dt=np.float32):
    h, w = raw.shape
    out = np.zeros((h, w, 3), dt)
    fn("debayer_full", dt)(w, h, ptr(np.ascontiguousarray(raw, np.uint16)), ptr(out), ptr(arr(black, dt)),
                           ptr(arr(white, dt)), ptr(arr(camwhite, dt)), ptr(arr(cfa, np.int32)))
    return out


def tracking_image(raw: np.ndarray, params, use_hp: int, sigma: float, dt=np.float32):
    h, w = raw.shape
    out = np.zeros((h, w), dt)
    fn("tracking_image", dt)(w, h, ptr(np.ascontiguousarray(raw, np.uint16)), ptr(out), C.addressof(params), use_hp,
                             sigma)
    return out


def fourier_highpass(img: np.ndarray, hp=0.01, hps=0.0025, clear_axis=0, dt=np.float32):
    out = np.ascontiguousarray(img, dt).copy()
    h, w = out.shape
    fn("fourier_highpass", dt)(w, h, ptr(out), hp, hps, clear_axis)
    return out


def gaussian_taps(sigma: float, dt=np.float32):
    buf = np.zeros(128, dt)
    n = fn("gaussian_taps", dt)(sigma, ptr(buf), 128)
    return buf[:n].copy()


def gauss_filter(img: np.ndarray, taps: np.ndarray, dt=np.float32):
    a = np.ascontiguousarray(img, dt)
    h, w = a.shape[:2]
    ch = 1 if a.ndim == 2 else a.shape[2]
    out = np.zeros_like(a)
    t = np.ascontiguousarray(taps, dt)
    fn("gauss_filter", dt)(w, h, ch, ptr(a), ptr(out), ptr(t), len(t))
    return out


def downscale_box(img: np.ndarray, level: int, dt=np.float32):
    a = np.ascontiguousarray(img, dt)
    h, w = a.shape
    out = np.zeros((h // level, w // level), dt)
    fn("downscale_box", dt)(w, h, ptr(a), level, ptr(out))
    return out


def tiles_border(img, M, tile, tcx, tcy, bs=(0, 0), rot=0.0, dt=np.float32):
    a = np.ascontiguousarray(img, dt)
    h, w = a.shape
    B = tile + 2 * M
    out = np.zeros((tcx * tcy, B, B), dt)
    fn("tiles_border", dt)(w, h, ptr(a), ptr(out), M, tile, tcx, tcy, bs[0], bs[1], rot)
    return out


def tiles_preshift(img, pre, M, tile, tcx, tcy, bs=(0, 0), rot=0.0, dt=np.float32):
    a = np.ascontiguousarray(img, dt)
    h, w = a.shape
    B = tile + 2 * M
    out = np.zeros((tcx * tcy, B, B), dt)
    p = np.ascontiguousarray(pre, dt)
    fn("tiles_preshift", dt)(w, h, ptr(a), ptr(out), ptr(p), M, tile, tcx, tcy, bs[0], bs[1], rot)
    return out


def tile_ssd(rt, tt, M, tile, dt=np.float32):
    rt = np.ascontiguousarray(rt, dt)
    tt = np.ascontiguousarray(tt, dt)
    n = rt.shape[0]
    S = 2 * M + 1
    out = np.zeros((n, S, S), dt)
    fn("tile_ssd", dt)(ptr(rt), ptr(tt), ptr(out), M, tile, n)
    return out


def find_minimum(simg, M, thr=1.0, dt=np.float32):
    s = np.ascontiguousarray(simg, dt)
    n = s.shape[0]
    out = np.zeros((n, 2), dt)
    fn("find_minimum", dt)(ptr(s), ptr(out), M, n, thr)
    return out


def upsample_shifts(inp, ol, nl, ocx, ocy, ncx, ncy, ot, nt, dt=np.float32):
    a = np.ascontiguousarray(inp, dt)
    out = np.zeros((ncy, ncx, 2), dt)
    fn("upsample_shifts", dt)(ptr(a), ptr(out), ol, nl, ocx, ocy, ncx, ncy, ot, nt)
    return out


def shift_tables(n, ref, strategy, block):
    m = lib().ora_shift_tables(n, ref, strategy, block, None, None, None, None)
    if m < 0:
        raise ValueError("invalid strategy")
    A = np.zeros(m * (n - 1), np.float32)
    pairs = np.zeros((m, 2), np.int32)
    index = np.zeros((n, n), np.int32)
    rows = np.zeros((m, 2), np.int32)
    lib().ora_shift_tables(n, ref, strategy, block, ptr(A), ptr(pairs), ptr(index), ptr(rows))
    return A.reshape(n - 1, m).T.copy(), pairs, index, rows  # A as (m, n1)


def minimize_shifts(d: np.ndarray, n, ref, strategy, block):
    """d: (tiles, m, 2) float64 measurements. Returns (dhat, r, status, iters)."""
    d = np.ascontiguousarray(d, np.float64)
    tiles, m, _ = d.shape
    dout = np.zeros_like(d)
    r = np.zeros((tiles, n - 1, 2), np.float64)
    st = np.zeros(tiles, np.int32)
    it = np.zeros(tiles, np.int32)
    rc = lib().ora_minimize_shifts(n, ref, strategy, block, tiles, ptr(d), ptr(dout), ptr(r), ptr(st), ptr(it))
    if rc < 0:
        raise ValueError("invalid strategy")
    return dout, r, st, it


def get_optimal_shifts(r, n, ref, trk, dt=np.float32):
    r = np.ascontiguousarray(r, dt)
    tiles = r.shape[0]
    out = np.zeros((tiles, 2), dt)
    fn("get_optimal_shifts", dt)(ptr(r), ptr(out), n, tiles, ref, trk)
    return out


def merge(params, raw: np.ndarray, base_shift=None, base_rot=None, dt=np.float32, want=()):
    """Full oracle pipeline. raw: (n, h, w) uint16. Returns dict with 'out' and the requested intermediates."""
    raw = np.ascontiguousarray(raw, np.uint16)
    n, h, w = raw.shape
    s2 = 2 if params.super_resolution == 2 else 1  # full-frame 2x merge: 2w x 2h output
    res = {"out": np.zeros((s2 * h, s2 * w, 3), dt)}
    bufs = {
        "one_to_one": None, "flows": (n, h, w, 2), "masks": (n, h // 2, w // 2, 4), "kern": (h, w, 3),
        "track": (n, h, w), "lk": (n, h, w), "sum": (s2 * h, s2 * w, 3), "wsum": (s2 * h, s2 * w, 3),
        "lk_cond": (n, h, w, 2),
    }
    for k in want:
        if k == "one_to_one":
            L = params.n_levels
            cx = (w // params.resize_level[0] - 2 * params.max_shift[0]) // params.tile_size[0]
            cy = (h // params.resize_level[0] - 2 * params.max_shift[0]) // params.tile_size[0]
            res[k] = np.zeros((cy * cx, n - 1, 2), dt)
            del L
        elif k == "lk_cond":
            res[k] = np.zeros(bufs[k], np.float32)
        else:
            res[k] = np.zeros(bufs[k], dt)
    if base_shift is None and getattr(params, "prealign_max_shift", 0) > 0:
        # ComputePreAlignment (R:PEFStudioDX/ImageStackAlignatorController.cs:1451-1528): initial shifts from the
        # tracking images (DeBayerBWGaussWB with the high-pass), global maximum of the FFT cross-correlation
        trk = [tracking_image(raw[f], params, params.use_high_pass, params.sigma_debayer_tracking, dt) for f in range(n)]
        init = prealign(trk)
        res["prealign"] = init
        base_shift = init[:, :2].ravel()
        base_rot = init[:, 2]
    bs = None if base_shift is None else np.ascontiguousarray(base_shift, np.float32)
    br = None if base_rot is None else np.ascontiguousarray(base_rot, np.float32)
    ref = C.c_int(-1)
    g = res.get
    rc = fn("merge", dt)(C.addressof(params), w, h, n, ptr(raw), ptr(bs), ptr(br), ptr(res["out"]), ptr(g("one_to_one")),
                         ptr(g("flows")), ptr(g("masks")), ptr(g("kern")), ptr(g("track")), ptr(g("lk")),
                         ptr(g("sum")), ptr(g("wsum")), C.byref(ref), ptr(g("lk_cond")))
    if rc != 0:
        raise RuntimeError(f"oracle merge failed ({rc})")
    res["ref"] = ref.value
    return res


# ---------------------------------------------------------------------------
# Global pre-alignment (translation): PreAlignment.ScanAngles at a single angle
# (R:PEFStudioDX/PreAlignment.cs:177-261) as ComputePreAlignment drives it
# (R:PEFStudioDX/ImageStackAlignatorController.cs:1451-1528), with numpy's FFT in double.
def cross_correlation(ref: np.ndarray, trk: np.ndarray) -> np.ndarray:
    """C2R(conj(R2C(ref)) * R2C(trk)) / (W H): c[s] = sum_x ref(x) trk((x + s) mod (W, H)), indexed [sy, sx]
    (conjugateComplexMulKernel conjugates its first operand, R:Kernels/kernel.cu:394-411)."""
    h, w = ref.shape
    a = np.fft.rfft2(np.asarray(ref, np.float64))
    b = np.fft.rfft2(np.asarray(trk, np.float64))
    return np.fft.irfft2(np.conj(a) * b, s=(h, w))


def prealign(frames) -> np.ndarray:
    """initialShifts of every frame against frame 0: (n, 4) = (-sx, -sy, angle 0, c[s]) for the first maximum s
    of the circular cross-correlation in memory order, wrapped into (-W/2, W/2] x (-H/2, H/2]
    (R:PEFStudioDX/PreAlignment.cs:248-258); frame 0 gets zeros."""
    frames = [np.asarray(f, np.float64) for f in frames]
    h, w = frames[0].shape
    out = np.zeros((len(frames), 4))
    for i in range(1, len(frames)):
        cc = cross_correlation(frames[0], frames[i])
        idx = int(np.argmax(cc))
        y, x = divmod(idx, w)
        v = cc[y, x]
        if x > w // 2:
            x -= w
        if y > h // 2:
            y -= h
        out[i] = (-x, -y, 0.0, v)
    return out


def prealignment_store(initial_shifts) -> tuple:
    """PreAlignmentStore (R:PEFStudioDX/PreAlignmentStore.cs:36-110): the reference index (the frame minimising the
    length of the summed shifts seen from it) and each frame's shift relative to it, GetShift(ref, f)."""
    s = np.asarray(initial_shifts, np.float64)[:, :2]
    n = len(s)
    basic = s[1:] - s[:-1]
    best, best_d = -1, np.inf
    for i in range(n - 1):
        tot = -basic[:i].sum(axis=0) + basic[i:].sum(axis=0)
        d = float(np.hypot(*tot))
        if d < best_d:
            best, best_d = i, d
    best = max(best, 0)
    per = np.zeros((n, 2))
    for i in range(n):
        per[i] = -basic[i:best].sum(axis=0) + basic[best:i].sum(axis=0)
    return best, per.astype(np.float32)


# ---------------------------------------------------------------------------
# Differentiable (relaxed) merge in FP64: forward loss/output and its gradient.
def _diff_args(prob):
    L = lib()
    for name in ("ora_merge_diff_forward", "ora_merge_diff_backward"):
        f = getattr(L, name)
        f.restype = C.c_double if name.endswith("forward") else None
    f64 = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
    keep = dict(raw=np.ascontiguousarray(prob["raw"], np.uint16), half=f64(prob["half"]), tensor=f64(prob["tensor"]),
                anchor=f64(prob["anchor"]), black=f64(prob["black"]), white=f64(prob["white"]),
                cfa=np.ascontiguousarray(prob["cfa"], np.int32))
    n, h, w = keep["raw"].shape
    args = [C.c_int(w), C.c_int(h), C.c_int(n), C.c_int(prob["reference"]), ptr(keep["raw"]), ptr(keep["half"]),
            ptr(keep["tensor"]), ptr(keep["anchor"]), ptr(keep["black"]), ptr(keep["white"]), ptr(keep["cfa"]),
            C.c_double(prob["alpha"]), C.c_double(prob["beta"]), C.c_double(prob["k_stretch"]),
            C.c_double(prob["k_shrink"])]
    return args, keep, (n, h, w)


def merge_diff_forward(prob, flow, theta, gt=None):
    """Relaxed merge in FP64. prob: dict(raw, half, tensor, anchor, black, white, cfa, reference, alpha, beta,
    k_stretch, k_shrink). Returns (loss, out (2h, 2w, 3), wsum, mask (n, h/2, w/2, 3))."""
    args, keep, (n, h, w) = _diff_args(prob)
    fl = np.ascontiguousarray(flow, np.float64)
    th = np.ascontiguousarray(theta, np.float64)
    out = np.zeros((2 * h, 2 * w, 3))
    ws = np.zeros_like(out)
    mask = np.zeros((n, h // 2, w // 2, 3))
    g = None if gt is None else np.ascontiguousarray(gt, np.float64)
    loss = lib().ora_merge_diff_forward(*args, ptr(fl), ptr(th), ptr(out), ptr(ws), ptr(mask), ptr(g))
    return loss, out, ws, mask


def merge_diff_backward(prob, flow, theta, gt):
    """Gradient of the relaxed-merge loss: (dflow (n, h, w, 2), dtheta (4,))."""
    args, keep, (n, h, w) = _diff_args(prob)
    fl = np.ascontiguousarray(flow, np.float64)
    th = np.ascontiguousarray(theta, np.float64)
    g = np.ascontiguousarray(gt, np.float64)
    dflow = np.zeros((n, h, w, 2))
    dth = np.zeros(4)
    lib().ora_merge_diff_backward(*args, ptr(fl), ptr(th), ptr(g), ptr(dflow), ptr(dth))
    return dflow, dth
