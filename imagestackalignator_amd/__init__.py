"""MI355X (gfx950) burst-merge hot path of ImageStackAlignator.

The product is the C-ABI library ``_lib/libisa_hip.so`` (declared in
``include/isa_hip.h``); this module loads it with ctypes and exposes thin
Python bindings used by the tests, ``bench.py`` and ``__graft_entry__``.
There is no CPU fallback: if the library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "_lib" / "libisa_hip.so"
HEADER = _PKG.parent / "include" / "isa_hip.h"

ISA_MAX_LEVELS = 8
ISA_OK = 0


class IsaError(RuntimeError):
    """A non-zero status from libisa_hip (the reference threw CudaException /
    ArgumentOutOfRangeException in these places)."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


class IsaParams(C.Structure):
    _fields_ = [
        ("sigma_debayer_tracking", C.c_float),
        ("sigma_debayer_accumulation", C.c_float),
        ("high_pass", C.c_float),
        ("high_pass_sigma", C.c_float),
        ("clear_axis", C.c_int),
        ("use_high_pass", C.c_int),
        ("matching_threshold", C.c_float),
        ("n_levels", C.c_int),
        ("resize_level", C.c_int * ISA_MAX_LEVELS),
        ("tile_size", C.c_int * ISA_MAX_LEVELS),
        ("max_shift", C.c_int * ISA_MAX_LEVELS),
        ("tracking_strategy", C.c_int),
        ("tracking_block_size", C.c_int),
        ("sigma_structure_tensor", C.c_float),
        ("dth", C.c_float),
        ("dtr", C.c_float),
        ("k_detail", C.c_float),
        ("k_denoise", C.c_float),
        ("k_stretch", C.c_float),
        ("k_shrink", C.c_float),
        ("iterations_lk", C.c_int),
        ("window_size_lk", C.c_int),
        ("min_det_lk", C.c_float),
        ("erode_size", C.c_int),
        ("threshold_m", C.c_float),
        ("super_resolution", C.c_int),
        ("green_channel_only", C.c_int),
        ("noise_alpha", C.c_float),
        ("noise_beta", C.c_float),
        ("black", C.c_float * 3),
        ("white", C.c_float * 3),
        ("camera_white", C.c_float * 3),
        ("cfa", C.c_int * 4),
        ("reference_index", C.c_int),
        ("tex_quant", C.c_int),
        ("xc_precision", C.c_int),
        ("lk_ill_ratio", C.c_float),
        ("prealign_max_shift", C.c_int),
        ("xc_levels", C.c_int),
    ]

    def set_levels(self, levels):
        """levels: list of (resize_level, tile_size, max_shift), finest first."""
        self.n_levels = len(levels)
        for i, (l, t, m) in enumerate(levels):
            self.resize_level[i] = l
            self.tile_size[i] = t
            self.max_shift[i] = m
        return self


class IsaSynthConfig(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("width", C.c_int),
        ("height", C.c_int),
        ("frames", C.c_int),
        ("global_shift_max", C.c_float),
        ("flow_max", C.c_float),
        ("flow_scale", C.c_float),
        ("occluded_fraction", C.c_float),
        ("gain", C.c_float),
        ("read_sigma", C.c_float),
        ("n_blobs", C.c_int),
        ("n_edges", C.c_int),
        ("n_waves", C.c_int),
        ("white", C.c_float),
    ]


_CTYPE = {
    "int": C.c_int,
    "float": C.c_float,
    "size_t": C.c_size_t,
    "void": None,
    "const char*": C.c_char_p,
    "char*": C.c_char_p,
    "isa_params*": C.POINTER(IsaParams),
    "const isa_params*": C.POINTER(IsaParams),
    "isa_synth_config*": C.POINTER(IsaSynthConfig),
    "const isa_synth_config*": C.POINTER(IsaSynthConfig),
    "int*": C.POINTER(C.c_int),
    "const int*": C.POINTER(C.c_int),
    "float*": C.c_void_p,
}


def _arg_ctype(decl: str):
    decl = decl.strip()
    arr = re.search(r"\[\s*\d*\s*\]$", decl)
    base = re.sub(r"\s*\[\s*\d*\s*\]$", "", decl)
    m = re.match(r"(.*?)([A-Za-z_]\w*)$", base)
    typ = re.sub(r"\s+", " ", m.group(1)).replace(" *", "*").strip()
    if arr:
        # host arrays of small fixed size: float[2], float[3], int[4]
        return C.POINTER(C.c_int) if typ.startswith(("int", "const int")) else C.POINTER(C.c_float)
    if typ in ("isa_stream_t",):
        return C.c_void_p
    if typ in ("const isa_params*", "isa_params*", "const isa_synth_config*", "isa_synth_config*"):
        return _CTYPE[typ]
    if typ in ("int*", "const int*") and not m.group(2).endswith("_dev"):
        return C.POINTER(C.c_int)
    if typ in ("float*", "const float*") and not m.group(2).endswith("_dev"):
        return C.POINTER(C.c_float)
    if "*" in typ:
        return C.c_void_p
    return _CTYPE[typ]


def _parse_header(path: Path):
    text = path.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    protos = {}
    for m in re.finditer(r"\n\s*((?:const\s+)?[A-Za-z_]\w*\s*\**)\s*(isa_\w+)\s*\(([^;{]*?)\)\s*;", text):
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3).strip()
        ret = re.sub(r"\s+", " ", ret).replace(" *", "*")
        argt = [] if args in ("", "void") else [_arg_ctype(a) for a in args.split(",")]
        if ret in ("const char*", "char*"):
            rt = C.c_char_p
        elif ret == "size_t":
            rt = C.c_size_t
        elif ret == "void":
            rt = None
        else:
            rt = C.c_int
        protos[name] = (rt, argt)
    return protos


PROTOTYPES = _parse_header(HEADER)


def load(path: Path | str | None = None) -> C.CDLL:
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} is missing: build it with `python -m imagestackalignator_amd.build` "
            "(this package has no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (rt, argt) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = rt
        fn.argtypes = argt
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(name: str, rc: int) -> None:
    if rc != ISA_OK:
        msg = lib().isa_last_error().decode(errors="replace")
        raise IsaError(name, rc, msg)


def call(name: str, *args) -> int:
    """Calls an isa_* function and raises IsaError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    check(name, rc)
    return rc


def f32s(*vals):
    return (C.c_float * len(vals))(*vals)


def i32s(*vals):
    return (C.c_int * len(vals))(*vals)


def default_params(levels=None, **kw) -> IsaParams:
    p = IsaParams()
    lib().isa_default_params(C.byref(p))
    if levels:
        p.set_levels(levels)
    for k, v in kw.items():
        if k in ("black", "white", "camera_white"):
            getattr(p, k)[:] = list(v)
        elif k == "cfa":
            p.cfa[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def default_synth(**kw) -> IsaSynthConfig:
    c = IsaSynthConfig()
    lib().isa_default_synth_config(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


from .merge import BurstMerger, synth_burst, ground_truth  # noqa: E402

__all__ = ["lib", "call", "check", "IsaError", "IsaParams", "IsaSynthConfig", "default_params", "default_synth",
           "BurstMerger", "synth_burst", "ground_truth", "f32s", "i32s", "PROTOTYPES", "LIB_PATH"]
