"""Raw ingest over the C ABI: the RawFile / PEFFile / DNGFile readers of the reference
(R:PefFile/RawFile.cs:28-279, R:PefFile/PEFFile.cs:28-316, R:PefFile/DNGFile.cs:36-950).

``RawFile`` parses a file's container with isa_raw_open and exposes the reference's properties;
``decode_files`` decodes the pixel data of many files in one isa_raw_decode call on the GPU, as the
controller loads a burst (one task per file, R:PEFStudioDX/ImageStackAlignatorController.cs:1066-1119).
``ExtraCameraProfiles`` reads the reference's profile XML (R:PefFile/ExtraCameraProfiles.cs) for
``RawFile.load_extra_camera_profile`` (R:PefFile/RawFile.cs:232-278).
"""
from __future__ import annotations

import ctypes as C
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

from . import call, lib
from .merge import _stream_ptr, _torch, check_tensor, raw_dtypes

FORMATS = {"auto": 0, "pef": 1, "dng": 2}
COMPRESSION = {1: "uncompressed", 7: "lossless_jpeg", 65535: "pentax_huffman"}


class IsaRawInfo(C.Structure):
    _fields_ = [
        ("format", C.c_int),
        ("compression", C.c_int),
        ("width", C.c_int),
        ("height", C.c_int),
        ("bit_depth", C.c_int),
        ("crop_left", C.c_int),
        ("crop_top", C.c_int),
        ("cropped_width", C.c_int),
        ("cropped_height", C.c_int),
        ("bayer_pattern", C.c_int * 4),
        ("white_level", C.c_float * 3),
        ("black_level", C.c_float * 3),
        ("white_balance", C.c_float * 3),
        ("color_twist", C.c_float * 12),
        ("color_twist_is_identity", C.c_int),
        ("roll_angle", C.c_float),
        ("roll_angle_present", C.c_int),
        ("noise_alpha", C.c_float),
        ("noise_beta", C.c_float),
        ("iso", C.c_int),
        ("exposure_num", C.c_uint),
        ("exposure_den", C.c_uint),
        ("orientation", C.c_int),
        ("max_black_delta", C.c_int),
        ("make", C.c_char * 64),
        ("unique_model", C.c_char * 64),
        ("recording_date", C.c_char * 32),
    ]


class RawFile:
    """A PEF or DNG file: its container parsed on the host, its RawImage decoded on the GPU.

    Properties follow RawFile (R:PefFile/RawFile.cs:121-229): raw_width / raw_height, crop_left / crop_top /
    cropped_width / cropped_height, bayer_pattern, white_level / black_level / white_balance (R, G, B),
    color_twist (3 x 4), roll_angle, iso, exposure_time, noise_model_alpha / _beta, make, unique_model_name.
    """

    def __init__(self, data, fmt: str = "auto"):
        if isinstance(data, (str, Path)):
            data = Path(data).read_bytes()
        self.data = bytes(data)
        self._h = C.c_void_p()
        call("isa_raw_open", self.data, len(self.data), FORMATS[fmt], C.byref(self._h))
        info = IsaRawInfo()
        call("isa_raw_get_info", self._h, C.byref(info))
        self.info = info
        self.raw_image = None
        self.noise_model_alpha = float(info.noise_alpha)
        self.noise_model_beta = float(info.noise_beta)
        self.crop_left, self.crop_top = info.crop_left, info.crop_top
        self.cropped_width, self.cropped_height = info.cropped_width, info.cropped_height

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().isa_raw_close(h)
            self._h = None

    format = property(lambda s: "pef" if s.info.format == 1 else "dng")
    compression = property(lambda s: COMPRESSION.get(s.info.compression, str(s.info.compression)))
    raw_width = property(lambda s: s.info.width)
    raw_height = property(lambda s: s.info.height)
    bit_depth = property(lambda s: s.info.bit_depth)
    bayer_pattern = property(lambda s: list(s.info.bayer_pattern))
    white_level = property(lambda s: np.array(s.info.white_level, np.float32))
    black_level = property(lambda s: np.array(s.info.black_level, np.float32))
    white_balance = property(lambda s: np.array(s.info.white_balance, np.float32))
    color_twist = property(lambda s: np.array(s.info.color_twist, np.float32).reshape(3, 4))
    color_twist_is_identity = property(lambda s: bool(s.info.color_twist_is_identity))
    roll_angle = property(lambda s: float(s.info.roll_angle))
    roll_angle_present = property(lambda s: bool(s.info.roll_angle_present))
    iso = property(lambda s: s.info.iso)
    exposure_time = property(lambda s: (s.info.exposure_num, s.info.exposure_den))
    orientation = property(lambda s: s.info.orientation)
    recording_date = property(lambda s: s.info.recording_date.decode(errors="replace"))
    make = property(lambda s: s.info.make.decode(errors="replace"))
    unique_model_name = property(lambda s: s.info.unique_model.decode(errors="replace"))
    max_black_delta = property(lambda s: s.info.max_black_delta)

    def decode(self, out=None):
        """Decodes RawImage on the GPU: an (H, W) 16-bit tensor (torch.int16 holding the u16 bits)."""
        return decode_files([self], [out] if out is not None else None)[0]

    def load_extra_camera_profile(self, xml_path) -> bool:
        """RawFile.LoadExtraCameraProfile: crop and noise model from the profile of this camera, if any.
        The colour matrices of a profile feed the colour specification, which is outside this library."""
        prof = ExtraCameraProfiles.load(xml_path).get_profile(self.make, self.unique_model_name)
        if prof is None:
            return False
        any_set = False
        if prof.get("crop") is not None:
            top, left, width, height = prof["crop"]
            w, h = self.raw_width, self.raw_height
            self.crop_left = max(2, left)
            self.crop_top = max(2, top)
            self.cropped_width = width - max(0, (left + width) - (w - 2))
            self.cropped_height = height - max(0, (top + height) - (h - 2))
            any_set = True
        if prof.get("color_matrix1") is not None:
            any_set = True
        if prof.get("noise") is not None:
            a, b = noise_model_value(prof["noise"], self.iso)
            self.noise_model_alpha = float(np.float32(a))
            self.noise_model_beta = float(np.float32(b))
            any_set = True
        return any_set


def noise_model_value(noise, iso: int):
    """NoiseModel.GetValue (R:PefFile/ExtraCameraProfiles.cs:171-205): the last ISO entry not above `iso`,
    linear interpolation towards the next one."""
    isos, alpha, beta = noise
    if len(isos) == 0:
        return 0.0, 0.0
    index = -1
    for i in range(len(isos) - 1, -1, -1):
        if isos[i] <= iso:
            index = i
            break
    if index == len(isos) - 1:
        return alpha[-1], beta[-1]
    if index < 0:
        return alpha[0], beta[0]
    w = (float(iso) - float(isos[index + 1])) / (float(isos[index]) - float(isos[index + 1]))
    return w * alpha[index] + (1.0 - w) * alpha[index + 1], w * beta[index] + (1.0 - w) * beta[index + 1]


class ExtraCameraProfiles:
    """The XmlSerializer form of ExtraCameraProfiles: Profiles / ExtraCameraProfile with Make, UniqueModel,
    ColorMatrix1/2, Illuminant1/2, NoiseModel (Iso, Alpha, Beta) and CropInfo (Top, Left, Width, Height)."""

    def __init__(self, profiles):
        self.profiles = profiles

    @staticmethod
    def load(path) -> "ExtraCameraProfiles":
        root = ET.parse(str(path)).getroot()
        out = []
        for p in root.iter("ExtraCameraProfile"):
            d = {"make": (p.findtext("Make") or ""), "unique_model": (p.findtext("UniqueModel") or "")}
            cm = p.find("ColorMatrix1")
            d["color_matrix1"] = [float(x.text) for x in cm] if cm is not None else None
            nm = p.find("NoiseModel")
            if nm is not None:
                isos = [int(x.text) for x in nm.find("Iso")] if nm.find("Iso") is not None else []
                al = [float(x.text) for x in nm.find("Alpha")] if nm.find("Alpha") is not None else []
                be = [float(x.text) for x in nm.find("Beta")] if nm.find("Beta") is not None else []
                d["noise"] = (isos, al, be)
            else:
                d["noise"] = None
            ci = p.find("CropInfo")
            d["crop"] = (int(ci.findtext("Top")), int(ci.findtext("Left")), int(ci.findtext("Width")),
                         int(ci.findtext("Height"))) if ci is not None else None
            out.append(d)
        return ExtraCameraProfiles(out)

    def get_profile(self, make: str, unique_model: str):
        for p in self.profiles:
            if p["make"] == make and p["unique_model"] == unique_model:
                return p
        return None


def decode_files(files, outs=None, stream=None):
    """Decodes the RawImage of every file in one isa_raw_decode call; returns (H, W) int16 tensors holding
    the u16 samples (or fills `outs`)."""
    torch = _torch()
    n = len(files)
    dev = torch.device("cuda")
    bufs = []
    for f in files:
        t = torch.empty(((len(f.data) + 3) // 4 * 4,), dtype=torch.uint8, device=dev)
        t[: len(f.data)].copy_(torch.frombuffer(bytearray(f.data), dtype=torch.uint8))
        bufs.append(t)
    if outs is None:
        outs = [torch.empty((f.raw_height, f.raw_width), dtype=torch.int16, device=dev) for f in files]
    for f, o in zip(files, outs):
        check_tensor("out", o, (f.raw_height, f.raw_width), raw_dtypes())
    return decode_device(files, bufs, outs, stream)


def decode_device(files, bufs, outs, stream=None, workspace=None):
    """isa_raw_decode on device copies `bufs` of the files (uint8 tensors, at least the file's size)."""
    torch = _torch()
    n = len(files)
    hs = (C.c_void_p * n)(*[f._h.value for f in files])
    need = C.c_size_t()
    call("isa_raw_decode_workspace", hs, n, C.byref(need))
    if workspace is None or workspace.numel() < need.value:
        workspace = torch.empty((max(1, need.value),), dtype=torch.uint8, device=bufs[0].device)
    for f, b in zip(files, bufs):
        if not (b.is_cuda and b.dtype == torch.uint8 and b.is_contiguous() and b.numel() >= len(f.data)):
            raise ValueError("file buffers: contiguous uint8 GPU tensors holding the whole file")
    fp = (C.c_void_p * n)(*[b.data_ptr() for b in bufs])
    op = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
    pitch = (C.c_int * n)(*[o.shape[1] * 2 for o in outs])
    call("isa_raw_decode", hs, n, fp, op, pitch, C.c_void_p(workspace.data_ptr()), workspace.numel(),
         stream if stream is not None else _stream_ptr())
    for f, o in zip(files, outs):
        f.raw_image = o
    return outs
