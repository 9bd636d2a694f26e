"""Oracle of the raw-file readers -- TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's cpu_baseline).

A restatement of RawFile / PEFFile / DNGFile of the reference: the TIFF container (R:PefFile/RawFile.cs:42-90),
the Pentax maker note (R:PefFile/PentaxMakerNotes.cs:42-199, Huffman table R:PefFile/PentaxMakerNotesEntry.cs:
4555-4613), the properties PEFFile.SetMembers (R:PefFile/PEFFile.cs:180-315) and DNGFile.SetMembers /
the DNGFile constructor (R:PefFile/DNGFile.cs:300-372, 605-950) compute, and the pixel decoders of
oracle/isa_oracle_raw.cpp. Values follow the reference's number types (float32 for the levels, C# casts).

Where the reference derives the DNG camera white through its colour-matrix machinery (DNGColorSpec), this
oracle uses the normalised AsShotNeutral, the fixed point that machinery returns for it.
"""
from __future__ import annotations

import ctypes as C
import math
import struct

import numpy as np

from . import oracle as _O

_SIZES = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 6: 1, 7: 1, 8: 2, 9: 4, 10: 8, 11: 4, 12: 8}
CFA_TYPE = 32803


class RawError(Exception):
    """Where the reference throws."""


class Entry:
    def __init__(self, tag, typ, count, values, offset, raw):
        self.tag, self.type, self.count, self.values, self.offset, self.raw = tag, typ, count, values, offset, raw


def _values(data: bytes, o: str, typ: int, count: int, at: int):
    n = _SIZES.get(typ, 1) * count
    raw = data[at:at + n]
    if len(raw) < n:
        raise RawError("entry value beyond the end of the file")
    if typ == 2:
        return raw.split(b"\x00")[0].decode("latin-1"), raw
    if typ in (1, 7):
        return list(raw), raw
    if typ == 6:
        return list(struct.unpack(f"{count}b", raw)), raw
    if typ in (5, 10):
        f = "I" if typ == 5 else "i"
        return [struct.unpack_from(o + f + f, raw, 8 * i) for i in range(count)], raw
    f = {3: "H", 4: "I", 8: "h", 9: "i", 11: "f", 12: "d"}.get(typ)
    if f is None:
        return list(raw), raw
    return list(struct.unpack(f"{o}{count}{f}", raw)), raw


def read_entries(data: bytes, o: str, off: int, base: int = 0):
    """One directory at `off`: {tag: Entry} and the next-directory offset. Offsets in entries are taken
    relative to `base` (0 for the file, the maker-note rule for maker notes)."""
    if off + 2 > len(data):
        raise RawError("directory beyond the end of the file")
    n = struct.unpack_from(o + "H", data, off)[0]
    out = {}
    for k in range(n):
        e = off + 2 + 12 * k
        tag, typ, count = struct.unpack_from(o + "HHI", data, e)
        voff = struct.unpack_from(o + "I", data, e + 8)[0]
        size = _SIZES.get(typ, 1) * count
        at = e + 8 if size <= 4 else voff - base
        vals, raw = _values(data, o, typ, count, at)
        out.setdefault(tag, Entry(tag, typ, count, vals, voff, raw))
    nxt = struct.unpack_from(o + "I", data, off + 2 + 12 * n)[0]
    return out, nxt


class IFD:
    def __init__(self, data, o, off):
        self.entries, self.next = read_entries(data, o, off)
        self.sub = []
        if 0x14A in self.entries:
            self.sub = [IFD(data, o, x) for x in self.entries[0x14A].values]
        self.exif = IFD(data, o, self.entries[0x8769].values[0]) if 0x8769 in self.entries else None

    def get(self, tag):
        e = self.entries.get(tag)
        return None if e is None else e.values


def rational(v) -> float:
    return v[0] / v[1] if isinstance(v, tuple) else float(v)


def maker_note(raw: bytes, offset: int):
    """PentaxMakerNotes(aData, offset): returns ({tag: Entry}, byte order) or (None, None)."""
    def order_of(b):
        if b == b"II":
            return "<"
        if b == b"MM":
            return ">"
        raise RawError("maker-note byte order")
    if raw[:4] == b"AOC\x00":
        o, add, pos = order_of(raw[4:6]), offset, 6
    elif raw[:8] == b"PENTAX \x00":
        o, add, pos = order_of(raw[8:10]), 0, 10
    elif raw[:10] == b"Adobe\x00MakN":
        o = order_of(raw[14:16])
        add = struct.unpack_from(o + "I", raw, 16)[0]
        if raw[20:24] == b"AOC\x00":
            o, add, pos = order_of(raw[24:26]), add - 20, 26
        elif raw[20:28] == b"PENTAX \x00":
            o, add, pos = order_of(raw[28:30]), -20, 30
        else:
            pos = 20
    else:
        return None, None
    entries, _ = read_entries(raw, o, pos, base=add)
    return entries, o


def pentax_huffman_table(raw: bytes, o: str) -> np.ndarray:
    """MNHuffmanTable: the 4097-entry decode table (entry 0 = 12 lookahead bits)."""
    dep = (struct.unpack_from(o + "H", raw, 0)[0] + 12) & 15
    bit0 = struct.unpack_from(f"{o}{dep}H", raw, 14)
    bit1 = raw[14 + 2 * dep:14 + 3 * dep]
    table = np.zeros(4097, np.uint16)
    for c in range(dep):
        i = bit0[c]
        end = (bit0[c] + (4096 >> bit1[c]) - 1) & 4095
        while i <= end:
            i += 1
            table[i] = (bit1[c] << 8 | c) & 0xFFFF
        table[0] = 12
    return table


def _black_from_pattern(vals, bayer):
    """Four black levels in CFA order: red, blue, and the larger of the first and last green."""
    ir = next(i for i, b in enumerate(bayer) if b == 0)
    ib = next(i for i, b in enumerate(bayer) if b == 2)
    g1 = g2 = -1
    for i, b in enumerate(bayer):
        if b == 1 and g1 == -1:
            g1 = i
        if b == 1 and g1 != -1:
            g2 = i
    return [np.float32(vals[ir]), np.float32(max(np.float32(vals[g1]), np.float32(vals[g2]))), np.float32(vals[ib])]


def _ushort_of_double(x: float) -> int:
    return int(math.trunc(x)) & 0xFFFF


class RawFile:
    """Metadata and pixels of a PEF or DNG file held in memory, as the reference's reader computes them."""

    def __init__(self, data: bytes, decode: bool = True, kind: str | None = None):
        self.data = bytes(data)
        if data[:2] == b"II":
            self.o = "<"
        elif data[:2] == b"MM":
            self.o = ">"
        else:
            raise RawError("endianness")
        if struct.unpack_from(self.o + "H", data, 2)[0] != 42:
            raise RawError("magic number")
        off = struct.unpack_from(self.o + "I", data, 4)[0]
        self.ifds = []
        while True:
            ifd = IFD(data, self.o, off)
            self.ifds.append(ifd)
            if ifd.next == 0:
                break
            off = ifd.next
        self.color_twist = np.zeros((3, 4), np.float32)
        self.color_twist[0, 0] = self.color_twist[1, 1] = self.color_twist[2, 2] = 1
        self.color_twist_is_identity = True
        self.roll_angle, self.roll_angle_present = 0.0, False
        self.noise_alpha = self.noise_beta = np.float32(0)
        self.recording_date = ""
        self.raw_image = None
        if kind is None:
            kind = "dng" if self.ifds[0].get(0xC612) is not None or self.ifds[0].get(0xC614) is not None else "pef"
        self.is_dng = kind == "dng"
        if self.is_dng:
            self._dng(decode)
        else:
            self._pef(decode)

    # -- PEF: R:PefFile/PEFFile.cs:180-315 and the constructor :83-170 ------------------------------------------
    def _pef(self, decode):
        raw = self.ifds[0]
        exif = raw.exif
        if exif is None:
            raise RawError("PEF without Exif IFD")
        self.iso = exif.get(0x8827)[0]
        self.exposure = exif.get(0x829A)[0]
        self.recording_date = exif.get(0x9004)
        self.bit_depth = raw.get(0x102)[0]
        mn_entry = exif.entries[0x927C]
        mn, mo = maker_note(bytes(mn_entry.raw), mn_entry.offset)
        if mn is None:
            raise RawError("no Pentax maker note")
        huff = pentax_huffman_table(bytes(mn[0x0220].raw), mo) if 0x0220 in mn else None
        cfa = exif.get(0xA302)
        xc, yc = cfa[1], cfa[3]
        self.bayer = [0] * (xc * yc)
        for x in range(xc):
            for y in range(yc):
                self.bayer[y * yc + x] = cfa[y * yc + x + 4]
        white_all = np.float32(mn[0x007E].values[0]) if 0x007E in mn else np.float32(2.0 ** self.bit_depth)
        black = [np.float32(0)] * 3
        if 0x0200 in mn:
            bp = mn[0x0200].values
            if len(bp) == 1:
                black = [np.float32(bp[0])] * 3
            if len(bp) == 3:
                black = [np.float32(v) for v in bp]
            if len(bp) == 4:
                black = _black_from_pattern(bp, self.bayer)
        self.black_level = np.array(black, np.float32)
        self.white_level = np.array([white_all - b for b in black], np.float32)
        scale = np.float32(mn[0x003D].values[0])
        wp = mn[0x0201].values
        self.white_balance = np.array([np.float32(wp[0]) / scale, np.float32(wp[1]) / scale,
                                       np.float32(wp[3]) / scale], np.float32)
        if 0x022B in mn:
            self.roll_angle = float(np.float32(-mn[0x022B].values[1]) / np.float32(2.0))
            self.roll_angle_present = True
        self.width, self.height = raw.get(0x100)[0], raw.get(0x101)[0]
        self.orientation = raw.get(0x112)[0] if raw.get(0x112) is not None else 1
        self.crop = (2, 2, self.width - 4, self.height - 4)
        self.make, self.unique_model = raw.get(0x10F), raw.get(0x110)
        self.huffman = huff
        self.compression = "pentax" if huff is not None else "none"
        if not decode:
            return
        out = np.zeros((self.height, self.width), np.uint16)
        offs = np.asarray(raw.get(0x111), np.uint32)
        if huff is not None and len(offs) == 1:
            rc = _lib().ora_pef_huffman(self.data, len(self.data), int(offs[0]), huff.ctypes.data, self.width,
                                          self.height, out.ctypes.data)
        elif huff is not None:
            raise RawError("compressed multiple strips")
        else:
            counts = np.asarray(raw.get(0x117), np.uint32)
            rc = _lib().ora_pef_uncompressed(self.data, len(self.data), offs.ctypes.data, counts.ctypes.data,
                                               len(offs), raw.get(0x116)[0], self.width, self.height,
                                               self.bit_depth, out.ctypes.data)
        if rc != 0:
            raise RawError("pixel data")
        self.raw_image = out

    # -- DNG: R:PefFile/DNGFile.cs:300-372, 605-950 and the decode loops :372-611 --------------------------------
    def _dng(self, decode):
        ifd0 = self.ifds[0]
        raw = None
        for ifd in self.ifds:
            if ifd.get(0x106) is not None and ifd.get(0x106)[0] == CFA_TYPE:
                raw = ifd
                break
        if raw is None:
            for ifd in self.ifds:
                for s in ifd.sub:
                    if s.get(0x106) is None:
                        raise RawError("SubIFD without PhotometricInterpretation")
                    if s.get(0x106)[0] == CFA_TYPE:
                        raw = s
                        break
        if raw is None:
            raise RawError("no CFA image")
        ver = ifd0.get(0xC612)
        self.large = ver is None or ver[1] > 1
        self.lin = np.asarray(raw.get(0xC618), np.uint16) if raw.get(0xC618) is not None else None
        self.width, self.height = raw.get(0x100)[0], raw.get(0x101)[0]
        if ifd0.get(0x132) is not None:
            self.recording_date = ifd0.get(0x132)
        priv = ifd0.entries.get(0xC634)
        if priv is not None:
            mn, _ = maker_note(bytes(priv.raw), 0)
            if mn is not None and 0x022B in mn:
                self.roll_angle = float(np.float32(-mn[0x022B].values[1]) / np.float32(2.0))
                self.roll_angle_present = True
        exif = ifd0.exif
        if exif is not None:
            self.iso = exif.get(0x8827)[0]
            self.exposure = exif.get(0x829A)[0]
            if exif.get(0x9004) is not None:
                self.recording_date = exif.get(0x9004)
        elif raw.get(0x8827) is not None:
            self.iso, self.exposure = raw.get(0x8827)[0], raw.get(0x829A)[0]
        elif ifd0.get(0x8827) is not None:
            self.iso, self.exposure = ifd0.get(0x8827)[0], ifd0.get(0x829A)[0]
        else:
            self.iso, self.exposure = 0, (0, 0)
        self.bit_depth = raw.get(0x102)[0]
        dim = raw.get(0x828D)
        if dim[0] != 2 or dim[1] != 2:
            raise RawError("CFA repeat pattern other than 2x2")
        self.bayer = list(raw.get(0x828E))
        helper = [0, 1, 2]
        pc = raw.get(0xC616)
        if pc is not None:
            self.color_twist = np.zeros((3, 4), np.float32)
            if len(pc) != 3:
                raise RawError("CFAPlaneColor with other than three planes")
            for i, c in enumerate(pc):
                helper[i] = c
                if c > 2:
                    raise RawError("CFAPlaneColor with a colour other than red/green/blue")
                self.color_twist[c, i] = 1
                if c != i:
                    self.color_twist_is_identity = False
        wl = raw.get(0xC61D)
        bl = raw.get(0xC61A)
        black = [np.float32(0)] * 3
        if bl is not None:
            v = [np.float32(rational(x)) for x in bl]
            if len(v) == 1:
                black = [v[0]] * 3
            if len(v) == 3:
                for i in range(3):
                    black[helper[i]] = v[i]
            if len(v) == 4:
                black = _black_from_pattern(v, self.bayer)
        self.black_level = np.array(black, np.float32)
        white = np.float32(wl[0])
        self.white_level = np.array([white - b for b in black], np.float32)
        neutral = ifd0.get(0xC628) or raw.get(0xC628)
        if neutral is None:
            raise RawError("AsShotNeutral missing")
        nv = [rational(x) for x in neutral]
        mx = max(nv)
        self.white_balance = np.zeros(3, np.float32)
        for c in range(3):
            self.white_balance[helper[c]] = np.float32(1.0) / np.float32(nv[c] / mx)
        if raw.get(0x112) is not None:
            self.orientation = raw.get(0x112)[0]
        elif ifd0.get(0x112) is not None:
            self.orientation = ifd0.get(0x112)[0]
        else:
            self.orientation = 1
        cl, ct, cw, ch = 0, 0, self.width, self.height
        aa = raw.get(0xC68D)
        if aa is not None:
            top, left, bottom, right = (int(x) for x in aa[:4])
            cl += left
            ct += top
            cw, ch = right - left, bottom - top
            b = self.bayer
            if top % 2:
                b[0], b[1], b[2], b[3] = b[2], b[3], b[0], b[1]
            if left % 2:
                b[0], b[1], b[2], b[3] = b[1], b[0], b[3], b[2]
        co, cs = raw.get(0xC61F), raw.get(0xC620)
        if co is not None and cs is not None:
            left = int(rational(co[0]))
            top = int(rational(co[0]))
            width = int(rational(cs[0]))
            height = int(rational(cs[0]))
            cl += left
            ct += top
            cw, ch = width, height
        self.crop = (max(2, cl), max(2, ct), cw - max(0, (cl + cw) - (self.width - 2)),
                     ch - max(0, (ct + ch) - (self.height - 2)))
        noise = raw.get(0xC761) or ifd0.get(0xC761)
        if noise is not None:
            if len(noise) > 2:
                self.noise_alpha, self.noise_beta = np.float32(noise[helper[1] * 2]), np.float32(noise[helper[1] * 2 + 1])
            else:
                self.noise_alpha, self.noise_beta = np.float32(noise[0]), np.float32(noise[1])
        self.make, self.unique_model = ifd0.get(0x10F), ifd0.get(0xC614)
        self.compression = {7: "ljpeg", 1: "none"}.get(raw.get(0x103)[0], str(raw.get(0x103)[0]))
        # black deltas: R:PefFile/DNGFile.cs:318-372
        w, h = self.width, self.height
        bh, bv = np.zeros(w, np.uint16), np.zeros(h, np.uint16)
        off_v = off_h = bottom = right = 0
        if aa is not None:
            off_v, off_h, bottom, right = int(aa[0]), int(aa[1]), h - int(aa[2]), w - int(aa[3])
        dv, dh = raw.get(0xC61C), raw.get(0xC61B)
        if dv is not None:
            if len(dv) + off_v + bottom != h:
                raise RawError("BlackLevelDeltaV count")
            for i, x in enumerate(dv):
                bv[off_v + i] = _ushort_of_double(rational(x))
        if dh is not None:
            if len(dh) + off_h + right != w:
                raise RawError("BlackLevelDeltaH count")
            for i, x in enumerate(dh):
                bh[off_h + i] = _ushort_of_double(rational(x))
        self.black_delta_h, self.black_delta_v = bh, bv
        self.raw_ifd = raw
        if not decode:
            return
        out = np.zeros((h, w), np.uint16)
        mb = C.c_int(0)
        lin = self.lin
        lin_p, lin_n = (lin.ctypes.data, len(lin)) if lin is not None else (None, 0)
        comp = raw.get(0x103)[0]
        L = _lib()
        if raw.get(0x142) is not None:
            if comp != 7:
                raise RawError("tiles other than lossless JPEG")
            offs = np.asarray(raw.get(0x144), np.uint32)
            cnts = np.asarray(raw.get(0x145), np.uint32)
            rc = L.ora_dng_ljpeg(self.data, len(self.data), offs.ctypes.data, cnts.ctypes.data, len(offs), 1,
                                 raw.get(0x142)[0], raw.get(0x143)[0], 0, w, h, int(self.large), lin_p, lin_n,
                                 bh.ctypes.data, bv.ctypes.data, out.ctypes.data, C.byref(mb))
        elif raw.get(0x111) is not None:
            offs = np.asarray(raw.get(0x111), np.uint32)
            cnts = np.asarray(raw.get(0x117), np.uint32)
            rps = raw.get(0x116)[0]
            if comp == 7:
                rc = L.ora_dng_ljpeg(self.data, len(self.data), offs.ctypes.data, cnts.ctypes.data, len(offs), 0,
                                     0, 0, rps, w, h, int(self.large), lin_p, lin_n, bh.ctypes.data, bv.ctypes.data,
                                     out.ctypes.data, C.byref(mb))
            elif comp == 1:
                rc = L.ora_dng_uncompressed(self.data, len(self.data), offs.ctypes.data, cnts.ctypes.data, len(offs),
                                            rps, w, h, self.bit_depth, lin_p, lin_n, bh.ctypes.data, bv.ctypes.data,
                                            out.ctypes.data, C.byref(mb))
            else:
                raise RawError("strip compression")
        else:
            rc = 0
        if rc != 0:
            raise RawError("pixel data")
        self.white_level = (self.white_level - np.float32(mb.value)).astype(np.float32)
        self.max_black_delta = mb.value
        self.raw_image = out


def _lib():
    L = _O.lib()
    if not getattr(L, "_raw_declared", False):
        declare(L)
        L._raw_declared = True
    return L


def declare(L) -> None:
    P, I, S = C.c_void_p, C.c_int, C.c_size_t
    L.ora_pef_huffman.argtypes = [C.c_char_p, S, S, P, I, I, P]
    L.ora_pef_uncompressed.argtypes = [C.c_char_p, S, P, P, I, I, I, I, I, P]
    L.ora_dng_ljpeg.argtypes = [C.c_char_p, S, P, P, I, I, I, I, I, I, I, I, P, I, P, P, P, C.POINTER(C.c_int)]
    L.ora_dng_uncompressed.argtypes = [C.c_char_p, S, P, P, I, I, I, I, I, P, I, P, P, P, C.POINTER(C.c_int)]
    for f in ("ora_pef_huffman", "ora_pef_uncompressed", "ora_dng_ljpeg", "ora_dng_uncompressed"):
        getattr(L, f).restype = C.c_int
