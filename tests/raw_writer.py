"""Writer of PEF- and DNG-style raw files for the ingest tests (test infrastructure only).

It is written from the published formats, not from the reader under test:
- TIFF 6.0 (byte order, IFD chains, field types, strips and tiles, SubIFDs, the Exif IFD);
- DNG 1.4 (CFA tags, BlackLevel / BlackLevelDeltaH/V, WhiteLevel, LinearizationTable, ActiveArea,
  DefaultCrop*, AsShotNeutral, NoiseProfile, DNGPrivateData);
- ITU-T T.81 Annex H (lossless JPEG: SOF3, DHT, DRI, SOS, predictors 1-7, byte stuffing, RSTn markers,
  the difference categories and their extra bits) with canonical Huffman codes (Annex C);
- the Pentax compressed raw layout as the readers consume it: one strip of Huffman-coded differences,
  category code + extra bits, predictors from the same-parity neighbour to the left (columns >= 2) or
  two rows above (columns 0 and 1), and the code table in maker-note tag 0x0220.

Each encoder takes a known uint16 image and returns file bytes, so a reader that decodes them back to the
same image is checked against an independent statement of the format (a round trip).
"""
from __future__ import annotations

import heapq
import struct

import numpy as np

BYTE, ASCII, SHORT, LONG, RATIONAL, SBYTE, UNDEFINED, SSHORT, SLONG, SRATIONAL, FLOAT, DOUBLE = range(1, 13)
_SIZE = {BYTE: 1, ASCII: 1, SHORT: 2, LONG: 4, RATIONAL: 8, SBYTE: 1, UNDEFINED: 1, SSHORT: 2, SLONG: 4,
         SRATIONAL: 8, FLOAT: 4, DOUBLE: 8}


class TiffWriter:
    """Appends blobs and IFDs to a growing buffer; IFDs are written after the data they point to."""

    def __init__(self, order: str = "<"):
        self.o = order
        self.buf = bytearray((b"II*\x00" if order == "<" else b"MM\x00*") + b"\x00\x00\x00\x00")

    def align(self, a: int = 2) -> None:
        while len(self.buf) % a:
            self.buf.append(0)

    def blob(self, data: bytes, align: int = 2) -> int:
        self.align(align)
        off = len(self.buf)
        self.buf += data
        return off

    def pack(self, typ: int, vals) -> bytes:
        o = self.o
        if typ == ASCII:
            s = vals.encode() if isinstance(vals, str) else bytes(vals)
            return s + b"\x00"
        if typ in (BYTE, UNDEFINED, SBYTE):
            return bytes(vals) if isinstance(vals, (bytes, bytearray)) else struct.pack(
                f"{len(vals)}{'b' if typ == SBYTE else 'B'}", *vals)
        if typ in (RATIONAL, SRATIONAL):
            f = "I" if typ == RATIONAL else "i"
            return b"".join(struct.pack(o + f + f, int(n), int(d)) for n, d in vals)
        f = {SHORT: "H", LONG: "I", SSHORT: "h", SLONG: "i", FLOAT: "f", DOUBLE: "d"}[typ]
        return struct.pack(f"{o}{len(vals)}{f}", *vals)

    def ifd(self, entries, next_off: int = 0) -> int:
        """entries: (tag, type, values). Values longer than 4 bytes go to a blob written first."""
        packed = []
        for tag, typ, vals in sorted(entries, key=lambda e: e[0]):
            data = self.pack(typ, vals)
            count = len(data) // _SIZE[typ]
            if len(data) > 4:
                field = struct.pack(self.o + "I", self.blob(data))
            else:
                field = data + b"\x00" * (4 - len(data))
            packed.append(struct.pack(self.o + "HHI", tag, typ, count) + field)
        self.align(2)
        off = len(self.buf)
        self.buf += struct.pack(self.o + "H", len(packed)) + b"".join(packed) + struct.pack(self.o + "I", next_off)
        return off

    def set_first_ifd(self, off: int) -> None:
        self.buf[4:8] = struct.pack(self.o + "I", off)

    def set_next_ifd(self, ifd_off: int, next_off: int) -> None:
        n = struct.unpack_from(self.o + "H", self.buf, ifd_off)[0]
        struct.pack_into(self.o + "I", self.buf, ifd_off + 2 + 12 * n, next_off)

    def bytes(self) -> bytes:
        return bytes(self.buf)


# ---------------------------------------------------------------------------------------------------------------
# Huffman codes (canonical, T.81 Annex C), with lengths limited as in T.81 Annex K.2.

def code_lengths(freq, limit: int) -> dict:
    """Huffman code lengths for the symbols with non-zero frequency, no longer than `limit`."""
    syms = [s for s in sorted(freq) if freq[s] > 0]
    if len(syms) == 1:
        return {syms[0]: 1}
    heap = [(freq[s], i, [s]) for i, s in enumerate(syms)]
    heapq.heapify(heap)
    depth = {s: 0 for s in syms}
    k = len(syms)
    while len(heap) > 1:
        f1, _, a = heapq.heappop(heap)
        f2, _, b = heapq.heappop(heap)
        for s in a + b:
            depth[s] += 1
        heapq.heappush(heap, (f1 + f2, k, a + b))
        k += 1
    bits = [0] * 64
    for s in syms:
        bits[depth[s]] += 1
    # Annex K.2 adjustment: move the longest codes up until none exceeds the limit.
    i = max(depth.values())
    while i > limit:
        while bits[i] > 0:
            j = i - 2
            while bits[j] == 0:
                j -= 1
            bits[i] -= 2
            bits[i - 1] += 1
            bits[j + 1] += 2
            bits[j] -= 1
        i -= 1
    order = sorted(syms, key=lambda s: (depth[s], s))
    out, n = {}, 1
    for s in order:
        while bits[n] == 0:
            n += 1
        out[s] = n
        bits[n] -= 1
    return out


def canonical(lengths: dict):
    """{symbol: (code, length)} in canonical order (by length, then symbol)."""
    codes, code, prev = {}, 0, 0
    for s in sorted(lengths, key=lambda s: (lengths[s], s)):
        n = lengths[s]
        code <<= n - prev
        prev = n
        codes[s] = (code, n)
        code += 1
    return codes


def pack_fields(values: np.ndarray, nbits: np.ndarray) -> bytes:
    """Concatenates fields (value, nbits <= 32) MSB-first into bytes; the last byte is padded with zeros."""
    values = np.asarray(values, np.uint64)
    nbits = np.asarray(nbits, np.int64)
    ends = np.cumsum(nbits)
    total = int(ends[-1]) if ends.size else 0
    nwords = (total + 31) // 32 + 1
    start = ends - nbits
    w = start >> 5
    o = start & 31
    # field bits sit at [o, o + n) of the 64-bit window starting at word w (MSB first)
    sh = 64 - o - nbits
    val = (values & ((np.uint64(1) << nbits.astype(np.uint64)) - np.uint64(1))) << sh.astype(np.uint64)
    hi = (val >> np.uint64(32)).astype(np.float64)
    lo = (val & np.uint64(0xFFFFFFFF)).astype(np.float64)
    words = np.bincount(w, weights=hi, minlength=nwords) + np.bincount(w + 1, weights=lo, minlength=nwords + 1)[:nwords]
    words = words.astype(np.uint64).astype(">u4")
    return words.tobytes()[: (total + 7) // 8]


def stuff_ff(data: bytes) -> bytes:
    """Inserts a 0x00 after every 0xFF (T.81 F.1.2.3)."""
    a = np.frombuffer(data, np.uint8)
    ff = np.nonzero(a == 0xFF)[0]
    return np.insert(a, ff + 1, 0).tobytes() if ff.size else data


def category(diff: int) -> int:
    return 0 if diff == 0 else int(abs(diff)).bit_length()


def categories(d: np.ndarray) -> np.ndarray:
    a = np.abs(np.asarray(d, np.int64))
    c = np.zeros(a.shape, np.int64)
    nz = a > 0
    c[nz] = np.floor(np.log2(a[nz].astype(np.float64))).astype(np.int64) + 1
    return c


def extra_bits(d: np.ndarray, ssss: np.ndarray) -> np.ndarray:
    d = np.asarray(d, np.int64)
    return np.where(d >= 0, d, d + (np.int64(1) << ssss) - 1)


# ---------------------------------------------------------------------------------------------------------------
# Pentax compressed raw

def pentax_table(lengths: dict, order: str) -> bytes:
    """Maker-note tag 0x0220: a 16-bit depth field (depth - 12 mod 16), 12 reserved bytes, depth 16-bit codes
    left-aligned to 12 bits, then depth 8-bit code lengths; symbol c = category c."""
    dep = max(lengths) + 1
    codes = canonical(lengths)
    out = struct.pack(order + "H", (dep - 12) & 15) + b"\x00" * 12
    out += b"".join(struct.pack(order + "H", codes[c][0] << (12 - codes[c][1]) if c in codes else 0) for c in range(dep))
    out += bytes(codes[c][1] if c in codes else 12 for c in range(dep))
    return out


def pentax_diffs(img: np.ndarray) -> np.ndarray:
    """Differences to the Pentax predictors: column 0/1 from the same column two rows above (0 in rows 0/1),
    other columns from the pixel two columns to the left."""
    v = img.astype(np.int64)
    d = np.empty_like(v)
    d[:, 2:] = v[:, 2:] - v[:, :-2]
    d[:, :2] = v[:, :2]
    d[2:, :2] = v[2:, :2] - v[:-2, :2]
    return d


def pentax_encode(img: np.ndarray, lengths: dict | None = None):
    """Returns (stream bytes, code lengths) for a <= 14-bit image."""
    d = pentax_diffs(img).ravel()
    cats = categories(d)
    if cats.max() > 14:
        raise ValueError("Pentax tables hold categories 0..14")
    if lengths is None:
        freq = {c: max(int((cats == c).sum()), 1) for c in range(int(cats.max()) + 1)}
        lengths = code_lengths(freq, 12)
    codes = canonical(lengths)
    cval = np.array([codes[c][0] if c in codes else 0 for c in range(15)], np.int64)
    clen = np.array([codes[c][1] if c in codes else 0 for c in range(15)], np.int64)
    vals = (cval[cats] << cats) | extra_bits(d, cats)
    return pack_fields(vals, clen[cats] + cats), lengths


def pack_bits(img: np.ndarray, bits: int) -> bytes:
    """Row-major samples as one MSB-first bit stream (no row padding), as the readers consume it."""
    v = img.astype(np.uint64).ravel()
    n = v.size * bits
    out = np.zeros((n + 7) // 8 + 8, np.uint8)
    pos = np.arange(v.size, dtype=np.uint64) * np.uint64(bits)
    for b in range(bits):
        bit = (v >> np.uint64(bits - 1 - b)) & np.uint64(1)
        p = pos + np.uint64(b)
        np.bitwise_or.at(out, (p >> np.uint64(3)).astype(np.int64),
                         (bit << (np.uint64(7) - (p & np.uint64(7)))).astype(np.uint8))
    return bytes(out[: (n + 7) // 8])


def maker_note_entries(order, start, entries, header: str = "AOC"):
    """A Pentax maker note placed at file offset `start`: "AOC\\0" + byte order (offsets absolute in the file)
    or "PENTAX \\0" + byte order (offsets relative to the maker note)."""
    hdr = (b"AOC\x00" if header == "AOC" else b"PENTAX \x00") + (b"II" if order == "<" else b"MM")
    tmp = TiffWriter(order)
    n = len(entries)
    fixed = len(hdr) + 2 + 12 * n + 4
    data = bytearray()
    recs = []
    base = start if header == "AOC" else 0
    for tag, typ, vals in sorted(entries, key=lambda e: e[0]):
        payload = tmp.pack(typ, vals)
        count = len(payload) // _SIZE[typ]
        if len(payload) > 4:
            if len(data) % 2:
                data.append(0)
            field = struct.pack(order + "I", base + fixed + len(data))
            data += payload
        else:
            field = payload + b"\x00" * (4 - len(payload))
        recs.append(struct.pack(order + "HHI", tag, typ, count) + field)
    return hdr + struct.pack(order + "H", n) + b"".join(recs) + b"\x00\x00\x00\x00" + bytes(data)


def write_pef(img: np.ndarray, *, order="<", bits=14, compressed=True, black=(512,), white_level=None,
              white_point=(8704, 8192, 8192, 11520), scaling=8192, roll=None, cfa=(0, 1, 1, 2), iso=200,
              exposure=(1, 125), make="PENTAX", model="PENTAX K-3", orientation=1, date="2020:05:06 07:08:09",
              rows_per_strip=None, mn_header="AOC", lengths=None) -> bytes:
    """A PEF-style file: IFD0 with one strip, Exif IFD with CFA pattern and the Pentax maker note."""
    h, w = img.shape
    tw = TiffWriter(order)
    mn = [(0x0201, SHORT, list(white_point)), (0x003D, SHORT, [scaling])]
    if black is not None:
        mn.append((0x0200, SHORT, list(black)))
    if white_level is not None:
        mn.append((0x007E, LONG, [white_level]))
    if roll is not None:
        mn.append((0x022B, SBYTE, [1, int(round(-2 * roll)), 0, 0, 0, 0, 0, 0]))
    strips = []
    if compressed:
        stream, lengths = pentax_encode(img, lengths)
        mn.append((0x0220, UNDEFINED, pentax_table(lengths, order)))
        strips.append(stream)
        rps = h
    else:
        rps = rows_per_strip or h
        for r0 in range(0, h, rps):
            strips.append(pack_bits(img[r0:r0 + rps], bits))
    offs = [tw.blob(s + b"\x00" * 16) for s in strips]  # slack after each strip: readers look ahead
    tw.align(4)
    mn_start = len(tw.buf) + 0
    note = maker_note_entries(order, mn_start, mn, mn_header)
    assert tw.blob(note, 4) == mn_start
    exif = tw.ifd([(0x8827, SHORT, [iso]), (0x829A, RATIONAL, [exposure]), (0x9004, ASCII, date),
                   (0xA302, UNDEFINED, bytes([0, 2, 0, 2, *cfa])),
                   (0x927C, UNDEFINED, b"")])
    # patch the maker note entry (UNDEFINED, count = len(note), offset = mn_start)
    n = struct.unpack_from(order + "H", tw.buf, exif)[0]
    for k in range(n):
        e = exif + 2 + 12 * k
        if struct.unpack_from(order + "H", tw.buf, e)[0] == 0x927C:
            struct.pack_into(order + "HHII", tw.buf, e, 0x927C, UNDEFINED, len(note), mn_start)
    ifd0 = tw.ifd([(0x100, LONG, [w]), (0x101, LONG, [h]), (0x102, SHORT, [bits]),
                   (0x103, SHORT, [65535 if compressed else 1]), (0x106, SHORT, [32803]),
                   (0x10F, ASCII, make), (0x110, ASCII, model), (0x111, LONG, offs),
                   (0x112, SHORT, [orientation]), (0x116, LONG, [rps]),
                   (0x117, LONG, [len(s) for s in strips]), (0x8769, LONG, [exif])])
    tw.set_first_ifd(ifd0)
    return tw.bytes()


# ---------------------------------------------------------------------------------------------------------------
# Lossless JPEG (T.81 Annex H)

def ljpeg_predict(psv: int, ra: int, rb: int, rc: int) -> int:
    return {1: ra, 2: rb, 3: rc, 4: ra + rb - rc, 5: ra + ((rb - rc) >> 1), 6: rb + ((ra - rc) >> 1),
            7: (ra + rb) >> 1}[psv]


def ljpeg_diffs(v: np.ndarray, precision: int, psv: int, restart_rows: int) -> np.ndarray:
    """Differences (T.81 H.1.2.1, modulo 2^16 into [-32767, 32768]) of v: (rows, wide, comps)."""
    rows, wide, comps = v.shape
    v = v.astype(np.int64)
    p = np.empty_like(v)
    p[:, 1:] = v[:, :-1]                       # first line (and every line for psv 1): Ra
    p[1:, 0] = v[:-1, 0]                       # first column: Rb
    if psv != 1:
        ra, rb, rc = v[1:, :-1], v[:-1, 1:], v[:-1, :-1]
        p[1:, 1:] = {2: rb, 3: rc, 4: ra + rb - rc, 5: ra + ((rb - rc) >> 1), 6: rb + ((ra - rc) >> 1),
                     7: (ra + rb) >> 1}[psv]
    starts = [0] + (list(range(restart_rows, rows, restart_rows)) if restart_rows else [])
    for r in starts:                           # first line after a restart: Ra, first sample 2^(P-1)
        p[r, 1:] = v[r, :-1]
        p[r, 0] = 1 << (precision - 1)
    d = (v - p) & 0xFFFF
    return np.where(d > 0x8000, d - 0x10000, d)


def ljpeg_encode(samples: np.ndarray, precision: int, comps: int, psv: int = 1, restart_rows: int = 0,
                 tables: str = "per_component") -> bytes:
    """samples: (rows, cols * comps) uint16, component c of MCU x at column x * comps + c. One scan, all
    components interleaved, Huffman tables fitted to the data (one per component, or one shared)."""
    rows, wide_c = samples.shape
    wide = wide_c // comps
    diffs = ljpeg_diffs(samples.reshape(rows, wide, comps), precision, psv, restart_rows)
    cats = np.where(diffs == 32768, 16, categories(diffs))
    ntab = comps if tables == "per_component" else 1
    tabs = []
    for c in range(ntab):
        sel = cats[:, :, c] if ntab > 1 else cats
        cnt = np.bincount(sel.ravel(), minlength=17)
        tabs.append(canonical(code_lengths({s: int(cnt[s]) for s in range(17) if cnt[s]}, 16)))
    out = bytearray(b"\xFF\xD8")
    out += b"\xFF\xC3" + struct.pack(">HBHHB", 8 + 3 * comps, precision, rows, wide, comps)
    out += b"".join(bytes([c + 1, 0x11, 0]) for c in range(comps))
    for th, codes in enumerate(tabs):
        counts = [0] * 16
        for s, (code, n) in codes.items():
            counts[n - 1] += 1
        syms = sorted(codes, key=lambda s: (codes[s][1], s))
        body = bytes([th]) + bytes(counts) + bytes(syms)
        out += b"\xFF\xC4" + struct.pack(">H", 2 + len(body)) + body
    if restart_rows:
        out += b"\xFF\xDD" + struct.pack(">HH", 4, restart_rows * wide)
    sos = bytes([comps]) + b"".join(bytes([c + 1, (c if ntab > 1 else 0) << 4]) for c in range(comps))
    sos += bytes([psv, 0, 0])
    out += b"\xFF\xDA" + struct.pack(">H", 2 + len(sos)) + sos
    cval = np.zeros((ntab, 17), np.int64)
    clen = np.zeros((ntab, 17), np.int64)
    for t, codes in enumerate(tabs):
        for s, (code, n) in codes.items():
            cval[t, s], clen[t, s] = code, n
    tsel = (np.arange(comps) if ntab > 1 else np.zeros(comps, np.int64))[None, None, :]
    extra = np.where(cats == 16, 0, cats)
    vals = (cval[tsel, cats] << extra) | np.where(cats == 16, 0, extra_bits(diffs, extra))
    lens = clen[tsel, cats] + extra
    step = restart_rows or rows
    for k, r0 in enumerate(range(0, rows, step)):
        if k:
            out += bytes([0xFF, 0xD0 + ((k - 1) & 7)])
        seg_v, seg_l = vals[r0:r0 + step].ravel(), lens[r0:r0 + step].ravel()
        data = pack_fields(seg_v, seg_l)
        pad = (-int(seg_l.sum())) % 8
        if pad:  # fill the last byte with 1-bits (T.81 F.1.2.3)
            data = data[:-1] + bytes([data[-1] | ((1 << pad) - 1)])
        out += stuff_ff(data)
    out += b"\xFF\xD9"
    return bytes(out)


def dng_tiles(img, tw_, th_, precision, comps, psv, restart_rows=0, tables="per_component"):
    """Lossless JPEG tiles in raster order; edge tiles padded by repeating the last column / row."""
    h, w = img.shape
    out = []
    for r0 in range(0, h, th_):
        for c0 in range(0, w, tw_):
            t = img[r0:r0 + th_, c0:c0 + tw_]
            t = np.pad(t, ((0, th_ - t.shape[0]), (0, tw_ - t.shape[1])), mode="edge")
            out.append(ljpeg_encode(t, precision, comps, psv, restart_rows, tables))
    return out


def write_dng(img: np.ndarray, *, order="<", bits=14, layout="tiles", compression=7, tile=(64, 32), comps=2,
              psv=1, restart_rows=0, tables="per_component", rows_per_strip=None, version=(1, 4, 0, 0),
              cfa=(0, 1, 1, 2), plane_color=None, black=(256,), black_type=SHORT, white=16383,
              linearization=None, delta_h=None, delta_v=None, active_area=None, crop=None,
              neutral=((1, 2), (1, 1), (2, 3)), noise=None, iso=100, exposure=(1, 60), exif_iso=True,
              make="RICOH IMAGING COMPANY, LTD.", unique_model="PENTAX K-3", orientation=1, sub_ifd=True,
              date="2021:02:03 04:05:06") -> bytes:
    """A DNG-style file. With sub_ifd the raw image is in a SubIFD of a small preview IFD0, else in IFD0."""
    h, w = img.shape
    tw = TiffWriter(order)
    raw = [(0x100, LONG, [w]), (0x101, LONG, [h]), (0x102, SHORT, [bits]), (0x103, SHORT, [compression]),
           (0x106, SHORT, [32803]), (0x115, SHORT, [1]), (0x828D, SHORT, [2, 2]), (0x828E, BYTE, list(cfa)),
           (0xC61D, LONG if white > 65535 else SHORT, [white])]
    if black is not None:
        raw.append((0xC61A, black_type, [(b, 1) for b in black] if black_type == RATIONAL else list(black)))
    if plane_color is not None:
        raw.append((0xC616, BYTE, list(plane_color)))
    if linearization is not None:
        raw.append((0xC618, SHORT, [int(x) for x in linearization]))
    if delta_h is not None:
        raw.append((0xC61B, SRATIONAL, [(int(x), 1) for x in delta_h]))
    if delta_v is not None:
        raw.append((0xC61C, SRATIONAL, [(int(x), 1) for x in delta_v]))
    if active_area is not None:
        raw.append((0xC68D, LONG, list(active_area)))
    if crop is not None:
        raw.append((0xC61F, RATIONAL, [(crop[0], 1), (crop[1], 1)]))
        raw.append((0xC620, RATIONAL, [(crop[2], 1), (crop[3], 1)]))
    if noise is not None:
        raw.append((0xC761, DOUBLE, list(noise)))
    if not exif_iso:
        raw.append((0x8827, SHORT, [iso]))
        raw.append((0x829A, RATIONAL, [exposure]))
    if layout == "tiles":
        segs = dng_tiles(img, tile[0], tile[1], bits, comps, psv, restart_rows, tables)
        offs = [tw.blob(s) for s in segs]
        raw += [(0x142, LONG, [tile[0]]), (0x143, LONG, [tile[1]]), (0x144, LONG, offs),
                (0x145, LONG, [len(s) for s in segs])]
    else:
        rps = rows_per_strip or h
        segs = []
        for r0 in range(0, h, rps):
            part = img[r0:r0 + rps]
            if compression == 7:
                segs.append(ljpeg_encode(part, bits, comps, psv, restart_rows, tables))
            elif bits == 16:
                segs.append(part.astype("<u2").tobytes())
            else:
                segs.append(pack_bits(part, bits))
        offs = [tw.blob(s) for s in segs]
        raw += [(0x111, LONG, offs), (0x116, LONG, [rps]), (0x117, LONG, [len(s) for s in segs])]
    ifd0_tags = [(0xC612, BYTE, list(version)) if version else None, (0x10F, ASCII, make),
                 (0xC614, ASCII, unique_model), (0x112, SHORT, [orientation]), (0x132, ASCII, date),
                 (0xC621, SRATIONAL, [(10000, 10000), (0, 1), (0, 1), (0, 1), (10000, 10000), (0, 1), (0, 1),
                                      (0, 1), (10000, 10000)]),
                 (0xC65A, SHORT, [21]), (0xC628, RATIONAL, list(neutral))]
    ifd0_tags = [t for t in ifd0_tags if t is not None]
    if exif_iso:
        exif = tw.ifd([(0x8827, SHORT, [iso]), (0x829A, RATIONAL, [exposure]), (0x9004, ASCII, date)])
        ifd0_tags.append((0x8769, LONG, [exif]))
    if sub_ifd:
        raw_off = tw.ifd(raw + [(0xFE, LONG, [0])])
        prev = tw.blob(bytes(8 * 8 * 3))
        ifd0 = tw.ifd(ifd0_tags + [(0xFE, LONG, [1]), (0x100, LONG, [8]), (0x101, LONG, [8]),
                                   (0x102, SHORT, [8, 8, 8]), (0x103, SHORT, [1]), (0x106, SHORT, [2]),
                                   (0x111, LONG, [prev]), (0x115, SHORT, [3]), (0x116, LONG, [8]),
                                   (0x117, LONG, [192]), (0x14A, LONG, [raw_off])])
    else:
        ifd0 = tw.ifd(ifd0_tags + raw)
    tw.set_first_ifd(ifd0)
    return tw.bytes()
