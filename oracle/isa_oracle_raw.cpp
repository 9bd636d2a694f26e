// CPU oracle of the raw-file decoders -- TEST INFRASTRUCTURE ONLY.
//
// A sequential restatement of the reference's pixel decoding, kept to its reading order and integer
// semantics (C# unchecked int arithmetic, shift counts taken modulo 32, EOF behaviour of each reader):
//   PEF Pentax Huffman stream      R:PefFile/PEFFile.cs:35-130  (getbithuff, ljpeg_diff, the predictor loop)
//   PEF uncompressed strips        R:PefFile/PEFFile.cs:131-168
//   DNG lossless JPEG tiles/strips R:PefFile/DNGFile.cs:42-300  (jhead, make_decoder_ref, ljpeg_start,
//                                  ljpeg_row, getbithuff, ljpeg_diff), loops :372-511
//   DNG uncompressed strips        R:PefFile/DNGFile.cs:512-596
// Returns 0, or -1 where the reference throws (index out of range, end of stream, a decoded value out of
// range). The metadata (IFDs, maker notes, levels) is restated in oracle/raw.py.
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint32_t shl(uint32_t v, int n) { return v << (n & 31); }
inline uint32_t shr(uint32_t v, int n) { return v >> (n & 31); }
inline int32_t ishl1(int n) { return (int32_t)(1u << (n & 31)); }

// getbithuff of R:PefFile/PEFFile.cs:38-71 (pef = true: bytes straight from the file, end of file throws)
// and of R:PefFile/DNGFile.cs:235-268 (pef = false: a 0xFF byte is followed by a skipped byte, end of the
// tile's stream reads as 0xFFFFFFFF, whose low byte 0xFF enters the buffer without a skip).
struct Bits {
    const uint8_t* d;
    size_t n, pos;
    bool pef;
    uint32_t bitbuf = 0;
    int vbits = 0;
    bool err = false;

    int read_byte() {  // -1 at the end
        if (pos >= n) return -1;
        return d[pos++];
    }
    uint32_t get(int nbits, const uint16_t* huff) {
        if (nbits > 25) return 0;
        if (nbits < 0) {
            vbits = 0;
            bitbuf = 0;
            return 0;
        }
        if (nbits == 0 || vbits < 0) return 0;
        while (vbits < nbits) {
            int c = read_byte();
            if (pef) {
                if (c < 0) {
                    err = true;
                    return 0;
                }
            } else if (c == 255) {
                read_byte();
            }
            bitbuf = (bitbuf << 8) + (uint8_t)c;
            vbits += 8;
        }
        uint32_t c = shr(shl(bitbuf, 32 - vbits), 32 - nbits);
        if (huff) {
            vbits -= huff[c] >> 8;
            c = (uint8_t)huff[c];
        } else {
            vbits -= nbits;
        }
        return c;
    }
    // ljpeg_diff: R:PefFile/PEFFile.cs:73-81, R:PefFile/DNGFile.cs:270-281 (large: DNG version > 1.1)
    int32_t diff(const uint16_t* huff, bool large) {
        int len = (int)get(huff[0], huff + 1);
        if (large && len == 16) return -32768;
        int32_t d = (int32_t)get(len, nullptr);
        if ((d & ishl1(len - 1)) == 0) d = (int32_t)((uint32_t)d - (uint32_t)(ishl1(len) - 1));
        return d;
    }
};

}  // namespace

extern "C" {

// PEF Huffman stream: width x height differences from file[offset...], huff = the 4097-entry table of
// MNHuffmanTable (R:PefFile/PentaxMakerNotesEntry.cs:4555-4613).
int ora_pef_huffman(const uint8_t* file, size_t size, size_t offset, const uint16_t* huff, int width, int height,
                    uint16_t* out) {
    if (offset > size) return -1;
    Bits b{file, size, offset, true};
    b.get(-1, nullptr);
    int32_t vpred[2][2] = {{0, 0}, {0, 0}}, hpred[2] = {0, 0};
    for (int row = 0; row < height; ++row)
        for (int col = 0; col < width; ++col) {
            int32_t d = b.diff(huff, false);
            if (b.err) return -1;
            if (col < 2) {
                vpred[row & 1][col] = (int32_t)((uint32_t)vpred[row & 1][col] + (uint32_t)d);
                hpred[col] = vpred[row & 1][col];
            } else {
                hpred[col & 1] = (int32_t)((uint32_t)hpred[col & 1] + (uint32_t)d);
            }
            out[(size_t)row * width + col] = (uint16_t)hpred[col & 1];
        }
    return 0;
}

// PEF uncompressed strips: every strip holds rows_per_strip lines of width samples, MSB first, no padding.
int ora_pef_uncompressed(const uint8_t* file, size_t size, const uint32_t* offsets, const uint32_t* counts, int nstrips,
                         int rows_per_strip, int width, int height, int bits, uint16_t* out) {
    uint32_t bitbuf = 0;  // a member of the reader in the reference: carried over between strips
    for (int strip = 0; strip < nstrips; ++strip) {
        if ((size_t)offsets[strip] + counts[strip] > size) return -1;
        const uint8_t* data = file + offsets[strip];
        size_t len = counts[strip], pos = 0;
        int off = 0;
        for (int line = 0; line < rows_per_strip; ++line)
            for (int pixel = 0; pixel < width; ++pixel) {
                int row = strip * rows_per_strip + line;
                if (pos >= len) return -1;
                while (off < bits) {
                    if (pos >= len) return -1;
                    bitbuf = (bitbuf << 8) + data[pos];
                    off += 8;
                    ++pos;
                }
                uint32_t v = shr(shl(bitbuf, 32 - off), 32 - bits);
                off -= bits;
                if (row >= height) return -1;
                out[(size_t)row * width + pixel] = (uint16_t)v;
            }
    }
    return 0;
}

namespace {

struct JHead {
    int algo = 0, bits = 0, high = 0, wide = 0, clrs = 0, sraw = 0, psv = 0, restart = 0;
    int vpred[6] = {0, 0, 0, 0, 0, 0};
    std::vector<uint16_t> tab[20];
    const uint16_t* huff[20] = {};
    std::vector<uint16_t> row;
};

// make_decoder_ref: R:PefFile/DNGFile.cs:72-92
void make_decoder(const uint8_t*& src, std::vector<uint16_t>& huff) {
    const uint8_t* count = (src += 16) - 17;
    int max;
    for (max = 16; max > 0 && count[max] == 0; --max) {
    }
    huff.assign(1 + (1u << max), 0);
    huff[0] = (uint16_t)max;
    int h = 1;
    for (int len = 1; len <= max; ++len)
        for (int i = 0; i < count[len]; ++i, ++src)
            for (int j = 0; j < 1 << (max - len); ++j)
                if (h <= 1 << max) huff[h++] = (uint16_t)(len << 8 | *src);
}

// A MemoryStream over one tile / strip.
struct Stream {
    const uint8_t* d;
    size_t n, pos = 0;
    int read_byte() { return pos < n ? d[pos++] : -1; }
    size_t read(uint8_t* dst, size_t k) {
        size_t m = pos + k <= n ? k : n - pos;
        memcpy(dst, d + pos, m);
        pos += m;
        return m;
    }
};

// ljpeg_start: R:PefFile/DNGFile.cs:94-171. Returns 0 when the reference returns 0, -1 where it throws.
int ljpeg_start(JHead& jh, Stream& s) {
    std::vector<uint8_t> data(0x10000, 0);
    jh.restart = 0x7FFFFFFF;
    s.read_byte();
    if (s.read_byte() != 0xd8) return 0;
    int tag;
    do {
        if (s.read(data.data(), 4) != 4) return 0;
        tag = data[0] << 8 | data[1];
        int len = (data[2] << 8 | data[3]) - 2;
        if (tag <= 0xff00) return 0;
        if (len < 0) return -1;  // Stream.Read with a negative count throws
        s.read(data.data(), (size_t)len);
        switch (tag) {
            case 0xffc3:
                jh.sraw = ((data[7] >> 4) * (data[7] & 15) - 1) & 3;
                jh.algo = tag & 0xff;
                jh.bits = data[0];
                jh.high = data[1] << 8 | data[2];
                jh.wide = data[3] << 8 | data[4];
                jh.clrs = data[5] + jh.sraw;
                break;
            case 0xffc1:
            case 0xffc0:
                jh.algo = tag & 0xff;
                jh.bits = data[0];
                jh.high = data[1] << 8 | data[2];
                jh.wide = data[3] << 8 | data[4];
                jh.clrs = data[5] + jh.sraw;
                break;
            case 0xffc4: {
                const uint8_t* dp = data.data();
                int c;
                while (dp < data.data() + len && ((c = *dp++) & -20) == 0) {
                    make_decoder(dp, jh.tab[c]);
                    jh.huff[c] = jh.tab[c].data();
                }
                break;
            }
            case 0xffda:
                jh.psv = data[1 + data[0] * 2];
                jh.bits -= data[3 + data[0] * 2] & 15;
                break;
            case 0xffdd:
                jh.restart = data[0] << 8 | data[1];
                break;
        }
    } while (tag != 0xffda);
    for (int c = 0; c < 19; ++c)
        if (!jh.huff[c + 1]) jh.huff[c + 1] = jh.huff[c];
    jh.row.assign((size_t)jh.wide * jh.clrs * 2, 0);
    return 1;
}

// ljpeg_row: R:PefFile/DNGFile.cs:173-230. Returns the row, or nullptr where the reference returns null.
const uint16_t* ljpeg_row(int jrow, JHead& jh, Stream& s, Bits& b, bool large, bool& err) {
    int spred = 0;
    if (jh.restart == 0) {  // the reference's modulo by zero throws
        err = true;
        return nullptr;
    }
    if ((int)((uint32_t)jrow * (uint32_t)jh.wide % (uint32_t)jh.restart) == 0) {
        for (int c = 0; c < 6; ++c) jh.vpred[c] = ishl1(jh.bits - 1);
        if (jrow > 0) {
            s.pos = s.pos >= 2 ? s.pos - 2 : 0;
            uint16_t mark = 0;
            int c;
            do {
                c = s.read_byte();
                mark = (uint16_t)((mark << 8) + c);
            } while (s.pos + 1 < s.n && mark >> 4 != 0xffd);
        }
        b.pos = s.pos;
        b.get(-1, nullptr);
    }
    uint16_t* row0 = jh.row.data() + (size_t)jh.wide * jh.clrs * (jrow & 1);
    uint16_t* row1 = jh.row.data() + (size_t)jh.wide * jh.clrs * ((jrow + 1) & 1);
    uint16_t* ret = row0;
    for (int col = 0; col < jh.wide; ++col)
        for (int c = 0; c < jh.clrs; ++c) {
            if (!jh.huff[c]) {
                err = true;
                return nullptr;
            }
            b.pos = s.pos;
            int32_t diff = b.diff(jh.huff[c], large);
            s.pos = b.pos;
            int pred;
            if (jh.sraw > 0 && c <= jh.sraw && (col | c) > 0)
                pred = spred;
            else if (col > 0)
                pred = row0[-jh.clrs];
            else {
                jh.vpred[c] = (int32_t)((uint32_t)jh.vpred[c] + (uint32_t)diff);
                pred = (int32_t)((uint32_t)jh.vpred[c] - (uint32_t)diff);
            }
            if (jrow > 0 && col > 0) switch (jh.psv) {
                    case 1: break;
                    case 2: pred = row1[0]; break;
                    case 3: pred = row1[-jh.clrs]; break;
                    case 4: pred = pred + row1[0] - row1[-jh.clrs]; break;
                    case 5: pred = pred + ((row1[0] - row1[-jh.clrs]) >> 1); break;
                    case 6: pred = row1[0] + ((pred - row1[-jh.clrs]) >> 1); break;
                    case 7: pred = (pred + row1[0]) >> 1; break;
                    default: pred = 0; break;
                }
            *row0 = (uint16_t)(pred + diff);
            if (shr(*row0, jh.bits) != 0) return nullptr;
            if (c <= jh.sraw) spred = *row0;
            ++row0;
            ++row1;
        }
    return ret;
}

inline uint16_t linearize_black(int v, const uint16_t* lin, int lin_n, int black) {
    int x = lin ? (int)lin[v < lin_n ? v : lin_n - 1] : v;
    x -= black;
    return (uint16_t)(x > 0 ? x : 0);
}

}  // namespace

// DNG lossless JPEG segments. tiled = 1: segments are tiles in raster order (tile_w x tile_h); else strips of
// rows_per_strip rows. bh / bv: the per-column / per-row black deltas (width / height entries). large: the
// DNG version test of R:PefFile/DNGFile.cs:306-309. max_black_out: the largest bh + bv met.
int ora_dng_ljpeg(const uint8_t* file, size_t size, const uint32_t* offsets, const uint32_t* counts, int nseg, int tiled,
                  int tile_w, int tile_h, int rows_per_strip, int width, int height, int large, const uint16_t* lin,
                  int lin_n, const uint16_t* bh, const uint16_t* bv, uint16_t* out, int* max_black_out) {
    int max_black = 0;
    int row = 0, col = 0;
    for (int seg = 0; seg < nseg; ++seg) {
        if ((size_t)offsets[seg] + counts[seg] > size) return -1;
        Stream s{file + offsets[seg], counts[seg]};
        JHead jh;
        int ret = ljpeg_start(jh, s);
        if (ret < 0) return -1;
        if (ret > 0 && jh.algo == 0xc3) {
            Bits b{s.d, s.n, s.pos, false};
            int jwide = jh.wide * jh.clrs;
            for (int jrow = 0; jrow < jh.high; ++jrow) {
                bool err = false;
                const uint16_t* rp = ljpeg_row(jrow, jh, s, b, large != 0, err);
                if (!rp) return -1;
                for (int jcol = 0; jcol < jwide; ++jcol) {
                    int r, c;
                    if (tiled) {
                        r = row + jrow;
                        c = col + jcol;
                    } else {
                        r = seg * rows_per_strip + jrow;
                        c = jcol;
                    }
                    if (c < width && r < height) {
                        int black = bh[c] + bv[r];
                        if (black > max_black) max_black = black;
                        out[(size_t)r * width + c] = linearize_black(rp[jcol], lin, lin_n, black);
                    }
                }
            }
        }
        if (tiled) {
            col += tile_w;
            if (col >= width) {
                col = 0;
                row += tile_h;
            }
        }
    }
    *max_black_out = max_black;
    return 0;
}

// DNG uncompressed strips: 16-bit samples read as little-endian words (whatever the file's byte order), or
// 8..15-bit samples MSB first without row padding.
int ora_dng_uncompressed(const uint8_t* file, size_t size, const uint32_t* offsets, const uint32_t* counts, int nstrips,
                         int rows_per_strip, int width, int height, int bits, const uint16_t* lin, int lin_n,
                         const uint16_t* bh, const uint16_t* bv, uint16_t* out, int* max_black_out) {
    int max_black = 0;
    uint32_t bitbuf = 0;
    for (int strip = 0; strip < nstrips; ++strip) {
        if ((size_t)offsets[strip] + counts[strip] > size) return -1;
        const uint8_t* data = file + offsets[strip];
        size_t len = counts[strip];
        if (bits == 16) {
            for (size_t pixel = 0; pixel < len / 2; ++pixel) {
                size_t r = (size_t)strip * rows_per_strip, c = pixel;
                if (c >= (size_t)width) {
                    r += pixel / width;
                    c = pixel % width;
                }
                if (r >= (size_t)height) return -1;
                int black = bh[c] + bv[r];
                if (black > max_black) max_black = black;
                int v = data[2 * pixel] | data[2 * pixel + 1] << 8;
                out[r * width + c] = linearize_black(v, lin, lin_n, black);
            }
        } else if (bits < 16) {
            size_t pos = 0;
            int off = 0;
            for (int line = 0; line < rows_per_strip; ++line)
                for (int pixel = 0; pixel < width; ++pixel) {
                    int r = strip * rows_per_strip + line;
                    if (pos >= len) return -1;
                    while (off < bits) {
                        if (pos >= len) return -1;
                        bitbuf = (bitbuf << 8) + data[pos];
                        off += 8;
                        ++pos;
                    }
                    uint32_t v = shr(shl(bitbuf, 32 - off), 32 - bits);
                    off -= bits;
                    if (r >= height) return -1;
                    int black = bh[pixel] + bv[r];
                    if (black > max_black) max_black = black;
                    out[(size_t)r * width + pixel] = linearize_black((int)v, lin, lin_n, black);
                }
        }
    }
    *max_black_out = max_black;
    return 0;
}

}  // extern "C"
