// Raw ingest: PEF / DNG container parsing on the host, pixel decoding on the device.
//
// Reference: RawFile (R:PefFile/RawFile.cs:42-90), PEFFile (R:PefFile/PEFFile.cs:28-316), DNGFile
// (R:PefFile/DNGFile.cs:36-950), PentaxMakerNotes (R:PefFile/PentaxMakerNotes.cs:42-199) and the Huffman table of
// MNHuffmanTable (R:PefFile/PentaxMakerNotesEntry.cs:4555-4613). The reference decodes one file per host thread,
// symbol after symbol. Here every entropy-coded stream (a PEF strip, a lossless-JPEG tile, strip or restart
// interval) is one work-group: its bits are cut into up to 1024 chunks, every chunk is decoded from a guessed
// start for each possible component phase, one lane chains the chunks' exits to the true starts (re-decoding
// chunks whose start was wrong until the chain is consistent: prefix codes resynchronise within a few symbols),
// and the chunks then write their differences in parallel. The predictors are undone by scans (row starts down
// the columns, then along each row); lossless-JPEG predictors 2-7 by one lane per tile component.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "isa_host.h"

namespace {

// =====================================================================================================
// Host: TIFF container
// =====================================================================================================

struct DataError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Reader {
    const uint8_t* d;
    size_t n;
    bool le;
    void need(size_t off, size_t k) const {
        if (off > n || k > n - off) throw DataError("read beyond the end of the file");
    }
    uint8_t u8(size_t o) const {
        need(o, 1);
        return d[o];
    }
    uint16_t u16(size_t o) const {
        need(o, 2);
        return le ? (uint16_t)(d[o] | d[o + 1] << 8) : (uint16_t)(d[o] << 8 | d[o + 1]);
    }
    uint32_t u32(size_t o) const {
        need(o, 4);
        return le ? ((uint32_t)d[o] | (uint32_t)d[o + 1] << 8 | (uint32_t)d[o + 2] << 16 | (uint32_t)d[o + 3] << 24)
                  : ((uint32_t)d[o] << 24 | (uint32_t)d[o + 1] << 16 | (uint32_t)d[o + 2] << 8 | (uint32_t)d[o + 3]);
    }
    uint64_t u64(size_t o) const {
        uint64_t a = u32(o), b = u32(o + 4);
        return le ? (b << 32 | a) : (a << 32 | b);
    }
};

int type_size(int t) {
    static const int s[13] = {1, 1, 1, 2, 4, 8, 1, 1, 2, 4, 8, 4, 8};
    return t >= 1 && t <= 12 ? s[t] : 1;
}

struct Tag {
    uint16_t tag = 0, type = 0;
    uint32_t count = 0;
    uint32_t field = 0;  // the 4-byte value/offset field as read
    size_t voff = 0;     // absolute offset of the value bytes in the buffer that holds them
};

// Directory entries (first entry of a tag wins, as GetEntry returns the first match).
struct Dir {
    std::vector<Tag> tags;
    std::vector<Dir> sub;
    std::vector<Dir> exif;
    uint32_t next = 0;
    const Tag* find(uint16_t t) const {
        for (const Tag& x : tags)
            if (x.tag == t) return &x;
        return nullptr;
    }
};

Dir read_dir(const Reader& r, size_t off, long base, int depth) {
    if (depth > 8) throw DataError("IFD nesting too deep");
    Dir dir;
    uint16_t n = r.u16(off);
    for (int k = 0; k < n; ++k) {
        size_t e = off + 2 + 12 * (size_t)k;
        Tag t;
        t.tag = r.u16(e);
        t.type = r.u16(e + 2);
        t.count = r.u32(e + 4);
        t.field = r.u32(e + 8);
        size_t bytes = (size_t)type_size(t.type) * t.count;
        t.voff = bytes <= 4 ? e + 8 : (size_t)((long)t.field - base);
        r.need(t.voff, bytes);
        dir.tags.push_back(t);
    }
    dir.next = r.u32(off + 2 + 12 * (size_t)n);
    return dir;
}

// Number k of an entry as a double (integers, rationals as numerator / denominator, floats).
double num(const Reader& r, const Tag& t, uint32_t k) {
    if (k >= t.count) throw DataError("tag value index out of range");
    size_t o = t.voff + (size_t)type_size(t.type) * k;
    switch (t.type) {
        case 1: case 7: return r.u8(o);
        case 6: return (int8_t)r.u8(o);
        case 3: return r.u16(o);
        case 8: return (int16_t)r.u16(o);
        case 4: return r.u32(o);
        case 9: return (int32_t)r.u32(o);
        case 5: return (double)r.u32(o) / (double)r.u32(o + 4);
        case 10: return (double)(int32_t)r.u32(o) / (double)(int32_t)r.u32(o + 4);
        case 11: {
            uint32_t b = r.u32(o);
            float f;
            memcpy(&f, &b, 4);
            return f;
        }
        case 12: {
            uint64_t b = r.u64(o);
            double f;
            memcpy(&f, &b, 8);
            return f;
        }
    }
    return r.u8(o);
}
uint32_t unum(const Reader& r, const Tag& t, uint32_t k) { return (uint32_t)(int64_t)num(r, t, k); }
std::string ascii(const Reader& r, const Tag& t) {
    r.need(t.voff, t.count);
    std::string s((const char*)r.d + t.voff, t.count);
    size_t z = s.find('\0');
    return z == std::string::npos ? s : s.substr(0, z);
}

Dir read_ifd(const Reader& r, size_t off, int depth = 0) {
    Dir d = read_dir(r, off, 0, depth);
    if (const Tag* s = d.find(0x14A))  // SubIFDs
        for (uint32_t k = 0; k < s->count; ++k) d.sub.push_back(read_ifd(r, unum(r, *s, k), depth + 1));
    if (const Tag* e = d.find(0x8769)) d.exif.push_back(read_ifd(r, unum(r, *e, 0), depth + 1));
    return d;
}

// A Pentax maker note (R:PefFile/PentaxMakerNotes.cs:42-199): its bytes, byte order and entries.
struct MakerNote {
    std::vector<uint8_t> bytes;
    bool le = true;
    Dir dir;
    bool valid = false;
};

MakerNote read_maker_note(const uint8_t* p, size_t n, long offset) {
    MakerNote m;
    m.bytes.assign(p, p + n);
    const uint8_t* b = m.bytes.data();
    auto order = [&](size_t at) {
        if (at + 2 > n) throw DataError("maker note too short");
        if (b[at] == 'I' && b[at + 1] == 'I') return true;
        if (b[at] == 'M' && b[at + 1] == 'M') return false;
        throw DataError("Could not determine file endianess for maker notes");
    };
    long add = 0;
    size_t pos;
    if (n >= 4 && !memcmp(b, "AOC\0", 4)) {
        m.le = order(4);
        add = offset;
        pos = 6;
    } else if (n >= 8 && !memcmp(b, "PENTAX \0", 8)) {
        m.le = order(8);
        pos = 10;
    } else if (n >= 10 && !memcmp(b, "Adobe\0MakN", 10)) {
        m.le = order(14);
        Reader rr{b, n, m.le};
        add = (long)rr.u32(16);
        if (n >= 24 && !memcmp(b + 20, "AOC\0", 4)) {
            m.le = order(24);
            add -= 20;
            pos = 26;
        } else if (n >= 28 && !memcmp(b + 20, "PENTAX \0", 8)) {
            m.le = order(28);
            add = -20;
            pos = 30;
        } else {
            pos = 20;
        }
    } else {
        return m;  // not a Pentax maker note: no entries
    }
    Reader rr{b, n, m.le};
    m.dir = read_dir(rr, pos, add, 0);
    m.valid = true;
    return m;
}

// MNHuffmanTable: 4097 entries, entry 0 = 12 lookahead bits, entry 1 + c = (code length << 8) | category.
std::vector<uint16_t> pentax_table(const MakerNote& m, const Tag& t) {
    Reader r{m.bytes.data(), m.bytes.size(), m.le};
    r.need(t.voff, t.count);
    Reader v{m.bytes.data() + t.voff, t.count, m.le};
    int dep = (v.u16(0) + 12) & 15;
    std::vector<uint16_t> tab(4097, 0);
    for (int c = 0; c < dep; ++c) {
        int b0 = v.u16(14 + 2 * (size_t)c);
        int b1 = v.u8(14 + 2 * (size_t)dep + c);
        for (int i = b0; i <= ((b0 + (4096 >> (b1 & 31)) - 1) & 4095);) tab[++i] = (uint16_t)(b1 << 8 | c);
        tab[0] = 12;
    }
    return tab;
}

// Lossless JPEG stream header (ljpeg_start, R:PefFile/DNGFile.cs:94-171).
struct LJHead {
    int ok = 0;  // 0: the reference skips the tile, 1: decodable
    int algo = 0, bits = 0, high = 0, wide = 0, clrs = 0, sraw = 0, psv = 0, restart = 0x7FFFFFFF;
    size_t data_start = 0;  // first entropy-coded byte, relative to the segment
    int tab[20] = {};       // table ids per DC/AC index (-1: none)
};

// A Huffman table as make_decoder_ref (R:PefFile/DNGFile.cs:72-92) reads it: the longest code length `max`
// and the codes per length with their symbols. The device expands it into the 1 + 2^max lookup entries.
struct TableSpec {
    int max = 0;
    uint8_t count[17] = {};
    std::vector<uint8_t> sym;
};

void make_spec(const uint8_t*& src, const uint8_t* end, TableSpec& t) {
    if (end - src < 16) throw DataError("DHT segment too short");
    const uint8_t* count = (src += 16) - 17;
    int max;
    for (max = 16; max > 0 && count[max] == 0; --max) {
    }
    t.max = max;
    for (int len = 1; len <= 16; ++len) t.count[len] = count[len];
    for (int len = 1; len <= max; ++len)
        for (int i = 0; i < count[len]; ++i, ++src) t.sym.push_back(src < end ? *src : 0);
}

// Parses the marker segments up to SOS; tables are appended to `specs` (their indices go to head.tab).
LJHead ljpeg_header(const uint8_t* seg, size_t n, std::vector<TableSpec>& specs) {
    LJHead jh;
    for (int& t : jh.tab) t = -1;
    std::vector<uint8_t> data(0x10000, 0);
    size_t pos = 0;
    auto rd = [&](uint8_t* dst, size_t k) {
        size_t m = std::min(k, n - pos);
        memcpy(dst, seg + pos, m);
        pos += m;
        return m;
    };
    if (pos < n) ++pos;
    if (pos >= n || seg[pos++] != 0xd8) return jh;
    int tag;
    do {
        if (rd(data.data(), 4) != 4) return jh;
        tag = data[0] << 8 | data[1];
        int len = (data[2] << 8 | data[3]) - 2;
        if (tag <= 0xff00) return jh;
        if (len < 0) throw DataError("lossless JPEG marker segment length < 2");
        rd(data.data(), (size_t)len);
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
                const uint8_t* end = data.data() + len;
                int c;
                while (dp < end && ((c = *dp++) & -20) == 0) {
                    TableSpec t;
                    make_spec(dp, end, t);
                    jh.tab[c] = (int)specs.size();
                    specs.push_back(std::move(t));
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
        if (jh.tab[c + 1] < 0) jh.tab[c + 1] = jh.tab[c];
    jh.data_start = pos;
    jh.ok = 1;
    return jh;
}

// =====================================================================================================
// Decode plan
// =====================================================================================================

enum UnitKind { U_PEF_HUFF = 0, U_LJPEG = 1, U_PACKED = 2, U_LE16 = 3 };

struct Unit {                 // a tile, a strip, or a PEF image
    int kind = 0;
    int rows = 0, cols = 0;   // decoded grid (lossless JPEG: high x wide * clrs)
    int r0 = 0, c0 = 0;       // placement of row 0 / column 0 in the image
    int clrs = 1, bits = 16, psv = 1, restart_rows = 0;
    int transform = 0;        // 1: linearisation, black deltas and clamp (DNG)
    size_t byte_off = 0, byte_len = 0;  // uncompressed strips: the strip's bytes in the file
    uint64_t sym = 0;         // entropy-coded units: first difference in this file's difference range
};

struct Seg {                  // one entropy-coded stream
    size_t off = 0, len = 0;  // bytes in the file (lossless JPEG: stuffed, up to the tile's end)
    uint64_t sym = 0;         // first difference in this file's range
    uint32_t n_sym = 0;
    int unit = 0;
    int clrs = 1;
    int tab[4] = {0, 0, 0, 0};  // offsets into the file's table array
};

}  // namespace

struct isa_raw_file {
    isa_raw_info info{};
    size_t size = 0;
    bool large = false;
    std::vector<Unit> units;
    std::vector<Seg> segs;
    std::vector<uint16_t> tables;    // PEF: the Pentax table
    std::vector<TableSpec> ljtabs;   // lossless JPEG: the tables of every tile / strip
    std::vector<uint16_t> lin, bh, bv;
    bool has_lin = false;
    uint64_t n_sym = 0;
};

namespace {

void copy_str(char* dst, size_t cap, const std::string& s) {
    size_t m = std::min(cap - 1, s.size());
    memcpy(dst, s.data(), m);
    dst[m] = 0;
}

// Black levels from a 4-value CFA-ordered list: red, max(first green, last green), blue
// (R:PefFile/PEFFile.cs:229-281, R:PefFile/DNGFile.cs:780-824).
void black_from_pattern(const double* v, const int* bayer, float out[3]) {
    int ir = -1, ib = -1, g1 = -1, g2 = -1;
    for (int i = 0; i < 4; ++i)
        if (bayer[i] == 0) {
            ir = i;
            break;
        }
    for (int i = 0; i < 4; ++i)
        if (bayer[i] == 2) {
            ib = i;
            break;
        }
    for (int i = 0; i < 4; ++i) {
        if (bayer[i] == 1 && g1 == -1) g1 = i;
        if (bayer[i] == 1 && g1 != -1) g2 = i;
    }
    if (ir < 0 || ib < 0 || g1 < 0) throw DataError("black levels per CFA site without red, green and blue");
    out[0] = (float)v[ir];
    out[2] = (float)v[ib];
    out[1] = std::max((float)v[g1], (float)v[g2]);
}

void identity_twist(isa_raw_info& in) {
    for (int i = 0; i < 12; ++i) in.color_twist[i] = 0.f;
    in.color_twist[0] = in.color_twist[5] = in.color_twist[10] = 1.f;
    in.color_twist_is_identity = 1;
}

// Plans the packed / little-endian strips and checks them where the reference's loops would throw.
void plan_uncompressed(isa_raw_file& f, const Reader& r, const Dir& raw, int kind, bool dng) {
    const Tag* so = raw.find(0x111);
    const Tag* sc = raw.find(0x117);
    const Tag* rp = raw.find(0x116);
    if (!so || !sc || !rp) throw DataError("strips without StripOffsets / StripByteCounts / RowsPerStrip");
    const int W = f.info.width, H = f.info.height, bits = f.info.bit_depth;
    long rps = (long)unum(r, *rp, 0);
    if (kind == U_PACKED && (bits < 8 || bits > 16))
        throw std::invalid_argument("uncompressed samples of other than 8 to 16 bits");
    for (uint32_t s = 0; s < so->count; ++s) {
        Unit u;
        u.kind = kind;
        u.byte_off = unum(r, *so, s);
        u.byte_len = unum(r, *sc, std::min(s, sc->count - 1));
        if (u.byte_off > f.size || u.byte_len > f.size - u.byte_off) throw DataError("strip beyond the end of the file");
        u.r0 = (int)std::min<long>((long)s * rps, 1L << 30);
        u.c0 = 0;
        u.cols = W;
        u.bits = bits;
        u.transform = dng ? 1 : 0;
        if (kind == U_LE16) {
            size_t px = u.byte_len / 2;
            if (px == 0) continue;
            // row of the last sample: strip * rps + (px - 1) / W must lie inside the image
            if ((long)u.r0 + (long)((px - 1) / W) >= H) throw DataError("strip holds rows beyond the image");
            u.rows = (int)((px + W - 1) / W);
        } else {
            if (rps <= 0) continue;
            if ((long)s * rps + rps - 1 >= H) throw DataError("strip rows beyond the image height");
            uint64_t total = (uint64_t)rps * W * bits;  // every sample needs its bytes inside the strip
            if ((total + 7) / 8 > u.byte_len) throw DataError("strip shorter than its samples");
            u.rows = (int)rps;
        }
        f.units.push_back(u);
    }
}

void finish_dng_black(isa_raw_file& f) {
    // largest black delta over the decoded pixels (R:PefFile/DNGFile.cs:424, 496, 542, 585; applied at :607-609)
    int mx = 0;
    const int W = f.info.width, H = f.info.height;
    for (const Unit& u : f.units) {
        int rb = std::min(H, u.r0 + u.rows), cb = std::min(W, u.c0 + u.cols);
        if (u.kind == U_LE16) {  // samples fill rows from the strip's first row
            size_t px = u.byte_len / 2;
            if (px < (size_t)W) cb = std::min<int>(cb, (int)px);
        }
        if (u.r0 >= rb || u.c0 >= cb) continue;
        int mh = *std::max_element(f.bh.begin() + u.c0, f.bh.begin() + cb);
        int mv = *std::max_element(f.bv.begin() + u.r0, f.bv.begin() + rb);
        mx = std::max(mx, mh + mv);
    }
    f.info.max_black_delta = mx;
    for (int c = 0; c < 3; ++c) f.info.white_level[c] -= (float)mx;
}

void parse_pef(isa_raw_file& f, const Reader& r, const std::vector<Dir>& ifds) {
    isa_raw_info& in = f.info;
    const Dir& raw = ifds[0];
    if (raw.exif.empty()) throw DataError("PEF without an Exif IFD");
    const Dir& exif = raw.exif[0];
    auto need = [&](const Dir& d, uint16_t t) {
        const Tag* x = d.find(t);
        if (!x) throw DataError("PEF tag missing");
        return x;
    };
    in.iso = (int)unum(r, *need(exif, 0x8827), 0);
    const Tag* et = need(exif, 0x829A);
    r.need(et->voff, 8);
    in.exposure_num = r.u32(et->voff);
    in.exposure_den = r.u32(et->voff + 4);
    copy_str(in.recording_date, sizeof in.recording_date, ascii(r, *need(exif, 0x9004)));
    in.bit_depth = (int)unum(r, *need(raw, 0x102), 0);
    const Tag* mt = need(exif, 0x927C);
    r.need(mt->voff, mt->count);
    MakerNote mn = read_maker_note(r.d + mt->voff, mt->count, (long)mt->field);
    const Tag* cfa = need(exif, 0xA302);
    if (cfa->count < 8) throw DataError("Exif CFA pattern too short");
    int xc = r.u8(cfa->voff + 1), yc = r.u8(cfa->voff + 3);
    if (xc * yc != 4) throw DataError("CFA pattern other than 2x2");
    for (int x = 0; x < xc; ++x)
        for (int y = 0; y < yc; ++y) in.bayer_pattern[y * yc + x] = r.u8(cfa->voff + 4 + (size_t)(y * yc + x));
    Reader mr{mn.bytes.data(), mn.bytes.size(), mn.le};
    const Tag* wl = mn.dir.find(0x007E);
    const Tag* wp = mn.dir.find(0x0201);
    const Tag* bp = mn.dir.find(0x0200);
    const Tag* ds = mn.dir.find(0x003D);
    if (!wp || !ds) throw DataError("PEF maker note without WhitePoint / DataScaling");
    float white_all = (float)std::pow(2.0, (double)in.bit_depth);
    if (wl) {
        // MNWhiteLevel keeps the raw 4-byte field for inline values (R:PefFile/PentaxMakerNotesEntry.cs:1206-1226)
        white_all = (float)((size_t)type_size(wl->type) * wl->count <= 4 ? wl->field : unum(mr, *wl, 0));
    }
    float black[3] = {0.f, 0.f, 0.f};
    if (bp) {
        if (bp->count == 1) black[0] = black[1] = black[2] = (float)num(mr, *bp, 0);
        if (bp->count == 3)
            for (int c = 0; c < 3; ++c) black[c] = (float)num(mr, *bp, c);
        if (bp->count == 4) {
            double v[4];
            for (int c = 0; c < 4; ++c) v[c] = num(mr, *bp, c);
            black_from_pattern(v, in.bayer_pattern, black);
        }
    }
    for (int c = 0; c < 3; ++c) {
        in.black_level[c] = black[c];
        in.white_level[c] = white_all - black[c];
    }
    float scale = (float)num(mr, *ds, 0);
    if (wp->count < 4) throw DataError("WhitePoint with fewer than 4 values");
    in.white_balance[0] = (float)num(mr, *wp, 0) / scale;
    in.white_balance[1] = (float)num(mr, *wp, 1) / scale;
    in.white_balance[2] = (float)num(mr, *wp, 3) / scale;
    if (const Tag* li = mn.dir.find(0x022B)) {
        in.roll_angle = -(float)(int8_t)mr.u8(li->voff + 1) / 2.0f;
        in.roll_angle_present = 1;
    }
    in.width = (int)unum(r, *need(raw, 0x100), 0);
    in.height = (int)unum(r, *need(raw, 0x101), 0);
    in.orientation = raw.find(0x112) ? (int)unum(r, *raw.find(0x112), 0) : 1;
    in.crop_left = 2;
    in.crop_top = 2;
    in.cropped_width = in.width - 4;
    in.cropped_height = in.height - 4;
    copy_str(in.make, sizeof in.make, ascii(r, *need(raw, 0x10F)));
    copy_str(in.unique_model, sizeof in.unique_model, ascii(r, *need(raw, 0x110)));
    identity_twist(in);
    if (in.width <= 0 || in.height <= 0 || in.width > 65535 || in.height > 65535)
        throw DataError("PEF image size out of range");
    const Tag* hu = mn.dir.find(0x0220);
    if (hu) {
        in.compression = ISA_RAW_PENTAX_HUFFMAN;
        f.tables = pentax_table(mn, *hu);
        const Tag* so = need(raw, 0x111);
        if (so->count != 1) throw DataError("TODO: implement compressed multiple stripes");
        Unit u;
        u.kind = U_PEF_HUFF;
        u.rows = in.height;
        u.cols = in.width;
        u.clrs = std::min(2, in.width);
        u.sym = 0;
        f.units.push_back(u);
        Seg s;
        s.off = unum(r, *so, 0);
        if (s.off > f.size) throw DataError("strip beyond the end of the file");
        // chunks cover the strip; the reader may continue into whatever follows it, up to the end of the file
        const Tag* sc = raw.find(0x117);
        s.len = f.size - s.off;
        if (sc) s.len = std::min<size_t>(s.len, unum(r, *sc, 0));
        s.n_sym = (uint32_t)((uint64_t)in.width * in.height);
        s.unit = 0;
        s.clrs = 1;
        f.segs.push_back(s);
        f.n_sym = s.n_sym;
    } else {
        in.compression = ISA_RAW_UNCOMPRESSED;
        plan_uncompressed(f, r, raw, U_PACKED, false);
    }
}

void parse_dng(isa_raw_file& f, const Reader& r, const std::vector<Dir>& ifds) {
    isa_raw_info& in = f.info;
    const Dir& ifd0 = ifds[0];
    const Dir* raw = nullptr;
    for (const Dir& d : ifds)
        if (const Tag* p = d.find(0x106))
            if (unum(r, *p, 0) == 32803) {
                raw = &d;
                break;
            }
    if (!raw)
        for (const Dir& d : ifds)
            for (const Dir& s : d.sub) {
                const Tag* p = s.find(0x106);
                if (!p) throw DataError("SubIFD without PhotometricInterpretation");
                if (unum(r, *p, 0) == 32803) {
                    raw = &s;
                    break;
                }
            }
    if (!raw) throw DataError("Can't find IFD with Bayer RAW image.");
    const Tag* ver = ifd0.find(0xC612);
    f.large = !ver || unum(r, *ver, 1) > 1;
    auto need = [&](const Dir& d, uint16_t t) {
        const Tag* x = d.find(t);
        if (!x) throw DataError("DNG tag missing");
        return x;
    };
    if (const Tag* lt = raw->find(0xC618)) {
        if (lt->count == 0) throw DataError("empty LinearizationTable");
        for (uint32_t k = 0; k < lt->count; ++k) f.lin.push_back((uint16_t)unum(r, *lt, k));
        f.has_lin = true;
    }
    in.width = (int)unum(r, *need(*raw, 0x100), 0);
    in.height = (int)unum(r, *need(*raw, 0x101), 0);
    if (in.width <= 0 || in.height <= 0 || in.width > 65535 || in.height > 65535)
        throw DataError("DNG image size out of range");
    if (const Tag* dt = ifd0.find(0x132)) copy_str(in.recording_date, sizeof in.recording_date, ascii(r, *dt));
    if (const Tag* pd = ifd0.find(0xC634)) {
        r.need(pd->voff, pd->count);
        MakerNote mn = read_maker_note(r.d + pd->voff, pd->count, 0);
        if (mn.valid)
            if (const Tag* li = mn.dir.find(0x022B)) {
                Reader mr{mn.bytes.data(), mn.bytes.size(), mn.le};
                in.roll_angle = -(float)(int8_t)mr.u8(li->voff + 1) / 2.0f;
                in.roll_angle_present = 1;
            }
    }
    auto exposure = [&](const Tag* t) {
        r.need(t->voff, 8);
        in.exposure_num = r.u32(t->voff);
        in.exposure_den = r.u32(t->voff + 4);
    };
    if (!ifd0.exif.empty()) {
        const Dir& ex = ifd0.exif[0];
        in.iso = (int)unum(r, *need(ex, 0x8827), 0);
        exposure(need(ex, 0x829A));
        if (const Tag* dd = ex.find(0x9004)) copy_str(in.recording_date, sizeof in.recording_date, ascii(r, *dd));
    } else if (raw->find(0x8827)) {
        in.iso = (int)unum(r, *raw->find(0x8827), 0);
        exposure(need(*raw, 0x829A));
    } else if (ifd0.find(0x8827)) {
        in.iso = (int)unum(r, *ifd0.find(0x8827), 0);
        exposure(need(ifd0, 0x829A));
    }
    in.bit_depth = (int)unum(r, *need(*raw, 0x102), 0);
    const Tag* dim = need(*raw, 0x828D);
    if (dim->count < 2 || unum(r, *dim, 0) != 2 || unum(r, *dim, 1) != 2)
        throw DataError("This file has a bayer pattern size different than 2x2. Can't decode that.");
    const Tag* cfa = need(*raw, 0x828E);
    if (cfa->count < 4) throw DataError("CFAPattern with fewer than 4 values");
    for (int i = 0; i < 4; ++i) in.bayer_pattern[i] = (int)unum(r, *cfa, i);
    identity_twist(in);
    int helper[3] = {0, 1, 2};
    if (const Tag* pc = raw->find(0xC616)) {
        for (int i = 0; i < 12; ++i) in.color_twist[i] = 0.f;
        if (pc->count != 3) throw DataError("This image doesn't contain three color planes.");
        for (int i = 0; i < 3; ++i) {
            int c = (int)unum(r, *pc, i);
            helper[i] = c;
            if (c > 2) throw DataError("This image contains colors different than red/green/blue.");
            in.color_twist[c * 4 + i] = 1.f;
            if (c != i) in.color_twist_is_identity = 0;
        }
    }
    const Tag* wl = need(*raw, 0xC61D);
    float black[3] = {0.f, 0.f, 0.f};
    if (const Tag* bl = raw->find(0xC61A)) {
        if (bl->count == 1) black[0] = black[1] = black[2] = (float)num(r, *bl, 0);
        if (bl->count == 3)
            for (int i = 0; i < 3; ++i) black[helper[i]] = (float)num(r, *bl, i);
        if (bl->count == 4) {
            double v[4];
            for (int c = 0; c < 4; ++c) v[c] = num(r, *bl, c);
            black_from_pattern(v, in.bayer_pattern, black);
        }
    }
    float white = (float)unum(r, *wl, 0);
    for (int c = 0; c < 3; ++c) {
        in.black_level[c] = black[c];
        in.white_level[c] = white - black[c];
    }
    // Camera white: the normalised AsShotNeutral (IFD0 first, then the raw IFD).
    const Tag* ne = ifd0.find(0xC628);
    if (!ne) ne = raw->find(0xC628);
    if (!ne || ne->count < 3) throw DataError("AsShotNeutral missing (AsShotWhiteXY alone is not supported)");
    double nv[3], mx = 0;
    for (int c = 0; c < 3; ++c) mx = std::max(mx, nv[c] = num(r, *ne, c));
    for (int c = 0; c < 3; ++c) in.white_balance[helper[c]] = 1.0f / (float)(nv[c] / mx);
    if (const Tag* o = raw->find(0x112))
        in.orientation = (int)unum(r, *o, 0);
    else if (const Tag* o0 = ifd0.find(0x112))
        in.orientation = (int)unum(r, *o0, 0);
    else
        in.orientation = 1;
    int cl = 0, ct = 0, cw = in.width, ch = in.height;
    const Tag* aa = raw->find(0xC68D);
    int aa_v[4] = {0, 0, 0, 0};
    if (aa) {
        if (aa->count < 4) throw DataError("ActiveArea with fewer than 4 values");
        for (int i = 0; i < 4; ++i) aa_v[i] = (int)unum(r, *aa, i);
        int top = aa_v[0], left = aa_v[1], bottom = aa_v[2], right = aa_v[3];
        cl += left;
        ct += top;
        cw = right - left;
        ch = bottom - top;
        int* b = in.bayer_pattern;
        if (top % 2 != 0) {
            int b0 = b[0], b1 = b[1];
            b[0] = b[2];
            b[1] = b[3];
            b[2] = b0;
            b[3] = b1;
        }
        if (left % 2 != 0) {
            int b0 = b[0], b2 = b[2];
            b[0] = b[1];
            b[2] = b[3];
            b[1] = b0;
            b[3] = b2;
        }
    }
    const Tag* co = raw->find(0xC61F);
    const Tag* cs = raw->find(0xC620);
    if (co && cs) {  // the reference reads element 0 for both coordinates and both sizes
        int left = (int)num(r, *co, 0), top = (int)num(r, *co, 0);
        int width = (int)num(r, *cs, 0), height = (int)num(r, *cs, 0);
        cl += left;
        ct += top;
        cw = width;
        ch = height;
    }
    in.crop_left = std::max(2, cl);
    in.crop_top = std::max(2, ct);
    in.cropped_width = cw - std::max(0, (cl + cw) - (in.width - 2));
    in.cropped_height = ch - std::max(0, (ct + ch) - (in.height - 2));
    const Tag* np = raw->find(0xC761);
    if (!np) np = ifd0.find(0xC761);
    if (np) {
        if (np->count > 2) {
            in.noise_alpha = (float)num(r, *np, (uint32_t)helper[1] * 2);
            in.noise_beta = (float)num(r, *np, (uint32_t)helper[1] * 2 + 1);
        } else {
            in.noise_alpha = (float)num(r, *np, 0);
            in.noise_beta = (float)num(r, *np, 1);
        }
    }
    copy_str(in.make, sizeof in.make, ascii(r, *need(ifd0, 0x10F)));
    copy_str(in.unique_model, sizeof in.unique_model, ascii(r, *need(ifd0, 0xC614)));

    // black deltas (R:PefFile/DNGFile.cs:318-372), cast to ushort as C# does (truncation, modulo 2^16)
    const int W = in.width, H = in.height;
    f.bh.assign(W, 0);
    f.bv.assign(H, 0);
    int off_v = 0, off_h = 0, bottom = 0, right = 0;
    if (aa) {
        off_v = aa_v[0];
        off_h = aa_v[1];
        bottom = H - aa_v[2];
        right = W - aa_v[3];
    }
    auto to_u16 = [](double x) { return (uint16_t)(int64_t)std::trunc(x); };
    if (const Tag* dv = raw->find(0xC61C)) {
        if ((long)dv->count + off_v + bottom != H) throw DataError("Count in IFDDNGBlackLevelDeltaV doesn't fit image height");
        for (uint32_t i = 0; i < dv->count; ++i) f.bv.at(off_v + i) = to_u16(num(r, *dv, i));
    }
    if (const Tag* dh = raw->find(0xC61B)) {
        if ((long)dh->count + off_h + right != W) throw DataError("Count in IFDDNGBlackLevelDeltaH doesn't fit image width");
        for (uint32_t i = 0; i < dh->count; ++i) f.bh.at(off_h + i) = to_u16(num(r, *dh, i));
    }

    // pixel data
    int comp = (int)unum(r, *need(*raw, 0x103), 0);
    in.compression = comp == 7 ? ISA_RAW_LOSSLESS_JPEG : comp;
    auto add_ljpeg = [&](size_t off, size_t len, int r0, int c0, int max_rows, int max_cols) {
        if (off > f.size || len > f.size - off) throw DataError("tile or strip beyond the end of the file");
        LJHead jh = ljpeg_header(r.d + off, len, f.ljtabs);
        if (!jh.ok || jh.algo != 0xc3) return;  // the reference leaves such a tile at 0
        if (jh.sraw != 0) throw std::invalid_argument("lossless JPEG with subsampled components (sraw)");
        if (jh.clrs < 1 || jh.clrs > 4) throw std::invalid_argument("lossless JPEG with other than 1 to 4 components");
        if (jh.restart == 0) throw DataError("lossless JPEG restart interval 0");
        if (jh.high <= 0 || jh.wide <= 0) return;  // nothing to decode
        // a stream larger than its tile or strip would overwrite its neighbours in file order
        if (jh.high > max_rows || (long)jh.wide * jh.clrs > max_cols)
            throw std::invalid_argument("lossless JPEG stream larger than its tile or strip");
        Unit u;
        u.kind = U_LJPEG;
        u.rows = jh.high;
        u.cols = jh.wide * jh.clrs;
        u.r0 = r0;
        u.c0 = c0;
        u.clrs = jh.clrs;
        u.bits = jh.bits;
        u.psv = jh.psv;
        u.transform = 1;
        u.sym = f.n_sym;
        for (int c = 0; c < jh.clrs; ++c)
            if (jh.tab[c] < 0) throw DataError("lossless JPEG component without a Huffman table");
        // restart intervals: rows where jrow * wide % restart == 0 re-synchronise on the next marker
        std::vector<int> starts{0};
        for (int jr = 1; jr < jh.high; ++jr)
            if ((uint32_t)jr * (uint32_t)jh.wide % (uint32_t)jh.restart == 0) starts.push_back(jr);
        if (starts.size() > 1) {
            u.restart_rows = starts[1];
            for (size_t k = 1; k < starts.size(); ++k)
                if (starts[k] != (int)k * u.restart_rows) throw std::invalid_argument("irregular restart rows");
        }
        size_t pos = jh.data_start;
        const uint8_t* p = r.d + off;
        f.units.push_back(u);
        for (size_t k = 0; k < starts.size(); ++k) {
            size_t end = len;
            size_t next = len;
            if (k + 1 < starts.size()) {  // the marker that ends this interval: the first FF Dx from here on
                size_t q = pos;
                while (q + 1 < len && !(p[q] == 0xFF && (p[q + 1] >> 4) == 0xD)) ++q;
                end = std::min(len, q + 2);
                next = end;
            }
            Seg s;
            s.off = off + pos;
            s.len = end - pos;
            int rows = (k + 1 < starts.size() ? starts[k + 1] : jh.high) - starts[k];
            s.n_sym = (uint32_t)((uint64_t)rows * u.cols);
            s.sym = f.n_sym;
            s.unit = (int)f.units.size() - 1;
            s.clrs = jh.clrs;
            for (int c = 0; c < jh.clrs; ++c) s.tab[c] = jh.tab[c];
            f.segs.push_back(s);
            f.n_sym += s.n_sym;
            pos = next;
        }
    };
    if (const Tag* tw = raw->find(0x142)) {
        const Tag* tl = need(*raw, 0x143);
        const Tag* to = need(*raw, 0x144);
        const Tag* tc = need(*raw, 0x145);
        if (comp != 7) throw DataError("I don't know how to read that file :(");
        int tile_w = (int)unum(r, *tw, 0), tile_h = (int)unum(r, *tl, 0);
        if (tile_w <= 0 || tile_h <= 0) throw DataError("tile size 0");
        int row = 0, col = 0;
        for (uint32_t t = 0; t < to->count; ++t) {
            add_ljpeg(unum(r, *to, t), unum(r, *tc, std::min(t, tc->count - 1)), row, col, tile_h, tile_w);
            col += tile_w;
            if (col >= W) {
                col = 0;
                row += tile_h;
            }
        }
    } else if (const Tag* so = raw->find(0x111)) {
        const Tag* sc = need(*raw, 0x117);
        long rps = (long)unum(r, *need(*raw, 0x116), 0);
        if (comp == 7) {
            for (uint32_t s = 0; s < so->count; ++s)
                add_ljpeg(unum(r, *so, s), unum(r, *sc, std::min(s, sc->count - 1)),
                          (int)std::min<long>((long)s * rps, 1L << 30), 0, (int)std::min<long>(rps, 1L << 30),
                          1 << 30);
        } else if (comp == 1) {
            if (in.bit_depth == 16)
                plan_uncompressed(f, r, *raw, U_LE16, true);
            else if (in.bit_depth < 16)
                plan_uncompressed(f, r, *raw, U_PACKED, true);
        } else {
            throw DataError("I don't know how to read that file :(");
        }
    }
    if (f.lin.empty()) f.lin.push_back(0);  // placeholder; unused without a table
    finish_dng_black(f);
}

// =====================================================================================================
// Device: plan structures
// =====================================================================================================

struct DUnit {
    uint64_t out_px;     // not used on the device (kept for alignment)
    uint64_t sym;        // first difference (absolute in the batch difference buffer)
    uint64_t byte_off;   // uncompressed strips: offset in the file
    uint64_t byte_len;
    uint64_t vchain;     // first row-start value (absolute in the batch chain buffer), linear units
    uint64_t vals;       // predictors 2..7: first value in the batch value buffer
    int32_t kind, file, rows, cols, r0, c0, clrs, bits, psv, restart_rows, transform, lin_n;
    int32_t lin, bh, bv;  // element offsets into the batch level arrays
    int32_t pad[3];
};

struct DSeg {
    uint64_t off;     // byte offset of the stream in its file
    uint64_t len;     // stream bytes in the file (stuffed for lossless JPEG)
    uint64_t ubuf;    // lossless JPEG: byte offset of its unstuffed copy in the stream buffer
    uint64_t ucap;    // capacity of that copy (multiple of 16)
    uint64_t diff;    // first difference (absolute)
    uint32_t n_sym;
    int32_t file, unit, clrs, large, ljpeg;
    int32_t tab[4];   // absolute element offsets of the component tables
    int32_t pad[2];
};

struct DSpec {        // a lossless-JPEG Huffman table to expand on the device
    int32_t dst;      // element offset of its 1 + 2^max entries in the table array
    int32_t max;
    int32_t sym;      // offset of its symbols in the symbol array
    int32_t nsym;
    uint8_t count[20];
};

struct DFile {
    const uint8_t* bytes;
    uint16_t* out;
    uint64_t size;
    int32_t pitch, width, height, pad;
};

enum { E_EOF = 1, E_RANGE = 2, E_SYNC = 3 };

__device__ inline void raise_error(int* status, int code, int file) {
    if (atomicCAS(status, 0, code) == 0) status[1] = file;
}

__device__ inline uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ inline int imin(int a, int b) { return a < b ? a : b; }
__device__ inline int imax(int a, int b) { return a > b ? a : b; }

__device__ inline uint32_t bswap32(uint32_t x) {
    return (x >> 24) | ((x >> 8) & 0xFF00u) | ((x << 8) & 0xFF0000u) | (x << 24);
}

// ---------------------------------------------------------------------------------------------------
// Output fill: every image starts at 0 (pixels no tile or strip covers stay 0).
// make_decoder_ref's table (R:PefFile/DNGFile.cs:72-92): entry 0 = max; entries are filled in code order, each
// code of length len covering 2^(max - len) entries; entries past the last code stay 0. One block per table.
__global__ void k_build_tables(const DSpec* specs, const uint8_t* syms, uint16_t* tables) {
    const DSpec& t = specs[blockIdx.x];
    const int max = t.max;
    uint16_t* out = tables + t.dst;
    const uint32_t n = 1u << max;
    if (threadIdx.x == 0) out[0] = (uint16_t)max;
    for (uint32_t h = threadIdx.x + 1; h <= n; h += blockDim.x) {
        // entries before length len: sum over shorter lengths of count * 2^(max - l)
        uint32_t cum = 0, before = 0;
        uint16_t e = 0;
        for (int len = 1; len <= max; ++len) {
            const uint32_t span = (uint32_t)t.count[len] << (max - len);
            if (h <= cum + span) {
                const uint32_t k = before + (h - cum - 1) / (1u << (max - len));
                e = (uint16_t)(len << 8 | (k < (uint32_t)t.nsym ? syms[t.sym + k] : 0));
                break;
            }
            cum += span;
            before += t.count[len];
        }
        out[h] = e;
    }
}

__global__ void k_zero_int(int* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void k_fill_out(const DFile* files, int n, int* status) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4) status[threadIdx.x] = 0;
    const DFile& f = files[blockIdx.y];
    const size_t row_words = (size_t)f.width;  // u16 per row
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)f.height * row_words;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t y = i / row_words, x = i - y * row_words;
        reinterpret_cast<uint16_t*>(reinterpret_cast<uint8_t*>(f.out) + y * f.pitch)[x] = 0;
    }
}

// ---------------------------------------------------------------------------------------------------
// Lossless-JPEG streams without their stuffing: the reference's reader drops the byte after every 0xFF it
// takes (R:PefFile/DNGFile.cs:251-255), so byte i is dropped when the run of 0xFF bytes just before it has
// odd length. The copy is followed by 0xFF up to its capacity (the reader sees 0xFF past the end).
constexpr int kUnstuffThreads = 256;
constexpr int kUnstuffPer = 16;

__global__ __launch_bounds__(kUnstuffThreads) void k_unstuff(const DSeg* segs, const int* ljseg, const DFile* files,
                                                            uint8_t* ubuf, uint64_t* ulen) {
    const DSeg& s = segs[ljseg[blockIdx.x]];
    const uint8_t* src = files[s.file].bytes + s.off;
    uint8_t* dst = ubuf + s.ubuf;
    __shared__ uint32_t scan[kUnstuffThreads];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int t = threadIdx.x;
    for (uint64_t base = 0; base < s.len; base += (uint64_t)kUnstuffThreads * kUnstuffPer) {
        uint64_t b0 = base + (uint64_t)t * kUnstuffPer;
        uint32_t keep = 0;  // bit k: byte b0 + k kept
        int run = 0;        // 0xFF bytes just before b0
        if (b0 < s.len)
            for (uint64_t q = b0; q > 0 && src[q - 1] == 0xFF; --q) ++run;
        for (int k = 0; k < kUnstuffPer; ++k) {
            uint64_t i = b0 + k;
            if (i >= s.len) break;
            uint8_t v = src[i];
            if ((run & 1) == 0) keep |= 1u << k;
            run = v == 0xFF ? run + 1 : 0;
        }
        uint32_t cnt = __popc(keep);
        scan[t] = cnt;
        __syncthreads();
        for (int o = 1; o < kUnstuffThreads; o <<= 1) {
            uint32_t v = t >= o ? scan[t - o] : 0;
            __syncthreads();
            scan[t] += v;
            __syncthreads();
        }
        uint64_t w = carry + scan[t] - cnt;
        for (int k = 0; k < kUnstuffPer; ++k)
            if (keep >> k & 1) dst[w++] = src[b0 + k];
        __syncthreads();
        if (t == kUnstuffThreads - 1) carry += scan[t];
        __syncthreads();
    }
    const uint64_t n = carry;
    for (uint64_t i = n + t; i < s.ucap; i += kUnstuffThreads) dst[i] = 0xFF;
    if (t == 0) ulen[ljseg[blockIdx.x]] = n;
}

// ---------------------------------------------------------------------------------------------------
// Bit cursor over a stream: MSB-first 64-bit window fed from 16-byte blocks that are loaded three blocks
// (384 bits, some 40 codes) before they are needed, so the decode rarely waits on memory.
struct Cursor {
    const uint8_t* base;  // 16-byte aligned origin of the block grid
    uint64_t avail;       // readable bytes from base (PEF: the file; JPEG: the padded copy)
    uint64_t start;       // bit offset of the stream's first bit from base
    int fill_ff;          // bytes past `avail` read as 0xFF (JPEG) or 0 with an EOF check (PEF)
    uint64_t buf;
    int nb;
    uint4 q0, q1, q2, q3;  // q0: block being spliced (word j next), q1..q3 loaded ahead
    int j;
    uint64_t next_block;
    uint32_t pos;          // bits consumed from the stream start

    __device__ uint32_t word_at(uint64_t b) const {
        uint32_t w = 0;
        for (int i = 0; i < 4; ++i) {
            const uint32_t v = b + i < avail ? base[b + i] : (fill_ff ? 0xFFu : 0u);
            w = w << 8 | v;
        }
        return w;
    }
    __device__ uint4 block(uint64_t k) const {
        const uint64_t b = k * 16;
        uint4 r;
        if (b + 16 <= avail) {
            r = reinterpret_cast<const uint4*>(base)[k];
            r.x = bswap32(r.x);
            r.y = bswap32(r.y);
            r.z = bswap32(r.z);
            r.w = bswap32(r.w);
        } else {
            r.x = word_at(b);
            r.y = word_at(b + 4);
            r.z = word_at(b + 8);
            r.w = word_at(b + 12);
        }
        return r;
    }
    __device__ uint32_t take() {
        const uint32_t w = j == 0 ? q0.x : j == 1 ? q0.y : j == 2 ? q0.z : q0.w;
        if (++j == 4) {
            q0 = q1;
            q1 = q2;
            q2 = q3;
            q3 = block(next_block++);
            j = 0;
        }
        return w;
    }
    __device__ void seek(uint32_t p) {
        pos = p;
        const uint64_t a = start + p;
        const uint64_t k = a >> 7;
        q0 = block(k);
        q1 = block(k + 1);
        q2 = block(k + 2);
        q3 = block(k + 3);
        next_block = k + 4;
        j = (int)((a >> 5) & 3);
        const int sh = (int)(a & 31);
        const uint64_t hi = take();
        const uint64_t lo = take();
        buf = (hi << 32 | lo) << sh;
        nb = 64 - sh;
    }
    __device__ uint32_t peek(int n) const { return n ? (uint32_t)(buf >> (64 - n)) : 0u; }
    __device__ void skip(int n) {
        buf <<= n;
        nb -= n;
        pos += n;
        if (nb <= 32) {
            buf |= (uint64_t)take() << (32 - nb);
            nb += 32;
        }
    }
    // PEF: the reference's reader throws when a read needs a byte past the end of the file.
    __device__ bool past_end(int n) const { return !fill_ff && start + pos + n > avail * 8; }
};

constexpr uint32_t kChunkBits = 4096;  // chunk of a stream decoded by one lane
constexpr uint32_t kCheckBits = 512;   // distance of a chunk's checkpoint from its start
constexpr uint32_t kNone = 0xFFFFFFFFu;

// One difference (ljpeg_diff, R:PefFile/PEFFile.cs:73-81, R:PefFile/DNGFile.cs:270-281) with lookahead m and the
// table at tab[1 + code]. Returns false at the end of a PEF file. stuck: the code consumed no bits (the
// reference then repeats this symbol forever).
__device__ inline bool decode_diff(Cursor& c, const uint16_t* tab, int m, int large, int32_t& diff, bool& stuck) {
    int len = 0, clen = 0;
    if (m > 0) {
        if (c.past_end(m)) return false;
        const uint32_t e = tab[1 + c.peek(m)];
        clen = e >> 8;
        len = e & 0xFF;
        c.skip(clen);
    }
    int xbits = 0;
    int32_t d;
    if (large && len == 16) {
        d = -32768;
    } else if (len == 0 || len > 25) {
        d = (int32_t)(0u - ((1u << (len & 31)) - 1u));
        if (len == 0) d = 0;
    } else {
        if (c.past_end(len)) return false;
        const uint32_t v = c.peek(len);
        c.skip(len);
        xbits = len;
        d = (int32_t)v;
        if ((v & (1u << (len - 1))) == 0) d = (int32_t)(v - ((1u << len) - 1u));
    }
    stuck = clen == 0 && xbits == 0;
    diff = d;
    return true;
}

// State of one chunk of a stream. A chunk is decoded from its start `st` once per possible component phase p
// at that start (streams with C components decode component k mod C at symbol k); ex / cnt: where the decode
// passes the chunk's end and how many codes it took; ckp / ckc: the first code boundary at least kCheckBits
// past the start and the codes before it. A chunk decoded again from another start stops as soon as it lands
// on that boundary in the same phase: from there on both decodes are the same.
struct Chunk {
    uint32_t st;
    uint32_t ex[4], cnt[4], ckp[4], ckc[4];
    uint32_t base, nwrite;
    int32_t seg;
    uint32_t dec_st;  // the start the stored decodes belong to
    uint8_t stuck, dirty, par, used, decoded, valid, pad[2];  // valid: bit p = phase p decoded from dec_st
};

struct StreamPlan {    // chunks [first, first + n) of segment seg
    int32_t seg, first, n, chunk_bits;
};

// The stream's cursor: PEF streams read the file (16-byte aligned origin), lossless-JPEG streams their
// unstuffed copy.
__device__ inline Cursor stream_cursor(const DSeg& s, int si, const DFile* files, const uint8_t* ubuf,
                                       const uint64_t* ulen, uint64_t& nbits) {
    Cursor cur;
    if (s.ljpeg) {
        cur.base = ubuf + s.ubuf;
        cur.avail = s.ucap;
        cur.start = 0;
        cur.fill_ff = 1;
        nbits = ulen[si] * 8;
    } else {
        const DFile& f = files[s.file];
        const uint64_t a0 = s.off & ~(uint64_t)15;
        cur.base = f.bytes + a0;
        cur.avail = f.size - a0;
        cur.start = (s.off - a0) * 8;
        cur.fill_ff = 0;
        nbits = s.len * 8;
    }
    if (nbits > 0xFFFFFFF0ull) nbits = 0xFFFFFFF0ull;
    return cur;
}

// Decodes every dirty chunk for each start phase, up to its end (or its old checkpoint). One work-group per range
// of at most 256 chunks of one stream; the stream's tables are read from LDS (kLds) so that a table lookup never
// waits behind the stream blocks loaded ahead (vector loads complete in order).
struct BlockPlan {
    int32_t stream, first, n, pad;
};

constexpr int kLdsTable = 12288;  // u16 table entries a work-group stages

template <bool kLds>
__device__ inline const uint16_t* stage_tables(const DSeg& s, const uint16_t* tables, uint16_t* tlds, int (&to)[4]) {
    if constexpr (!kLds) {
        for (int c = 0; c < 4; ++c) to[c] = s.tab[c];
        return tables;
    } else {
        int o = 0;
        for (int c = 0; c < 4; ++c) {
            to[c] = o;
            if (c >= s.clrs) continue;
            const int len = 1 + (1 << __ldg(tables + s.tab[c]));
            for (int i = threadIdx.x; i < len; i += blockDim.x) tlds[o + i] = tables[s.tab[c] + i];
            o += len;
        }
        __syncthreads();
        return tlds;
    }
}

template <bool kLds>
__global__ __launch_bounds__(256) void k_chunk_decode(Chunk* chunks, const BlockPlan* blocks, const StreamPlan* streams,
                                                      const DSeg* segs, const DFile* files, const uint8_t* ubuf,
                                                      const uint64_t* ulen, const uint16_t* tables, int* redecodes) {
    __shared__ uint16_t tlds[kLds ? kLdsTable : 1];
    const BlockPlan& bp = blocks[blockIdx.x];
    __shared__ int any_dirty;
    if (threadIdx.x == 0) any_dirty = 0;
    __syncthreads();
    if ((int)threadIdx.x < bp.n && chunks[bp.first + threadIdx.x].dirty) any_dirty = 1;
    __syncthreads();
    if (!any_dirty) return;
    const StreamPlan& sp = streams[bp.stream];
    const DSeg& s = segs[sp.seg];
    int to[4];
    const uint16_t* tb = stage_tables<kLds>(s, tables, tlds, to);
    if ((int)threadIdx.x >= bp.n) return;
    const int g = bp.first + threadIdx.x;
    Chunk& k = chunks[g];
    if (!k.dirty) return;
    if (k.decoded) atomicAdd(redecodes, 1);
    uint64_t nbits;
    Cursor cur = stream_cursor(s, sp.seg, files, ubuf, ulen, nbits);
    const int j = g - sp.first;
    const int C = s.clrs;
    const uint32_t cb = (uint32_t)sp.chunk_bits;
    const uint32_t end = j == sp.n - 1 ? (uint32_t)nbits : (uint32_t)umin64((uint64_t)(j + 1) * cb, nbits);
    const uint32_t st = (uint32_t)umin64(k.st, nbits);
    const int large = s.large;
    const int m0 = tb[to[0]], m1 = C > 1 ? tb[to[1]] : 0, m2 = C > 2 ? tb[to[2]] : 0, m3 = C > 3 ? tb[to[3]] : 0;
    const int to0 = to[0], to1 = to[1], to2 = to[2], to3 = to[3];
    // Snapshot of the previous decodes (one per entry phase): a new decode that lands on one of their
    // checkpoints in the same component phase follows that decode from there on.
    const bool had = k.decoded;
    uint32_t o_ex[4], o_cnt[4], o_ckp[4], o_ckc[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const bool live = had && p < C;
        o_ex[p] = live ? k.ex[p] : 0;
        o_cnt[p] = live ? k.cnt[p] : 0;
        o_ckp[p] = live ? k.ckp[p] : kNone;
        o_ckc[p] = live ? k.ckc[p] : 0;
    }
    const uint8_t o_stuck = k.stuck;
    uint8_t stuck_bits = 0;
    for (int p = 0; p < C; ++p) {
        cur.seek(st);
        uint32_t n = 0, ckp = kNone, ckc = 0;
        uint32_t lo = umin64(umin64(o_ckp[0], o_ckp[1]), umin64(o_ckp[2], o_ckp[3]));
        int q = p;
        int from = -1;  // the old phase this decode joined
        bool stk = false;
        while (cur.pos < end) {
            if (cur.pos >= lo) {
                uint32_t nlo = kNone;
#pragma unroll
                for (int pp = 0; pp < 4; ++pp) {
                    if (from < 0 && o_ckp[pp] == cur.pos && (p + n) % (uint32_t)C == (pp + o_ckc[pp]) % (uint32_t)C)
                        from = pp;
                    if (o_ckp[pp] > cur.pos && o_ckp[pp] < nlo) nlo = o_ckp[pp];
                }
                if (from >= 0) break;
                lo = nlo;
            }
            if (ckp == kNone && cur.pos >= st + kCheckBits) {
                ckp = cur.pos;
                ckc = n;
            }
            const int tq = q == 0 ? to0 : q == 1 ? to1 : q == 2 ? to2 : to3;
            const int m = q == 0 ? m0 : q == 1 ? m1 : q == 2 ? m2 : m3;
            int32_t d;
            if (!decode_diff(cur, tb + tq, m, large, d, stk)) break;
            if (stk) break;
            ++n;
            if (++q == C) q = 0;
        }
        if (from >= 0) {  // the rest is the old decode's (stuck bit included)
            const uint32_t fe = from == 0 ? o_ex[0] : from == 1 ? o_ex[1] : from == 2 ? o_ex[2] : o_ex[3];
            const uint32_t fc = from == 0 ? o_cnt[0] : from == 1 ? o_cnt[1] : from == 2 ? o_cnt[2] : o_cnt[3];
            const uint32_t fk = from == 0 ? o_ckc[0] : from == 1 ? o_ckc[1] : from == 2 ? o_ckc[2] : o_ckc[3];
            k.ex[p] = fe;
            k.cnt[p] = n + (fc - fk);
            k.ckp[p] = cur.pos;
            k.ckc[p] = n;
            if (o_stuck >> from & 1) stuck_bits |= 1 << p;
        } else {
            k.ex[p] = cur.pos;
            k.cnt[p] = n;
            k.ckp[p] = ckp;
            k.ckc[p] = ckc;
            if (stk) stuck_bits |= 1 << p;
        }
    }
    k.stuck = stuck_bits;
    k.valid = (uint8_t)((1 << C) - 1);
    k.dec_st = k.st;
    k.decoded = 1;
    k.dirty = 0;
}

// Chain walk of a stream with C > 1 components (one thread per stream): chunk j must start where chunk j - 1
// ends for the phase it is entered with; a mismatch marks chunk j dirty with the new start. Also the base,
// entry phase and number of codes each chunk writes; the walk stops where the stream's N codes are complete.
__global__ void k_walk_serial(Chunk* chunks, const StreamPlan* streams, const int* list, int n_list,
                              const DSeg* segs, int* any_dirty) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_list) return;
    const StreamPlan& sp = streams[list[g]];
    const DSeg& s = segs[sp.seg];
    const int C = s.clrs;
    const uint32_t N = s.n_sym;
    uint32_t pos = 0, base = 0;
    int par = 0;
    bool done = false, a = false;
    for (int j = 0; j < sp.n; ++j) {
        Chunk& k = chunks[sp.first + j];
        k.used = 0;
        if (done) continue;
        if (k.st != pos) {
            k.st = pos;
            k.dirty = 1;
            a = true;
        } else if (k.dec_st != k.st || !(k.valid >> par & 1)) {  // this phase is not decoded from here yet
            k.dirty = 1;
            a = true;
        }
        k.used = 1;
        k.base = base;
        k.par = (uint8_t)par;
        const uint32_t n = k.cnt[par];
        if ((k.stuck >> par & 1) || j == sp.n - 1 || base + n >= N) {
            k.nwrite = N - base;
            done = true;
            continue;
        }
        k.nwrite = n;
        base += n;
        pos = k.ex[par];
        par = (int)((par + n) % (uint32_t)C);
    }
    if (a) atomicAdd(any_dirty, 1);
}

// Chain check of a single-table stream, one thread per chunk: chunk j starts where chunk j - 1 ends.
__global__ __launch_bounds__(256) void k_walk_parallel(Chunk* chunks, const int* par_chunks, int n,
                                                       const int* chunk_stream, const StreamPlan* streams,
                                                       int* any_dirty) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = par_chunks[i];
    const StreamPlan& sp = streams[chunk_stream[g]];
    const uint32_t want = g == sp.first ? 0u : chunks[g - 1].ex[0];
    Chunk& k = chunks[g];
    if (k.st != want) {
        k.st = want;
        k.dirty = 1;
        atomicAdd(any_dirty, 1);
    }
}

// Bases of a single-table stream once its chunks agree: an exclusive scan of the counts; the chunk that
// completes the N codes (or meets a code that consumes nothing, or is the last) writes the remainder.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_scan_parallel(Chunk* chunks, const StreamPlan* streams,
                                                                const int* list, const DSeg* segs) {
    const StreamPlan& sp = streams[list[blockIdx.x]];
    const uint32_t N = segs[sp.seg].n_sym;
    const int t = threadIdx.x;
    __shared__ uint64_t ssum[kScanThreads];
    __shared__ int first_end;
    const int per = (sp.n + kScanThreads - 1) / kScanThreads;
    const int j0 = t * per, j1 = imin(sp.n, j0 + per);
    uint64_t acc = 0;
    for (int j = j0; j < j1; ++j) acc += chunks[sp.first + j].cnt[0];
    ssum[t] = acc;
    if (t == 0) first_end = sp.n - 1;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {
        const uint64_t v = t >= o ? ssum[t - o] : 0;
        __syncthreads();
        ssum[t] += v;
        __syncthreads();
    }
    uint64_t base = t > 0 ? ssum[t - 1] : 0;
    for (int j = j0; j < j1; ++j) {
        Chunk& k = chunks[sp.first + j];
        if ((k.stuck & 1) || base + k.cnt[0] >= N) atomicMin(&first_end, j);
        k.base = (uint32_t)umin64(base, N);
        base += k.cnt[0];
    }
    __syncthreads();
    const int e = first_end;
    for (int j = j0; j < j1; ++j) {
        Chunk& k = chunks[sp.first + j];
        k.par = 0;
        k.used = j <= e;
        k.nwrite = j < e ? k.cnt[0] : j == e ? N - k.base : 0;
    }
}

// Every used chunk writes its codes' differences from its true start and entry phase.
template <bool kLds>
__global__ __launch_bounds__(256) void k_chunk_write(const Chunk* chunks, const BlockPlan* blocks,
                                                     const StreamPlan* streams, const DSeg* segs, const DFile* files,
                                                     const uint8_t* ubuf, const uint64_t* ulen,
                                                     const uint16_t* tables, uint16_t* diffs, int* status) {
    __shared__ uint16_t tlds[kLds ? kLdsTable : 1];
    const BlockPlan& bp = blocks[blockIdx.x];
    const StreamPlan& sp = streams[bp.stream];
    const DSeg& s = segs[sp.seg];
    int to[4];
    const uint16_t* tb = stage_tables<kLds>(s, tables, tlds, to);
    if ((int)threadIdx.x >= bp.n) return;
    const Chunk& k = chunks[bp.first + threadIdx.x];
    if (!k.used || k.nwrite == 0) return;
    uint64_t nbits;
    Cursor cur = stream_cursor(s, sp.seg, files, ubuf, ulen, nbits);
    const int C = s.clrs;
    const int large = s.large;
    cur.seek(k.st);
    int q = k.par;
    const uint32_t n = k.nwrite;
    // Differences are kept modulo 2^16: every later step (running sums, predictors on stored samples) reduces
    // modulo 2^16 before a value is used. Eight at a time go out as one 16-byte store.
    const uint64_t o0 = s.diff + k.base;
    uint16_t* out = diffs + o0;
    const uint32_t head = (uint32_t)((8 - (o0 & 7)) & 7);
    uint32_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    int32_t d = 0;
    bool stk = false;
    uint32_t i = 0;
    auto put = [&](uint32_t idx, uint16_t v) {
        if (idx < head || idx + (8 - ((idx - head) & 7)) > n) {
            out[idx] = v;  // unaligned head and the tail that does not fill a 16-byte group
            return;
        }
        const uint32_t w = (idx - head) & 7;
        const uint32_t hi = (uint32_t)v << 16;
        switch (w) {
            case 0: p0 = v; break;
            case 1: p0 |= hi; break;
            case 2: p1 = v; break;
            case 3: p1 |= hi; break;
            case 4: p2 = v; break;
            case 5: p2 |= hi; break;
            case 6: p3 = v; break;
            default: *reinterpret_cast<uint4*>(out + idx - 7) = make_uint4(p0, p1, p2, p3 | hi); break;
        }
    };
    for (; i < n; ++i) {
        const int tq = q == 0 ? to[0] : q == 1 ? to[1] : q == 2 ? to[2] : to[3];
        if (!decode_diff(cur, tb + tq, tb[tq], large, d, stk)) {
            raise_error(status, E_EOF, s.file);
            return;
        }
        put(i, (uint16_t)d);
        if (stk) break;
        if (++q == C) q = 0;
    }
    for (++i; i < n && stk; ++i) put(i, (uint16_t)d);  // the reference repeats a code that consumes nothing
}

// Chunk set-up: guessed starts every chunk_bits bits, everything dirty.
__global__ __launch_bounds__(256) void k_chunk_init(Chunk* chunks, int n_chunks, const StreamPlan* streams,
                                                    const int* chunk_stream) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_chunks) return;
    const StreamPlan& sp = streams[chunk_stream[g]];
    Chunk& k = chunks[g];
    k.st = (uint32_t)(g - sp.first) * (uint32_t)sp.chunk_bits;
    k.seg = sp.seg;
    k.stuck = 0;
    k.dirty = 1;
    k.par = 0;
    k.used = 0;
    k.decoded = 0;
    k.valid = 0;
    k.dec_st = 0;
    k.base = 0;
    k.nwrite = 0;
}

// ---------------------------------------------------------------------------------------------------
// Undo the linear predictors. Row starts: column x < C of row r is a running sum of the differences at
// (r', x) over the rows r' of its chain (PEF: rows of the same parity, R:PefFile/PEFFile.cs:111-113; lossless
// JPEG: the rows of its restart interval from 2^(bits-1), R:PefFile/DNGFile.cs:178-214). Each row is then a
// running sum per column phase x mod C (PEF :114-115; lossless JPEG predictor 1).
constexpr int kChainThreads = 1024;

__global__ __launch_bounds__(kChainThreads) void k_row_starts(const DUnit* units, const int* lin_units,
                                                             const uint16_t* diffs, int32_t* vchain) {
    const DUnit& u = units[lin_units[blockIdx.x]];
    const int t = threadIdx.x;
    const int C = u.clrs;
    const bool pef = u.kind == U_PEF_HUFF;
    const int R = pef ? 2 : 1;
    const int32_t v0 = pef ? 0 : (int32_t)(1u << ((u.bits - 1) & 31));
    __shared__ int32_t ssum[kChainThreads];
    __shared__ uint8_t sflag[kChainThreads];
    for (int c = 0; c < C && c < u.cols; ++c)
        for (int q = 0; q < R; ++q) {
            const int n = (u.rows - q + R - 1) / R;  // rows q, q + R, ...
            if (n <= 0) continue;
            const int per = (n + kChainThreads - 1) / kChainThreads;
            const int i0 = t * per, i1 = imin(n, i0 + per);
            // local segmented sums: a flag starts a new chain (value v0 + d)
            int32_t acc = 0;
            uint8_t fl = 0;
            for (int i = i0; i < i1; ++i) {
                const int r = q + i * R;
                const int32_t d = diffs[u.sym + (uint64_t)r * u.cols + c];
                const bool start = pef ? (i == 0) : (u.restart_rows ? r % u.restart_rows == 0 : r == 0);
                if (start) {
                    acc = (int32_t)((uint32_t)v0 + (uint32_t)d);
                    fl = 1;
                } else {
                    acc = (int32_t)((uint32_t)acc + (uint32_t)d);
                }
            }
            ssum[t] = acc;
            sflag[t] = fl;
            __syncthreads();
            for (int o = 1; o < kChainThreads; o <<= 1) {
                int32_t ps = 0;
                uint8_t pf = 0;
                if (t >= o) {
                    ps = ssum[t - o];
                    pf = sflag[t - o];
                }
                __syncthreads();
                if (t >= o && !sflag[t]) {
                    ssum[t] = (int32_t)((uint32_t)ssum[t] + (uint32_t)ps);
                    sflag[t] = pf;
                }
                __syncthreads();
            }
            int32_t run = t > 0 ? ssum[t - 1] : 0;
            for (int i = i0; i < i1; ++i) {
                const int r = q + i * R;
                const int32_t d = diffs[u.sym + (uint64_t)r * u.cols + c];
                const bool start = pef ? (i == 0) : (u.restart_rows ? r % u.restart_rows == 0 : r == 0);
                run = start ? (int32_t)((uint32_t)v0 + (uint32_t)d) : (int32_t)((uint32_t)run + (uint32_t)d);
                vchain[u.vchain + (uint64_t)r * C + c] = run;
            }
            __syncthreads();
        }
}

__device__ inline uint16_t level_transform(const DUnit& u, const uint16_t* lin, const uint16_t* bh,
                                           const uint16_t* bv, int v, int x, int y) {
    int o = v;
    if (u.lin_n > 0) o = lin[u.lin + (v < u.lin_n ? v : u.lin_n - 1)];
    o -= (int)bh[u.bh + x] + (int)bv[u.bv + y];
    return (uint16_t)(o > 0 ? o : 0);
}

constexpr int kRowWaves = 4;  // rows per 256-thread block, one wave per row

__device__ inline void row_scan(const DUnit& u, int r, int lane, const uint16_t* diffs, const int32_t* vchain,
                                const DFile* files, const uint16_t* lin, const uint16_t* bh, const uint16_t* bv,
                                int* status);

// One row of a linear unit per wave. The row is read in rounds of R = 64 - 64 mod C samples, lane l taking
// sample k R + l, whose column phase l mod C is then the same in every round; a strided scan over the lanes of
// each phase and a carry per phase give the running sums. Sample x < C starts its phase at the row start value.
__global__ __launch_bounds__(64 * kRowWaves) void k_rows(const DUnit* units, const int* row_unit,
                                                        const int* row_index, int n_rows, const uint16_t* diffs,
                                                        const int32_t* vchain, const DFile* files,
                                                        const uint16_t* lin, const uint16_t* bh, const uint16_t* bv,
                                                        int* status) {
    const int lane = threadIdx.x & 63;
    for (int g = blockIdx.x * kRowWaves + (threadIdx.x >> 6); g < n_rows; g += gridDim.x * kRowWaves)
        row_scan(units[row_unit[g]], row_index[g], lane, diffs, vchain, files, lin, bh, bv, status);
}

__device__ inline void row_scan(const DUnit& u, int r, int lane, const uint16_t* diffs, const int32_t* vchain,
                                const DFile* files, const uint16_t* lin, const uint16_t* bh, const uint16_t* bv,
                                int* status) {
    const int C = u.clrs;
    const int n = u.cols;
    const int R = 64 - 64 % C;
    const int c = lane % C;
    const uint16_t* d = diffs + u.sym + (uint64_t)r * n;
    const DFile& f = files[u.file];
    const int y = u.r0 + r;
    uint16_t* orow = reinterpret_cast<uint16_t*>(reinterpret_cast<uint8_t*>(f.out) + (size_t)y * f.pitch);
    const bool lj = u.kind == U_LJPEG;
    const int32_t start = lane < C && lane < n ? vchain[u.vchain + (uint64_t)r * C + lane] : 0;
    uint32_t carry = 0;
    bool bad = false;
    for (int x0 = 0; x0 < n; x0 += R) {
        const int x = x0 + lane;
        const bool in = lane < R && x < n;
        uint32_t v = !in ? 0u : x < C ? (uint32_t)start : (uint32_t)d[x];
        for (int o = C; o < R; o <<= 1) {
            const uint32_t w = __shfl_up(v, o, 64);
            if (lane >= o) v += w;
        }
        v += carry;
        carry = __shfl(v, R - C + c, 64);
        if (in) {
            const uint16_t v16 = (uint16_t)v;
            if (lj && (v16 >> (u.bits & 31)) != 0) bad = true;
            const int xo = u.c0 + x;
            if (xo < f.width && y < f.height)
                orow[xo] = u.transform ? level_transform(u, lin, bh, bv, v16, xo, y) : v16;
        }
    }
    if (bad) raise_error(status, E_RANGE, u.file);
}

// Lossless-JPEG predictors 2..7 (ljpeg_row, R:PefFile/DNGFile.cs:199-222): one lane per (unit, component).
__global__ void k_rows_nonlinear(const DUnit* units, const int* nl_units, int n_nl, const uint16_t* diffs,
                                 uint16_t* vals, const DFile* files, const uint16_t* lin, const uint16_t* bh,
                                 const uint16_t* bv, int* status) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_nl * 4) return;
    const DUnit& u = units[nl_units[g >> 2]];
    const int c = g & 3;
    if (c >= u.clrs) return;
    const int C = u.clrs, n = u.cols;
    const DFile& f = files[u.file];
    uint16_t* val = vals + u.vals;
    int32_t vpred = 0;
    for (int r = 0; r < u.rows; ++r) {
        if (u.restart_rows ? r % u.restart_rows == 0 : r == 0) vpred = (int32_t)(1u << ((u.bits - 1) & 31));
        const int y = u.r0 + r;
        uint16_t* orow = reinterpret_cast<uint16_t*>(reinterpret_cast<uint8_t*>(f.out) + (size_t)y * f.pitch);
        for (int x = c; x < n; x += C) {
            const int32_t d = diffs[u.sym + (uint64_t)r * n + x];
            int pred;
            if (x >= C) {
                pred = val[(uint64_t)r * n + x - C];
            } else {
                vpred = (int32_t)((uint32_t)vpred + (uint32_t)d);
                pred = (int32_t)((uint32_t)vpred - (uint32_t)d);
            }
            if (r > 0 && x >= C) {
                const int up = val[(uint64_t)(r - 1) * n + x], ul = val[(uint64_t)(r - 1) * n + x - C];
                switch (u.psv) {
                    case 1: break;
                    case 2: pred = up; break;
                    case 3: pred = ul; break;
                    case 4: pred = pred + up - ul; break;
                    case 5: pred = pred + ((up - ul) >> 1); break;
                    case 6: pred = up + ((pred - ul) >> 1); break;
                    case 7: pred = (pred + up) >> 1; break;
                    default: pred = 0; break;
                }
            }
            const uint16_t v16 = (uint16_t)(pred + d);
            val[(uint64_t)r * n + x] = v16;
            if ((v16 >> (u.bits & 31)) != 0) {
                raise_error(status, E_RANGE, u.file);
                return;
            }
            const int xo = u.c0 + x;
            if (xo < f.width && y < f.height) orow[xo] = level_transform(u, lin, bh, bv, v16, xo, y);
        }
    }
}

// Uncompressed strips: MSB-first packed samples (PEF R:PefFile/PEFFile.cs:131-168, DNG R:PefFile/DNGFile.cs:561-596)
// or little-endian 16-bit words (DNG R:PefFile/DNGFile.cs:518-551). One thread per sample.
__global__ void k_uncompressed(const DUnit* units, const int* un_units, const uint64_t* un_first, int n_un,
                               uint64_t total, const DFile* files, const uint16_t* lin, const uint16_t* bh,
                               const uint16_t* bv) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    int lo = 0, hi = n_un - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (un_first[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const DUnit& u = units[un_units[lo]];
    const uint64_t k = g - un_first[lo];  // sample index within the strip
    const DFile& f = files[u.file];
    const uint8_t* p = f.bytes + u.byte_off;
    const int y = u.r0 + (int)(k / u.cols), x = (int)(k % u.cols);
    uint32_t v;
    if (u.kind == U_LE16) {
        v = (uint32_t)p[2 * k] | (uint32_t)p[2 * k + 1] << 8;
    } else {
        const uint64_t b = k * u.bits;
        const uint64_t i0 = b >> 3;
        uint32_t w = 0;
        for (int i = 0; i < 4; ++i) w = w << 8 | (i0 + i < u.byte_len ? p[i0 + i] : 0u);
        v = (w << (b & 7)) >> (32 - u.bits);
    }
    if (y >= f.height || x >= f.width) return;
    uint16_t* orow = reinterpret_cast<uint16_t*>(reinterpret_cast<uint8_t*>(f.out) + (size_t)y * f.pitch);
    orow[x] = u.transform ? level_transform(u, lin, bh, bv, (int)v, x, y) : (uint16_t)v;
}

// =====================================================================================================
// Batch layout
// =====================================================================================================

struct Layout {
    size_t files, units, segs, tables, lin, bh, bv, lists, status, ulen, ubuf, diffs, vchain, vals;
    size_t chunks, chunk_stream, streams, walk_lists, dirty, total;
    size_t n_units = 0, n_segs = 0, n_tables = 0, n_lin = 0, n_bh = 0, n_bv = 0, n_lists = 0;
    size_t ubuf_bytes = 0, n_diffs = 0, n_vchain = 0, n_vals = 0, n_chunks = 0, n_specs = 0, n_spec_syms = 0;
    size_t specs = 0, spec_syms = 0, blocks = 0;
};

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Chunking of a stream: single-table streams in chunks of kChunkBits (chained by a parallel check and scan);
// streams that alternate tables in at most 2048 chunks (chained by one thread). The stream's bits are bounded
// by its stored bytes (a lossless-JPEG stream loses its stuffing bytes on the device).
void chunk_plan(const Seg& s, int& chunk_bits, int& n_chunks) {
    const uint64_t bits = std::max<uint64_t>(1, (uint64_t)s.len * 8);
    uint64_t cb = kChunkBits;
    if (bits <= 256ull * kChunkBits)
        cb = std::max<uint64_t>(2048, (bits + 255) / 256);  // a small stream fills one work-group
    else if (s.clrs > 1)
        cb = std::max<uint64_t>(cb, (bits + 2047) / 2048);
    cb = std::min<uint64_t>(cb, 1u << 30);
    chunk_bits = (int)cb;
    n_chunks = (int)std::max<uint64_t>(1, (bits + cb - 1) / cb);
}

Layout layout_of(const isa_raw_file* const* fs, int n) {
    Layout L{};
    size_t rows = 0;
    for (int i = 0; i < n; ++i) {
        const isa_raw_file& f = *fs[i];
        L.n_units += f.units.size();
        L.n_segs += f.segs.size();
        L.n_tables += f.tables.size();
        for (const TableSpec& t : f.ljtabs) {
            L.n_tables += 1 + ((size_t)1 << t.max);
            L.n_specs += 1;
            L.n_spec_syms += t.sym.size();
        }
        L.n_lin += f.lin.size();
        L.n_bh += f.bh.size();
        L.n_bv += f.bv.size();
        L.n_diffs += f.n_sym;
        for (const Seg& s : f.segs) {
            if (f.units[s.unit].kind == U_LJPEG) L.ubuf_bytes += align_up(s.len + 16, 16);
            int cb, nc;
            chunk_plan(s, cb, nc);
            L.n_chunks += nc;
        }
        for (const Unit& u : f.units) {
            if (u.kind == U_PEF_HUFF || (u.kind == U_LJPEG && u.psv == 1)) {
                L.n_vchain += (size_t)u.rows * u.clrs;
                rows += u.rows;
            }
            if (u.kind == U_LJPEG && u.psv != 1) L.n_vals += (size_t)u.rows * u.cols;
        }
    }
    // index lists: ljpeg segments, linear units, row->unit and row->index, nonlinear units, uncompressed units
    L.n_lists = 2 * L.n_segs + L.n_units * 3 + 2 * rows + 2;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    L.files = take(sizeof(DFile) * n);
    L.units = take(sizeof(DUnit) * L.n_units);
    L.segs = take(sizeof(DSeg) * L.n_segs);
    L.tables = take(2 * L.n_tables + 64);
    L.lin = take(2 * L.n_lin + 16);
    L.bh = take(2 * L.n_bh + 16);
    L.bv = take(2 * L.n_bv + 16);
    L.lists = take(8 * L.n_lists + 64);
    L.status = take(16);
    L.ulen = take(8 * L.n_segs + 8);
    L.ubuf = take(L.ubuf_bytes + 16);
    L.diffs = take(2 * L.n_diffs + 64);
    L.vchain = take(4 * L.n_vchain + 16);
    L.chunks = take(sizeof(Chunk) * L.n_chunks + 16);
    L.chunk_stream = take(4 * L.n_chunks + 16);
    L.streams = take(sizeof(StreamPlan) * L.n_segs + 16);
    L.walk_lists = take(4 * (2 * L.n_segs + L.n_chunks) + 64);
    L.blocks = take(sizeof(BlockPlan) * (L.n_chunks + L.n_segs) + 16);
    L.dirty = take(16);
    L.specs = take(sizeof(DSpec) * L.n_specs + 16);
    L.spec_syms = take(L.n_spec_syms + 16);
    L.vals = take(2 * L.n_vals + 16);
    L.total = o;
    return L;
}

thread_local int g_stats[2] = {0, 0};  // the last decode on this thread: chain walks, chunk re-decodes
thread_local int g_walks = 0;

int open_file(const void* file, size_t size, int format, isa_raw_file** out) {
    ISA_REQUIRE(file && out, "isa_raw_open: null argument");
    ISA_REQUIRE(format >= ISA_RAW_AUTO && format <= ISA_RAW_DNG, "isa_raw_open: format %d", format);
    *out = nullptr;
    auto f = std::make_unique<isa_raw_file>();
    f->size = size;
    const uint8_t* d = static_cast<const uint8_t*>(file);
    try {
        if (size < 8) throw DataError("file shorter than a TIFF header");
        bool le;
        if (d[0] == 'I' && d[1] == 'I')
            le = true;
        else if (d[0] == 'M' && d[1] == 'M')
            le = false;
        else
            throw DataError("Could not determine file endianess. Is this a proper TIFF/PEF/DNG file?");
        Reader r{d, size, le};
        if (r.u16(2) != 42) throw DataError("This is not a valid TIFF/PEF/DNG file: Magic number is not 42.");
        std::vector<Dir> ifds;
        uint32_t off = r.u32(4);
        for (int guard = 0;; ++guard) {
            if (guard > 64) throw DataError("IFD chain too long");
            ifds.push_back(read_ifd(r, off));
            if (ifds.back().next == 0) break;
            off = ifds.back().next;
        }
        if (format == ISA_RAW_AUTO)
            format = ifds[0].find(0xC612) || ifds[0].find(0xC614) ? ISA_RAW_DNG : ISA_RAW_PEF;
        f->info.format = format;
        if (format == ISA_RAW_PEF)
            parse_pef(*f, r, ifds);
        else
            parse_dng(*f, r, ifds);
        if (f->n_sym > 0xFFFFFFFFull) throw std::invalid_argument("more than 2^32 samples");
    } catch (const DataError& e) {
        isa::set_error("isa_raw_open: %s", e.what());
        return ISA_ERR_DATA;
    } catch (const std::invalid_argument& e) {
        isa::set_error("isa_raw_open: %s", e.what());
        return ISA_ERR_UNSUPPORTED;
    } catch (const std::exception& e) {
        isa::set_error("isa_raw_open: %s", e.what());
        return ISA_ERR_DATA;
    }
    *out = f.release();
    return ISA_OK;
}

}  // namespace

extern "C" {

int isa_raw_open(const void* file, size_t size, int format, isa_raw_file** out) {
    return open_file(file, size, format, out);
}

int isa_raw_get_info(const isa_raw_file* f, isa_raw_info* info) {
    ISA_REQUIRE(f && info, "isa_raw_get_info: null argument");
    *info = f->info;
    return ISA_OK;
}

void isa_raw_close(isa_raw_file* f) { delete f; }

int isa_raw_decode_stats(int* max_sync_walks, int* chunk_redecodes) {
    ISA_REQUIRE(max_sync_walks && chunk_redecodes, "isa_raw_decode_stats: null argument");
    *max_sync_walks = g_stats[0];
    *chunk_redecodes = g_stats[1];
    return ISA_OK;
}

int isa_raw_decode_workspace(const isa_raw_file* const* files, int n, size_t* bytes) {
    ISA_REQUIRE(files && bytes && n > 0, "isa_raw_decode_workspace: bad arguments");
    for (int i = 0; i < n; ++i) ISA_REQUIRE(files[i], "isa_raw_decode_workspace: file %d is null", i);
    *bytes = layout_of(files, n).total;
    return ISA_OK;
}

int isa_raw_decode(const isa_raw_file* const* files, int n, const void* const* file_dev, void* const* raw_dev,
                   const int* raw_pitch, void* workspace_dev, size_t workspace_bytes, isa_stream_t stream) {
    ISA_REQUIRE(files && file_dev && raw_dev && raw_pitch && n > 0, "isa_raw_decode: bad arguments");
    for (int i = 0; i < n; ++i) {
        ISA_REQUIRE(files[i] && file_dev[i] && raw_dev[i], "isa_raw_decode: file %d has a null pointer", i);
        ISA_REQUIRE(raw_pitch[i] >= 2 * files[i]->info.width && raw_pitch[i] % 2 == 0,
                    "isa_raw_decode: pitch of file %d", i);
        ISA_REQUIRE(reinterpret_cast<uintptr_t>(file_dev[i]) % 16 == 0, "isa_raw_decode: file %d not 16-byte aligned", i);
    }
    const Layout L = layout_of(files, n);
    ISA_REQUIRE(workspace_dev && workspace_bytes >= L.total, "isa_raw_decode: workspace %zu < %zu bytes",
                workspace_bytes, L.total);
    hipStream_t st = isa::to_stream(stream);
    uint8_t* ws = static_cast<uint8_t*>(workspace_dev);

    // host staging of the plan
    std::vector<DFile> hf(n);
    std::vector<DUnit> hu;
    std::vector<DSeg> hs;
    std::vector<uint16_t> ht, hl, hbh, hbv;
    std::vector<int> lj_segs, lin_units, row_unit, row_index, nl_units, un_units;
    std::vector<StreamPlan> hsp;
    std::vector<int> hcs, serial_streams, par_chunks, scan_streams;
    std::vector<BlockPlan> blk_lds, blk_glb;
    std::vector<uint64_t> un_first;
    std::vector<DSpec> hspec;
    std::vector<uint8_t> hsyms;
    uint64_t sym_base = 0, vchain_base = 0, vals_base = 0, ubuf_base = 0, un_total = 0;
    int max_rows = 0;
    // table array: the host-built (Pentax) tables of every file first, then the expanded lossless-JPEG tables
    size_t host_tab = 0;
    for (int i = 0; i < n; ++i) host_tab += files[i]->tables.size();
    size_t lj_tab = host_tab;
    for (int i = 0; i < n; ++i) {
        const isa_raw_file& f = *files[i];
        hf[i] = DFile{static_cast<const uint8_t*>(file_dev[i]), static_cast<uint16_t*>(raw_dev[i]), f.size,
                      raw_pitch[i], f.info.width, f.info.height, 0};
        const int unit0 = (int)hu.size(), tab0 = (int)ht.size();
        const int spec0 = (int)hspec.size();
        for (const TableSpec& t : f.ljtabs) {
            DSpec d{};
            d.dst = (int32_t)lj_tab;
            d.max = t.max;
            d.sym = (int32_t)hsyms.size();
            d.nsym = (int32_t)t.sym.size();
            for (int l = 0; l <= 16; ++l) d.count[l] = t.count[l];
            hsyms.insert(hsyms.end(), t.sym.begin(), t.sym.end());
            hspec.push_back(d);
            lj_tab += 1 + ((size_t)1 << t.max);
        }
        const int lin0 = (int)hl.size(), bh0 = (int)hbh.size(), bv0 = (int)hbv.size();
        for (const Unit& u : f.units) {
            DUnit d{};
            d.sym = sym_base + u.sym;
            d.byte_off = u.byte_off;
            d.byte_len = u.byte_len;
            d.kind = u.kind;
            d.file = i;
            d.rows = u.rows;
            d.cols = u.cols;
            d.r0 = u.r0;
            d.c0 = u.c0;
            d.clrs = u.clrs;
            d.bits = u.bits;
            d.psv = u.psv;
            d.restart_rows = u.restart_rows;
            d.transform = u.transform;
            d.lin_n = f.has_lin ? (int)f.lin.size() : 0;
            d.lin = lin0;
            d.bh = bh0;
            d.bv = bv0;
            const int idx = (int)hu.size();
            if (u.kind == U_PEF_HUFF || (u.kind == U_LJPEG && u.psv == 1)) {
                d.vchain = vchain_base;
                vchain_base += (uint64_t)u.rows * u.clrs;
                lin_units.push_back(idx);
                for (int r = 0; r < u.rows; ++r) {
                    row_unit.push_back(idx);
                    row_index.push_back(r);
                }
                max_rows = std::max(max_rows, u.rows);
            } else if (u.kind == U_LJPEG) {
                d.vals = vals_base;
                vals_base += (uint64_t)u.rows * u.cols;
                nl_units.push_back(idx);
            } else {
                un_units.push_back(idx);
                un_first.push_back(un_total);
                un_total += u.kind == U_LE16 ? u.byte_len / 2 : (uint64_t)u.rows * u.cols;
            }
            hu.push_back(d);
        }
        // lossless JPEG tables in their file (f.info.format == DNG; for PEF a bit-exact Pentax
        // table of 4097 entries with a 12-bit lookahead)
        for (const Seg& sg : f.segs) {
            DSeg d{};
            d.off = sg.off;
            d.len = sg.len;
            d.diff = sym_base + sg.sym;
            d.n_sym = sg.n_sym;
            d.file = i;
            d.unit = unit0 + sg.unit;
            d.clrs = sg.clrs;
            d.large = f.info.format == ISA_RAW_DNG && f.large ? 1 : 0;
            d.ljpeg = f.units[sg.unit].kind == U_LJPEG ? 1 : 0;
            for (int c = 0; c < 4; ++c)
                d.tab[c] = d.ljpeg ? (c < sg.clrs ? hspec[spec0 + sg.tab[c]].dst : 0) : tab0 + sg.tab[c];
            {
                StreamPlan sp{};
                sp.seg = (int)hs.size();
                sp.first = (int)hcs.size();
                chunk_plan(sg, sp.chunk_bits, sp.n);
                const int si = (int)hsp.size();
                for (int c = 0; c < sp.n; ++c) {
                    if (sg.clrs == 1) par_chunks.push_back((int)hcs.size());
                    hcs.push_back(si);
                }
                (sg.clrs == 1 ? scan_streams : serial_streams).push_back(si);
                hsp.push_back(sp);
                int entries = 0;
                for (int c = 0; c < sg.clrs; ++c)
                    entries += f.units[sg.unit].kind == U_LJPEG ? 1 + (1 << f.ljtabs.at(sg.tab[c]).max)
                                                               : (int)f.tables.size();
                for (int b0 = 0; b0 < sp.n; b0 += 256)
                    (entries <= kLdsTable ? blk_lds : blk_glb)
                        .push_back(BlockPlan{si, sp.first + b0, std::min(256, sp.n - b0), 0});
            }
            if (d.ljpeg) {
                d.ubuf = ubuf_base;
                d.ucap = align_up(sg.len + 16, 16);
                ubuf_base += d.ucap;
                lj_segs.push_back((int)hs.size());
            }
            hs.push_back(d);
        }
        ht.insert(ht.end(), f.tables.begin(), f.tables.end());
        hl.insert(hl.end(), f.lin.begin(), f.lin.end());
        hbh.insert(hbh.end(), f.bh.begin(), f.bh.end());
        hbv.insert(hbv.end(), f.bv.begin(), f.bv.end());
        sym_base += f.n_sym;
    }
    (void)max_rows;
    std::vector<int> lists;
    auto put_list = [&](const std::vector<int>& v) {
        size_t at = lists.size();
        lists.insert(lists.end(), v.begin(), v.end());
        return at;
    };
    const size_t o_lj = put_list(lj_segs), o_lin = put_list(lin_units), o_ru = put_list(row_unit),
                 o_ri = put_list(row_index), o_nl = put_list(nl_units), o_un = put_list(un_units);
    std::vector<int> walk;
    walk.insert(walk.end(), serial_streams.begin(), serial_streams.end());
    walk.insert(walk.end(), scan_streams.begin(), scan_streams.end());
    walk.insert(walk.end(), par_chunks.begin(), par_chunks.end());
    ISA_REQUIRE(hcs.size() == L.n_chunks && hsp.size() == L.n_segs, "isa_raw_decode: internal chunk plan");
    // un_first as a separate u64 array after the int lists
    std::vector<uint8_t> list_bytes(lists.size() * 4 + 8 + un_first.size() * 8);
    memcpy(list_bytes.data(), lists.data(), lists.size() * 4);
    const size_t o_unf = align_up(lists.size() * 4, 8);
    list_bytes.resize(o_unf + un_first.size() * 8 + 8);
    if (!un_first.empty()) memcpy(list_bytes.data() + o_unf, un_first.data(), un_first.size() * 8);
    ISA_REQUIRE(list_bytes.size() <= 8 * L.n_lists + 64, "isa_raw_decode: internal list size");

    auto up = [&](size_t at, const void* src, size_t bytes) {
        return bytes ? isa::check_hip(hipMemcpyAsync(ws + at, src, bytes, hipMemcpyHostToDevice, st), "raw plan upload")
                     : ISA_OK;
    };
    ISA_TRY(up(L.files, hf.data(), sizeof(DFile) * hf.size()));
    ISA_TRY(up(L.units, hu.data(), sizeof(DUnit) * hu.size()));
    ISA_TRY(up(L.segs, hs.data(), sizeof(DSeg) * hs.size()));
    ISA_TRY(up(L.tables, ht.data(), 2 * ht.size()));
    ISA_TRY(up(L.specs, hspec.data(), sizeof(DSpec) * hspec.size()));
    ISA_TRY(up(L.spec_syms, hsyms.data(), hsyms.size()));
    ISA_TRY(up(L.lin, hl.data(), 2 * hl.size()));
    ISA_TRY(up(L.bh, hbh.data(), 2 * hbh.size()));
    ISA_TRY(up(L.bv, hbv.data(), 2 * hbv.size()));
    ISA_TRY(up(L.lists, list_bytes.data(), list_bytes.size()));
    ISA_TRY(up(L.chunk_stream, hcs.data(), 4 * hcs.size()));
    ISA_TRY(up(L.streams, hsp.data(), sizeof(StreamPlan) * hsp.size()));
    ISA_TRY(up(L.walk_lists, walk.data(), 4 * walk.size()));
    std::vector<BlockPlan> blks(blk_lds);
    blks.insert(blks.end(), blk_glb.begin(), blk_glb.end());
    ISA_TRY(up(L.blocks, blks.data(), sizeof(BlockPlan) * blks.size()));

    const DFile* dF = reinterpret_cast<const DFile*>(ws + L.files);
    const DUnit* dU = reinterpret_cast<const DUnit*>(ws + L.units);
    const DSeg* dS = reinterpret_cast<const DSeg*>(ws + L.segs);
    const uint16_t* dT = reinterpret_cast<const uint16_t*>(ws + L.tables);
    const uint16_t* dL = reinterpret_cast<const uint16_t*>(ws + L.lin);
    const uint16_t* dBH = reinterpret_cast<const uint16_t*>(ws + L.bh);
    const uint16_t* dBV = reinterpret_cast<const uint16_t*>(ws + L.bv);
    const int* dLists = reinterpret_cast<const int*>(ws + L.lists);
    const uint64_t* dUnf = reinterpret_cast<const uint64_t*>(ws + L.lists + o_unf);
    int* dStatus = reinterpret_cast<int*>(ws + L.status);
    uint64_t* dUlen = reinterpret_cast<uint64_t*>(ws + L.ulen);
    uint8_t* dUbuf = ws + L.ubuf;
    uint16_t* dDiff = reinterpret_cast<uint16_t*>(ws + L.diffs);
    int32_t* dChain = reinterpret_cast<int32_t*>(ws + L.vchain);
    uint16_t* dVals = reinterpret_cast<uint16_t*>(ws + L.vals);

    int max_h = 0;
    for (int i = 0; i < n; ++i) max_h = std::max(max_h, files[i]->info.height);
    k_fill_out<<<dim3(std::max(1, std::min(1024, max_h)), n), 256, 0, st>>>(dF, n, dStatus);
    ISA_TRY(isa::check_launch("k_fill_out"));
    if (!hspec.empty()) {
        k_build_tables<<<(unsigned)hspec.size(), 256, 0, st>>>(reinterpret_cast<const DSpec*>(ws + L.specs),
                                                               ws + L.spec_syms, reinterpret_cast<uint16_t*>(ws + L.tables));
        ISA_TRY(isa::check_launch("k_build_tables"));
    }
    if (!lj_segs.empty()) {
        k_unstuff<<<(unsigned)lj_segs.size(), kUnstuffThreads, 0, st>>>(dS, dLists + o_lj, dF, dUbuf, dUlen);
        ISA_TRY(isa::check_launch("k_unstuff"));
    }
    // Chunked decode of the entropy-coded streams: decode the dirty chunks, walk the chains, until every
    // chunk starts where its predecessor ends (ulen, by segment, comes from k_unstuff).
    if (!hs.empty()) {
        Chunk* dC = reinterpret_cast<Chunk*>(ws + L.chunks);
        const int* dCS = reinterpret_cast<const int*>(ws + L.chunk_stream);
        const StreamPlan* dSP = reinterpret_cast<const StreamPlan*>(ws + L.streams);
        const int* dW = reinterpret_cast<const int*>(ws + L.walk_lists);
        int* dDirty = reinterpret_cast<int*>(ws + L.dirty);
        const BlockPlan* dB = reinterpret_cast<const BlockPlan*>(ws + L.blocks);
        const int nc = (int)hcs.size(), nser = (int)serial_streams.size(), nscan = (int)scan_streams.size();
        const int npar = (int)par_chunks.size();
        const unsigned cgrid = (unsigned)((nc + 255) / 256);
        k_chunk_init<<<cgrid, 256, 0, st>>>(dC, nc, dSP, dCS);
        ISA_TRY(isa::check_launch("k_chunk_init"));
        int max_chunks = 1;
        for (const StreamPlan& sp : hsp) max_chunks = std::max(max_chunks, sp.n);
        int walks = 0;
        for (;;) {
            if (!blk_lds.empty())
                k_chunk_decode<true><<<(unsigned)blk_lds.size(), 256, 0, st>>>(dC, dB, dSP, dS, dF, dUbuf, dUlen, dT,
                                                                               dStatus + 3);
            if (!blk_glb.empty())
                k_chunk_decode<false><<<(unsigned)blk_glb.size(), 256, 0, st>>>(dC, dB + blk_lds.size(), dSP, dS, dF,
                                                                                dUbuf, dUlen, dT, dStatus + 3);
            ISA_TRY(isa::check_launch("k_chunk_decode"));
            k_zero_int<<<1, 64, 0, st>>>(dDirty, 4);
            if (nser > 0) k_walk_serial<<<(unsigned)((nser + 63) / 64), 64, 0, st>>>(dC, dSP, dW, nser, dS, dDirty);
            if (npar > 0)
                k_walk_parallel<<<(unsigned)((npar + 255) / 256), 256, 0, st>>>(dC, dW + nser + nscan, npar, dCS,
                                                                                 dSP, dDirty);
            ISA_TRY(isa::check_launch("chunk walk"));
            int dirty = 0;
            ISA_TRY(isa::check_hip(hipMemcpyAsync(&dirty, dDirty, 4, hipMemcpyDeviceToHost, st), "raw walk read"));
            ISA_TRY(isa::check_hip(hipStreamSynchronize(st), "raw walk"));
            ++walks;
            if (dirty == 0) break;
            if (walks > max_chunks + 2) {
                isa::set_error("isa_raw_decode: the chunks of a stream did not synchronise");
                return ISA_ERR_DATA;
            }
        }
        g_walks = walks;
        if (nscan > 0) {
            k_scan_parallel<<<(unsigned)nscan, kScanThreads, 0, st>>>(dC, dSP, dW + nser, dS);
            ISA_TRY(isa::check_launch("k_scan_parallel"));
        }
        if (!blk_lds.empty())
            k_chunk_write<true><<<(unsigned)blk_lds.size(), 256, 0, st>>>(dC, dB, dSP, dS, dF, dUbuf, dUlen, dT,
                                                                          dDiff, dStatus);
        if (!blk_glb.empty())
            k_chunk_write<false><<<(unsigned)blk_glb.size(), 256, 0, st>>>(dC, dB + blk_lds.size(), dSP, dS, dF,
                                                                           dUbuf, dUlen, dT, dDiff, dStatus);
        ISA_TRY(isa::check_launch("k_chunk_write"));
    }
    if (!lin_units.empty()) {
        k_row_starts<<<(unsigned)lin_units.size(), kChainThreads, 0, st>>>(dU, dLists + o_lin, dDiff, dChain);
        ISA_TRY(isa::check_launch("k_row_starts"));
        const int n_rows = (int)row_unit.size();
        k_rows<<<(unsigned)std::min((n_rows + kRowWaves - 1) / kRowWaves, 4096), 64 * kRowWaves, 0, st>>>(
            dU, dLists + o_ru, dLists + o_ri, n_rows, dDiff, dChain, dF, dL, dBH, dBV, dStatus);
        ISA_TRY(isa::check_launch("k_rows"));
    }
    if (!nl_units.empty()) {
        const int lanes = (int)nl_units.size() * 4;
        k_rows_nonlinear<<<(lanes + 63) / 64, 64, 0, st>>>(dU, dLists + o_nl, (int)nl_units.size(), dDiff, dVals, dF,
                                                           dL, dBH, dBV, dStatus);
        ISA_TRY(isa::check_launch("k_rows_nonlinear"));
    }
    if (un_total > 0) {
        k_uncompressed<<<(unsigned)((un_total + 255) / 256), 256, 0, st>>>(dU, dLists + o_un, dUnf,
                                                                           (int)un_units.size(), un_total, dF, dL,
                                                                           dBH, dBV);
        ISA_TRY(isa::check_launch("k_uncompressed"));
    }
    int status[4] = {0, 0, 0, 0};
    ISA_TRY(isa::check_hip(hipMemcpyAsync(status, dStatus, 16, hipMemcpyDeviceToHost, st), "raw status read"));
    ISA_TRY(isa::check_hip(hipStreamSynchronize(st), "isa_raw_decode"));
    g_stats[0] = g_walks;
    g_stats[1] = status[3];
    if (status[0] != 0) {
        const char* what = status[0] == E_EOF     ? "the stream ends before the image is complete"
                           : status[0] == E_RANGE ? "a decoded value exceeds the stream's bit depth"
                                                  : "the chunks of a stream did not synchronise";
        isa::set_error("isa_raw_decode: file %d: %s", status[1], what);
        return ISA_ERR_DATA;
    }
    return ISA_OK;
}

}  // extern "C"
