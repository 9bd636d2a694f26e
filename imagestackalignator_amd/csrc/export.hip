// 16-bit RGB TIFF export (R:PefFile/ImageFileDirectory.cs:76-123, the writer SaveAs16BitTiff calls,
// R:PEFStudioDX/ImageStackAlignatorController.cs:2258-2259). Host code only.
//
// The reference writes a little-endian header with the IFD at offset 8, then its 13 entries in construction
// order (not sorted by tag), a zero next-IFD offset, then in a second pass the values that do not fit in an
// entry's 4 bytes (BitsPerSample's three SHORTs, the two RATIONAL resolutions, the "MK TIFF" string), each
// appended at the end of the file with its entry's offset field patched, and last the pixel data, whose
// position patches StripOffsets. With 13 entries the appended values sit at 170, 176, 184 and 192, and the
// pixels at 200.
#include <cstdio>
#include <cstring>
#include <vector>

#include "isa_host.h"

namespace {

struct TiffBytes {
    std::vector<unsigned char> b;
    void u16(unsigned v) {
        b.push_back((unsigned char)(v & 0xff));
        b.push_back((unsigned char)((v >> 8) & 0xff));
    }
    void u32(unsigned v) {
        for (int i = 0; i < 4; ++i) b.push_back((unsigned char)((v >> (8 * i)) & 0xff));
    }
    void put32(size_t at, unsigned v) {
        for (int i = 0; i < 4; ++i) b[at + i] = (unsigned char)((v >> (8 * i)) & 0xff);
    }
    // one 12-byte entry; returns the position of its value / offset field
    size_t entry(unsigned tag, unsigned type, unsigned count, unsigned value) {
        u16(tag);
        u16(type);
        u32(count);
        const size_t at = b.size();
        u32(value);
        return at;
    }
};

enum : unsigned { kAscii = 2, kShort = 3, kLong = 4, kRational = 5 };

// The header, the directory and its out-of-entry values: everything before the pixel data.
std::vector<unsigned char> tiff16_header(unsigned w, unsigned h) {
    TiffBytes t;
    const unsigned char head[8] = {0x49, 0x49, 0x2A, 0, 8, 0, 0, 0};
    t.b.assign(head, head + 8);
    t.u16(13);
    t.entry(256, kLong, 1, w);                              // ImageWidth
    t.entry(257, kLong, 1, h);                              // ImageLength
    const size_t bits = t.entry(258, kShort, 3, 0);          // BitsPerSample 16, 16, 16 (appended)
    t.entry(259, kShort, 1, 1);                             // Compression: none
    t.entry(262, kShort, 1, 2);                             // PhotometricInterpretation: RGB
    const size_t strip = t.entry(273, kLong, 1, 0);          // StripOffsets (patched last)
    t.entry(277, kShort, 1, 3);                             // SamplesPerPixel
    t.entry(278, kLong, 1, h);                              // RowsPerStrip
    t.entry(279, kLong, 1, w * h * 3u * 2u);                // StripByteCounts
    const size_t xres = t.entry(282, kRational, 1, 0);       // XResolution 720/10 (appended)
    const size_t yres = t.entry(283, kRational, 1, 0);       // YResolution 720/10 (appended)
    t.entry(296, kShort, 1, 1);                             // ResolutionUnit: none
    const size_t make = t.entry(271, kAscii, 8, 0);          // Make "MK TIFF" (appended)
    t.u32(0);                                               // no next IFD
    t.put32(bits, (unsigned)t.b.size());
    for (int i = 0; i < 3; ++i) t.u16(16);
    t.put32(xres, (unsigned)t.b.size());
    t.u32(720);
    t.u32(10);
    t.put32(yres, (unsigned)t.b.size());
    t.u32(720);
    t.u32(10);
    t.put32(make, (unsigned)t.b.size());
    const char* mk = "MK TIFF";
    for (int i = 0; i < 8; ++i) t.b.push_back((unsigned char)mk[i]);  // with its NUL
    t.put32(strip, (unsigned)t.b.size());
    return t.b;
}

}  // namespace

extern "C" {

size_t isa_tiff16_encode(void* dst, size_t capacity, const uint16_t* rgb_host, int width, int height) {
    if (width <= 0 || height <= 0 || width > 65535 || height > 65535) {
        isa::set_error("isa_tiff16_encode: bad size %d x %d", width, height);
        return 0;
    }
    const std::vector<unsigned char> head = tiff16_header((unsigned)width, (unsigned)height);
    const size_t pix = (size_t)width * height * 3;
    const size_t total = head.size() + 2 * pix;
    if (!dst) return total;
    if (capacity < total || !rgb_host) {
        isa::set_error("isa_tiff16_encode: %zu bytes needed, %zu given", total, capacity);
        return 0;
    }
    unsigned char* o = static_cast<unsigned char*>(dst);
    std::memcpy(o, head.data(), head.size());
    o += head.size();
    for (size_t i = 0; i < pix; ++i) {  // little endian on any host
        o[2 * i] = (unsigned char)(rgb_host[i] & 0xff);
        o[2 * i + 1] = (unsigned char)(rgb_host[i] >> 8);
    }
    return total;
}

int isa_write_tiff16(const char* path, const uint16_t* rgb_host, int width, int height) {
    ISA_REQUIRE(path && rgb_host, "isa_write_tiff16: null argument");
    const size_t n = isa_tiff16_encode(nullptr, 0, rgb_host, width, height);
    if (n == 0) return ISA_ERR_ARG;
    std::vector<unsigned char> buf(n);
    if (isa_tiff16_encode(buf.data(), n, rgb_host, width, height) != n) return ISA_ERR_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        isa::set_error("isa_write_tiff16: cannot open %s", path);
        return ISA_ERR_ARG;
    }
    const size_t wr = std::fwrite(buf.data(), 1, n, f);
    const int closed = std::fclose(f);
    if (wr != n || closed != 0) {
        isa::set_error("isa_write_tiff16: short write to %s", path);
        return ISA_ERR_ARG;
    }
    return ISA_OK;
}

}  // extern "C"
