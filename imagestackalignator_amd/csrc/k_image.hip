// Image-domain kernels: debayer, gamma/grey, Gaussian, box resize, structure
// tensor, kernel parameters, erosion, weighting. gfx950, FP32.
#include <vector>

#include "isa_device.h"
#include "isa_host.h"

namespace {

struct DebayerArgs {
    Cfa cfa;
    float black[3], scale[3], div[3];
    double inv_div[3];  // rd(1 / div[c])
    int fast_div;       // every div[c] in [2^-26, 2^26]: div_c may multiply by inv_div
};

// v / div[c], bit-identical to the IEEE float division. Let x = v = X 2^p, d = div[c] = D 2^q (X, D
// 24-bit integers) and m = M 2^k (M odd, 25 bits) a float rounding midpoint. x = d m is impossible
// (X would hold the odd 25-bit factor M), so x - d m is a nonzero multiple of 2^min(p, q + k) and
// |x/d - m| / m >= min(1/X, 1/(D M)) > 2^-49. The double product v * rd(1/d) is within 2^-52
// (relative) of x/d, so it rounds to the same float. The bound needs a normal quotient: |v| >= 2^-100
// (or v = 0) with d in [2^-26, 2^26] (fast_div); other v take the division.
__device__ __forceinline__ float div_c(float v, const DebayerArgs& a, int c) {
    const float av = fabsf(v);
    if (!a.fast_div || (av != 0.0f && av < 0x1p-100f)) return v / pick3(c, a.div[0], a.div[1], a.div[2]);
    return (float)((double)v * pick3(c, a.inv_div[0], a.inv_div[1], a.inv_div[2]));
}

// Green estimate at a pixel of CFA colour `site` (R:Kernels/DeBayerKernels.cu:53-149).
// RAW(x, y) returns the raw value as float.
template <class RawF>
__device__ __forceinline__ float green_at_site(int site, int x, int y, const DebayerArgs& a, RawF raw) {
    if (site == ISA_CFA_GREEN) return (raw(x, y) - a.black[1]) * a.scale[1];
    if (site != ISA_CFA_RED && site != ISA_CFA_BLUE) return 0.0f;
    int c = site == ISA_CFA_RED ? 0 : 2;
    float bc = pick3(c, a.black[0], a.black[1], a.black[2]), sc = pick3(c, a.scale[0], a.scale[1], a.scale[2]);
    float bg = a.black[1], sg = a.scale[1];
    float p = (raw(x, y) - bc) * sc;
    float xm2 = (raw(x - 2, y) - bc) * sc, xm1 = (raw(x - 1, y) - bg) * sg;
    float xp1 = (raw(x + 1, y) - bg) * sg, xp2 = (raw(x + 2, y) - bc) * sc;
    float ym2 = (raw(x, y - 2) - bc) * sc, ym1 = (raw(x, y - 1) - bg) * sg;
    float yp1 = (raw(x, y + 1) - bg) * sg, yp2 = (raw(x, y + 2) - bc) * sc;
    float gx = 0.5f * fabsf(xp1 - xm1);
    float gy = 0.5f * fabsf(yp1 - ym1);
    float lx = 0.25f * fabsf(2.0f * p - xm2 - xp2);
    float ly = 0.25f * fabsf(2.0f * p - ym2 - yp2);
    float ix = 0.125f * (-xm2 + 4.0f * xm1 + 2.0f * p + 4.0f * xp1 - xp2);
    float iy = 0.125f * (-ym2 + 4.0f * ym1 + 2.0f * p + 4.0f * yp1 - yp2);
    float w = (gy + ly) / (gx + gy + lx + ly + 0.000000001f);
    return w * ix + (1.0f - w) * iy;
}

template <class RawF>
__device__ __forceinline__ float green_at(int x, int y, const DebayerArgs& a, RawF raw) {
    return green_at_site(a.cfa.at(y, x), x, y, a, raw);
}

// Red and blue by colour-difference interpolation against green
// (R:Kernels/DeBayerKernels.cu:151-231) at a pixel of CFA colour `site` whose right-hand
// neighbour has colour `row`. G(x, y) returns the green plane.
template <class RawF, class GreenF>
__device__ __forceinline__ void redblue_at_site(int site, int row, int x, int y, const DebayerArgs& a, RawF raw,
                                                GreenF G, float& r, float& b) {
    float g = G(x, y);
    auto RC = [&](int c, int xx, int yy) {
        return (raw(xx, yy) - pick3(c, a.black[0], a.black[1], a.black[2])) * pick3(c, a.scale[0], a.scale[1], a.scale[2]);
    };
    r = 0.0f;
    b = 0.0f;
    if (site == ISA_CFA_GREEN) {
        int ch = row == ISA_CFA_RED ? 0 : 2, cv = row == ISA_CFA_RED ? 2 : 0;
        float hx = g + 0.5f * ((RC(ch, x - 1, y) - G(x - 1, y)) + (RC(ch, x + 1, y) - G(x + 1, y)));
        float vy = g + 0.5f * ((RC(cv, x, y - 1) - G(x, y - 1)) + (RC(cv, x, y + 1) - G(x, y + 1)));
        if (row == ISA_CFA_RED) { r = hx; b = vy; } else { b = hx; r = vy; }
    } else if (site == ISA_CFA_RED || site == ISA_CFA_BLUE) {
        int own = site == ISA_CFA_RED ? 0 : 2, other = site == ISA_CFA_RED ? 2 : 0;
        float v = RC(own, x, y);
        float d = g + 0.25f * ((RC(other, x - 1, y - 1) - G(x - 1, y - 1)) + (RC(other, x + 1, y - 1) - G(x + 1, y - 1)) +
                               (RC(other, x + 1, y + 1) - G(x + 1, y + 1)) + (RC(other, x - 1, y + 1) - G(x - 1, y + 1)));
        if (site == ISA_CFA_RED) { r = v; b = d; } else { b = v; r = d; }
    }
}

template <class RawF, class GreenF>
__device__ __forceinline__ void redblue_at(int x, int y, const DebayerArgs& a, RawF raw, GreenF G, float& r, float& b) {
    redblue_at_site(a.cfa.at(y, x), a.cfa.at(y, x + 1), x, y, a, raw, G, r, b);
}

__global__ void k_debayer_green(int w, int h, const float* __restrict__ in, int in_pitch, float* __restrict__ out,
                                int out_pitch, DebayerArgs a) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w - 2 || x < 2 || y >= h - 2 || y < 2) return;
    auto raw = [&](int xx, int yy) { return row_ptr(in, in_pitch, yy)[xx]; };
    row_ptr(out, out_pitch, y)[x * 3 + 1] = green_at(x, y, a, raw);
}

__global__ void k_debayer_redblue(int w, int h, const float* __restrict__ in, int in_pitch, float* out, int out_pitch,
                                  DebayerArgs a) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w - 2 || x < 2 || y >= h - 2 || y < 2) return;
    auto raw = [&](int xx, int yy) { return row_ptr(in, in_pitch, yy)[xx]; };
    auto G = [&](int xx, int yy) { return row_ptr(out, out_pitch, yy)[xx * 3 + 1]; };
    float r, b;
    redblue_at(x, y, a, raw, G, r, b);
    float* o = row_ptr(out, out_pitch, y) + x * 3;
    o[0] = r;
    o[2] = b;
}

// ---------------------------------------------------------------------------
// sRGB OETF power x^y, y = float(1/2.4), for x in (2^-9, 1] (the OETF's power branch starts at
// 0.0031308 > 2^-9). With x = 2^-e m, m in [1, 2), and c = 1 + (j + 1/2)/128 the centre of m's
// 1/128 interval: x^y = 2^(-e y) c^y (1 + r)^y with r = (m - c)/c, |r| <= 2^-8. The first two
// factors come from tables (double, within a few ulp), the last from its binomial series to r^5
// (remainder below 2^-53), so the double z is within 2^-48 (relative) of x^y. When z lies farther
// than 2^-40 from every rounding midpoint of its float, x^y and pow_r's double (within 2^-52 of
// x^y) both round to that float; otherwise pow_r itself is evaluated.
namespace srgb_pow {
constexpr float kY = 1.0f / 2.4f;
constexpr int kJ = 128, kE = 10;

constexpr double ln_1to2(double c) {  // ln c = 2 atanh((c - 1)/(c + 1)) for c in [1, 2]
    const double u = (c - 1.0) / (c + 1.0), u2 = u * u;
    double s = 0.0, t = u;
    for (int k = 0; k < 40; ++k) {
        s += t / (2 * k + 1);
        t *= u2;
    }
    return 2.0 * s;
}

constexpr double exp_small(double t) {  // e^t for |t| <= 1, Taylor series
    double s = 1.0, term = 1.0;
    for (int k = 1; k < 30; ++k) {
        term = term * t / k;
        s += term;
    }
    return s;
}

struct Tables {
    double mant[kJ][2];  // (c_j^y, 1 / c_j)
    double expo[kE];     // 2^(-e y)
};

constexpr Tables make_tables() {
    Tables T{};
    const double y = (double)kY;
    for (int j = 0; j < kJ; ++j) {
        const double c = 1.0 + (j + 0.5) / kJ;
        T.mant[j][0] = exp_small(y * ln_1to2(c));
        T.mant[j][1] = 1.0 / c;
    }
    for (int e = 0; e < kE; ++e) {  // 2^(-e y) = 2^-n 2^(n - e y), n = ceil(e y)
        const double t = e * y;
        int n = (int)t;
        if (n < t) ++n;
        double v = exp_small((n - t) * 0.69314718055994530942);
        for (int i = 0; i < n; ++i) v *= 0.5;
        T.expo[e] = v;
    }
    return T;
}

// binomial coefficients of (1 + r)^y
constexpr double kB1 = (double)kY, kB2 = kB1 * ((double)kY - 1.0) / 2.0, kB3 = kB2 * ((double)kY - 2.0) / 3.0,
                 kB4 = kB3 * ((double)kY - 3.0) / 4.0, kB5 = kB4 * ((double)kY - 4.0) / 5.0;
}  // namespace srgb_pow

__constant__ srgb_pow::Tables kSrgbPow = srgb_pow::make_tables();

// T: the tables (a per-block LDS copy where the power is evaluated per pixel, so that the
// per-lane lookups are LDS reads rather than vector-memory gathers).
__device__ __forceinline__ float pow_srgb(float x, const srgb_pow::Tables& T) {
    using namespace srgb_pow;
    const unsigned ux = __float_as_uint(x);
    const int e = 127 - (int)(ux >> 23);
    if (e < 0 || e >= kE) return pow_r(x, kY);
    const unsigned mb = ux & 0x7fffffu;
    const int j = (int)(mb >> 16);
    const double d = (double)((int)(mb & 0xffffu) - 0x8000) * 0x1p-23;  // m - c_j, exact
    const double r = d * T.mant[j][1];
    double q = fma(kB5, r, kB4);
    q = fma(q, r, kB3);
    q = fma(q, r, kB2);
    q = fma(q, r, kB1);
    q = fma(q, r, 1.0);
    const double z = T.expo[e] * T.mant[j][0] * q;
    const float f = (float)z;
    const unsigned uf = __float_as_uint(f);
    double half = __longlong_as_double((long long)((uf >> 23) + 872u) << 52);  // half an ulp of f (normal)
    if ((uf & 0x7fffffu) == 0 && z < (double)f) half *= 0.5;                   // the gap below 2^k is halved
    if (fabs(z - (double)f) < half - 0x1p-40 * z) return f;
    return pow_r(x, kY);
}

// sRGB OETF after NaN->0 and clamp to [0,1] (R:Kernels/kernel.cu:271-313). The power is pow_r's
// float (FAST evaluates it through pow_srgb).
template <bool FAST = true>
__device__ __forceinline__ float srgb_prepared(float v, const srgb_pow::Tables& T = kSrgbPow) {
    if (isnan(v)) v = 0.0f;
    v = fmaxf(fminf(v, 1.0f), 0.0f);
    if (v <= 0.0031308f) return 12.92f * v;
    return (1.0f + 0.055f) * (FAST ? pow_srgb(v, T) : pow_r(v, 1.0f / 2.4f)) - 0.055f;
}

// Fused DeBayerFullRes / tracking-image front end, one thread per 2x2 Bayer quad: the four CFA
// sites of a quad are the same in every thread, so the site branches of the debayer are uniform.
// A block covers DQ_X x DQ_Y quads; the green plane is formed in LDS for those quads and one ring
// of quads around them (the red/blue step reads green one pixel out).
// MODE 0: write the decoded float3 image (debayer / div, 0 on the 2-px border).
// MODE 1: write grey = coef . sRGB(decoded) (gamma + ColorToGray fused).
constexpr int DQ_X = 64, DQ_Y = 8;
constexpr int DQ_RW = 2 * DQ_X + 8, DQ_RH = 2 * DQ_Y + 8;  // raw tile: pixels x0-4 .. x0+2 DQ_X+3
constexpr int DQ_GW = 2 * DQ_X + 4, DQ_GH = 2 * DQ_Y + 4;  // green tile: pixels x0-2 .. x0+2 DQ_X+1
template <int MODE>
__global__ __launch_bounds__(DQ_X* DQ_Y) void k_debayer_quad(int w, int h, const uint16_t* __restrict__ raw,
                                                              int raw_pitch, float* __restrict__ out, int out_pitch,
                                                              DebayerArgs a, float c0, float c1, float c2) {
    __shared__ float s_raw[DQ_RH][DQ_RW + 1];
    __shared__ float s_g[DQ_GH][DQ_GW + 1];
    __shared__ srgb_pow::Tables s_pow;
    const int x0 = blockIdx.x * 2 * DQ_X, y0 = blockIdx.y * 2 * DQ_Y;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * DQ_X + tx;
    if (MODE == 1) {
        constexpr int nd = (int)(sizeof(srgb_pow::Tables) / sizeof(double));
        static_assert(nd <= DQ_X * DQ_Y, "one table word per thread");
        if (tid < nd) reinterpret_cast<double*>(&s_pow)[tid] = reinterpret_cast<const double*>(&kSrgbPow)[tid];
    }
    for (int ly = ty; ly < DQ_RH; ly += DQ_Y) {
        const uint16_t* r = row_ptr(raw, raw_pitch, clampi(y0 - 4 + ly, 0, h - 1));
        for (int lx = tx; lx < DQ_RW; lx += DQ_X) s_raw[ly][lx] = (float)r[clampi(x0 - 4 + lx, 0, w - 1)];
    }
    __syncthreads();
    auto R = [&](int xx, int yy) { return s_raw[yy - (y0 - 4)][xx - (x0 - 4)]; };
    for (int q = tid; q < (DQ_X + 2) * (DQ_Y + 2); q += DQ_X * DQ_Y) {
        const int gqy = q / (DQ_X + 2), gqx = q - gqy * (DQ_X + 2);
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int gx = x0 - 2 + 2 * gqx + dx, gy = y0 - 2 + 2 * gqy + dy;
                float g = 0.0f;  // never-written border of the decoded image reads as the Set(0) value
                if (gx >= 2 && gx < w - 2 && gy >= 2 && gy < h - 2) g = green_at_site(a.cfa.c[dy * 2 + dx], gx, gy, a, R);
                s_g[2 * gqy + dy][2 * gqx + dx] = g;
            }
    }
    __syncthreads();
    auto G = [&](int xx, int yy) { return s_g[yy - (y0 - 2)][xx - (x0 - 2)]; };
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
        const int y = y0 + 2 * ty + dy;
        float res[2][3];
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const int x = x0 + 2 * tx + dx;
            float v[3] = {0.0f, 0.0f, 0.0f};
            if (x >= 2 && x < w - 2 && y >= 2 && y < h - 2) {
                redblue_at_site(a.cfa.c[dy * 2 + dx], a.cfa.c[dy * 2 + (dx ^ 1)], x, y, a, R, G, v[0], v[2]);
                v[1] = G(x, y);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) res[dx][c] = div_c(v[c], a, c);
        }
        if (y >= h) continue;
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const int x = x0 + 2 * tx + dx;
            if (x >= w) continue;
            if (MODE == 0) {
                float* o = row_ptr(out, out_pitch, y) + x * 3;
                o[0] = res[dx][0];
                o[1] = res[dx][1];
                o[2] = res[dx][2];
            } else {
                float r = srgb_prepared(res[dx][0], s_pow), g = srgb_prepared(res[dx][1], s_pow),
                      b = srgb_prepared(res[dx][2], s_pow);
                row_ptr(out, out_pitch, y)[x] = c0 * r + c1 * g + c2 * b;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// Colour pipeline of the merged image (R:PEFStudioDX/ImageStackAlignatorController.cs:2198-2262,
// SaveAs16BitTiff; 2264-2322, ProcessColorsAndShow), one pass per output pixel instead of nine NPP calls:
// ThresholdLTGT to [0, CameraWhite], ColorTwist camera -> ProPhoto, ThresholdLTGT to [0, 1], LUTCubic (the
// tone curve), ColorTwist ProPhoto -> sRGB, Mul 2^Exposure, GammasRGB, Mul scale, the crop ROI and Convert
// with round-to-nearest-even and saturation, to RGB u16 (the TIFF) or BGRA u8 with alpha 255 (the display
// image). Every step is float arithmetic in the written order.
struct ColorArgs {
    float wb[3];     // upper bounds of the first threshold (CameraWhite)
    float m1[12];    // camera -> ProPhoto, 3 x 4 row-major: out_r = ((m[0] v0 + m[1] v1) + m[2] v2) + m[3]
    float m2[12];    // ProPhoto -> sRGB
    float amp;       // 2^Exposure
    float scale;     // 65535 (u16) or 255 (u8)
    int lut_n;       // tone-curve levels (4 .. kLutMax), x strictly increasing
    int crop_x, crop_y, w, h;  // output = the crop w x h at (crop_x, crop_y) of the input
    int bgra8;       // 0: RGB u16, 1: BGRA u8
};
constexpr int kLutMax = 1024;

__device__ __forceinline__ float twist_row(const float* m, float a, float b, float c) {
    return ((m[0] * a + m[1] * b) + m[2] * c) + m[3];
}

// Cubic through the four levels around v (Lagrange form, weights in float in a fixed order): the window
// starts one level left of v's interval and is clamped into [0, n - 4]; v is clamped to [x0, x(n-1)].
__device__ __forceinline__ float lut_cubic(float v, const float* lx, const float* ly, int n) {
    v = fminf(fmaxf(v, lx[0]), lx[n - 1]);
    int lo = 0, hi = n - 1;  // largest i with lx[i] <= v, i <= n - 2
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lx[mid] <= v) lo = mid;
        else hi = mid;
    }
    const int j0 = min(max(lo - 1, 0), n - 4);
    float r = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float wk = 1.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m)
            if (m != k) wk = wk * ((v - lx[j0 + m]) / (lx[j0 + k] - lx[j0 + m]));
        r = k == 0 ? ly[j0] * wk : r + ly[j0 + k] * wk;
    }
    return r;
}

__global__ __launch_bounds__(256) void k_color_pipeline(const float* __restrict__ img, int pitch, const float* lut_x,
                                                        const float* lut_y, void* out, int out_pitch, ColorArgs a) {
    __shared__ float sx[kLutMax], sy[kLutMax];
    for (int i = threadIdx.y * blockDim.x + threadIdx.x; i < a.lut_n; i += blockDim.x * blockDim.y) {
        sx[i] = lut_x[i];
        sy[i] = lut_y[i];
    }
    __syncthreads();
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= a.w || y >= a.h) return;
    const float* p = row_ptr(img, pitch, y + a.crop_y) + (size_t)(x + a.crop_x) * 3;
    float v[3] = {p[0], p[1], p[2]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (v[c] < 0.0f) v[c] = 0.0f;
        else if (v[c] > a.wb[c]) v[c] = a.wb[c];
    }
    float t[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) t[r] = twist_row(a.m1 + 4 * r, v[0], v[1], v[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (t[c] < 0.0f) t[c] = 0.0f;
        else if (t[c] > 1.0f) t[c] = 1.0f;
        t[c] = lut_cubic(t[c], sx, sy, a.lut_n);
    }
    float o[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) o[r] = twist_row(a.m2 + 4 * r, t[0], t[1], t[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = srgb_prepared(o[c] * a.amp) * a.scale;
    if (a.bgra8) {
        unsigned char* q = (unsigned char*)out + (size_t)y * out_pitch + (size_t)x * 4;
#pragma unroll
        for (int c = 0; c < 3; ++c) q[2 - c] = (unsigned char)fminf(fmaxf(rintf(o[c]), 0.0f), 255.0f);
        q[3] = 255;
    } else {
        unsigned short* q = (unsigned short*)((char*)out + (size_t)y * out_pitch) + (size_t)x * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) q[c] = (unsigned short)fminf(fmaxf(rintf(o[c]), 0.0f), 65535.0f);
    }
}

__global__ void k_gamma(float* img, int w, int h, int pitch) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    float* p = row_ptr(img, pitch, y) + x * 3;
    p[0] = srgb_prepared(p[0]);
    p[1] = srgb_prepared(p[1]);
    p[2] = srgb_prepared(p[2]);
}

template <bool FAST>
__global__ void k_srgb_1d(const float* in, float* out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = srgb_prepared<FAST>(in[i]);
}

__global__ void k_color_to_gray(const float* in, int in_pitch, float* out, int out_pitch, int w, int h, float c0,
                                float c1, float c2) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const float* p = row_ptr(in, in_pitch, y) + x * 3;
    row_ptr(out, out_pitch, y)[x] = c0 * p[0] + c1 * p[1] + c2 * p[2];
}

struct Taps {
    float t[99];
    int n;
};

// Separable Gaussian, replicate border: horizontal pass into tmp, vertical pass into out.
template <int CH>
__global__ void k_gauss_h(const float* in, int in_pitch, float* tmp, int w, int h, Taps taps) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const float* r = row_ptr(in, in_pitch, y);
    int c0 = taps.n / 2;
    // every channel's sum in tap order; the channels of a tap are read together
    float s[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) s[c] = 0.0f;
    for (int i = 0; i < taps.n; ++i) {
        const float* q = r + clampi(x + i - c0, 0, w - 1) * CH;
        const float t = taps.t[i];
#pragma unroll
        for (int c = 0; c < CH; ++c) s[c] += t * q[c];
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) tmp[((size_t)y * w + x) * CH + c] = s[c];
}

template <int CH>
__global__ void k_gauss_v(const float* tmp, float* out, int out_pitch, int w, int h, Taps taps) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    int c0 = taps.n / 2;
    float s[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) s[c] = 0.0f;
    for (int i = 0; i < taps.n; ++i) {
        const float* q = tmp + ((size_t)clampi(y + i - c0, 0, h - 1) * w + x) * CH;
        const float t = taps.t[i];
#pragma unroll
        for (int c = 0; c < CH; ++c) s[c] += t * q[c];
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) row_ptr(out, out_pitch, y)[x * CH + c] = s[c];
}

// Single-channel Gaussian fused in LDS (taps <= 15): same operation order as the two-pass form.
constexpr int GF_BX = 64, GF_BY = 16, GF_R = 7;
__global__ __launch_bounds__(GF_BX* GF_BY) void k_gauss_fused1(const float* __restrict__ in, int in_pitch,
                                                                float* __restrict__ out, int out_pitch, int w, int h,
                                                                Taps taps) {
    __shared__ float s_in[GF_BY + 2 * GF_R][GF_BX + 2 * GF_R];
    __shared__ float s_h[GF_BY + 2 * GF_R][GF_BX];
    const int c0 = taps.n / 2;
    const int x0 = blockIdx.x * GF_BX, y0 = blockIdx.y * GF_BY;
    const int tid = threadIdx.y * GF_BX + threadIdx.x;
    const int rows = GF_BY + 2 * c0, cols = GF_BX + 2 * c0;
    for (int i = tid; i < rows * cols; i += GF_BX * GF_BY) {
        int ly = i / cols, lx = i - ly * cols;
        int gx = clampi(x0 - c0 + lx, 0, w - 1), gy = clampi(y0 - c0 + ly, 0, h - 1);
        s_in[ly][lx] = row_ptr(in, in_pitch, gy)[gx];
    }
    __syncthreads();
    for (int i = tid; i < rows * GF_BX; i += GF_BX * GF_BY) {
        int ly = i / GF_BX, lx = i - ly * GF_BX;
        int gx = x0 + lx;
        float s = 0.0f;
        for (int k = 0; k < taps.n; ++k) {
            int xx = clampi(gx + k - c0, 0, w - 1);  // replicate at the image edge
            s += taps.t[k] * s_in[ly][xx - (x0 - c0)];
        }
        s_h[ly][lx] = s;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= w || y >= h) return;
    float s = 0.0f;
    for (int k = 0; k < taps.n; ++k) {
        int yy = clampi(y + k - c0, 0, h - 1);
        s += taps.t[k] * s_h[yy - (y0 - c0)][threadIdx.x];
    }
    row_ptr(out, out_pitch, y)[x] = s;
}

// The same filter for a compile-time tap count: the block's rows and columns (plus the tap
// radius, clamped at the image edge) are staged once, so no per-tap clamping or index division
// remains. Same products and sums in the same order as k_gauss_fused1.
template <int N>
__global__ __launch_bounds__(GF_BX* GF_BY) void k_gauss_fixed(const float* __restrict__ in, int in_pitch,
                                                              float* __restrict__ out, int out_pitch, int w, int h,
                                                              Taps taps) {
    constexpr int C = N / 2, RH = GF_BY + 2 * C, RW = GF_BX + 2 * C;
    __shared__ float s_in[RH][RW + 1];
    __shared__ float s_h[RH][GF_BX];
    const int x0 = blockIdx.x * GF_BX, y0 = blockIdx.y * GF_BY;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int ly = ty; ly < RH; ly += GF_BY) {
        const float* r = row_ptr(in, in_pitch, clampi(y0 - C + ly, 0, h - 1));
        for (int lx = tx; lx < RW; lx += GF_BX) s_in[ly][lx] = r[clampi(x0 - C + lx, 0, w - 1)];
    }
    float t[N];
#pragma unroll
    for (int k = 0; k < N; ++k) t[k] = taps.t[k];
    __syncthreads();
    for (int ly = ty; ly < RH; ly += GF_BY) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < N; ++k) acc += t[k] * s_in[ly][tx + k];
        s_h[ly][tx] = acc;
    }
    __syncthreads();
    const int x = x0 + tx, y = y0 + ty;
    if (x >= w || y >= h) return;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < N; ++k) acc += t[k] * s_h[ty + k][tx];
    row_ptr(out, out_pitch, y)[x] = acc;
}

// The same filter in registers for widths that are a multiple of 4 (16-byte aligned rows): lane L of a
// wave holds columns x..x+3 (x = 4 (64 blockIdx.x + L)) of one row as a float4 and takes its C
// horizontal neighbours on each side from lanes L - 1 and L + 1 (the wave's edge lanes load one more
// float4); a wave walks GR_ROWS output rows, keeping the horizontal results of the last N input rows.
// Rows and columns are clamped as in k_gauss_fixed, and the products and sums are the same in the
// same order, so the output is identical.
constexpr int GR_ROWS = 16, GR_WAVES = 4;

template <int N>
__global__ __launch_bounds__(64 * GR_WAVES) void k_gauss_rows(const float* __restrict__ in, int in_pitch,
                                                             float* __restrict__ out, int out_pitch, int w, int h,
                                                             Taps taps) {
    constexpr int C = N / 2;
    static_assert(C <= 4, "neighbours come from the adjacent lanes only");
    const int lane = threadIdx.x, x = 4 * (blockIdx.x * 64 + lane);
    const int y0 = (blockIdx.y * GR_WAVES + threadIdx.y) * GR_ROWS;
    if (y0 >= h) return;  // whole wave
    const bool live = x < w;
    const int xl = live ? x : w - 4;  // lanes past the row end load a valid vector, write nothing
    float t[N];
#pragma unroll
    for (int k = 0; k < N; ++k) t[k] = taps.t[k];
    const unsigned pf = (unsigned)in_pitch / 4u;
    auto row_h = [&](int ry, float hv[4]) {
        const unsigned ro = (unsigned)clampi(ry, 0, h - 1) * pf;
        const float4 v = *(const float4*)(in + ro + xl);
        float4 lv = v, rv = v;  // neighbour vectors x-4..x-1 and x+4..x+7
        if (lane == 0 && xl >= 4) lv = *(const float4*)(in + ro + xl - 4);
        if (lane == 63 && xl + 4 < w) rv = *(const float4*)(in + ro + xl + 4);
        float c[4 + 2 * C];
        c[C + 0] = v.x;
        c[C + 1] = v.y;
        c[C + 2] = v.z;
        c[C + 3] = v.w;
        const float u[4] = {__shfl_up(v.x, 1), __shfl_up(v.y, 1), __shfl_up(v.z, 1), __shfl_up(v.w, 1)};
        const float d[4] = {__shfl_down(v.x, 1), __shfl_down(v.y, 1), __shfl_down(v.z, 1), __shfl_down(v.w, 1)};
        const float lvv[4] = {lv.x, lv.y, lv.z, lv.w}, rvv[4] = {rv.x, rv.y, rv.z, rv.w};
#pragma unroll
        for (int k = 1; k <= C; ++k) {
            // column x - k (clamped to 0) and x + 3 + k (clamped to w - 1)
            c[C - k] = xl - k < 0 ? v.x : (lane == 0 ? lvv[4 - k] : u[4 - k]);
            c[C + 3 + k] = xl + 3 + k > w - 1 ? v.w : (lane == 63 ? rvv[k - 1] : d[k - 1]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < N; ++k) acc += t[k] * c[j + k];
            hv[j] = acc;
        }
    };
    float ring[N][4];
#pragma unroll
    for (int k = 0; k < N - 1; ++k) row_h(y0 - C + k, ring[k]);
    const int y1 = min(y0 + GR_ROWS, h);
    for (int y = y0; y < y1; ++y) {
        // ring[N - 1] <- input row y + C; ring[k] holds input row y - C + k
        row_h(y + C, ring[N - 1]);
        if (live) {
            float o[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float acc = 0.0f;
#pragma unroll
                for (int k = 0; k < N; ++k) acc += t[k] * ring[k][j];
                o[j] = acc;
            }
            *(float4*)(out + (unsigned)y * ((unsigned)out_pitch / 4u) + x) = make_float4(o[0], o[1], o[2], o[3]);
        }
#pragma unroll
        for (int k = 0; k + 1 < N; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) ring[k][j] = ring[k + 1][j];
    }
}

__global__ void k_resize_box(const float* in, int in_pitch, float* out, int out_pitch, int ow, int oh, int l) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= ow || y >= oh) return;
    float s = 0.0f;
    for (int j = 0; j < l; ++j) {
        const float* r = row_ptr(in, in_pitch, y * l + j) + x * l;
        for (int i = 0; i < l; ++i) s += r[i];
    }
    row_ptr(out, out_pitch, y)[x] = s / (float)(l * l);
}

// Box means of 2x2, 4x4 and 8x8 blocks in one pass (pyramid levels 2, 4 and 8; any output may be
// null): one thread per 8x8 block, loaded as 16 float4; every mean sums its pixels row by row in
// increasing x and divides by l^2, as k_resize_box, so the values are identical.
__global__ void k_resize_box248(const float* __restrict__ in, int in_pitch, float* o2, int p2, float* o4, int p4,
                                float* o8, int p8, int bw, int bh) {
    const int bx = blockIdx.x * blockDim.x + threadIdx.x, by = blockIdx.y * blockDim.y + threadIdx.y;
    if (bx >= bw || by >= bh) return;
    float v[8][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float4* r = (const float4*)(row_ptr(in, in_pitch, by * 8 + j) + bx * 8);
        const float4 a = r[0], b = r[1];
        v[j][0] = a.x; v[j][1] = a.y; v[j][2] = a.z; v[j][3] = a.w;
        v[j][4] = b.x; v[j][5] = b.y; v[j][6] = b.z; v[j][7] = b.w;
    }
    if (o8) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int i = 0; i < 8; ++i) s += v[j][i];
        row_ptr(o8, p8, by)[bx] = s / 64.0f;
    }
    if (o4)
#pragma unroll
        for (int qy = 0; qy < 2; ++qy) {
            float q[2];
#pragma unroll
            for (int qx = 0; qx < 2; ++qx) {
                float s = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i) s += v[qy * 4 + j][qx * 4 + i];
                q[qx] = s / 16.0f;
            }
            *(float2*)(row_ptr(o4, p4, by * 2 + qy) + bx * 2) = make_float2(q[0], q[1]);
        }
    if (o2)
#pragma unroll
        for (int qy = 0; qy < 4; ++qy) {
            float q[4];
#pragma unroll
            for (int qx = 0; qx < 4; ++qx) {
                float s = 0.0f;
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int i = 0; i < 2; ++i) s += v[qy * 2 + j][qx * 2 + i];
                q[qx] = s / 4.0f;
            }
            *(float4*)(row_ptr(o2, p2, by * 4 + qy) + bx * 4) = make_float4(q[0], q[1], q[2], q[3]);
        }
}

__global__ void k_structure_tensor(const float* dx, const float* dy, int dpitch, float* out, int out_pitch, int w,
                                   int h) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    float a = row_ptr(dx, dpitch, y)[x], b = row_ptr(dy, dpitch, y)[x];
    float* o = row_ptr(out, out_pitch, y) + x * 3;
    o[0] = a * a;
    o[1] = b * b;
    o[2] = a * b;
}

// ComputeKernelParam (R:Kernels/kernel.cu:627-700), in place; CH = 3 or 4 (the 4th channel is left alone).
template <int CH>
__global__ void k_kernel_param(float* img, int w, int h, int pitch, float Dth, float Dtr, float kDetail,
                               float kDenoise, float kStretch, float kShrink) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    float* p = row_ptr(img, pitch, y) + x * CH;
    float a11 = p[0], a22 = p[1], a12 = p[2];
    float help = sqrtf((a22 - a11) * (a22 - a11) + 4.0f * a12 * a12);
    float c = 2.0f * a12, s = a22 - a11 + help;
    float norm = sqrtf(c * c + s * s);
    if (norm > 0) { c /= norm; s /= norm; }
    else { c = 1.0f; s = 0.0f; }
    float lam1 = (a11 + a22 + help) / 2.0f, lam2 = (a11 + a22 - help) / 2.0f;
    float A = 1.0f + sqrtf((lam1 - lam2) * (lam1 - lam2) / ((lam1 + lam2) * (lam1 + lam2)));
    float D = 1.0f - sqrtf(lam1) / Dtr + Dth;
    D = fmaxf(fminf(1.0f, D), 0.0f);
    float k1h = kDetail * kStretch * A, k2h = kDetail / kShrink * A;
    float k1 = (1.0f - D) * k1h + D * kDetail * kDenoise;
    float k2 = (1.0f - D) * k2h + D * kDetail * kDenoise;
    k1 *= k1;
    k2 *= k2;
    float x2 = c, y2 = s, x1 = s, y1 = -c;
    float b11 = k1 * x1 * x1 + x2 * x2 * k2;
    float b12 = k1 * x1 * y1 + x2 * y2 * k2;
    float b22 = k1 * y1 * y1 + y2 * y2 * k2;
    float det = b11 * b22 - b12 * b12 + 0.0000000001f;
    p[0] = b22 / det;
    p[1] = b11 / det;
    p[2] = -b12 / det;
}

__global__ void k_erode4(const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int size) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    int r = size / 2;
    float m[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
    for (int j = -r; j <= size - 1 - r; ++j) {
        const float* row = row_ptr(in, in_pitch, clampi(y + j, 0, h - 1));
        for (int i = -r; i <= size - 1 - r; ++i) {
            const float* p = row + clampi(x + i, 0, w - 1) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) m[c] = fminf(m[c], p[c]);
        }
    }
    float* o = row_ptr(out, out_pitch, y) + x * 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = m[c];
}

// The same erosion (replicate border), separable (a min over the window's rows of mins over its
// columns is the min over the window; fminf ignores NaN in any order) and staged in LDS.
constexpr int ER_BX = 32, ER_BY = 16, ER_RMAX = 7;
__global__ __launch_bounds__(ER_BX* ER_BY) void k_erode_staged(const float* in, int in_pitch, float* out, int out_pitch,
                                                               int w, int h, int size) {
    __shared__ float4 s_in[ER_BY + 2 * ER_RMAX][ER_BX + 2 * ER_RMAX];
    __shared__ float4 s_h[ER_BY + 2 * ER_RMAX][ER_BX];
    const int r = size / 2, lo = -r, hi = size - 1 - r;  // window offsets lo..hi (as k_erode4)
    const int x0 = blockIdx.x * ER_BX, y0 = blockIdx.y * ER_BY, tx = threadIdx.x, ty = threadIdx.y;
    const int rows = ER_BY + hi - lo, cols = ER_BX + hi - lo;
    for (int ly = ty; ly < rows; ly += ER_BY) {
        const float4* row = (const float4*)row_ptr(in, in_pitch, clampi(y0 + lo + ly, 0, h - 1));
        for (int lx = tx; lx < cols; lx += ER_BX) s_in[ly][lx] = row[clampi(x0 + lo + lx, 0, w - 1)];
    }
    __syncthreads();
    auto min4 = [](float4 m, float4 v) {
        return make_float4(fminf(m.x, v.x), fminf(m.y, v.y), fminf(m.z, v.z), fminf(m.w, v.w));
    };
    for (int ly = ty; ly < rows; ly += ER_BY) {
        float4 m = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
        for (int k = 0; k <= hi - lo; ++k) m = min4(m, s_in[ly][tx + k]);
        s_h[ly][tx] = m;
    }
    __syncthreads();
    const int x = x0 + tx, y = y0 + ty;
    if (x >= w || y >= h) return;
    float4 m = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
    for (int k = 0; k <= hi - lo; ++k) m = min4(m, s_h[ty + k][tx]);
    ((float4*)row_ptr(out, out_pitch, y))[x] = m;
}

__global__ void k_apply_weighting(float* inout, const float* sum, const float* weight, int w, int h, int pitch,
                                  float thr) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    float* io = row_ptr(inout, pitch, y) + x * 3;
    const float* s = row_ptr(sum, pitch, y) + x * 3;
    const float* wt = row_ptr(weight, pitch, y) + x * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float val = s[c], ww = wt[c], v = io[c];
        if (ww < thr) { val += v; ww += 1.0f; }
        v = 0.0f;
        if (ww != 0.0f) v = val / ww;
        io[c] = v;
    }
}

// The same, elementwise (each channel value is weighted on its own): one float4 of a row's 3 w floats
// per thread, so a wave reads and writes 1 KB runs. Needs w % 4 == 0, 16-byte aligned rows and pointers.
__global__ void k_apply_weighting_x4(float* inout, const float* sum, const float* weight, int w, int h, int pitch,
                                     float thr) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (4 * j >= 3 * w || y >= h) return;
    float4* io = reinterpret_cast<float4*>(row_ptr(inout, pitch, y)) + j;
    const float4 s4 = reinterpret_cast<const float4*>(row_ptr(sum, pitch, y))[j];
    const float4 w4 = reinterpret_cast<const float4*>(row_ptr(weight, pitch, y))[j];
    const float4 v4 = *io;
    const float sv[4] = {s4.x, s4.y, s4.z, s4.w}, wv[4] = {w4.x, w4.y, w4.z, w4.w};
    float ov[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {  // the per-element operations of k_apply_weighting
        float val = sv[e], ww = wv[e], v = ov[e];
        if (ww < thr) { val += v; ww += 1.0f; }
        v = 0.0f;
        if (ww != 0.0f) v = val / ww;
        ov[e] = v;
    }
    *io = make_float4(ov[0], ov[1], ov[2], ov[3]);
}

DebayerArgs make_debayer_args(const float black[3], const float scale[3], const float div[3], const int cfa[4]) {
    DebayerArgs a;
    for (int i = 0; i < 4; ++i) a.cfa.c[i] = cfa[i];
    a.fast_div = 1;
    for (int c = 0; c < 3; ++c) {
        a.black[c] = black[c];
        a.scale[c] = scale[c];
        a.div[c] = div ? div[c] : 1.0f;
        a.inv_div[c] = 1.0 / (double)a.div[c];
        if (!(a.div[c] >= 0x1p-26f && a.div[c] <= 0x1p26f)) a.fast_div = 0;
    }
    return a;
}

bool cfa_ok(const int cfa[4]) {
    for (int i = 0; i < 4; ++i)
        if (cfa[i] < 0 || cfa[i] > 2) return false;
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers shared with the merger.
namespace isa {

int launch_raw_to_gray(const uint16_t* raw, int raw_pitch, float* gray, int gray_pitch, int w, int h,
                       const float black[3], const float white[3], const float camwhite[3], const int cfa[4],
                       int green_only, hipStream_t s) {
    float scale[3], div[3];
    for (int c = 0; c < 3; ++c) {
        scale[c] = 1.0f / camwhite[c];
        div[c] = white[c] / camwhite[c];
    }
    DebayerArgs a = make_debayer_args(black, scale, div, cfa);
    float c0 = green_only ? 0.0f : 0.2989f, c1 = green_only ? 1.0f : 0.5870f, c2 = green_only ? 0.0f : 0.1141f;
    dim3 grid(div_up(w, 2 * DQ_X), div_up(h, 2 * DQ_Y));
    k_debayer_quad<1><<<grid, dim3(DQ_X, DQ_Y), 0, s>>>(w, h, raw, raw_pitch, gray, gray_pitch, a, c0, c1, c2);
    return check_launch("raw_to_gray");
}

int launch_gauss1(const float* in, int in_pitch, float* out, int out_pitch, float* tmp, int w, int h,
                  const float* taps, int n, hipStream_t s);
bool resize_box248_ok(const float* in, int in_pitch, int w, int h);
int launch_resize_box248(const float* in, int in_pitch, float* o2, int p2, float* o4, int p4, float* o8, int p8, int w,
                         int h, hipStream_t s);

int launch_resize_box(const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int l, hipStream_t s) {
    int ow = w / l, oh = h / l;
    k_resize_box<<<dim3(div_up(ow, 32), div_up(oh, 8)), dim3(32, 8), 0, s>>>(in, in_pitch, out, out_pitch, ow, oh, l);
    return check_launch("resize_box");
}

int launch_kernel_param4(float* img, int w, int h, int pitch, float dth, float dtr, float kd, float kn, float ks,
                         float kh, hipStream_t s) {
    k_kernel_param<4><<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(img, w, h, pitch, dth, dtr, kd, kn, ks,
                                                                                kh);
    return check_launch("kernel_param4");
}

}  // namespace isa

static Taps make_taps(const float* t, int n) {
    Taps tp;
    tp.n = n;
    for (int i = 0; i < 99; ++i) tp.t[i] = i < n ? t[i] : 0.0f;
    return tp;
}

bool isa::resize_box248_ok(const float* in, int in_pitch, int w, int h) {
    return w % 8 == 0 && h % 8 == 0 && in_pitch % 16 == 0 && ((size_t)in & 15) == 0;
}

// Levels 2, 4 and 8 of a w x h image in one pass (resize_box248_ok; outputs may be null, their pitches
// multiples of 16, 8 and 4 bytes).
int isa::launch_resize_box248(const float* in, int in_pitch, float* o2, int p2, float* o4, int p4, float* o8, int p8,
                              int w, int h, hipStream_t s) {
    const int bw = w / 8, bh = h / 8;
    k_resize_box248<<<dim3(div_up(bw, 64), div_up(bh, 4)), dim3(64, 4), 0, s>>>(in, in_pitch, o2, p2, o4, p4, o8, p8,
                                                                                bw, bh);
    return check_launch("resize_box248");
}

int isa::launch_gauss1(const float* in, int in_pitch, float* out, int out_pitch, float* tmp, int w, int h,
                       const float* taps, int n, hipStream_t s) {
    Taps tp = make_taps(taps, n);
    const bool vec = w % 4 == 0 && w >= 8 && in_pitch % 16 == 0 && out_pitch % 16 == 0 && ((size_t)in & 15) == 0 &&
                     ((size_t)out & 15) == 0 && (size_t)in_pitch * h < ((size_t)1 << 32) &&
                     (size_t)out_pitch * h < ((size_t)1 << 32);
    if (vec && (n == 3 || n == 5 || n == 7)) {
        const dim3 g(div_up(w / 4, 64), div_up(h, GR_ROWS * GR_WAVES)), b(64, GR_WAVES);
        if (n == 3) k_gauss_rows<3><<<g, b, 0, s>>>(in, in_pitch, out, out_pitch, w, h, tp);
        if (n == 5) k_gauss_rows<5><<<g, b, 0, s>>>(in, in_pitch, out, out_pitch, w, h, tp);
        if (n == 7) k_gauss_rows<7><<<g, b, 0, s>>>(in, in_pitch, out, out_pitch, w, h, tp);
        return check_launch("gauss_rows");
    }
    const dim3 grid(div_up(w, GF_BX), div_up(h, GF_BY)), block(GF_BX, GF_BY);
    switch (n) {
        case 3: k_gauss_fixed<3><<<grid, block, 0, s>>>(in, in_pitch, out, out_pitch, w, h, tp); return check_launch("gauss3");
        case 5: k_gauss_fixed<5><<<grid, block, 0, s>>>(in, in_pitch, out, out_pitch, w, h, tp); return check_launch("gauss5");
        case 7: k_gauss_fixed<7><<<grid, block, 0, s>>>(in, in_pitch, out, out_pitch, w, h, tp); return check_launch("gauss7");
        default: break;
    }
    if (n <= 2 * GF_R + 1) {
        k_gauss_fused1<<<grid, block, 0, s>>>(in, in_pitch, out, out_pitch, w, h, tp);
        return check_launch("gauss_fused1");
    }
    dim3 b(32, 8), g(div_up(w, 32), div_up(h, 8));
    k_gauss_h<1><<<g, b, 0, s>>>(in, in_pitch, tmp, w, h, tp);
    k_gauss_v<1><<<g, b, 0, s>>>(tmp, out, out_pitch, w, h, tp);
    return check_launch("gauss1");
}

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

int isa_debayer_green(int width, int height, const float* in_dev, int in_pitch, float* out3_dev, int out_pitch,
                      const float black[3], const float scale[3], const int cfa[4], isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && in_dev && out3_dev && cfa_ok(cfa), "isa_debayer_green: bad arguments");
    DebayerArgs a = make_debayer_args(black, scale, nullptr, cfa);
    k_debayer_green<<<dim3(isa::div_up(width, 32), isa::div_up(height, 16)), dim3(32, 16), 0, isa::to_stream(stream)>>>(
        width, height, in_dev, in_pitch, out3_dev, out_pitch, a);
    return isa::check_launch("isa_debayer_green");
}

int isa_debayer_redblue(int width, int height, const float* in_dev, int in_pitch, float* out3_dev, int out_pitch,
                        const float black[3], const float scale[3], const int cfa[4], isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && in_dev && out3_dev && cfa_ok(cfa), "isa_debayer_redblue: bad arguments");
    DebayerArgs a = make_debayer_args(black, scale, nullptr, cfa);
    k_debayer_redblue<<<dim3(isa::div_up(width, 32), isa::div_up(height, 16)), dim3(32, 16), 0,
                        isa::to_stream(stream)>>>(width, height, in_dev, in_pitch, out3_dev, out_pitch, a);
    return isa::check_launch("isa_debayer_redblue");
}

int isa_debayer_full(int width, int height, const uint16_t* raw_dev, int raw_pitch, float* out3_dev, int out_pitch,
                     const float black[3], const float white[3], const float camera_white[3], const int cfa[4],
                     isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && raw_dev && out3_dev && cfa_ok(cfa), "isa_debayer_full: bad arguments");
    float scale[3], div[3];
    for (int c = 0; c < 3; ++c) {
        scale[c] = 1.0f / camera_white[c];
        div[c] = white[c] / camera_white[c];
    }
    DebayerArgs a = make_debayer_args(black, scale, div, cfa);
    dim3 grid(isa::div_up(width, 2 * DQ_X), isa::div_up(height, 2 * DQ_Y));
    k_debayer_quad<0><<<grid, dim3(DQ_X, DQ_Y), 0, isa::to_stream(stream)>>>(width, height, raw_dev, raw_pitch,
                                                                              out3_dev, out_pitch, a, 0.f, 0.f, 0.f);
    return isa::check_launch("isa_debayer_full");
}

int isa_gamma_srgb(float* img3_dev, int width, int height, int pitch, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && img3_dev, "isa_gamma_srgb: bad arguments");
    k_gamma<<<dim3(isa::div_up(width, 16), isa::div_up(height, 16)), dim3(16, 16), 0, isa::to_stream(stream)>>>(
        img3_dev, width, height, pitch);
    return isa::check_launch("isa_gamma_srgb");
}

int isa_srgb_oetf(const float* in_dev, float* out_dev, size_t count, int exact, isa_stream_t stream) {
    ISA_REQUIRE(in_dev && out_dev, "isa_srgb_oetf: bad arguments");
    if (count == 0) return ISA_OK;
    const unsigned blocks = (unsigned)((count + 255) / 256);
    if (exact)
        k_srgb_1d<false><<<blocks, 256, 0, isa::to_stream(stream)>>>(in_dev, out_dev, count);
    else
        k_srgb_1d<true><<<blocks, 256, 0, isa::to_stream(stream)>>>(in_dev, out_dev, count);
    return isa::check_launch("isa_srgb_oetf");
}

int isa_color_pipeline(const float* img3_dev, int img_pitch, int width, int height, const float camera_white[3],
                       const float camera_to_prophoto[12], const float* lut_x_dev, const float* lut_y_dev, int lut_n,
                       const float prophoto_to_srgb[12], float exposure, int crop_left, int crop_top, int crop_width,
                       int crop_height, int format, void* out_dev, int out_pitch, isa_stream_t stream) {
    ISA_REQUIRE(img3_dev && out_dev && lut_x_dev && lut_y_dev && camera_white && camera_to_prophoto && prophoto_to_srgb,
                "isa_color_pipeline: null argument");
    ISA_REQUIRE(width > 0 && height > 0 && img_pitch >= width * 12, "isa_color_pipeline: bad image size or pitch");
    ISA_REQUIRE(lut_n >= 4 && lut_n <= kLutMax, "isa_color_pipeline: the tone curve needs 4 .. %d levels", kLutMax);
    ISA_REQUIRE(format == ISA_COLOR_RGB16 || format == ISA_COLOR_BGRA8, "isa_color_pipeline: unknown format");
    ISA_REQUIRE(crop_left >= 0 && crop_top >= 0 && crop_width > 0 && crop_height > 0 &&
                    crop_left + crop_width <= width && crop_top + crop_height <= height,
                "isa_color_pipeline: crop outside the image");
    ISA_REQUIRE(out_pitch >= crop_width * (format == ISA_COLOR_RGB16 ? 6 : 4), "isa_color_pipeline: bad output pitch");
    ColorArgs a;
    for (int c = 0; c < 3; ++c) a.wb[c] = camera_white[c];
    for (int i = 0; i < 12; ++i) {
        a.m1[i] = camera_to_prophoto[i];
        a.m2[i] = prophoto_to_srgb[i];
    }
    a.amp = (float)std::pow(2.0, (double)exposure);
    a.scale = format == ISA_COLOR_RGB16 ? 65535.0f : 255.0f;
    a.lut_n = lut_n;
    a.crop_x = crop_left;
    a.crop_y = crop_top;
    a.w = crop_width;
    a.h = crop_height;
    a.bgra8 = format == ISA_COLOR_BGRA8;
    k_color_pipeline<<<dim3(isa::div_up(crop_width, 32), isa::div_up(crop_height, 8)), dim3(32, 8), 0,
                       isa::to_stream(stream)>>>(img3_dev, img_pitch, lut_x_dev, lut_y_dev, out_dev, out_pitch, a);
    return isa::check_launch("isa_color_pipeline");
}

int isa_color_to_gray(const float* in3_dev, int in_pitch, float* out_dev, int out_pitch, int width, int height,
                      const float coef[3], isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && in3_dev && out_dev, "isa_color_to_gray: bad arguments");
    k_color_to_gray<<<dim3(isa::div_up(width, 32), isa::div_up(height, 8)), dim3(32, 8), 0, isa::to_stream(stream)>>>(
        in3_dev, in_pitch, out_dev, out_pitch, width, height, coef[0], coef[1], coef[2]);
    return isa::check_launch("isa_color_to_gray");
}

int isa_gaussian_taps(float sigma, float* taps_out, int capacity) {
    std::vector<float> t;
    if (sigma <= 0) {
        t = {0, 0, 0, 0, 1, 0, 0, 0, 0};
    } else {
        int size = (int)(sigma / 0.6f - 0.4f) * 2 + 1 + 2;
        size = size < 99 ? size : 99;
        t.resize(size);
        int center = size / 2;
        for (int i = 0; i < size; ++i) {
            int x = i - center;
            float e = (float)(-(x * x)) / (2.0f * sigma * sigma);  // exponent in float, exp in double
            t[i] = (float)std::exp((double)e);
        }
        float sum = 0;
        for (int i = 0; i < size; ++i) sum += t[i];
        for (int i = 0; i < size; ++i) t[i] /= sum;
    }
    if ((int)t.size() > capacity) return -1;
    for (size_t i = 0; i < t.size(); ++i) taps_out[i] = t[i];
    return (int)t.size();
}

int isa_gauss_filter_border(const float* in_dev, int in_pitch, float* out_dev, int out_pitch, float* tmp_dev, int width,
                            int height, int channels, const float* taps, int n_taps, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && in_dev && out_dev && tmp_dev && taps && n_taps > 0 && n_taps <= 99,
                "isa_gauss_filter_border: bad arguments");
    ISA_REQUIRE(channels == 1 || channels == 3 || channels == 4, "isa_gauss_filter_border: channels must be 1, 3 or 4");
    hipStream_t s = isa::to_stream(stream);
    // one channel: the merger's filter (every form gives the same values)
    if (channels == 1) return isa::launch_gauss1(in_dev, in_pitch, out_dev, out_pitch, tmp_dev, width, height, taps,
                                                 n_taps, s);
    Taps tp = make_taps(taps, n_taps);
    dim3 b(32, 8), g(isa::div_up(width, 32), isa::div_up(height, 8));
    if (channels == 3) {
        k_gauss_h<3><<<g, b, 0, s>>>(in_dev, in_pitch, tmp_dev, width, height, tp);
        k_gauss_v<3><<<g, b, 0, s>>>(tmp_dev, out_dev, out_pitch, width, height, tp);
    } else {
        k_gauss_h<4><<<g, b, 0, s>>>(in_dev, in_pitch, tmp_dev, width, height, tp);
        k_gauss_v<4><<<g, b, 0, s>>>(tmp_dev, out_dev, out_pitch, width, height, tp);
    }
    return isa::check_launch("isa_gauss_filter_border");
}

int isa_resize_box(const float* in_dev, int in_pitch, float* out_dev, int out_pitch, int width, int height, int level,
                   isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && level >= 1 && in_dev && out_dev && width / level > 0 && height / level > 0,
                "isa_resize_box: bad arguments");
    return isa::launch_resize_box(in_dev, in_pitch, out_dev, out_pitch, width, height, level, isa::to_stream(stream));
}

int isa_resize_box_levels(const float* in_dev, int in_pitch, int width, int height, float* out2_dev, int pitch2,
                          float* out4_dev, int pitch4, float* out8_dev, int pitch8, isa_stream_t stream) {
    ISA_REQUIRE(in_dev && width > 0 && height > 0, "isa_resize_box_levels: bad arguments");
    ISA_REQUIRE(isa::resize_box248_ok(in_dev, in_pitch, width, height),
                "isa_resize_box_levels: width and height must be multiples of 8, the input 16-byte aligned");
    ISA_REQUIRE((!out2_dev || (pitch2 % 16 == 0 && ((size_t)out2_dev & 15) == 0)) &&
                    (!out4_dev || (pitch4 % 8 == 0 && ((size_t)out4_dev & 7) == 0)),
                "isa_resize_box_levels: level-2 rows must be 16-byte and level-4 rows 8-byte aligned");
    return isa::launch_resize_box248(in_dev, in_pitch, out2_dev, pitch2, out4_dev, pitch4, out8_dev, pitch8, width,
                                     height, isa::to_stream(stream));
}

int isa_compute_structure_tensor(const float* dx_dev, const float* dy_dev, int dxdy_pitch, float* out3_dev,
                                 int out_pitch, int width, int height, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && dx_dev && dy_dev && out3_dev, "isa_compute_structure_tensor: bad arguments");
    k_structure_tensor<<<dim3(isa::div_up(width, 32), isa::div_up(height, 8)), dim3(32, 8), 0,
                         isa::to_stream(stream)>>>(dx_dev, dy_dev, dxdy_pitch, out3_dev, out_pitch, width, height);
    return isa::check_launch("isa_compute_structure_tensor");
}

int isa_compute_kernel_param(float* kernel3_dev, int width, int height, int pitch, float dth, float dtr,
                             float k_detail, float k_denoise, float k_stretch, float k_shrink, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && kernel3_dev, "isa_compute_kernel_param: bad arguments");
    k_kernel_param<3><<<dim3(isa::div_up(width, 32), isa::div_up(height, 8)), dim3(32, 8), 0, isa::to_stream(stream)>>>(
        kernel3_dev, width, height, pitch, dth, dtr, k_detail, k_denoise, k_stretch, k_shrink);
    return isa::check_launch("isa_compute_kernel_param");
}

int isa_erode_border4(const float* in4_dev, int in_pitch, float* out4_dev, int out_pitch, int width, int height,
                      int size, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && size >= 1 && size <= 41 && in4_dev && out4_dev,
                "isa_erode_border4: bad arguments");
    if (size / 2 <= ER_RMAX && size - 1 - size / 2 <= ER_RMAX) {
        k_erode_staged<<<dim3(isa::div_up(width, ER_BX), isa::div_up(height, ER_BY)), dim3(ER_BX, ER_BY), 0,
                         isa::to_stream(stream)>>>(in4_dev, in_pitch, out4_dev, out_pitch, width, height, size);
        return isa::check_launch("isa_erode_border4");
    }
    k_erode4<<<dim3(isa::div_up(width, 32), isa::div_up(height, 8)), dim3(32, 8), 0, isa::to_stream(stream)>>>(
        in4_dev, in_pitch, out4_dev, out_pitch, width, height, size);
    return isa::check_launch("isa_erode_border4");
}


int isa_apply_weighting(float* inout3_dev, const float* sum3_dev, const float* weight3_dev, int width, int height,
                        int pitch, float threshold, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && inout3_dev && sum3_dev && weight3_dev, "isa_apply_weighting: bad arguments");
    auto al16 = [](const void* q) { return reinterpret_cast<uintptr_t>(q) % 16 == 0; };
    if (width % 4 == 0 && pitch % 16 == 0 && al16(inout3_dev) && al16(sum3_dev) && al16(weight3_dev))
        k_apply_weighting_x4<<<dim3(isa::div_up(3 * width / 4, 256), height), 256, 0, isa::to_stream(stream)>>>(
            inout3_dev, sum3_dev, weight3_dev, width, height, pitch, threshold);
    else
        k_apply_weighting<<<dim3(isa::div_up(width, 32), isa::div_up(height, 8)), dim3(32, 8), 0,
                            isa::to_stream(stream)>>>(inout3_dev, sum3_dev, weight3_dev, width, height, pitch,
                                                      threshold);
    return isa::check_launch("isa_apply_weighting");
}

}  // extern "C"
