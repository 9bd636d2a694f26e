// CPU oracle for the burst-merge hot path -- TEST INFRASTRUCTURE ONLY.
// See isa_oracle.hpp for the conventions (T = float reference semantics,
// T = double FP64 oracle). Build with -ffp-contract=off.
#include "isa_oracle.hpp"

#include <cstdio>

namespace ora {

int g_threads = 1;

// ---------------------------------------------------------------------------
// a1  Full-resolution debayer. R:Kernels/DeBayerKernels.cu:53-149 (green),
//     :151-231 (red/blue); driver R:PEFStudioDX/ImageStackAlignatorController.cs:1268-1311.
// in: W x H raw as T, out: W x H x 3 (caller pre-fills with 0; the 2-px
// border is never written).
template <class T>
void debayer_green(int w, int h, const T* in, T* out, const T black[3], const T scale[3], const int cfa[4]) {
    auto R = [&](int x, int y) { return (in[(size_t)y * w + x] - black[0]) * scale[0]; };
    auto G = [&](int x, int y) { return (in[(size_t)y * w + x] - black[1]) * scale[1]; };
    auto B = [&](int x, int y) { return (in[(size_t)y * w + x] - black[2]) * scale[2]; };
    parallel_for(h, [&](int y0, int y1) {
        for (int y = std::max(y0, 2); y < std::min(y1, h - 2); ++y)
            for (int x = 2; x < w - 2; ++x) {
                int site = cfa_at(cfa, y, x);
                T g = 0;
                if (site == CFA_GREEN) {
                    g = G(x, y);
                } else if (site == CFA_RED || site == CFA_BLUE) {
                    auto C = [&](int xx, int yy) { return site == CFA_RED ? R(xx, yy) : B(xx, yy); };
                    T p = C(x, y);
                    T xm2 = C(x - 2, y), xm1 = G(x - 1, y), xp1 = G(x + 1, y), xp2 = C(x + 2, y);
                    T ym2 = C(x, y - 2), ym1 = G(x, y - 1), yp1 = G(x, y + 1), yp2 = C(x, y + 2);
                    T gx = (T)0.5 * std::fabs(xp1 - xm1);
                    T gy = (T)0.5 * std::fabs(yp1 - ym1);
                    T lx = (T)0.25 * std::fabs((T)2 * p - xm2 - xp2);
                    T ly = (T)0.25 * std::fabs((T)2 * p - ym2 - yp2);
                    T ix = (T)0.125 * (-xm2 + (T)4 * xm1 + (T)2 * p + (T)4 * xp1 - xp2);
                    T iy = (T)0.125 * (-ym2 + (T)4 * ym1 + (T)2 * p + (T)4 * yp1 - yp2);
                    T wgt = (gy + ly) / (gx + gy + lx + ly + (T)0.000000001f);
                    g = wgt * ix + ((T)1 - wgt) * iy;
                }
                out[((size_t)y * w + x) * 3 + 1] = g;
            }
    });
}

template <class T>
void debayer_redblue(int w, int h, const T* in, T* out, const T black[3], const T scale[3], const int cfa[4]) {
    auto RAWC = [&](int c, int x, int y) { return (in[(size_t)y * w + x] - black[c]) * scale[c]; };
    auto GR = [&](int x, int y) { return out[((size_t)y * w + x) * 3 + 1]; };
    parallel_for(h, [&](int y0, int y1) {
        for (int y = std::max(y0, 2); y < std::min(y1, h - 2); ++y)
            for (int x = 2; x < w - 2; ++x) {
                int site = cfa_at(cfa, y, x);
                int row = cfa_at(cfa, y, x + 1);
                T g = GR(x, y);
                T r = 0, b = 0;
                if (site == CFA_GREEN) {
                    // horizontal neighbours carry the row's colour, vertical the other one
                    int ch = row == CFA_RED ? 0 : 2, cv = row == CFA_RED ? 2 : 0;
                    T hx = g + (T)0.5 * ((RAWC(ch, x - 1, y) - GR(x - 1, y)) + (RAWC(ch, x + 1, y) - GR(x + 1, y)));
                    T vy = g + (T)0.5 * ((RAWC(cv, x, y - 1) - GR(x, y - 1)) + (RAWC(cv, x, y + 1) - GR(x, y + 1)));
                    if (row == CFA_RED) { r = hx; b = vy; } else { b = hx; r = vy; }
                } else if (site == CFA_RED || site == CFA_BLUE) {
                    int own = site == CFA_RED ? 0 : 2, other = site == CFA_RED ? 2 : 0;
                    T v = RAWC(own, x, y);
                    T d = g + (T)0.25 * ((RAWC(other, x - 1, y - 1) - GR(x - 1, y - 1)) + (RAWC(other, x + 1, y - 1) - GR(x + 1, y - 1)) +
                                         (RAWC(other, x + 1, y + 1) - GR(x + 1, y + 1)) + (RAWC(other, x - 1, y + 1) - GR(x - 1, y + 1)));
                    if (site == CFA_RED) { r = v; b = d; } else { b = v; r = d; }
                }
                out[((size_t)y * w + x) * 3 + 0] = r;
                out[((size_t)y * w + x) * 3 + 2] = b;
            }
    });
}

// Whole DeBayerFullRes: decoded = 0; green; red/blue; per-channel division by
// white/cameraWhite (NPP Div). R:...Controller.cs:1268-1311.
template <class T>
void debayer_full(int w, int h, const uint16_t* raw, T* out, const T black[3], const T white[3], const T camwhite[3],
                  const int cfa[4]) {
    std::vector<T> in((size_t)w * h);
    for (size_t i = 0; i < in.size(); ++i) in[i] = (T)raw[i];
    std::fill(out, out + (size_t)w * h * 3, (T)0);
    T scale[3], div[3];
    for (int c = 0; c < 3; ++c) {
        scale[c] = (T)1 / camwhite[c];
        div[c] = white[c] / camwhite[c];
    }
    debayer_green(w, h, in.data(), out, black, scale, cfa);
    debayer_redblue(w, h, in.data(), out, black, scale, cfa);
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int c = 0; c < 3; ++c) out[i * 3 + c] = out[i * 3 + c] / div[c];
}

// ---------------------------------------------------------------------------
// a2  sRGB gamma (R:Kernels/kernel.cu:271-313) and NPP ColorToGray
//     (R:...Controller.cs:2404-2408).
template <class T> T srgb_oetf(T v) {
    if (v <= (T)0.0031308f) return (T)12.92f * v;
    return ((T)1.0f + (T)0.055f) * pow_r(v, (T)1.0f / (T)2.4f) - (T)0.055f;
}

template <class T> void gamma_srgb(int w, int h, T* img3) {
    parallel_for(h, [&](int y0, int y1) {
        for (size_t i = (size_t)y0 * w * 3; i < (size_t)y1 * w * 3; ++i) {
            T v = img3[i];
            if (std::isnan(v)) v = 0;
            v = std::fmax(std::fmin(v, (T)1), (T)0);
            img3[i] = srgb_oetf(v);
        }
    });
}

template <class T> void color_to_gray(int w, int h, const T* img3, T* gray, const T coef[3]) {
    for (size_t i = 0; i < (size_t)w * h; ++i)
        gray[i] = coef[0] * img3[i * 3] + coef[1] * img3[i * 3 + 1] + coef[2] * img3[i * 3 + 2];
}

// ---------------------------------------------------------------------------
// f3  Colour pipeline of the merged image (R:PEFStudioDX/ImageStackAlignatorController.cs:2198-2262,
//     SaveAs16BitTiff; 2264-2322, ProcessColorsAndShow): NPP ThresholdLTGT(0, 0, CameraWhite, CameraWhite),
//     ColorTwist camera -> ProPhoto, ThresholdLTGT(0, 0, 1, 1), LUTCubic through the tone curve, ColorTwist
//     ProPhoto -> sRGB, Mul 2^Exposure, GammasRGB, Mul 65535 (255), ROI, Convert (round to nearest even,
//     saturate). NPP's own rounding order is undocumented; the float operations run in the written order,
//     and LUTCubic is the cubic through the four levels around the value (Lagrange form).

// Tone curve: v clamped to [x0, x(n-1)], interval i = the largest with x[i] <= v (i <= n - 2), window of four
// levels from i - 1 clamped into [0, n - 4], weights prod_(m != k) (v - x_m) / (x_k - x_m) in float.
inline float lut_cubic(float v, const float* lx, const float* ly, int n) {
    v = std::fmin(std::fmax(v, lx[0]), lx[n - 1]);
    int i = 0;
    while (i + 1 < n - 1 && lx[i + 1] <= v) ++i;
    const int j0 = std::min(std::max(i - 1, 0), n - 4);
    float r = 0.0f;
    for (int k = 0; k < 4; ++k) {
        float wk = 1.0f;
        for (int m = 0; m < 4; ++m)
            if (m != k) wk = wk * ((v - lx[j0 + m]) / (lx[j0 + k] - lx[j0 + m]));
        r = k == 0 ? ly[j0] * wk : r + ly[j0 + k] * wk;
    }
    return r;
}

// Convert 32f -> integer with NPP_RND_NEAR (ties to even) and saturation to [0, maxv].
inline float convert_near(float v, float maxv) { return std::fmin(std::fmax(std::nearbyint(v), 0.0f), maxv); }

inline float twist_row(const float* m, float a, float b, float c) { return ((m[0] * a + m[1] * b) + m[2] * c) + m[3]; }

// img3: w x h float3; out: crop w x crop h, 3 x u16 (bgra8 = 0) or 4 x u8 (B, G, R, 255).
void color_pipeline(int w, int h, const float* img3, const float wb[3], const float m1[12], const float* lx,
                    const float* ly, int n, const float m2[12], float amp, int cx, int cy, int cw, int ch, int bgra8,
                    void* out) {
    const float scale = bgra8 ? 255.0f : 65535.0f;
    (void)h;
    for (int y = 0; y < ch; ++y)
        for (int x = 0; x < cw; ++x) {
            const float* p = img3 + ((size_t)(y + cy) * w + (x + cx)) * 3;
            float v[3] = {p[0], p[1], p[2]}, t[3], o[3];
            for (int c = 0; c < 3; ++c) {
                if (v[c] < 0.0f) v[c] = 0.0f;
                else if (v[c] > wb[c]) v[c] = wb[c];
            }
            for (int r = 0; r < 3; ++r) t[r] = twist_row(m1 + 4 * r, v[0], v[1], v[2]);
            for (int c = 0; c < 3; ++c) {
                if (t[c] < 0.0f) t[c] = 0.0f;
                else if (t[c] > 1.0f) t[c] = 1.0f;
                t[c] = lut_cubic(t[c], lx, ly, n);
            }
            for (int r = 0; r < 3; ++r) o[r] = twist_row(m2 + 4 * r, t[0], t[1], t[2]);
            for (int c = 0; c < 3; ++c) {
                float g = o[c] * amp;  // GammasRGB: NaN -> 0, clamp to [0, 1], OETF
                if (std::isnan(g)) g = 0.0f;
                g = std::fmax(std::fmin(g, 1.0f), 0.0f);
                o[c] = srgb_oetf(g) * scale;
            }
            if (bgra8) {
                unsigned char* q = (unsigned char*)out + ((size_t)y * cw + x) * 4;
                for (int c = 0; c < 3; ++c) q[2 - c] = (unsigned char)convert_near(o[c], 255.0f);
                q[3] = 255;
            } else {
                uint16_t* q = (uint16_t*)out + ((size_t)y * cw + x) * 3;
                for (int c = 0; c < 3; ++c) q[c] = (uint16_t)convert_near(o[c], 65535.0f);
            }
        }
}

// ---------------------------------------------------------------------------
// a3  Fourier high-pass (R:PEFStudioDX/PreAlignment.cs:268-281 with the
//     transfer function of R:Kernels/kernel.cu:702-779, lp = lps = 1).
// The transfer H(kx,ky) equals 1 exactly for every bin with dist >= hp+hps,
// so C2R(H X)/(W H) = x - C2R((1-H) X)/(W H), and (1-H) is supported on a
// small low-frequency block. The low-pass part is evaluated here by a
// partial DFT in double (exact up to double rounding), then clamped to [0,1].
template <class T> T fourier_filter_value(int kx, int ky, int w, int h, T lp, T hp, T lps, T hps, int clear_axis) {
    T mx = (T)kx, my = (T)ky;
    if (my > (T)h * (T)0.5f) my = ((T)h - my) * (T)-1.0f;
    mx /= (T)w;
    my /= (T)h;
    T dist = std::sqrt(mx * mx + my * my);
    T fil = 0;
    lp = lp - lps;
    hp = hp + hps;
    if (lp > 0) { if (dist <= lp) fil = 1; }
    else { if (dist <= (T)1.0f) fil = 1; }
    if (lps > 0) {
        T fil2 = (-fil + (T)1.0f) * exp_r(-((dist - lp) * (dist - lp) / ((T)2 * lps * lps)));
        if (fil2 > (T)0.001f) fil = fil2;
    }
    if (lps > 0 && lp == 0 && hp == 0 && hps == 0) fil = exp_r(-((dist - lp) * (dist - lp) / ((T)2 * lps * lps)));
    if (hp > 0) {
        T fil2 = 0;
        if (dist >= hp) fil2 = 1;
        fil *= fil2;
        if (hps > 0) {
            T fil3 = (-fil2 + (T)1.0f) * exp_r(-((dist - hp) * (dist - hp) / ((T)2 * hps * hps)));
            if (fil3 > (T)0.001f) fil = fil3;
        }
    }
    if (clear_axis > 0) {
        // the axis test of the reference runs in float whatever T is: (|ky| / h) * h rounds below ky for
        // many (h, ky), so bin |ky| = clear_axis is cleared as well there
        float myf = (float)ky;
        if (myf > (float)h * 0.5f) myf = ((float)h - myf) * -1.0f;
        myf /= (float)h;
        const float ay = std::fabs(myf) * (float)h;
        if (kx < clear_axis || ay < (float)clear_axis) fil = 0;
    }
    return fil;
}

// Returns the number of kx / |ky| bins that can carry (1 - H) != 0.
inline void highpass_band(int w, int h, double hp_total, int& kx_n, int& ky_n) {
    kx_n = std::min(w / 2 + 1, (int)std::ceil(hp_total * w) + 1);
    ky_n = std::min(h / 2 + 1, (int)std::ceil(hp_total * h) + 1);
}

template <class T> void fourier_highpass(int w, int h, T* img, T high_pass, T high_pass_sigma, int clear_axis) {
    const double two_pi = 6.283185307179586476925286766559;
    int kxn, kyn;
    highpass_band(w, h, (double)high_pass + (double)high_pass_sigma, kxn, kyn);
    if (clear_axis > 0) { kxn = w / 2 + 1; kyn = h / 2 + 1; }
    // ky set: 0..kyn-1 and h-kyn+1..h-1 (distinct)
    std::vector<int> kys;
    for (int k = 0; k < kyn && k < h; ++k) kys.push_back(k);
    for (int k = std::max(kyn, h - kyn + 1); k < h; ++k) kys.push_back(k);
    int ny = (int)kys.size();
    // row DFT: U[y][kx] = sum_x img[y][x] e^{-2 pi i kx x / w}
    std::vector<double> cw((size_t)w * kxn), sw((size_t)w * kxn);
    for (int x = 0; x < w; ++x)
        for (int k = 0; k < kxn; ++k) {
            double a = two_pi * (double)(((long long)k * x) % w) / (double)w;
            cw[(size_t)x * kxn + k] = std::cos(a);
            sw[(size_t)x * kxn + k] = std::sin(a);
        }
    std::vector<double> ur((size_t)h * kxn), ui((size_t)h * kxn);
    parallel_for(h, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int k = 0; k < kxn; ++k) {
                double sr = 0, si = 0;
                for (int x = 0; x < w; ++x) {
                    double v = (double)img[(size_t)y * w + x];
                    sr += v * cw[(size_t)x * kxn + k];
                    si -= v * sw[(size_t)x * kxn + k];
                }
                ur[(size_t)y * kxn + k] = sr;
                ui[(size_t)y * kxn + k] = si;
            }
    });
    // column DFT at the selected ky, times (1 - H) c_kx / (w h)
    std::vector<double> yr((size_t)ny * kxn), yi((size_t)ny * kxn);
    for (int j = 0; j < ny; ++j) {
        int ky = kys[j];
        for (int k = 0; k < kxn; ++k) {
            double sr = 0, si = 0;
            for (int y = 0; y < h; ++y) {
                double a = two_pi * (double)(((long long)ky * y) % h) / (double)h;
                double c = std::cos(a), s = std::sin(a);
                double vr = ur[(size_t)y * kxn + k], vi = ui[(size_t)y * kxn + k];
                sr += vr * c + vi * s;
                si += vi * c - vr * s;
            }
            double fil = (double)fourier_filter_value<T>(k, ky, w, h, (T)1, high_pass, (T)1, high_pass_sigma, clear_axis);
            double ck = (k == 0 || (w % 2 == 0 && k == w / 2)) ? 1.0 : 2.0;
            double g = (1.0 - fil) * ck / ((double)w * (double)h);
            yr[(size_t)j * kxn + k] = sr * g;
            yi[(size_t)j * kxn + k] = si * g;
        }
    }
    // inverse column: V[y][kx] = sum_j Y[j][kx] e^{+2 pi i ky y / h}
    std::vector<double> vr((size_t)h * kxn), vi((size_t)h * kxn);
    parallel_for(h, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int k = 0; k < kxn; ++k) {
                double sr = 0, si = 0;
                for (int j = 0; j < ny; ++j) {
                    double a = two_pi * (double)(((long long)kys[j] * y) % h) / (double)h;
                    double c = std::cos(a), s = std::sin(a);
                    sr += yr[(size_t)j * kxn + k] * c - yi[(size_t)j * kxn + k] * s;
                    si += yr[(size_t)j * kxn + k] * s + yi[(size_t)j * kxn + k] * c;
                }
                vr[(size_t)y * kxn + k] = sr;
                vi[(size_t)y * kxn + k] = si;
            }
    });
    // inverse row (real part) and subtraction, then clamp to [0,1] (NPP ThresholdLTGT(0,0,1,1))
    parallel_for(h, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < w; ++x) {
                double low = 0;
                for (int k = 0; k < kxn; ++k)
                    low += vr[(size_t)y * kxn + k] * cw[(size_t)x * kxn + k] - vi[(size_t)y * kxn + k] * sw[(size_t)x * kxn + k];
                double v = (double)img[(size_t)y * w + x] - low;
                T o = (T)v;
                o = o < (T)0 ? (T)0 : (o > (T)1 ? (T)1 : o);
                img[(size_t)y * w + x] = o;
            }
    });
}

// ---------------------------------------------------------------------------
// a4  Gaussian taps (R:...Controller.cs:2351-2382) and NPP FilterGaussBorder
//     with replicate border: horizontal pass, then vertical pass.
template <class T> std::vector<T> gaussian_taps(T sigma) {
    if (sigma <= 0) return std::vector<T>{0, 0, 0, 0, 1, 0, 0, 0, 0};
    int size = (int)(sigma / (T)0.6f - (T)0.4f) * 2 + 1 + 2;
    size = std::min(size, 99);
    std::vector<T> t(size);
    int center = size / 2;
    for (int i = 0; i < size; ++i) {
        int x = i - center;
        T e = (T)(-(x * x)) / ((T)2 * sigma * sigma);  // C# evaluates the exponent in float
        t[i] = (T)std::exp((double)e);
    }
    T sum = 0;
    for (int i = 0; i < size; ++i) sum += t[i];
    for (int i = 0; i < size; ++i) t[i] /= sum;
    return t;
}

template <class T> void gauss_filter(int w, int h, int ch, const T* in, T* out, const std::vector<T>& taps) {
    int n = (int)taps.size(), c0 = n / 2;
    std::vector<T> tmp((size_t)w * h * ch);
    parallel_for(h, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < ch; ++c) {
                    T s = 0;
                    for (int i = 0; i < n; ++i) {
                        int xx = std::min(std::max(x + i - c0, 0), w - 1);
                        s += taps[i] * in[((size_t)y * w + xx) * ch + c];
                    }
                    tmp[((size_t)y * w + x) * ch + c] = s;
                }
    });
    parallel_for(h, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < ch; ++c) {
                    T s = 0;
                    for (int i = 0; i < n; ++i) {
                        int yy = std::min(std::max(y + i - c0, 0), h - 1);
                        s += taps[i] * tmp[((size_t)yy * w + x) * ch + c];
                    }
                    out[((size_t)y * w + x) * ch + c] = s;
                }
    });
}

// E2: DeBayerBWGaussWB (R:...Controller.cs:2384-2418): debayer, gamma, grey,
// optional high-pass, Gaussian blur.
template <class T>
void tracking_image(int w, int h, const uint16_t* raw, T* out, const T black[3], const T white[3], const T camwhite[3],
                    const int cfa[4], int green_only, int use_high_pass, T high_pass, T high_pass_sigma, int clear_axis,
                    T sigma) {
    std::vector<T> dec((size_t)w * h * 3), gray((size_t)w * h);
    debayer_full(w, h, raw, dec.data(), black, white, camwhite, cfa);
    gamma_srgb(w, h, dec.data());
    T coef[3] = {(T)0.2989f, (T)0.5870f, (T)0.1141f};
    if (green_only) { coef[0] = 0; coef[1] = 1; coef[2] = 0; }
    color_to_gray(w, h, dec.data(), gray.data(), coef);
    if (use_high_pass) fourier_highpass(w, h, gray.data(), high_pass, high_pass_sigma, clear_axis);
    gauss_filter(w, h, 1, gray.data(), out, gaussian_taps(sigma));
}

// ---------------------------------------------------------------------------
// a5  NPP ResizeSqrPixel(1/l, SuperSampling) at an integer factor: l x l box
//     mean; row-major sum, then division by l*l.
template <class T> void downscale_box(int w, int h, const T* in, int level, T* out) {
    int ow = w / level, oh = h / level;
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            T s = 0;
            for (int j = 0; j < level; ++j)
                for (int i = 0; i < level; ++i) s += in[(size_t)(y * level + j) * w + x * level + i];
            out[(size_t)y * ow + x] = s / (T)(level * level);
        }
}

// ---------------------------------------------------------------------------
// a6  Tile extraction. R:Kernels/kernel.cu:171-228 (reference, zero border),
//     :230-288 (track, pre-shifted). Tiles are (T+2M)^2, tile-major.
template <class T>
static void tile_offset(int tx, int ty, int tile, int w, int h, T bsx, T bsy, T rot, T psx, T psy, int& ox, int& oy) {
    T sf = sin_r(rot), cf = cos_r(rot);
    T sx = psx + (cf * -bsx - sf * -bsy);
    T sy = psy + (sf * -bsx + cf * -bsy);
    T pcx = (T)(tx * tile + tile / 2 - w / 2);
    T pcy = (T)(ty * tile + tile / 2 - h / 2);
    sx += cf * pcx - sf * pcy - pcx;
    sy += sf * pcx + cf * pcy - pcy;
    ox = (int)std::round(sx);
    oy = (int)std::round(sy);
}

template <class T>
void tiles_border(int w, int h, const T* img, T* tiles, int M, int tile, int tcx, int tcy, T bsx, T bsy, T rot) {
    int B = tile + 2 * M;
    for (int t = 0; t < tcx * tcy; ++t) {
        int ty = t / tcx, tx = t - ty * tcx;
        int ox, oy;
        tile_offset<T>(tx, ty, tile, w, h, bsx, bsy, rot, 0, 0, ox, oy);
        for (int py = 0; py < B; ++py)
            for (int px = 0; px < B; ++px) {
                T v = 0;
                if (!(px < M || py < M || px >= tile + M || py >= tile + M)) {
                    int ix = std::min(std::max(tx * tile + px + ox, 0), w - 1);
                    int iy = std::min(std::max(ty * tile + py + oy, 0), h - 1);
                    v = img[(size_t)iy * w + ix];
                }
                tiles[(size_t)t * B * B + py * B + px] = v;
            }
    }
}

template <class T>
void tiles_preshift(int w, int h, const T* img, T* tiles, const T* preshift /*tcy x tcx x 2*/, int M, int tile, int tcx,
                    int tcy, T bsx, T bsy, T rot) {
    int B = tile + 2 * M;
    for (int t = 0; t < tcx * tcy; ++t) {
        int ty = t / tcx, tx = t - ty * tcx;
        int ox, oy;
        tile_offset<T>(tx, ty, tile, w, h, bsx, bsy, rot, preshift[(size_t)t * 2], preshift[(size_t)t * 2 + 1], ox, oy);
        for (int py = 0; py < B; ++py)
            for (int px = 0; px < B; ++px) {
                int ix = std::min(std::max(tx * tile + px + ox, 0), w - 1);
                int iy = std::min(std::max(ty * tile + py + oy, 0), h - 1);
                tiles[(size_t)t * B * B + py * B + px] = img[(size_t)iy * w + ix];
            }
    }
}

// ---------------------------------------------------------------------------
// a7+a8  Tile SSD. The reference forms D(s) = sum r^2 + box(t^2)(s) - 2 cc(s)
// through cuFFT (R:PEFStudioDX/PatchTracker.cs:327-348; R:Kernels/kernel.cu:27-169).
// Because the reference tile is zero outside its T x T interior and |s| <= M,
// the circular correlation never wraps, so every term is a plain sum over the
// interior. Evaluated in double in a canonical, column-wise order (r(y, x) is
// the interior, t the full (T+2M)^2 block, i = s + M):
//   R2[x]     = sum_y r(y,x)^2                       (fma chain in increasing y)
//   TT[x'][0] = sum_{y<T} t[y][x']^2                 (fma chain in increasing y)
//   TT[x'][k+1] = (TT[x'][k] - t[k][x']^2) + t[k+T][x']^2   (squares exact in double)
//   C[x][i]   = sum_y r(y,x) t[y+iy][x+ix]            (fma chain in increasing y)
//   Q[x][i]   = fma(-2, C[x][i], TT[x+ix][iy])
//   D(s)      = tree_x(R2) + tree_x(Q[.][i]), with the pairwise tree
//   for stride = 1, 2, 4, ...: v[x] += v[x + stride] for x a multiple of 2 stride.
// Output per tile: (2M+1)^2 values, index iy*(2M+1) + ix.
static void tree_sum(std::vector<double>& v, int n, int stride_elems, int count) {
    // v holds n rows of `count` values each (row stride stride_elems); sums rows into row 0
    for (int st = 1; st < n; st *= 2)
        for (int y = 0; y + st < n; y += 2 * st)
            for (int c = 0; c < count; ++c) v[(size_t)y * stride_elems + c] += v[(size_t)(y + st) * stride_elems + c];
}

template <class T>
void tile_ssd(const T* ref_tiles, const T* trk_tiles, T* out, int M, int tile, int tile_count) {
    int B = tile + 2 * M, S = 2 * M + 1;
    parallel_for(tile_count, [&](int t0, int t1) {
        std::vector<double> R2(tile), TT((size_t)B * S), Q((size_t)tile * S * S);
        for (int t = t0; t < t1; ++t) {
            const T* r = ref_tiles + (size_t)t * B * B;
            const T* k = trk_tiles + (size_t)t * B * B;
            auto rv = [&](int y, int x) { return (double)r[(M + y) * B + M + x]; };
            auto tv = [&](int j, int x) { return (double)k[j * B + x]; };
            for (int x = 0; x < tile; ++x) {
                double acc = 0;
                for (int y = 0; y < tile; ++y) acc = std::fma(rv(y, x), rv(y, x), acc);
                R2[x] = acc;
            }
            for (int xc = 0; xc < B; ++xc) {
                double acc = 0;
                for (int y = 0; y < tile; ++y) acc = std::fma(tv(y, xc), tv(y, xc), acc);
                TT[(size_t)xc * S] = acc;
                for (int q = 0; q + 1 < S; ++q) {
                    acc = (acc - tv(q, xc) * tv(q, xc)) + tv(q + tile, xc) * tv(q + tile, xc);
                    TT[(size_t)xc * S + q + 1] = acc;
                }
            }
            for (int x = 0; x < tile; ++x)
                for (int iy = 0; iy < S; ++iy)
                    for (int ix = 0; ix < S; ++ix) {
                        double acc = 0;
                        for (int y = 0; y < tile; ++y) acc = std::fma(rv(y, x), tv(y + iy, x + ix), acc);
                        Q[((size_t)x * S + iy) * S + ix] = std::fma(-2.0, acc, TT[(size_t)(x + ix) * S + iy]);
                    }
            tree_sum(R2, tile, 1, 1);
            tree_sum(Q, tile, S * S, S * S);
            for (int s = 0; s < S * S; ++s) out[(size_t)t * S * S + s] = (T)(R2[0] + Q[s]);
        }
    });
}

// ---------------------------------------------------------------------------
// a9  findMinimum (R:Kernels/kernel.cu:413-546): integer argmin (first, strict
//     <), 3x3 least-squares quadratic fit, border / |mu|>1 / threshold rules.
template <class T> void find_minimum(const T* shift_img, T* coords /*tiles x 2*/, int M, int tile_count, T threshold) {
    static const T FA11[9] = {(T)(1.0f / 4.0f), (T)(-2.0f / 4.0f), (T)(1.0f / 4.0f), (T)(2.0f / 4.0f), (T)(-4.0f / 4.0f),
                              (T)(2.0f / 4.0f), (T)(1.0f / 4.0f), (T)(-2.0f / 4.0f), (T)(1.0f / 4.0f)};
    static const T FA22[9] = {(T)(1.0f / 4.0f), (T)(2.0f / 4.0f), (T)(1.0f / 4.0f), (T)(-2.0f / 4.0f), (T)(-4.0f / 4.0f),
                              (T)(-2.0f / 4.0f), (T)(1.0f / 4.0f), (T)(2.0f / 4.0f), (T)(1.0f / 4.0f)};
    static const T FA12[9] = {(T)(1.0f / 4.0f), 0, (T)(-1.0f / 4.0f), 0, 0, 0, (T)(-1.0f / 4.0f), 0, (T)(1.0f / 4.0f)};
    static const T Fb1[9] = {(T)(-1.0f / 8.0f), 0, (T)(1.0f / 8.0f), (T)(-2.0f / 8.0f), 0, (T)(2.0f / 8.0f),
                             (T)(-1.0f / 8.0f), 0, (T)(1.0f / 8.0f)};
    static const T Fb2[9] = {(T)(-1.0f / 8.0f), (T)(-2.0f / 8.0f), (T)(-1.0f / 8.0f), 0, 0, 0,
                             (T)(1.0f / 8.0f), (T)(2.0f / 8.0f), (T)(1.0f / 8.0f)};
    int S = 2 * M + 1, n = S * S;
    for (int t = 0; t < tile_count; ++t) {
        const T* v = shift_img + (size_t)t * n;
        T mn = std::numeric_limits<T>::max(), mx = -std::numeric_limits<T>::max();
        int mi = -1;
        for (int i = 0; i < n; ++i) {
            mx = std::fmax(mx, v[i]);
            if (v[i] < mn) { mn = v[i]; mi = i; }
        }
        T cy = (T)(mi / S);
        T cx = (T)mi - cy * (T)S;
        if (cx < 1 || cy < 1 || cx >= (T)(2 * M) || cy >= (T)(2 * M)) {
            cx = 0;
            cy = 0;
        } else {
            T A11 = 0, A22 = 0, A12 = 0, b1 = 0, b2 = 0;
            for (int i = 0; i < 9; ++i) {
                int off = (i / 3 - 1) * S + (i % 3 - 1);
                T img = v[mi + off];
                A11 += FA11[i] * img;
                A22 += FA22[i] * img;
                A12 += FA12[i] * img;
                b1 += Fb1[i] * img;
                b2 += Fb2[i] * img;
            }
            A11 = std::fmax(A11, (T)0);
            A22 = std::fmax(A22, (T)0);
            T det = A11 * A22 - A12 * A12;
            if (det < 0) { A12 = 0; det = A11 * A22; }
            if (det != 0) {
                T mux = (A22 * b1 - A12 * b2) / det;
                T muy = (A11 * b2 - A12 * b1) / det;
                if (std::fabs(mux) > 1) mux = 0;
                if (std::fabs(muy) > 1) muy = 0;
                cx -= mux;
                cy -= muy;
            }
            cx -= (T)M;
            cy -= (T)M;
        }
        if (threshold + mn > mx) { cx = 0; cy = 0; }
        coords[(size_t)t * 2] = cx;
        coords[(size_t)t * 2 + 1] = cy;
    }
}

// ---------------------------------------------------------------------------
// a11  UpSampleShifts (R:Kernels/kernel.cu:551-598).
template <class T>
void upsample_shifts(const T* in, T* out, int old_level, int new_level, int ocx, int ocy, int ncx, int ncy, int old_tile,
                     int new_tile) {
    T factor = (T)old_level * (T)old_tile / (T)(new_level * new_tile);
    for (int by = 0; by < ncy; ++by)
        for (int bx = 0; bx < ncx; ++bx) {
            T ox = (T)bx / factor, oy = (T)by / factor;
            int x0 = (int)std::floor(ox), x1 = (int)std::ceil(ox), y0 = (int)std::floor(oy), y1 = (int)std::ceil(oy);
            x0 = std::min(x0, ocx - 1); x1 = std::min(x1, ocx - 1);
            y0 = std::min(y0, ocy - 1); y1 = std::min(y1, ocy - 1);
            for (int c = 0; c < 2; ++c) {
                T mm = in[((size_t)y0 * ocx + x0) * 2 + c], pm = in[((size_t)y0 * ocx + x1) * 2 + c];
                T mp = in[((size_t)y1 * ocx + x0) * 2 + c], pp = in[((size_t)y1 * ocx + x1) * 2 + c];
                T t1 = mm + (pm - mm) * ((T)1 - ((T)x1 - ox));
                T t2 = mp + (pp - mp) * ((T)1 - ((T)x1 - ox));
                T o = t1 + (t2 - t1) * ((T)1 - ((T)y1 - oy));
                o *= (T)old_level / (T)new_level;
                out[((size_t)by * ncx + bx) * 2 + c] = o;
            }
        }
}

// ---------------------------------------------------------------------------
// a12  Shift tables. R:PEFStudioDX/ShiftCollection.cs:113-496.
int shift_count(int n, int strategy, int block) {
    switch (strategy) {
        case STRAT_FULL: return (int)((n - 1) * ((float)n / 2.0f));
        case STRAT_ONLY_ON_REFERENCE: return n - 1;
        case STRAT_ON_REFERENCE_BLOCK: {
            int in_block = (int)((block - 1) * ((float)block / 2.0f));
            return in_block + block * (n - block);
        }
        case STRAT_BLOCKS: return (int)(-0.5 * block * (block - 2 * n + 1));
    }
    return 0;
}

static void ref_block_range(int n, int ref, int block, int& first, int& last) {
    if (ref < n / 2) {
        first = std::max(0, ref - block / 2);
        last = first + block - 1;
    } else {
        last = std::min(n - 1, ref + block / 2);
        first = last - block + 1;
    }
}

bool build_shift_tables(int n, int ref, int strategy, int block_in, ShiftTables& t) {
    int block = block_in >= n ? n - 1 : block_in;
    t.n = n; t.ref = ref; t.strategy = strategy; t.block = block;
    int m = shift_count(n, strategy, block), n1 = n - 1;
    t.m = m; t.n1 = n1;
    t.A.assign((size_t)m * n1, 0.f);
    t.pair_ref.clear(); t.pair_track.clear();
    t.index.assign((size_t)n * n, -1);
    auto A = [&](int row, int col) -> float& { return t.A[(size_t)col * m + row]; };
    if (strategy == STRAT_FULL || strategy == STRAT_BLOCKS) {
        int dists = strategy == STRAT_FULL ? n1 : block;
        int counter = 0;
        for (int s = 0; s < dists; ++s) {
            int count = n1 - s;
            for (int line = 0; line < count; ++line) {
                for (int l = line; l <= line + n1 - count; ++l) A(counter, l) = 1;
                ++counter;
            }
        }
        if (strategy == STRAT_FULL) {
            for (int a = 0; a < n; ++a)
                for (int b = a + 1; b < n; ++b) { t.pair_ref.push_back(a); t.pair_track.push_back(b); }
        } else {
            for (int s = 0; s < block; ++s) {
                int count = n1 - s;
                for (int line = 0; line < count; ++line) { t.pair_ref.push_back(line); t.pair_track.push_back(line + n - count); }
            }
        }
        int c = 0;
        for (int d = 1; d <= dists; ++d)
            for (int f = 0; f + d < n; ++f) t.index[(size_t)f * n + f + d] = c++;
    } else if (strategy == STRAT_ONLY_ON_REFERENCE) {
        for (int img = 0; img < n1; ++img) {
            if (img < ref)
                for (int i = img; i < ref; ++i) A(img, i) = -1;
            else
                for (int i = ref; i <= img; ++i) A(img, i) = 1;
        }
        int c = 0;
        for (int k = 0; k < n; ++k)
            if (k != ref) {
                t.pair_ref.push_back(ref); t.pair_track.push_back(k);
                t.index[(size_t)ref * n + k] = c++;
            }
    } else if (strategy == STRAT_ON_REFERENCE_BLOCK) {
        int first, last;
        ref_block_range(n, ref, block, first, last);
        int counter = 0;
        for (int s = 0; s < block - 1; ++s) {
            int count = block - 1 - s;
            for (int line = 0; line < count; ++line) {
                for (int l = line; l <= line + block - 1 - count; ++l) A(counter, l + first) = 1;
                ++counter;
            }
        }
        for (int i = 0; i < first; ++i)
            for (int tb = first; tb <= last; ++tb) {
                for (int j = i; j < tb; ++j) A(counter, j) = -1;
                ++counter;
            }
        for (int i = last + 1; i < n; ++i)
            for (int tb = first; tb <= last; ++tb) {
                for (int j = tb; j < i; ++j) A(counter, j) = 1;
                ++counter;
            }
        int c = 0;
        for (int s = 0; s < block - 1; ++s) {
            int count = block - 1 - s;
            for (int line = 0; line < count; ++line) {
                t.pair_ref.push_back(line + first); t.pair_track.push_back(line + first + block - count);
                t.index[(size_t)(line + first) * n + line + first + block - count] = c++;
            }
        }
        for (int i = 0; i < first; ++i)
            for (int tb = first; tb <= last; ++tb) {
                t.pair_ref.push_back(tb); t.pair_track.push_back(i);
                t.index[(size_t)tb * n + i] = c++;
            }
        for (int i = last + 1; i < n; ++i)
            for (int tb = first; tb <= last; ++tb) {
                t.pair_ref.push_back(tb); t.pair_track.push_back(i);
                t.index[(size_t)tb * n + i] = c++;
            }
    } else {
        return false;
    }
    if ((int)t.pair_ref.size() != m) return false;  // R:ShiftCollection.cs:522-525
    // derive (a, b) per row: the non-zero columns of a row form one interval
    // [lo, hi] with one sign s, i.e. the row measures s (x_{hi+1} - x_lo).
    t.row_a.assign(m, 0); t.row_b.assign(m, 0);
    for (int k = 0; k < m; ++k) {
        int lo = -1, hi = -1; float s = 0;
        for (int c = 0; c < n1; ++c) {
            float v = A(k, c);
            if (v != 0) {
                if (lo < 0) { lo = c; s = v; }
                else if (c != hi + 1 || v != s) return false;
                hi = c;
            }
        }
        if (lo < 0) return false;
        if (s > 0) { t.row_a[k] = lo; t.row_b[k] = hi + 1; }
        else { t.row_a[k] = hi + 1; t.row_b[k] = lo; }
    }
    return true;
}

// ---------------------------------------------------------------------------
// a13  ShiftMinimizer, the reference's dense formulation in double.
// R:PEFStudioDX/ShiftCollection.cs:738-822; R:Kernels/ShiftMinimizerKernels.cu:80-176.
// Per tile: up to 10 x {G = A^T A; G^-1; r = G^-1 A^T d; dhat = A r; drop the
// row with the largest |d - dhat|^2 > 1}. Singularity (status -1) is decided
// exactly: A has full column rank iff the graph whose edges are the active
// rows' frame pairs is connected. A singular tile gets r = 0 (the reference
// leaves the batched-inverse output undefined there).
// d, dout: [tile][m][2]; r: [tile][n1][2]; status: [tile].
static bool rows_connected(int n, const std::vector<int>& a, const std::vector<int>& b, const std::vector<char>& active) {
    std::vector<int> p(n);
    for (int i = 0; i < n; ++i) p[i] = i;
    std::function<int(int)> find = [&](int x) { return p[x] == x ? x : (p[x] = find(p[x])); };
    for (size_t k = 0; k < a.size(); ++k)
        if (active[k]) p[find(a[k])] = find(b[k]);
    int r0 = find(0);
    for (int i = 1; i < n; ++i)
        if (find(i) != r0) return false;
    return true;
}

static void invert_gauss_jordan(int n, std::vector<double>& g, std::vector<double>& inv) {
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(g[(size_t)r * n + c]) > std::fabs(g[(size_t)piv * n + c])) piv = r;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(g[(size_t)c * n + k], g[(size_t)piv * n + k]);
                std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
            }
        double d = g[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) { g[(size_t)c * n + k] /= d; inv[(size_t)c * n + k] /= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = g[(size_t)r * n + c];
            if (f == 0) continue;
            for (int k = 0; k < n; ++k) {
                g[(size_t)r * n + k] -= f * g[(size_t)c * n + k];
                inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
            }
        }
    }
}

void minimize_shifts(const ShiftTables& tb, int tiles, const double* d_in, double* d_out, double* r_out, int* status_out,
                     int* iters_out) {
    int m = tb.m, n1 = tb.n1;
    parallel_for(tiles, [&](int t0, int t1) {
        std::vector<double> A, d, G, Ginv, S, r((size_t)n1 * 2), dh((size_t)m * 2);
        std::vector<char> active;
        for (int t = t0; t < t1; ++t) {
            A.assign(tb.A.begin(), tb.A.end());
            d.assign(d_in + (size_t)t * m * 2, d_in + (size_t)(t + 1) * m * 2);
            active.assign(m, 1);
            int status = 0, it;
            std::fill(r.begin(), r.end(), 0.0);
            for (it = 0; it < 10; ++it) {
                if (!rows_connected(tb.n, tb.row_a, tb.row_b, active)) {
                    std::fill(r.begin(), r.end(), 0.0);
                    status = -1;
                    break;
                }
                G.assign((size_t)n1 * n1, 0.0);
                for (int i = 0; i < n1; ++i)
                    for (int j = 0; j < n1; ++j) {
                        double s = 0;
                        for (int k = 0; k < m; ++k) s += A[(size_t)i * m + k] * A[(size_t)j * m + k];
                        G[(size_t)i * n1 + j] = s;
                    }
                invert_gauss_jordan(n1, G, Ginv);
                S.assign((size_t)n1 * m, 0.0);
                for (int i = 0; i < n1; ++i)
                    for (int k = 0; k < m; ++k) {
                        double s = 0;
                        for (int j = 0; j < n1; ++j) s += Ginv[(size_t)i * n1 + j] * A[(size_t)j * m + k];
                        S[(size_t)i * m + k] = s;
                    }
                for (int i = 0; i < n1; ++i)
                    for (int c = 0; c < 2; ++c) {
                        double s = 0;
                        for (int k = 0; k < m; ++k) s += S[(size_t)i * m + k] * d[(size_t)k * 2 + c];
                        r[(size_t)i * 2 + c] = s;
                    }
                for (int k = 0; k < m; ++k)
                    for (int c = 0; c < 2; ++c) {
                        double s = 0;
                        for (int j = 0; j < n1; ++j) s += A[(size_t)j * m + k] * r[(size_t)j * 2 + c];
                        dh[(size_t)k * 2 + c] = s;
                    }
                double mx = 1;
                int idx = -1;
                for (int k = 0; k < m; ++k) {
                    double dx = d[(size_t)k * 2] - dh[(size_t)k * 2], dy = d[(size_t)k * 2 + 1] - dh[(size_t)k * 2 + 1];
                    double dist = dx * dx + dy * dy;
                    if (dist > mx) { idx = k; mx = dist; }
                }
                if (idx == -1) { status = -1; break; }
                status = idx;
                d[(size_t)idx * 2] = 0; d[(size_t)idx * 2 + 1] = 0;
                for (int j = 0; j < n1; ++j) A[(size_t)j * m + idx] = 0;
                active[idx] = 0;
            }
            for (int k = 0; k < m; ++k)
                for (int c = 0; c < 2; ++c) {
                    double s = 0;
                    for (int j = 0; j < n1; ++j) s += (double)tb.A[(size_t)j * m + k] * r[(size_t)j * 2 + c];
                    d_out[((size_t)t * m + k) * 2 + c] = s;
                }
            for (int i = 0; i < n1 * 2; ++i) r_out[(size_t)t * n1 * 2 + i] = r[i];
            if (status_out) status_out[t] = status;
            if (iters_out) iters_out[t] = std::min(it + 1, 10);
        }
    });
}

// a14  getOptimalShifts (R:Kernels/ShiftMinimizerKernels.cu:178-218).
template <class T> void get_optimal_shifts(const T* r /*tiles x n1 x 2*/, T* out, int n, int tiles, int ref, int trk) {
    int n1 = n - 1;
    for (int t = 0; t < tiles; ++t) {
        T sx = 0, sy = 0;
        if (ref < trk)
            for (int i = ref; i < trk; ++i) { sx += r[((size_t)t * n1 + i) * 2]; sy += r[((size_t)t * n1 + i) * 2 + 1]; }
        else if (trk < ref)
            for (int i = trk; i < ref; ++i) { sx -= r[((size_t)t * n1 + i) * 2]; sy -= r[((size_t)t * n1 + i) * 2 + 1]; }
        out[(size_t)t * 2] = sx;
        out[(size_t)t * 2 + 1] = sy;
    }
}

// ---------------------------------------------------------------------------
// a15  CreateFlowFieldFromTiles (R:Kernels/opticalFlow.cu:47-93); tile shift
//      texture: linear, clamp, normalised over the tile image (tw x th).
template <class T>
void flow_from_tiles(const T* tiles, int tw, int th, T* flow, int w, int h, T bsx, T bsy, T rot, int quant) {
    parallel_for(h, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < w; ++x) {
                T c = cos_r(rot), s = sin_r(rot);
                T fx = c * -bsx - s * -bsy;
                T fy = s * -bsx + c * -bsy;
                T pcx = (T)(x - w / 2), pcy = (T)(y - h / 2);
                fx += c * pcx - s * pcy - pcx;
                fy += s * pcx + c * pcy - pcy;
                T u = ((T)x + (T)0.5f) / (T)w, v = ((T)y + (T)0.5f) / (T)h;
                fx += tex_linear(tiles, tw, th, 2, 0, u, v, ADDR_CLAMP, quant);
                fy += tex_linear(tiles, tw, th, 2, 1, u, v, ADDR_CLAMP, quant);
                flow[((size_t)y * w + x) * 2] = fx;
                flow[((size_t)y * w + x) * 2 + 1] = fy;
            }
    });
}

// a16  Lucas-Kanade pieces (R:Kernels/opticalFlow.cu:27-44, 96-185, 189-325).
template <class T> void warp(const T* src, const T* flow, T* out, int w, int h, int quant) {
    parallel_for(h, [&](int y0, int y1) {
        for (int iy = y0; iy < y1; ++iy)
            for (int ix = 0; ix < w; ++ix) {
                T u = ((T)ix + (T)0.5f) / (T)w, v = ((T)iy + (T)0.5f) / (T)h;
                T sx = tex_point(flow, w, h, 2, 0, u, v, ADDR_CLAMP);
                T sy = tex_point(flow, w, h, 2, 1, u, v, ADDR_CLAMP);
                T x = ((T)ix + (T)0.5f + sx) / (T)w;
                T y = ((T)iy + (T)0.5f + sy) / (T)h;
                out[(size_t)iy * w + ix] = tex_linear(src, w, h, 1, 0, x, y, ADDR_MIRROR, quant);
            }
    });
}

template <class T> static T stencil5(const T* img, int w, int h, T x, T y, T dx, T dy, int quant) {
    T t = tex_linear(img, w, h, 1, 0, x + (T)2.0f * dx, y + (T)2.0f * dy, ADDR_MIRROR, quant);
    t -= tex_linear(img, w, h, 1, 0, x + (T)1.0f * dx, y + (T)1.0f * dy, ADDR_MIRROR, quant) * (T)8.0f;
    t += tex_linear(img, w, h, 1, 0, x - (T)1.0f * dx, y - (T)1.0f * dy, ADDR_MIRROR, quant) * (T)8.0f;
    t -= tex_linear(img, w, h, 1, 0, x - (T)2.0f * dx, y - (T)2.0f * dy, ADDR_MIRROR, quant);
    t /= (T)12.0f;
    return t;
}

template <class T> void derivatives(const T* src, const T* tgt, T* Ix, T* Iy, T* Iz, int w, int h, int quant) {
    parallel_for(h, [&](int y0, int y1) {
        for (int iy = y0; iy < y1; ++iy)
            for (int ix = 0; ix < w; ++ix) {
                T dx = (T)1.0f / (T)w, dy = (T)1.0f / (T)h;
                T x = ((T)ix + (T)0.5f) * dx, y = ((T)iy + (T)0.5f) * dy;
                size_t o = (size_t)iy * w + ix;
                T t0 = stencil5(src, w, h, x, y, dx, (T)0, quant);
                T t1 = stencil5(tgt, w, h, x, y, dx, (T)0, quant);
                Ix[o] = (t0 + t1) * (T)0.5f;
                Iz[o] = tex_linear(src, w, h, 1, 0, x, y, ADDR_MIRROR, quant) - tex_linear(tgt, w, h, 1, 0, x, y, ADDR_MIRROR, quant);
                t0 = stencil5(src, w, h, x, y, (T)0, dy, quant);
                t1 = stencil5(tgt, w, h, x, y, (T)0, dy, quant);
                Iy[o] = (t0 + t1) * (T)0.5f;
            }
    });
}

template <class T> void derivatives2(const T* img, T* Ix, T* Iy, int w, int h, int quant) {
    parallel_for(h, [&](int y0, int y1) {
        for (int iy = y0; iy < y1; ++iy)
            for (int ix = 0; ix < w; ++ix) {
                T dx = (T)1.0f / (T)w, dy = (T)1.0f / (T)h;
                T x = ((T)ix + (T)0.5f) * dx, y = ((T)iy + (T)0.5f) * dy;
                size_t o = (size_t)iy * w + ix;
                Ix[o] = stencil5(img, w, h, x, y, dx, (T)0, quant);
                Iy[o] = stencil5(img, w, h, x, y, (T)0, dy, quant);
            }
    });
}

// The closed-form solve of lucasKanadeOptim (R:Kernels/opticalFlow.cu:237-321) for the window sums
// (g0, g1, g3) of pixel (px, py), in T: returns false when the sigma1 gate skips the pixel, else the
// update (u0, u1), NaN mapped to 0. Derivatives of type D (float for the FP32 path, float or double
// for the FP64 solve of an FP32 pass).
template <class T, class D>
static bool lk_solve(T g0, T g1, T g3, const D* Fx, const D* Fy, const D* Ft, int w, int px, int py, int hw, T min_det,
                     T& u0, T& u1) {
    int ws = hw * 2 + 1;
    T a = g0, b = g1, c = g1, d = g3;
    T theta = (T)0.5f * atan2_r((T)2.0f * a * c + (T)2.0f * b * d, a * a + b * b - c * c - d * d);
    T ct = cos_r(theta), st = sin_r(theta);
    T UT0 = ct, UT2 = -st, UT1 = st, UT3 = ct;
    T S1 = a * a + b * b + c * c + d * d;
    T S2 = std::sqrt((a * a + b * b - c * c - d * d) * (a * a + b * b - c * c - d * d) +
                     (T)4 * (a * c + b * d) * (a * c + b * d));
    T sigma1 = std::sqrt((S1 + S2) / (T)2);
    T sigma2 = std::sqrt((S1 - S2) / (T)2);
    T smin = std::fmin(sigma1, sigma1);  // R:Kernels/opticalFlow.cu:255 gates on sigma1 (quirk kept)
    if (smin < min_det) return false;
    sigma1 = sigma1 != 0 ? (T)1.0f / sigma1 : (T)0;
    sigma2 = sigma2 != 0 ? (T)1.0f / sigma2 : (T)0;
    T S0 = sigma1, Sb = 0, Sc = 0, S3 = sigma2;
    T eps = (T)0.5f * atan2_r((T)2.0f * a * b + (T)2.0f * c * d, a * a - b * b + c * c - d * d);
    T ce = cos_r(eps), se = sin_r(eps);
    T s11 = (a * ct + c * st) * ce + (b * ct + d * st) * se;
    T s22 = (a * st - c * ct) * se + (-b * st + d * ct) * ce;
    s11 = s11 > (T)0.0f ? (T)1.0f : s11 < 0 ? (T)-1.0f : (T)0.0f;
    s22 = s22 > (T)0.0f ? (T)1.0f : s22 < 0 ? (T)-1.0f : (T)0.0f;
    T V0 = s11 * ce, V1 = -s22 * se, V2 = s11 * se, V3 = s22 * ce;
    T m0 = S0 * UT0 + Sb * UT2, m1 = S0 * UT1 + Sb * UT3, m2 = Sc * UT0 + S3 * UT2, m3 = Sc * UT1 + S3 * UT3;
    T i0 = V0 * m0 + V1 * m2, i1 = V0 * m1 + V1 * m3, i2 = V2 * m0 + V3 * m2, i3 = V2 * m1 + V3 * m3;
    u0 = 0;
    u1 = 0;
    for (int i = 0; i < ws * ws; ++i) {
        int yy = i / ws, xx = i - yy * ws;
        size_t o = (size_t)(py + yy - hw) * w + px + xx - hw;
        T dx = (T)Fx[o], dy = (T)Fy[o], dt = (T)Ft[o];
        u0 += (i0 * dx + i1 * dy) * dt;
        u1 += (i2 * dx + i3 * dy) * dt;
    }
    u0 = std::isnan(u0) ? (T)0 : u0;
    u1 = std::isnan(u1) ? (T)0 : u1;
    return true;
}

// Ill-conditioned window of an FP32 pass: lambda2 / lambda1 of the window matrix below ill_ratio,
// estimated in double from the FP32 sums as det / trace^2 (also true for det <= 0 and NaN). There the
// FP32 closed form loses lambda2 to cancellation (S1 - S2 = 2 lambda2^2 next to S1 ~ lambda1^2).
inline bool lk_ill_conditioned(float g0, float g1, float g3, double ill_ratio) {
    const double a = g0, b = g1, d = g3, tr = a + d, det = a * d - b * b;  // exact products of floats
    return !(det > ill_ratio * tr * tr);
}

// FP64 solve of an ill-conditioned window from its FP32 sums (a, b, d of the window matrix, b0, b1 of
// sum Ix It, sum Iy It): the same gate as the reference (sigma1 = lambda1 < min_det skips), then
// u = G^-1 (b0, b1) by the adjugate for det > 0 (the closed-form pseudo-inverse of the reference is the
// inverse there), and the rank-1 pseudo-inverse G / trace^2 when rounding left det <= 0. NaN -> 0.
inline bool lk_solve_fp64(float fa, float fb, float fd, float fb0, float fb1, double min_det, double& u0, double& u1) {
    const double a = fa, b = fb, d = fd, b0 = fb0, b1 = fb1, tr = a + d;
    const double lam1 = 0.5 * (tr + std::sqrt((a - d) * (a - d) + 4.0 * b * b));
    if (lam1 < min_det) return false;
    const double det = a * d - b * b;
    if (det > 0) {
        u0 = (d * b0 - b * b1) / det;
        u1 = (a * b1 - b * b0) / det;
    } else {
        const double t2 = tr * tr;
        u0 = (a * b0 + b * b1) / t2;
        u1 = (b * b0 + d * b1) / t2;
    }
    if (std::isnan(u0)) u0 = 0;
    if (std::isnan(u1)) u1 = 0;
    return true;
}

std::atomic<long long> g_lk_ill_windows{0}, g_lk_windows{0};  // windows solved in FP64 / all (FP32 passes)

// One lucasKanadeOptim pass; flow updated in place (each pixel reads only itself). For T = float and
// ill_ratio > 0, windows with lambda2 / lambda1 < ill_ratio (lk_ill_conditioned) are solved in FP64 from
// their FP32 sums (lk_solve_fp64, sums in the window's raster order) and the FP64 update is rounded once
// and added to the FP32 flow; ill_ratio = 0 is the reference's FP32 arithmetic everywhere.
// cond (optional, w x h x 2): per pixel the smallest det / trace^2 of its window's sums so far and the
// smallest distance |lambda1 / min_det - 1| from the sigma1 gate (lowered here).
template <class T>
void lucas_kanade_step(T* flow, const T* Fx, const T* Fy, const T* Ft, int w, int h, int hw, T min_det,
                       double ill_ratio, float* cond = nullptr) {
    const double ill = std::is_same<T, float>::value ? ill_ratio : 0.0;
    parallel_for(h, [&](int y0, int y1) {
        long long n_ill = 0, n_all = 0;
        for (int py = std::max(y0, hw); py < std::min(y1, h - hw); ++py)
            for (int px = hw; px < w - hw; ++px) {
                T g0 = 0, g1 = 0, g3 = 0;
                for (int y = -hw; y <= hw; ++y)
                    for (int x = -hw; x <= hw; ++x) {
                        size_t o = (size_t)(py + y) * w + px + x;
                        T dx = Fx[o], dy = Fy[o];
                        g0 += dx * dx;
                        g1 += dx * dy;
                        g3 += dy * dy;
                    }
                T u0, u1;
                ++n_all;
                if (cond) {
                    // (det / trace^2, |lambda1 / min_det - 1|): distance from singular and from the gate
                    const double a = (double)g0, b = (double)g1, d = (double)g3, tr = a + d;
                    const float r = (float)((a * d - b * b) / (tr * tr));
                    const double lam1 = 0.5 * (tr + std::sqrt((a - d) * (a - d) + 4.0 * b * b));
                    const float gm = (float)std::fabs(lam1 / (double)min_det - 1.0);
                    float* c = cond + ((size_t)py * w + px) * 2;
                    if (!(r >= c[0])) c[0] = r;  // NaN (flat window) sticks
                    if (!(gm >= c[1])) c[1] = gm;
                }
                if (ill > 0 && lk_ill_conditioned((float)g0, (float)g1, (float)g3, ill)) {
                    ++n_ill;
                    T b0 = 0, b1 = 0;
                    for (int y = -hw; y <= hw; ++y)
                        for (int x = -hw; x <= hw; ++x) {
                            size_t o = (size_t)(py + y) * w + px + x;
                            b0 += Fx[o] * Ft[o];
                            b1 += Fy[o] * Ft[o];
                        }
                    double v0, v1;
                    if (!lk_solve_fp64((float)g0, (float)g1, (float)g3, (float)b0, (float)b1, (double)min_det, v0, v1))
                        continue;
                    u0 = (T)v0;
                    u1 = (T)v1;
                } else if (!lk_solve<T, T>(g0, g1, g3, Fx, Fy, Ft, w, px, py, hw, min_det, u0, u1)) {
                    continue;
                }
                flow[((size_t)py * w + px) * 2] += u0;
                flow[((size_t)py * w + px) * 2 + 1] += u1;
            }
        if (ill > 0) {
            g_lk_ill_windows.fetch_add(n_ill, std::memory_order_relaxed);
            g_lk_windows.fetch_add(n_all, std::memory_order_relaxed);
        }
    });
}

// OpticalFlow.LucasKanade (R:PEFStudioDX/OpticalFlow.cs:147-164). src is
// replaced by the final warped image, as in the reference.
// cond (optional, w x h x 2): per pixel the smallest window conditioning det / trace^2 and gate distance
// |lambda1 / min_det - 1| over the iterations (+inf where no window is centred on the pixel).
template <class T>
void lucas_kanade(T* src, const T* tgt, const T* tiles, int tw, int th, T* flow, int w, int h, int iterations, T bsx,
                  T bsy, T rot, T min_det, int window, int quant, double ill_ratio, float* cond = nullptr) {
    if (cond) std::fill(cond, cond + (size_t)w * h * 2, std::numeric_limits<float>::infinity());
    flow_from_tiles(tiles, tw, th, flow, w, h, bsx, bsy, rot, quant);
    size_t n = (size_t)w * h;
    std::vector<T> tmp(n), Ix(n), Iy(n), Iz(n);
    for (int it = 0; it < iterations; ++it) {
        warp(src, flow, tmp.data(), w, h, quant);
        derivatives(tmp.data(), tgt, Ix.data(), Iy.data(), Iz.data(), w, h, quant);
        lucas_kanade_step(flow, Ix.data(), Iy.data(), Iz.data(), w, h, window / 2, min_det, ill_ratio, cond);
    }
    warp(src, flow, tmp.data(), w, h, quant);
    std::copy(tmp.begin(), tmp.end(), src);
}

// ---------------------------------------------------------------------------
// a17  ComputeStructureTensor + ComputeKernelParam (R:Kernels/kernel.cu:600-700).
template <class T> void structure_tensor(const T* dx, const T* dy, T* out3, int w, int h) {
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        out3[i * 3] = dx[i] * dx[i];
        out3[i * 3 + 1] = dy[i] * dy[i];
        out3[i * 3 + 2] = dx[i] * dy[i];
    }
}

template <class T> void kernel_param(T* k3, int w, int h, T Dth, T Dtr, T kDetail, T kDenoise, T kStretch, T kShrink) {
    parallel_for(h, [&](int y0, int y1) {
        for (size_t i = (size_t)y0 * w; i < (size_t)y1 * w; ++i) {
            T a11 = k3[i * 3], a22 = k3[i * 3 + 1], a12 = k3[i * 3 + 2];
            T help = std::sqrt((a22 - a11) * (a22 - a11) + (T)4.0f * a12 * a12);
            T c = (T)2.0f * a12, s = a22 - a11 + help;
            T norm = std::sqrt(c * c + s * s);
            if (norm > 0) { c /= norm; s /= norm; }
            else { c = 1; s = 0; }
            T lam1 = (a11 + a22 + help) / (T)2.0f, lam2 = (a11 + a22 - help) / (T)2.0f;
            T A = (T)1 + std::sqrt((lam1 - lam2) * (lam1 - lam2) / ((lam1 + lam2) * (lam1 + lam2)));
            T D = (T)1 - std::sqrt(lam1) / Dtr + Dth;
            D = std::fmax(std::fmin((T)1.0f, D), (T)0.0f);
            T k1h = kDetail * kStretch * A, k2h = kDetail / kShrink * A;
            T k1 = ((T)1.0f - D) * k1h + D * kDetail * kDenoise;
            T k2 = ((T)1.0f - D) * k2h + D * kDetail * kDenoise;
            k1 *= k1;
            k2 *= k2;
            T x2 = c, y2 = s, x1 = s, y1 = -c;
            T b11 = k1 * x1 * x1 + x2 * x2 * k2;
            T b12 = k1 * x1 * y1 + x2 * y2 * k2;
            T b22 = k1 * y1 * y1 + y2 * y2 * k2;
            T det = b11 * b22 - b12 * b12 + (T)0.0000000001f;
            k3[i * 3] = b22 / det;
            k3[i * 3 + 1] = b11 / det;
            k3[i * 3 + 2] = -b12 / det;
        }
    });
}

// ---------------------------------------------------------------------------
// a18  deBayersSubSample3 (R:Kernels/DeBayerKernels.cu:242-283).
template <class T> void subsample3(const uint16_t* raw, T* out3, T max_val, int hw, int hh, const int cfa[4]) {
    for (int y = 0; y < hh; ++y)
        for (int x = 0; x < hw; ++x) {
            T p[3] = {0, 0, 0};
            T factor = (T)1.0f / max_val;
            for (int ix = 0; ix < 2; ++ix)
                for (int iy = 0; iy < 2; ++iy) {
                    int site = cfa[iy * 2 + ix];
                    T v = (T)raw[(size_t)(2 * y + iy) * hw * 2 + 2 * x + ix];
                    if (site == CFA_GREEN) p[1] += v * factor * (T)0.5f;
                    else if (site == CFA_RED) p[0] = v * factor;
                    else if (site == CFA_BLUE) p[2] = v * factor;
                }
            for (int c = 0; c < 3; ++c) out3[((size_t)y * hw + x) * 3 + c] = p[c];
        }
}

// a19  ComputeRobustnessMask (R:Kernels/RobustnessModell.cu:28-158) on the
//      half-resolution grid; flow texture (full resolution, fw x fh) linear +
//      mirror. mask: hw x hh x 4, border never written.
template <class T>
void robustness_mask(const T* ref3, const T* mov3, T* mask4, const T* flow, int fw, int fh, int hw, int hh, T alpha,
                     T beta, T thr_m, int quant) {
    parallel_for(hh, [&](int y0, int y1) {
        for (int py = std::max(y0, 1); py < std::min(y1, hh - 1); ++py)
            for (int px = 1; px < hw - 1; ++px) {
                T u = ((T)px + (T)0.5f) / (T)hw, v = ((T)py + (T)0.5f) / (T)hh;
                T fx = tex_linear(flow, fw, fh, 2, 0, u, v, ADDR_MIRROR, quant);
                T fy = tex_linear(flow, fw, fh, 2, 1, u, v, ADDR_MIRROR, quant);
                T maxx = fx, maxy = fy, minx = fx, miny = fy;
                for (int y = -2; y <= 2; ++y)
                    for (int x = -2; x <= 2; ++x) {
                        T su = ((T)px + (T)x + (T)0.5f) / (T)hw, sv = ((T)py + (T)y + (T)0.5f) / (T)hh;
                        T sx = tex_linear(flow, fw, fh, 2, 0, su, sv, ADDR_MIRROR, quant);
                        T sy = tex_linear(flow, fw, fh, 2, 1, su, sv, ADDR_MIRROR, quant);
                        maxx = std::fmax(sx, fx); maxy = std::fmax(sy, fy);  // overwritten each time (quirk kept)
                        minx = std::fmin(sx, fx); miny = std::fmin(sy, fy);
                    }
                int shx = (int)std::round(fx * (T)0.5f), shy = (int)std::round(fy * (T)0.5f);
                T mr[3] = {0, 0, 0}, mm[3] = {0, 0, 0}, sr[3] = {0, 0, 0};
                T pr[9][3];
                for (int y = -1; y <= 1; ++y)
                    for (int x = -1; x <= 1; ++x) {
                        const T* p = ref3 + ((size_t)(py + y) * hw + px + x) * 3;
                        for (int c = 0; c < 3; ++c) { pr[(y + 1) * 3 + x + 1][c] = p[c]; mr[c] += p[c]; }
                        int yy = std::min(std::max(py + shy + y, 0), hh - 1);
                        int xx = std::min(std::max(px + shx + x, 0), hw - 1);
                        const T* q = mov3 + ((size_t)yy * hw + xx) * 3;
                        for (int c = 0; c < 3; ++c) mm[c] += q[c];
                    }
                for (int c = 0; c < 3; ++c) { mr[c] /= (T)9.0f; mm[c] /= (T)9.0f; }
                T meandist = std::fabs(mr[0] - mm[0]) + std::fabs(mr[1] - mm[1]) + std::fabs(mr[2] - mm[2]);
                meandist /= (T)3.0f;
                maxx *= (T)0.5f * meandist; maxy *= (T)0.5f * meandist;
                minx *= (T)0.5f * meandist; miny *= (T)0.5f * meandist;
                T Mv = std::sqrt((maxx - minx) * (maxx - minx) + (maxy - miny) * (maxy - miny));
                for (int i = 0; i < 9; ++i)
                    for (int c = 0; c < 3; ++c) sr[c] += (pr[i][c] - mr[c]) * (pr[i][c] - mr[c]);
                T out[4];
                T s = (T)1.5f;
                if (Mv > thr_m) s = 0;
                for (int c = 0; c < 3; ++c) {
                    T sd = std::sqrt(sr[c] / (T)9.0f);
                    T smd = std::sqrt(alpha * mr[c] + beta);
                    if (c == 1) smd = smd / std::sqrt((T)2.0f);
                    T dist = std::fabs(mr[c] - mm[c]);
                    T sig = std::fmax(smd, sd);
                    dist = dist * (sd * sd / (sd * sd + smd * smd));
                    out[c] = std::fmax(std::fmin(s * exp_r(-dist * dist / (sig * sig)) - (T)0.12f, (T)1.0f), (T)0.0f);
                }
                out[3] = Mv;
                for (int c = 0; c < 4; ++c) mask4[((size_t)py * hw + px) * 4 + c] = out[c];
            }
    });
}

// NPP ErodeBorder, size x size all-set mask, centred anchor, replicate border.
template <class T> void erode(const T* in4, T* out4, int w, int h, int size) {
    int r = size / 2;
    parallel_for(h, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 4; ++c) {
                    T m = std::numeric_limits<T>::infinity();
                    for (int j = -r; j <= size - 1 - r; ++j)
                        for (int i = -r; i <= size - 1 - r; ++i) {
                            int yy = std::min(std::max(y + j, 0), h - 1), xx = std::min(std::max(x + i, 0), w - 1);
                            m = std::fmin(m, in4[((size_t)yy * w + xx) * 4 + c]);
                        }
                    out4[((size_t)y * w + x) * 4 + c] = m;
                }
    });
}

// ---------------------------------------------------------------------------
// a20  accumulateImages / accumulateImagesSuperRes (R:Kernels/DeBayerKernels.cu:288-468).
template <class T> static inline T kernel_weight(int px, int py, T k0, T k1, T k2) {
    T w = (T)(px * px) * k0 + (T)(2 * px * py) * k2 + (T)(py * py) * k1;
    w = exp_r((T)-0.5f * w);
    if (!std::isfinite(w)) w = px * py == 0 ? (T)1 : (T)0;
    return w;
}

// One 1x tap of accumulateImages (R:Kernels/DeBayerKernels.cu:314-372): the sample at the clamped,
// shifted site, its CFA colour, (raw - black) / white, the certainty at the clamped unshifted
// half-resolution position (Q15), and the two running sums of that colour.
template <class T>
static inline void accumulate_tap(const uint16_t* raw, const T* mask4, int mw, int w, int h, int x, int y, int px,
                                  int py, int sx, int sy, T k0, T k1, T k2, const T white[3], const T black[3],
                                  const int cfa[4], T acc[3], T wt[3]) {
    int ppsx = std::min(std::max(x + px + sx, 0), w - 1), ppsy = std::min(std::max(y + py + sy, 0), h - 1);
    int ppx = std::min(std::max(x + px, 0), w - 1), ppy = std::min(std::max(y + py, 0), h - 1);
    int site = cfa_at(cfa, ppsy, ppsx);
    T wgt = kernel_weight(px, py, k0, k1, k2);
    T v = (T)raw[(size_t)ppsy * w + ppsx];
    int c = site == CFA_GREEN ? 1 : site == CFA_RED ? 0 : site == CFA_BLUE ? 2 : -1;
    if (c < 0) return;
    v = (v - black[c]) / white[c];
    T cert = mask4[((size_t)(ppy / 2) * mw + (ppx / 2)) * 4 + c];
    if (!std::isfinite(cert)) cert = 0;
    acc[c] += v * wgt * cert;
    wt[c] += wgt * cert;
}

// Tap orders of the 1x gather. Raster: py = -2..2, px = -2..2, the reference's order. By class:
// class = ((py & 1) << 1) | (px & 1) = 0..3, raster order within a class -- the order of the HIP
// kernels, which keep one colour per class away from the image edge. The two orders add the same
// terms to each colour's sums; they differ in how the two green classes interleave and, where
// clamping maps taps of several classes to one colour, at the image edge.
struct TapOrder {
    int pxy[50];
    explicit TapOrder(bool by_class) {
        int n = 0;
        for (int cls = 0; cls < (by_class ? 4 : 1); ++cls)
            for (int py = -2; py <= 2; ++py)
                for (int px = -2; px <= 2; ++px)
                    if (!by_class || (((py & 1) << 1) | (px & 1)) == cls) { pxy[2 * n] = px; pxy[2 * n + 1] = py; ++n; }
    }
};
static const int* tap_order(bool by_class) {
    static const TapOrder raster(false), classes(true);  // thread-safe initialisation
    return by_class ? classes.pxy : raster.pxy;
}

template <class T>
static void accumulate_ordered(const uint16_t* raw, T* sum3, T* wsum3, const T* mask4, const T* kern3, const T* flow,
                               const T white[3], const T black[3], int w, int h, const int cfa[4], bool by_class) {
    int mw = w / 2;
    const int* order = tap_order(by_class);
    parallel_for(h, [&](int y0, int y1) {
        for (int y = std::max(y0, 1); y < std::min(y1, h - 1); ++y)
            for (int x = 1; x < w - 1; ++x) {
                size_t o = (size_t)y * w + x;
                T acc[3] = {sum3[o * 3], sum3[o * 3 + 1], sum3[o * 3 + 2]};
                T wt[3] = {wsum3[o * 3], wsum3[o * 3 + 1], wsum3[o * 3 + 2]};
                T k0 = kern3[o * 3], k1 = kern3[o * 3 + 1], k2 = kern3[o * 3 + 2];
                int sx = (int)std::round(flow[o * 2]), sy = (int)std::round(flow[o * 2 + 1]);
                for (int t = 0; t < 25; ++t)
                    accumulate_tap(raw, mask4, mw, w, h, x, y, order[2 * t], order[2 * t + 1], sx, sy, k0, k1, k2, white,
                                   black, cfa, acc, wt);
                for (int c = 0; c < 3; ++c) { sum3[o * 3 + c] = acc[c]; wsum3[o * 3 + c] = wt[c]; }
            }
    });
}

// accumulateImages, the reference's raster tap order (R:Kernels/DeBayerKernels.cu:314-372).
template <class T>
void accumulate(const uint16_t* raw, T* sum3, T* wsum3, const T* mask4, const T* kern3, const T* flow, const T white[3],
                const T black[3], int w, int h, const int cfa[4]) {
    accumulate_ordered(raw, sum3, wsum3, mask4, kern3, flow, white, black, w, h, cfa, false);
}

// The same sums with the taps visited class by class (the HIP kernels' order).
template <class T>
void accumulate_class_order(const uint16_t* raw, T* sum3, T* wsum3, const T* mask4, const T* kern3, const T* flow,
                            const T white[3], const T black[3], int w, int h, const int cfa[4]) {
    accumulate_ordered(raw, sum3, wsum3, mask4, kern3, flow, white, black, w, h, cfa, true);
}

// kern4, flow: full-resolution W x H textures (clamp / mirror, linear). The output grid has ow x oh pixels at
// half the raw spacing whose pixel 0 lies (ox, oy) output pixels from the frame origin: the reference's
// central-half form is (w, h, w/2, h/2) (R:Kernels/DeBayerKernels.cu:395-423, which writes w / 2 and w / 4 for
// ox and ox / 2), the full-frame 2x merge (2w, 2h, 0, 0) -- the same per-pixel rule with zero offset.
template <class T>
void accumulate_sr_grid(const uint16_t* raw, T* sum3, T* wsum3, const T* mask4, const T* kern4, const T* flow,
                        const T white[3], const T black[3], int w, int h, const int cfa[4], int quant, int ow, int oh,
                        int ox, int oy) {
    int mw = w / 2;
    parallel_for(oh, [&](int y0, int y1) {
        for (int y = std::max(y0, 1); y < std::min(y1, oh - 1); ++y)
            for (int x = 1; x < ow - 1; ++x) {
                size_t o = (size_t)y * ow + x;
                T acc[3] = {sum3[o * 3], sum3[o * 3 + 1], sum3[o * 3 + 2]};
                T wt[3] = {wsum3[o * 3], wsum3[o * 3 + 1], wsum3[o * 3 + 2]};
                T posx = ((T)x + (T)0.5f + (T)ox) / (T)2.0f / (T)w;
                T posy = ((T)y + (T)0.5f + (T)oy) / (T)2.0f / (T)h;
                T k0 = tex_linear(kern4, w, h, 4, 0, posx, posy, ADDR_CLAMP, quant);
                T k1 = tex_linear(kern4, w, h, 4, 1, posx, posy, ADDR_CLAMP, quant);
                T k2 = tex_linear(kern4, w, h, 4, 2, posx, posy, ADDR_CLAMP, quant);
                T fx = tex_linear(flow, w, h, 2, 0, posx, posy, ADDR_MIRROR, quant);
                T fy = tex_linear(flow, w, h, 2, 1, posx, posy, ADDR_MIRROR, quant);
                int sx = (int)std::round(fx * (T)2), sy = (int)std::round(fy * (T)2);
                for (int py = -2; py <= 2; ++py)
                    for (int px = -2; px <= 2; ++px) {
                        int ppsx = x + px + sx + ox, ppsy = y + py + sy + oy;
                        int ppx = x + px + ox, ppy = y + py + oy;
                        ppsx = std::min(std::max(ppsx / 2, 0 + ox / 2), ow / 2 - 1 + ox / 2);
                        ppsy = std::min(std::max(ppsy / 2, 0 + oy / 2), oh / 2 - 1 + oy / 2);
                        ppx = std::min(std::max(ppx / 2, 0 + ox / 2), ow / 2 - 1 + ox / 2);
                        ppy = std::min(std::max(ppy / 2, 0 + oy / 2), oh / 2 - 1 + oy / 2);
                        int site = cfa_at(cfa, ppsy, ppsx);
                        T wgt = kernel_weight(px, py, k0, k1, k2);
                        T v = (T)raw[(size_t)ppsy * w + ppsx];
                        int c = site == CFA_GREEN ? 1 : site == CFA_RED ? 0 : site == CFA_BLUE ? 2 : -1;
                        if (c < 0) continue;
                        v = (v - black[c]) / white[c];
                        T cert = mask4[((size_t)(ppy / 2) * mw + (ppx / 2)) * 4 + c];
                        if (!std::isfinite(cert)) cert = 0;
                        acc[c] += v * wgt * cert;
                        wt[c] += wgt * cert;
                    }
                for (int c = 0; c < 3; ++c) { sum3[o * 3 + c] = acc[c]; wsum3[o * 3 + c] = wt[c]; }
            }
    });
}

// accumulateImagesSuperRes as the reference launches it: the central half at w x h.
template <class T>
void accumulate_sr(const uint16_t* raw, T* sum3, T* wsum3, const T* mask4, const T* kern4, const T* flow, const T white[3],
                   const T black[3], int w, int h, const int cfa[4], int quant) {
    accumulate_sr_grid(raw, sum3, wsum3, mask4, kern4, flow, white, black, w, h, cfa, quant, w, h, w / 2, h / 2);
}

// a21  ApplyWeighting (R:Kernels/kernel.cu:316-372).
template <class T> void apply_weighting(T* inout3, const T* sum3, const T* w3, int w, int h, T thr) {
    for (size_t i = 0; i < (size_t)w * h * 3; ++i) {
        T val = sum3[i], wt = w3[i], io = inout3[i];
        if (wt < thr) { val += io; wt += 1; }
        io = 0;
        if (wt != 0) io = val / wt;
        inout3[i] = io;
    }
}

// ---------------------------------------------------------------------------
// Pipeline: SkipPreAlignment -> TrackPatches -> PrepareAccumulation ->
// Accumulate -> final ApplyWeighting. R:...Controller.cs:1530-1569, 1601-1746,
// 1795-2002; R:PEFStudioDX/PatchTracker.cs:317-360; R:PEFStudioDX/OpticalFlow.cs:147-164.
}  // namespace ora

#include "isa_oracle_pipeline.inc"
#include "isa_oracle_diff.inc"
