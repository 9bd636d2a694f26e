// Device-side helpers shared by the gfx950 kernels of libisa_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ISA_CFA_RED 0
#define ISA_CFA_GREEN 1
#define ISA_CFA_BLUE 2

// One of three values by a run-time index, from registers: an array in the kernel arguments indexed by a
// per-lane value would be read from memory on every use.
template <class T> __device__ __forceinline__ T pick3(int i, T a0, T a1, T a2) { return i == 0 ? a0 : (i == 1 ? a1 : a2); }

// Bayer pattern passed by value (replaces the reference's __constant__ c_cfaPattern); coordinates >= 0.
struct Cfa {
    int c[4];
    __device__ __forceinline__ int at(int y, int x) const {
        const int r0 = (x & 1) ? c[1] : c[0], r1 = (x & 1) ? c[3] : c[2];
        return (y & 1) ? r1 : r0;
    }
};

template <class T> __device__ __forceinline__ T* row_ptr(T* base, int pitch, int y) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)pitch * (size_t)y);
}
template <class T> __device__ __forceinline__ const T* row_ptr(const T* base, int pitch, int y) {
    return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)pitch * (size_t)y);
}

// Global load at a 32-bit byte offset from a kernel-argument base: the address is the uniform base
// plus a zero-extended 32-bit offset, so it needs no per-lane 64-bit arithmetic. The caller keeps
// every offset below 2^32.
template <class T> __device__ __forceinline__ T ld_off(const T* base, unsigned byte_off) {
    typedef const __attribute__((address_space(1))) char* gchar;
    return *(const __attribute__((address_space(1))) T*)((gchar)base + byte_off);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// a / (float)n for an integer 1 <= n <= 2^24, bit-identical to the IEEE float division but without
// its ~10-instruction sequence. If n is a power of two the quotient is exact on both routes.
// Otherwise n has an odd factor > 1, so a/n (a float with a 24-bit significand) is never a float
// midpoint and lies at least 2^-25/n (relative) from one; the double product a * rd(1/n) is
// within 2^-52 of a/n, so its rounding to float is the correctly rounded quotient.
__device__ __forceinline__ float div_n(float a, double inv_n) { return (float)((double)a * inv_n); }

// ---------------------------------------------------------------------------
// Texture-unit emulation: normalised coordinates, x = u*W - 0.5, i = floor(x),
// alpha = frac(x) (optionally rounded to 8 fractional bits). Mirror addressing
// is the half-sample-symmetric reflection with period 2W.
#define ISA_ADDR_CLAMP 0
#define ISA_ADDR_MIRROR 1

__device__ __forceinline__ int tex_addr(int i, int n, int mode) {
    if ((unsigned)i < (unsigned)n) return i;  // in range: both modes are the identity
    if (mode == ISA_ADDR_CLAMP) return clampi(i, 0, n - 1);
    int p = 2 * n;
    i %= p;
    if (i < 0) i += p;
    return i < n ? i : p - 1 - i;
}

__device__ __forceinline__ float tex_frac(float a, int quant) { return quant ? rintf(a * 256.0f) / 256.0f : a; }

// Bilinear fetch of channel c of a `ch`-channel float image.
template <int CH>
__device__ __forceinline__ float tex_linear(const float* img, int pitch, int w, int h, int c, float u, float v, int mode,
                                            int quant) {
    float x = u * (float)w - 0.5f;
    float y = v * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = tex_frac(x - fx, quant), b = tex_frac(y - fy, quant);
    int x0 = (int)fx, y0 = (int)fy;
    int xa = tex_addr(x0, w, mode), xb = tex_addr(x0 + 1, w, mode);
    int ya = tex_addr(y0, h, mode), yb = tex_addr(y0 + 1, h, mode);
    const float* ra = row_ptr(img, pitch, ya);
    // With both weights 0 the bilinear sum below is t00 + (+-0) terms, i.e. exactly t00 for finite
    // texels; the other three fetches are skipped (the common case of fetches at texel centres
    // with quantised weights).
    if (a == 0.0f && b == 0.0f) return ra[xa * CH + c];
    const float* rb = row_ptr(img, pitch, yb);
    float t00 = ra[xa * CH + c], t10 = ra[xb * CH + c];
    float t01 = rb[xa * CH + c], t11 = rb[xb * CH + c];
    float oma = 1.0f - a, omb = 1.0f - b;
    return oma * omb * t00 + a * omb * t10 + oma * b * t01 + a * b * t11;
}

// Bilinear fetch of all CH channels at once (same weights for every channel).
template <int CH>
__device__ __forceinline__ void tex_linear_all(const float* img, int pitch, int w, int h, float u, float v, int mode,
                                               int quant, float* out) {
    float x = u * (float)w - 0.5f;
    float y = v * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = tex_frac(x - fx, quant), b = tex_frac(y - fy, quant);
    int x0 = (int)fx, y0 = (int)fy;
    int xa = tex_addr(x0, w, mode), xb = tex_addr(x0 + 1, w, mode);
    int ya = tex_addr(y0, h, mode), yb = tex_addr(y0 + 1, h, mode);
    const float* ra = row_ptr(img, pitch, ya);
    const float* rb = row_ptr(img, pitch, yb);
    float oma = 1.0f - a, omb = 1.0f - b;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        float t00 = ra[xa * CH + c], t10 = ra[xb * CH + c];
        float t01 = rb[xa * CH + c], t11 = rb[xb * CH + c];
        out[c] = oma * omb * t00 + a * omb * t10 + oma * b * t01 + a * b * t11;
    }
}

// ---------------------------------------------------------------------------
// Transcendentals evaluated in double and rounded once to float. The CPU
// oracle does the same, so both sides agree bit for bit except when the two
// double results straddle a float rounding boundary (probability ~2^-28 per call).
__device__ __forceinline__ float exp_r(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float pow_r(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ float atan2_r(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ float sin_r(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cos_r(float x) { return (float)cos((double)x); }

// ---------------------------------------------------------------------------
// Per-pixel formulas shared by the unfused (one-kernel-per-reference-kernel)
// and fused entry points, so that both evaluate identical operations.

// Accumulation weight of offset (px, py) for inverse covariance (k0, k1, k2)
// = (Omega11, Omega22, Omega12); non-finite -> cross fallback
// (R:Kernels/DeBayerKernels.cu:335-338).
__device__ __forceinline__ float kernel_weight(int px, int py, float k0, float k1, float k2) {
    float w = (float)(px * px) * k0 + (float)(2 * px * py) * k2 + (float)(py * py) * k1;
    w = exp_r(-0.5f * w);
    if (!isfinite(w)) w = (px * py == 0) ? 1.0f : 0.0f;
    return w;
}
