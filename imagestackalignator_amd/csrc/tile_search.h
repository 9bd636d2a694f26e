// Device-side pieces shared by the tile-search kernels (FP64 direct search and the MFMA variants):
// the per-tile integer offset, findMinimum and the per-level launch arguments.
#pragma once
#include "isa_device.h"

namespace {

// Tile offset with the rotation given as its (sin, cos) pair, each rounded once to float.
__device__ __forceinline__ void tile_offset_sc(int tx, int ty, int tile, int w, int h, float bsx, float bsy, float sf,
                                               float cf, float psx, float psy, int& ox, int& oy) {
    float sx = psx + (cf * -bsx - sf * -bsy);
    float sy = psy + (sf * -bsx + cf * -bsy);
    float pcx = (float)(tx * tile + tile / 2 - w / 2);
    float pcy = (float)(ty * tile + tile / 2 - h / 2);
    sx += cf * pcx - sf * pcy - pcx;
    sy += sf * pcx + cf * pcy - pcy;
    ox = (int)roundf(sx);
    oy = (int)roundf(sy);
}

// findMinimum (R:Kernels/kernel.cu:413-546) on an S x S map.
// Sub-pixel position of the minimum mi (value mn; mx the largest value) of an S x S shift image:
// the 3x3 quadratic fit of findMinimum, or 0 at the window edge / for a flat image.
// Returns the decision code 2 mi + z (z = 1 when the threshold test zeroes the shift), -1 without a
// minimum: the integer part of the search, which the matrix-core variants certify.
__device__ int find_min_fit(const float* v, int M, int mi, float mn, float mx, float threshold, float& cx,
                            float& cy) {
    const float FA11[9] = {1.0f / 4.0f, -2.0f / 4.0f, 1.0f / 4.0f, 2.0f / 4.0f, -4.0f / 4.0f, 2.0f / 4.0f, 1.0f / 4.0f,
                           -2.0f / 4.0f, 1.0f / 4.0f};
    const float FA22[9] = {1.0f / 4.0f, 2.0f / 4.0f, 1.0f / 4.0f, -2.0f / 4.0f, -4.0f / 4.0f, -2.0f / 4.0f, 1.0f / 4.0f,
                           2.0f / 4.0f, 1.0f / 4.0f};
    const float FA12[9] = {1.0f / 4.0f, 0.0f, -1.0f / 4.0f, 0.0f, 0.0f, 0.0f, -1.0f / 4.0f, 0.0f, 1.0f / 4.0f};
    const float Fb1[9] = {-1.0f / 8.0f, 0.0f, 1.0f / 8.0f, -2.0f / 8.0f, 0.0f, 2.0f / 8.0f, -1.0f / 8.0f, 0.0f,
                          1.0f / 8.0f};
    const float Fb2[9] = {-1.0f / 8.0f, -2.0f / 8.0f, -1.0f / 8.0f, 0.0f, 0.0f, 0.0f, 1.0f / 8.0f, 2.0f / 8.0f,
                          1.0f / 8.0f};
    const int S = 2 * M + 1;
    cy = (float)(mi / S);
    cx = (float)mi - cy * (float)S;
    if (cx < 1 || cy < 1 || cx >= (float)(2 * M) || cy >= (float)(2 * M)) {
        cx = 0.0f;
        cy = 0.0f;
    } else {
        float A11 = 0, A22 = 0, A12 = 0, b1 = 0, b2 = 0;
        for (int i = 0; i < 9; ++i) {
            int off = (i / 3 - 1) * S + (i % 3 - 1);
            float img = v[mi + off];
            A11 += FA11[i] * img;
            A22 += FA22[i] * img;
            A12 += FA12[i] * img;
            b1 += Fb1[i] * img;
            b2 += Fb2[i] * img;
        }
        A11 = fmaxf(A11, 0.0f);
        A22 = fmaxf(A22, 0.0f);
        float det = A11 * A22 - A12 * A12;
        if (det < 0) { A12 = 0; det = A11 * A22; }
        if (det != 0) {
            float mux = (A22 * b1 - A12 * b2) / det;
            float muy = (A11 * b2 - A12 * b1) / det;
            if (fabsf(mux) > 1) mux = 0;
            if (fabsf(muy) > 1) muy = 0;
            cx -= mux;
            cy -= muy;
        }
        cx -= (float)M;
        cy -= (float)M;
    }
    const bool zeroed = threshold + mn > mx;
    if (zeroed) { cx = 0.0f; cy = 0.0f; }
    return mi < 0 ? -1 : 2 * mi + (zeroed ? 1 : 0);
}

// findMinimum: the first index of the smallest value (values >= FLT_MAX and NaN never qualify),
// the largest value, then find_min_fit.
__device__ int find_min_core(const float* v, int M, float threshold, float& cx, float& cy) {
    const int S = 2 * M + 1, n = S * S;
    float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
    int mi = -1;
    for (int i = 0; i < n; ++i) {
        float val = v[i];
        mx = fmaxf(mx, val);
        if (val < mn) { mn = val; mi = i; }
    }
    return find_min_fit(v, M, mi, mn, mx, threshold, cx, cy);
}

// Order-preserving map of a non-NaN float to unsigned (for integer min / max).
__device__ __forceinline__ unsigned ordered_bits(float v) {
    const unsigned u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered_bits(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
// Argmin key of value v at index i: the smallest key is the first index of the smallest value, as
// in find_min_core's scan (-0 ties with +0; values >= FLT_MAX and NaN never qualify).
__device__ __forceinline__ unsigned long long argmin_key(float v, int i) {
    if (!(v < 3.402823466e+38f)) return ~0ull;
    return ((unsigned long long)ordered_bits(v + 0.0f) << 32) | (unsigned)i;
}

struct TrackArgs {
    const float* const* frames;
    int img_pitch, w, h, tile, cx, cy;
    const int* pa;  // reference frame of pair p
    const int* pb;  // tracked frame of pair p
    const int* pk;  // measurement index of pair p
    const float* base;  // per frame: bsx, bsy (already divided by the level), sin(rot), cos(rot)
    float* shifts;
    size_t meas_stride;
    float threshold;
    float* shift_images;  // optional: [pair][tile][S*S]
    int* decisions;       // optional: [pair][tile] decision code of find_min_fit
};


}  // namespace
