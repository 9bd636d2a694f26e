// Optical flow: flow field from tile shifts, warping, derivatives and the
// Lucas-Kanade step (R:Kernels/opticalFlow.cu; driver R:PEFStudioDX/OpticalFlow.cs:147-164).
#include <cmath>

#include "isa_device.h"
#include "isa_host.h"

namespace {

// CreateFlowFieldFromTiles (R:Kernels/opticalFlow.cu:47-93). c, s: cos and sin of the rotation, each
// evaluated in double and rounded once to float (by the launcher, once per field).
__global__ void k_flow_from_tiles(float* flow, int flow_pitch, const float* tiles, int tiles_pitch, int tw, int th,
                                  int w, int h, float bsx, float bsy, float c, float s, int quant) {
    // four consecutive pixels of one row per thread: the row's texture coordinate, weights and tile
    // rows are shared; every pixel's arithmetic is tex_linear_all<2>'s, in its order
    const int x0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x), y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x0 >= w || y >= h) return;
    const double iw = 1.0 / (double)w, ih = 1.0 / (double)h;
    const float gx = c * -bsx - s * -bsy, gy = s * -bsx + c * -bsy;
    const float pcy = (float)(y - h / 2);
    const float v = div_n((float)y + 0.5f, ih);
    const float yf = v * (float)th - 0.5f, fy = floorf(yf);
    const float b = tex_frac(yf - fy, quant), omb = 1.0f - b;
    const int ty0 = (int)fy;
    const float* ra = row_ptr(tiles, tiles_pitch, tex_addr(ty0, th, ISA_ADDR_CLAMP));
    const float* rb = row_ptr(tiles, tiles_pitch, tex_addr(ty0 + 1, th, ISA_ADDR_CLAMP));
    float out[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int x = min(x0 + i, w - 1);
        const float pcx = (float)(x - w / 2);
        const float fx = gx + (c * pcx - s * pcy - pcx), fyv = gy + (s * pcx + c * pcy - pcy);
        const float u = div_n((float)x + 0.5f, iw);
        const float xf = u * (float)tw - 0.5f, fxf = floorf(xf);
        const float a = tex_frac(xf - fxf, quant), oma = 1.0f - a;
        const int tx0 = (int)fxf;
        const int xa = tex_addr(tx0, tw, ISA_ADDR_CLAMP), xb = tex_addr(tx0 + 1, tw, ISA_ADDR_CLAMP);
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            const float t00 = ra[xa * 2 + ch], t10 = ra[xb * 2 + ch], t01 = rb[xa * 2 + ch], t11 = rb[xb * 2 + ch];
            out[2 * i + ch] = (ch == 0 ? fx : fyv) + (oma * omb * t00 + a * omb * t10 + oma * b * t01 + a * b * t11);
        }
    }
    float* o = row_ptr(flow, flow_pitch, y) + x0 * 2;
    if (x0 + 4 <= w && (reinterpret_cast<uintptr_t>(o) & 15) == 0) {
        reinterpret_cast<float4*>(o)[0] = make_float4(out[0], out[1], out[2], out[3]);
        reinterpret_cast<float4*>(o)[1] = make_float4(out[4], out[5], out[6], out[7]);
    } else {
        for (int i = 0; i < 4 && x0 + i < w; ++i) {
            o[2 * i] = out[2 * i];
            o[2 * i + 1] = out[2 * i + 1];
        }
    }
}

// WarpingKernel (R:Kernels/opticalFlow.cu:27-44): flow texture point/clamp, image linear/mirror.
__global__ void k_warp(const float* src, int src_pitch, float* out, int out_pitch, const float* flow, int flow_pitch,
                       int w, int h, int quant) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= w || iy >= h) return;
    const double iw = 1.0 / (double)w, ih = 1.0 / (double)h;
    float u = div_n((float)ix + 0.5f, iw), v = div_n((float)iy + 0.5f, ih);
    int fxi = tex_addr((int)floorf(u * (float)w), w, ISA_ADDR_CLAMP);
    int fyi = tex_addr((int)floorf(v * (float)h), h, ISA_ADDR_CLAMP);
    const float* f = row_ptr(flow, flow_pitch, fyi) + fxi * 2;
    float x = div_n((float)ix + 0.5f + f[0], iw);
    float y = div_n((float)iy + 0.5f + f[1], ih);
    row_ptr(out, out_pitch, iy)[ix] = tex_linear<1>(src, src_pitch, w, h, 0, x, y, ISA_ADDR_MIRROR, quant);
}

__device__ __forceinline__ float stencil5(const float* img, int pitch, int w, int h, float x, float y, float dx,
                                          float dy, int quant) {
    float t = tex_linear<1>(img, pitch, w, h, 0, x + 2.0f * dx, y + 2.0f * dy, ISA_ADDR_MIRROR, quant);
    t -= tex_linear<1>(img, pitch, w, h, 0, x + 1.0f * dx, y + 1.0f * dy, ISA_ADDR_MIRROR, quant) * 8.0f;
    t += tex_linear<1>(img, pitch, w, h, 0, x - 1.0f * dx, y - 1.0f * dy, ISA_ADDR_MIRROR, quant) * 8.0f;
    t -= tex_linear<1>(img, pitch, w, h, 0, x - 2.0f * dx, y - 2.0f * dy, ISA_ADDR_MIRROR, quant);
    t = div_n(t, 1.0 / 12.0);
    return t;
}

// Texture taps of the 5-point stencils, per column and per row. The stencil fetches at
// x + k dx (k = -2..2) depend only on the column (and y + k dy only on the row), so their texel
// indices (mirrored) and weights are tabulated once per image size; a fetch then combines one column
// entry with one row entry. Same float operations as tex_linear, so the values are identical.
struct Tap {
    int i0, i1;
    float a;
};

// Stencil tap k (-2..2) of column (or row) ix of an n-texel axis: the fetch at (ix + 0.5) / n + k / n.
__device__ __forceinline__ Tap stencil_tap(int ix, int n, int k, int quant) {
    float d = 1.0f / (float)n;
    float x = ((float)ix + 0.5f) * d;
    float u = k == 0 ? x : (k > 0 ? x + (float)k * d : x - (float)(-k) * d);
    float xf = u * (float)n - 0.5f;
    float fl = floorf(xf);
    Tap t;
    t.a = tex_frac(xf - fl, quant);
    t.i0 = tex_addr((int)fl, n, ISA_ADDR_MIRROR);
    t.i1 = tex_addr((int)fl + 1, n, ISA_ADDR_MIRROR);
    return t;
}

__global__ void k_tap_tables(Tap* col, Tap* row, int w, int h, int quant) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool is_col = i < w;
    const int n = is_col ? w : h;
    const int ix = is_col ? i : i - w;
    if (ix >= n || i >= w + h) return;
    Tap* out = (is_col ? col : row) + (size_t)ix * 5;
#pragma unroll
    for (int k = -2; k <= 2; ++k) out[k + 2] = stencil_tap(ix, n, k, quant);
}

// ComputeDerivativesKernel (R:Kernels/opticalFlow.cu:96-147): 5-point stencils of both images along x
// and y, averaged, and the temporal difference, with the stencil taps read from the tables.
// Every tap index of pixel (ix, iy) lies in [ix-3, ix+3] x [iy-3, iy+3] (mirroring at the edges
// only moves an index towards the pixel), so both images are staged in LDS as the block plus a
// 3-pixel halo and all fetches are served from there.
constexpr int DV_BX = 64, DV_BY = 8, DV_H = 3, DV_LW = DV_BX + 2 * DV_H, DV_LH = DV_BY + 2 * DV_H;

struct LTap {
    int c0, c1;  // local texel columns (or rows)
    float a;
    int p;       // the single texel when a is 0 or 1 (then the bilinear sum is that texel exactly)
};

template <int LW>
__device__ __forceinline__ float lds_fetch_w(const float (*L)[LW], const LTap& c, const LTap& r) {
    const float a = c.a, b = r.a;
    // weights at the corners reduce the bilinear sum to one texel exactly (finite data)
    if (b == 0.0f) {
        if (a == 0.0f) return L[r.c0][c.c0];
        if (a == 1.0f) return L[r.c0][c.c1];
    } else if (b == 1.0f) {
        if (a == 0.0f) return L[r.c1][c.c0];
        if (a == 1.0f) return L[r.c1][c.c1];
    }
    float t00 = L[r.c0][c.c0], t10 = L[r.c0][c.c1], t01 = L[r.c1][c.c0], t11 = L[r.c1][c.c1];
    float oma = 1.0f - a, omb = 1.0f - b;
    return oma * omb * t00 + a * omb * t10 + oma * b * t01 + a * b * t11;
}

__device__ __forceinline__ float lds_fetch(const float (*L)[DV_LW], const LTap& c, const LTap& r) {
    return lds_fetch_w<DV_LW>(L, c, r);
}

// TABLE: taps read from the tables of k_tap_tables; otherwise evaluated in place (the same floats).
template <bool TABLE>
__global__ __launch_bounds__(DV_BX* DV_BY) void k_derivatives_tab(const float* src, const float* tgt, int img_pitch,
                                                                   float* Ix, float* Iy, float* Iz, int pitch, int w,
                                                                   int h, const Tap* __restrict__ col,
                                                                   const Tap* __restrict__ row, int quant) {
    __shared__ float Ls[DV_LH][DV_LW], Lt[DV_LH][DV_LW];
    const int x0 = blockIdx.x * DV_BX - DV_H, y0 = blockIdx.y * DV_BY - DV_H;
    const int tid = threadIdx.y * DV_BX + threadIdx.x;
    for (int i = tid; i < DV_LH * DV_LW; i += DV_BX * DV_BY) {
        int ly = i / DV_LW, lx = i - ly * DV_LW;
        int gy = clampi(y0 + ly, 0, h - 1), gx = clampi(x0 + lx, 0, w - 1);
        Ls[ly][lx] = row_ptr(src, img_pitch, gy)[gx];
        Lt[ly][lx] = row_ptr(tgt, img_pitch, gy)[gx];
    }
    __syncthreads();
    const int ix = blockIdx.x * DV_BX + threadIdx.x, iy = blockIdx.y * DV_BY + threadIdx.y;
    if (ix >= w || iy >= h) return;
    LTap c[5], r[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        Tap tc = TABLE ? col[(size_t)ix * 5 + k] : stencil_tap(ix, w, k - 2, quant);
        Tap tr = TABLE ? row[(size_t)iy * 5 + k] : stencil_tap(iy, h, k - 2, quant);
        c[k] = {clampi(tc.i0 - x0, 0, DV_LW - 1), clampi(tc.i1 - x0, 0, DV_LW - 1), tc.a, 0};
        r[k] = {clampi(tr.i0 - y0, 0, DV_LH - 1), clampi(tr.i1 - y0, 0, DV_LH - 1), tr.a, 0};
    }
    auto sx = [&](const float (*L)[DV_LW]) {
        float t = lds_fetch(L, c[4], r[2]);
        t -= lds_fetch(L, c[3], r[2]) * 8.0f;
        t += lds_fetch(L, c[1], r[2]) * 8.0f;
        t -= lds_fetch(L, c[0], r[2]);
        t = div_n(t, 1.0 / 12.0);
        return t;
    };
    auto sy = [&](const float (*L)[DV_LW]) {
        float t = lds_fetch(L, c[2], r[4]);
        t -= lds_fetch(L, c[2], r[3]) * 8.0f;
        t += lds_fetch(L, c[2], r[1]) * 8.0f;
        t -= lds_fetch(L, c[2], r[0]);
        t = div_n(t, 1.0 / 12.0);
        return t;
    };
    float t0 = sx(Ls), t1 = sx(Lt);
    row_ptr(Ix, pitch, iy)[ix] = (t0 + t1) * 0.5f;
    row_ptr(Iz, pitch, iy)[ix] = lds_fetch(Ls, c[2], r[2]) - lds_fetch(Lt, c[2], r[2]);
    t0 = sy(Ls);
    t1 = sy(Lt);
    row_ptr(Iy, pitch, iy)[ix] = (t0 + t1) * 0.5f;
}

// ComputeDerivatives2Kernel (R:Kernels/opticalFlow.cu:150-185).
__global__ void k_derivatives2(const float* img, int img_pitch, float* Ix, float* Iy, int pitch, int w, int h,
                               int quant) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= w || iy >= h) return;
    float dx = 1.0f / (float)w, dy = 1.0f / (float)h;
    float x = ((float)ix + 0.5f) * dx, y = ((float)iy + 0.5f) * dy;
    row_ptr(Ix, pitch, iy)[ix] = stencil5(img, img_pitch, w, h, x, y, dx, 0.0f, quant);
    row_ptr(Iy, pitch, iy)[ix] = stencil5(img, img_pitch, w, h, x, y, 0.0f, dy, quant);
}

// Closed-form 2x2 SVD pseudo-inverse of [[a, b], [b, d]] (R:Kernels/opticalFlow.cu:236-292).
// Returns false when the gate sigma1 < minDet holds (the reference's fminf(sigma1, sigma1)).
// cos and sin of 0.5 * atan2(y, x) without trigonometric calls: with (X, Y) scaled to a unit vector
// (cos 2t, sin 2t), cos t = sqrt((1 + cos 2t) / 2) and sin t = sin 2t / (2 cos t) for x >= 0, and the
// mirrored pair for x < 0 (t in [-pi/2, pi/2], so cos t >= 0; the sign of sin t follows y, -0 included).
__device__ __forceinline__ void half_angle(float y, float x, float& c, float& s) {
    float m = fmaxf(fabsf(x), fabsf(y));
    if (!(m > 0.0f)) {  // atan2(+-0, +-0) and NaN inputs: t = 0 (x = -0 gives pi; only the sign of
        c = (signbit(x) && !isnan(x)) ? 0.0f : 1.0f;  // U matters there and no update survives the gate)
        s = (signbit(x) && !isnan(x)) ? copysignf(1.0f, y) : 0.0f;
        return;
    }
    const float rm = __builtin_amdgcn_rcpf(m);
    float xs = x * rm, ys = y * rm;
    float rr = rsqrtf(xs * xs + ys * ys);
    float c2 = xs * rr, s2 = ys * rr;
    if (c2 >= 0.0f) {
        c = sqrtf(0.5f * (1.0f + c2));
        s = s2 * 0.5f * __builtin_amdgcn_rcpf(c);
    } else {
        s = copysignf(sqrtf(0.5f * (1.0f - c2)), s2);
        c = s2 * 0.5f * __builtin_amdgcn_rcpf(s);
    }
}

// Ill-conditioned window: lambda2 / lambda1 below ill_ratio, as det / trace^2 of the FP32 sums in double
// (exact products of floats); also true for det <= 0 and NaN. The FP32 closed form loses lambda2 there
// (S1 - S2 = 2 lambda2^2 cancels next to S1 ~ lambda1^2).
__device__ __forceinline__ bool lk_ill(float fa, float fb, float fd, double ill_ratio) {
    const double a = fa, b = fb, d = fd, tr = a + d;
    return !(a * d - b * b > ill_ratio * tr * tr);
}

// FP64 solve of an ill-conditioned window from its FP32 sums: the reference's gate (lambda1 = sigma1 <
// min_det skips), u = G^-1 (b0, b1) by the adjugate for det > 0 (the reference's closed-form
// pseudo-inverse is the inverse there), the rank-1 pseudo-inverse G / trace^2 for det <= 0; NaN -> 0.
__device__ __forceinline__ bool lk_solve_fp64(float fa, float fb, float fd, float fb0, float fb1, float min_det,
                                              float& u0, float& u1) {
    const double a = fa, b = fb, d = fd, b0 = fb0, b1 = fb1, tr = a + d;
    const double lam1 = 0.5 * (tr + sqrt((a - d) * (a - d) + 4.0 * b * b));
    if (lam1 < (double)min_det) return false;
    const double det = a * d - b * b;
    double v0, v1;
    if (det > 0) {
        v0 = (d * b0 - b * b1) / det;
        v1 = (a * b1 - b * b0) / det;
    } else {
        const double t2 = tr * tr;
        v0 = (a * b0 + b * b1) / t2;
        v1 = (b * b0 + d * b1) / t2;
    }
    u0 = isnan(v0) ? 0.0f : (float)v0;
    u1 = isnan(v1) ? 0.0f : (float)v1;
    return true;
}

// FP32 solve of a window that lk_ill has passed (det > ill_ratio * trace^2 > 0 in exact products): G is
// invertible, so the reference's closed-form pseudo-inverse V diag(1 / sigma) U^T is G^-1 = adj(G) / det,
// formed here without the SVD's square roots and half angle. The gate is the same sigma1 = lambda1 of the
// symmetric G; det and the adjugate products take one rounding each (fma), the reciprocal and the square
// root are the hardware ones (about 1 ulp). NaN -> 0 as in the reference.
__device__ __forceinline__ bool lk_solve_wc(float a, float b, float d, float b0, float b1, float min_det, float& u0,
                                            float& u1) {
    const float e = a - d;
    const float lam1 = 0.5f * ((a + d) + __builtin_amdgcn_sqrtf(e * e + 4.0f * (b * b)));
    if (lam1 < min_det) return false;
    const float r = __builtin_amdgcn_rcpf(fmaf(a, d, -(b * b)));
    const float v0 = fmaf(d, b0, -(b * b1)) * r, v1 = fmaf(a, b1, -(b * b0)) * r;
    u0 = isnan(v0) ? 0.0f : v0;
    u1 = isnan(v1) ? 0.0f : v1;
    return true;
}

template <bool EXACT>
__device__ __forceinline__ bool lk_pinv(float a, float b, float d, float min_det, float inv[4]) {
    float c = b;
    float ct, st;
    if (EXACT) {
        float theta = 0.5f * atan2_r(2.0f * a * c + 2.0f * b * d, a * a + b * b - c * c - d * d);
        ct = cos_r(theta);
        st = sin_r(theta);
    } else {
        half_angle(2.0f * a * c + 2.0f * b * d, a * a + b * b - c * c - d * d, ct, st);
    }
    float UT0 = ct, UT2 = -st, UT1 = st, UT3 = ct;
    // the fast form (fused LK) takes the hardware square root and reciprocal (about 1 ulp) where the
    // exact form rounds correctly
    auto sqrt_ = [](float v) { return EXACT ? sqrtf(v) : __builtin_amdgcn_sqrtf(v); };
    auto rcp_ = [](float v) { return EXACT ? 1.0f / v : __builtin_amdgcn_rcpf(v); };
    float S1 = a * a + b * b + c * c + d * d;
    float S2 = sqrt_((a * a + b * b - c * c - d * d) * (a * a + b * b - c * c - d * d) +
                     4.0f * (a * c + b * d) * (a * c + b * d));
    float sigma1 = sqrt_((S1 + S2) / 2.0f);
    float sigma2 = sqrt_((S1 - S2) / 2.0f);
    float smin = fminf(sigma1, sigma1);
    if (smin < min_det) return false;
    sigma1 = sigma1 != 0 ? rcp_(sigma1) : 0.0f;
    sigma2 = sigma2 != 0 ? rcp_(sigma2) : 0.0f;
    float S0 = sigma1, Sb = 0.0f, Sc = 0.0f, S3 = sigma2;
    float ce, se;
    if (EXACT) {
        float eps = 0.5f * atan2_r(2.0f * a * b + 2.0f * c * d, a * a - b * b + c * c - d * d);
        ce = cos_r(eps);
        se = sin_r(eps);
    } else {
        half_angle(2.0f * a * b + 2.0f * c * d, a * a - b * b + c * c - d * d, ce, se);
    }
    float s11 = (a * ct + c * st) * ce + (b * ct + d * st) * se;
    float s22 = (a * st - c * ct) * se + (-b * st + d * ct) * ce;
    s11 = s11 > 0.0f ? 1.0f : s11 < 0 ? -1.0f : 0.0f;
    s22 = s22 > 0.0f ? 1.0f : s22 < 0 ? -1.0f : 0.0f;
    float V0 = s11 * ce, V1 = -s22 * se, V2 = s11 * se, V3 = s22 * ce;
    float m0 = S0 * UT0 + Sb * UT2, m1 = S0 * UT1 + Sb * UT3, m2 = Sc * UT0 + S3 * UT2, m3 = Sc * UT1 + S3 * UT3;
    inv[0] = V0 * m0 + V1 * m2;
    inv[1] = V0 * m1 + V1 * m3;
    inv[2] = V2 * m0 + V3 * m2;
    inv[3] = V2 * m1 + V3 * m3;
    return true;
}

// lucasKanadeOptim (R:Kernels/opticalFlow.cu:189-325): per-pixel window loops.
__global__ void k_lk_step(float* flow, int flow_pitch, const float* Fx, const float* Fy, const float* Ft, int pitch,
                          int w, int h, int hw, float min_det) {
    int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px < hw || px >= w - hw || py < hw || py >= h - hw) return;
    int ws = hw * 2 + 1;
    float g0 = 0, g1 = 0, g3 = 0;
    for (int y = -hw; y <= hw; ++y) {
        const float* rx = row_ptr(Fx, pitch, py + y);
        const float* ry = row_ptr(Fy, pitch, py + y);
        for (int x = -hw; x <= hw; ++x) {
            float dx = rx[px + x], dy = ry[px + x];
            g0 += dx * dx;
            g1 += dx * dy;
            g3 += dy * dy;
        }
    }
    float inv[4];
    if (!lk_pinv<true>(g0, g1, g3, min_det, inv)) return;
    float u0 = 0, u1 = 0;
    for (int i = 0; i < ws * ws; ++i) {
        int yy = i / ws, xx = i - yy * ws;
        int gy = py + yy - hw, gx = px + xx - hw;
        float dx = row_ptr(Fx, pitch, gy)[gx], dy = row_ptr(Fy, pitch, gy)[gx], dt = row_ptr(Ft, pitch, gy)[gx];
        u0 += (inv[0] * dx + inv[1] * dy) * dt;
        u1 += (inv[2] * dx + inv[3] * dy) * dt;
    }
    u0 = isnan(u0) ? 0.0f : u0;
    u1 = isnan(u1) ? 0.0f : u1;
    float* f = row_ptr(flow, flow_pitch, py) + px * 2;
    f[0] += u0;
    f[1] += u1;
}

// Lucas-Kanade step with the five window sums (Ix^2, IxIy, Iy^2, IxIt, IyIt) formed by separable
// box sums in LDS instead of 2 (2h+1)^2 loads per pixel: a plain (2h+1)-term sum per output of each
// horizontal and vertical pass. The solve is the same closed form (windows below ill_ratio in FP64,
// as in the fused kernel); the update is pinv . (sum Ix It, sum Iy It), which equals the reference's
// sum of pinv . (Ix, Iy) It.
constexpr int LK_BX = 32, LK_BY = 16, LK_HMAX = 10, LK_SEG = 8;
__global__ __launch_bounds__(LK_BX* LK_BY) void k_lk_step_box(float* flow, int flow_pitch, const float* Fx,
                                                               const float* Fy, const float* Ft, int pitch, int w,
                                                               int h, int hw, float min_det, double ill_ratio) {
    constexpr int RW = LK_BX + 2 * LK_HMAX;  // region width (max)
    __shared__ float s_p[LK_BY + 2 * LK_HMAX][RW][5];
    __shared__ float s_h[LK_BY + 2 * LK_HMAX][LK_BX][5];
    __shared__ float s_v[LK_BY][LK_BX][5];
    const int x0 = blockIdx.x * LK_BX, y0 = blockIdx.y * LK_BY;
    const int tid = threadIdx.y * LK_BX + threadIdx.x, nt = LK_BX * LK_BY;
    const int rows = LK_BY + 2 * hw, cols = LK_BX + 2 * hw, win = 2 * hw + 1;
    for (int i = tid; i < rows * cols; i += nt) {
        int ly = i / cols, lx = i - ly * cols;
        int gx = clampi(x0 - hw + lx, 0, w - 1), gy = clampi(y0 - hw + ly, 0, h - 1);
        float dx = row_ptr(Fx, pitch, gy)[gx], dy = row_ptr(Fy, pitch, gy)[gx], dt = row_ptr(Ft, pitch, gy)[gx];
        s_p[ly][lx][0] = dx * dx;
        s_p[ly][lx][1] = dx * dy;
        s_p[ly][lx][2] = dy * dy;
        s_p[ly][lx][3] = dx * dt;
        s_p[ly][lx][4] = dy * dt;
    }
    __syncthreads();
    // horizontal: task (row, segment, quantity)
    constexpr int HSEG = LK_BX / LK_SEG;
    for (int task = tid; task < rows * HSEG * 5; task += nt) {
        int q = task % 5, rs = task / 5, seg = rs % HSEG, ly = rs / HSEG;
        int xs = seg * LK_SEG;
        for (int i = 0; i < LK_SEG; ++i) {
            float acc = 0.0f;
            for (int k = 0; k < win; ++k) acc += s_p[ly][xs + i + k][q];
            s_h[ly][xs + i][q] = acc;
        }
    }
    __syncthreads();
    // vertical: task (column, segment, quantity)
    constexpr int VSEG = LK_BY / LK_SEG;
    for (int task = tid; task < LK_BX * VSEG * 5; task += nt) {
        int q = task % 5, cs = task / 5, seg = cs % VSEG, lx = cs / VSEG;
        int ys = seg * LK_SEG;
        for (int i = 0; i < LK_SEG; ++i) {
            float acc = 0.0f;
            for (int k = 0; k < win; ++k) acc += s_h[ys + i + k][lx][q];
            s_v[ys + i][lx][q] = acc;
        }
    }
    __syncthreads();
    const int px = x0 + threadIdx.x, py = y0 + threadIdx.y;
    if (px < hw || px >= w - hw || py < hw || py >= h - hw) return;
    float g[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) g[q] = s_v[threadIdx.y][threadIdx.x][q];
    float u0, u1;
    if (ill_ratio > 0.0 && lk_ill(g[0], g[1], g[2], ill_ratio)) {
        if (!lk_solve_fp64(g[0], g[1], g[2], g[3], g[4], min_det, u0, u1)) return;
    } else {
        float inv[4];
        if (!lk_pinv<false>(g[0], g[1], g[2], min_det, inv)) return;
        u0 = inv[0] * g[3] + inv[1] * g[4];
        u1 = inv[2] * g[3] + inv[3] * g[4];
        u0 = isnan(u0) ? 0.0f : u0;
        u1 = isnan(u1) ? 0.0f : u1;
    }
    float* f = row_ptr(flow, flow_pitch, py) + px * 2;
    f[0] += u0;
    f[1] += u1;
}

// One Lucas-Kanade iteration fused into one kernel (warp, derivatives, window sums, solve, update)
// for half-windows HW <= 8. Per 64 x 32 block of output pixels:
//  A. the warped source and the target over the block plus HW + 3 pixels of halo, into LDS
//     (same arithmetic as k_warp; the flow texel of each column and row is tabulated per block);
//  B. Ix, Iy, It over the block plus HW of halo from those tiles (same arithmetic as
//     k_derivatives_tab; pixels outside the image are never inside the window of an updated
//     pixel and are set to 0);
//  C. the five window sums as separable box sums: per region row, sums along x over segments of 8
//     output columns (from a register window of the row's derivatives), then per output column sums
//     along y over 4 output rows, both by window_sums (no running subtraction);
//  D. the solve and update, reading flow_in and writing flow_out (the halo of the warp reads
//     neighbouring blocks' flow, so the flow is double-buffered across iterations). With an ill ratio
//     set, windows that pass the FP64 conditioning test take the FP32 adjugate inverse (lk_solve_wc) and
//     the others the FP64 solve; without one, every window takes the closed-form pseudo-inverse.
constexpr int LKF_BX = 64, LKF_BY = 32, LKF_RPT = 4;  // 512 threads, 4 output rows each
constexpr int LKF_SEG = 8;                            // output columns per horizontal task

// Sums of W consecutive terms, out[i] = p[i] + ... + p[i + W - 1] for i = 0 .. N - 1, built from partial
// sums that only add (a running sum that subtracts the term leaving the window keeps the rounding error
// of every larger sum it passed through). Per group of G = min(N, W) outputs the terms shared by all
// its windows form one core sum; each window adds the suffix sum of the terms left of the core and the
// prefix sum of those right of it: out = (left + core) + right, so each is within gamma_W of the sum of
// its terms' magnitudes, as a plain loop is.
template <int N, int W>
__device__ __forceinline__ void window_sums(const float (&p)[N + W - 1], float (&out)[N]) {
    constexpr int G = N < W ? N : W;
#pragma unroll
    for (int g0 = 0; g0 < N; g0 += G) {
        const int gn = N - g0 < G ? N - g0 : G;     // outputs i = g0 .. g0 + gn - 1
        const int c0 = g0 + gn - 1, c1 = g0 + W - 1;  // terms c0 .. c1 lie in every window of the group
        float core = p[c0];
#pragma unroll
        for (int k = c0 + 1; k <= c1; ++k) core += p[k];
        float left[G], right[G];
        left[gn - 1] = 0.0f;
        right[0] = 0.0f;
#pragma unroll
        for (int j = gn - 2; j >= 0; --j) left[j] = j == gn - 2 ? p[g0 + j] : p[g0 + j] + left[j + 1];
#pragma unroll
        for (int j = 1; j < gn; ++j) right[j] = j == 1 ? p[c1 + 1] : right[j - 1] + p[c1 + j];
#pragma unroll
        for (int j = 0; j < gn; ++j) {
            float v = core;
            if (j < gn - 1) v = left[j] + v;
            if (j > 0) v = v + right[j];
            out[g0 + j] = v;
        }
    }
}

template <int HW>
struct LkfShared {
    static constexpr int DH = LKF_BY + 2 * HW, DW = LKF_BX + 2 * HW;  // derivative region
    static constexpr int WR = DH + 6, WC = DW + 6;                     // warped region
    static constexpr int LDW = DW | 1, HS = LKF_BX + 1;                // padded LDS row strides
    struct AB {                                                       // phases A and B
        float Lw[WR][WC], Lt[WR][WC];
        LTap Tc[DW][5], Tr[DH][5];
        int fx[WC], fy[WR];  // flow texel (point sampling, clamp) of each warped column and row
        float Ld[3][DH][LDW];
    };
    union {
        AB ab;
        float H[5][DH][HS];  // phase C: horizontal window sums, written once every thread has read Ld
    } u;
    int mixed;  // some tap of the block has a weight strictly between 0 and 1
    int ident;  // every tap of the block is the neighbouring texel itself (cleared otherwise)
};

template <int HW>
__global__ __launch_bounds__(LKF_BX*(LKF_BY / LKF_RPT)) void k_lk_fused(
    const float* src, const float* tgt, int img_pitch, const float* flow_in, float* flow_out, int flow_pitch, int w,
    int h, float min_det, int quant, const Tap* __restrict__ col, const Tap* __restrict__ row, double ill_ratio) {
    using S = LkfShared<HW>;
    constexpr int DH = S::DH, DW = S::DW, WR = S::WR, WC = S::WC;
    constexpr int NT = LKF_BX * (LKF_BY / LKF_RPT);
    __shared__ S sm;
    auto& Lw = sm.u.ab.Lw;
    auto& Lt = sm.u.ab.Lt;
    auto& Ld = sm.u.ab.Ld;
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * LKF_BX, y0 = blockIdx.y * LKF_BY;
    const int wx0 = x0 - HW - 3, wy0 = y0 - HW - 3;
    const double iw = 1.0 / (double)w, ih = 1.0 / (double)h;
    // flow_in at this thread's output pixels (read by phase D), loaded first so that it has arrived by then
    // (clamped to the image; rows and columns past it are not written)
    float2 fpre[LKF_RPT];
    {
        const int pxp = min(x0 + tid % LKF_BX, w - 1), oyp = y0 + (tid / LKF_BX) * LKF_RPT;
#pragma unroll
        for (int o = 0; o < LKF_RPT; ++o)
            fpre[o] = *(const float2*)(row_ptr(flow_in, flow_pitch, min(oyp + o, h - 1)) + pxp * 2);
    }
    // this thread's entries of the tap tables of the block's derivative region, loaded before the first
    // barrier; the flow texels (WarpingKernel reads the flow by point sampling at ((x + 0.5) / w,
    // (y + 0.5) / h)) go to LDS meanwhile
    constexpr int NTAP = ((DW + DH) * 5 + NT - 1) / NT;
    Tap tpre[NTAP];
#pragma unroll
    for (int j = 0; j < NTAP; ++j) {
        const int i = min(tid + j * NT, (DW + DH) * 5 - 1);
        const bool is_c = i < DW * 5;
        const int e = is_c ? i : i - DW * 5, pix = e / 5, k = e - pix * 5;
        const int g = (is_c ? x0 : y0) - HW + pix, n = is_c ? w : h;
        tpre[j] = (is_c ? col : row)[(size_t)clampi(g, 0, n - 1) * 5 + k];
    }
    for (int i = tid; i < WC + WR; i += NT) {
        const bool is_c = i < WC;
        const int g = is_c ? clampi(wx0 + i, 0, w - 1) : clampi(wy0 + i - WC, 0, h - 1);
        const int n = is_c ? w : h;
        const float u = div_n((float)g + 0.5f, is_c ? iw : ih);
        const int t = tex_addr((int)floorf(u * (float)n), n, ISA_ADDR_CLAMP);
        if (is_c) sm.u.ab.fx[i] = t;
        else sm.u.ab.fy[i - WC] = t;
    }
    if (tid == 0) {
        sm.mixed = 0;
        sm.ident = 1;
    }
    __syncthreads();
    // the tap tables, translated to warped-tile coordinates
#pragma unroll
    for (int j = 0; j < NTAP; ++j) {
        const int i = tid + j * NT;
        if (i >= (DW + DH) * 5) break;
        const Tap t = tpre[j];
        const bool is_c = i < DW * 5;
        const int e = is_c ? i : i - DW * 5, pix = e / 5, k = e - pix * 5;
        const int o = is_c ? wx0 : wy0, lim = is_c ? WC : WR;
        LTap l = {clampi(t.i0 - o, 0, lim - 1), clampi(t.i1 - o, 0, lim - 1), t.a, 0};
        l.p = t.a == 0.0f ? l.c0 : l.c1;
        if (t.a != 0.0f && t.a != 1.0f) sm.mixed = 1;
        if (l.p != pix + k + 1) sm.ident = 0;  // tap k of region pixel pix is warped-tile index pix + 1 + k
        if (is_c) sm.u.ab.Tc[pix][k] = l;
        else sm.u.ab.Tr[pix][k] = l;
    }
    // (phase A reads only the flow texels, complete at the barrier above; the tap tables and flags are read
    // after the barrier that ends phase A)
    // A. warped source and target. All flow loads first, then the texel coordinates of every element.
    // When the 2x2 footprint of every element of the wave lies inside the image (the common case)
    // the texel indices are the coordinates themselves and all gathers are issued together from
    // 32-bit element offsets; otherwise each index is mirrored as the texture unit does.
    constexpr int NA = (WR * WC + NT - 1) / NT;
    const unsigned pf = (unsigned)img_pitch / 4u, fpf = (unsigned)flow_pitch / 4u;
    float fa[NA], fb[NA];
    int gxs[NA], gys[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        const int i = min(tid + k * NT, WR * WC - 1);
        const int ly = i / WC, lx = i - ly * WC;
        gxs[k] = clampi(wx0 + lx, 0, w - 1);
        gys[k] = clampi(wy0 + ly, 0, h - 1);
        const unsigned fo = ((unsigned)sm.u.ab.fy[ly] * fpf + (unsigned)sm.u.ab.fx[lx] * 2u) * 4u;
        fa[k] = ld_off(flow_in, fo);
        fb[k] = ld_off(flow_in, fo + 4u);
    }
    float wa[NA], wb[NA];
    int ix0[NA], iy0[NA];
    bool in = true;
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        float x = div_n((float)gxs[k] + 0.5f + fa[k], iw);
        float y = div_n((float)gys[k] + 0.5f + fb[k], ih);
        // tex_linear<1> (mirror) without its texel-centre shortcut: same value for finite data
        float xf = x * (float)w - 0.5f, yf = y * (float)h - 0.5f;
        float fx = floorf(xf), fy = floorf(yf);
        wa[k] = tex_frac(xf - fx, quant);
        wb[k] = tex_frac(yf - fy, quant);
        ix0[k] = (int)fx;
        iy0[k] = (int)fy;
        in = in & ((unsigned)ix0[k] < (unsigned)(w - 1)) & ((unsigned)iy0[k] < (unsigned)(h - 1));
    }
    float val[NA], tv[NA];
    if (__all(in)) {
        float t00[NA], t10[NA], t01[NA], t11[NA];
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            const unsigned o = ((unsigned)iy0[k] * pf + (unsigned)ix0[k]) * 4u, o1 = o + pf * 4u;
            t00[k] = ld_off(src, o);
            t10[k] = ld_off(src, o + 4u);
            t01[k] = ld_off(src, o1);
            t11[k] = ld_off(src, o1 + 4u);
            tv[k] = ld_off(tgt, ((unsigned)gys[k] * pf + (unsigned)gxs[k]) * 4u);
        }
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            const float a = wa[k], b = wb[k], oma = 1.0f - a, omb = 1.0f - b;
            val[k] = oma * omb * t00[k] + a * omb * t10[k] + oma * b * t01[k] + a * b * t11[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            int xa = tex_addr(ix0[k], w, ISA_ADDR_MIRROR), xb = tex_addr(ix0[k] + 1, w, ISA_ADDR_MIRROR);
            int ya = tex_addr(iy0[k], h, ISA_ADDR_MIRROR), yb = tex_addr(iy0[k] + 1, h, ISA_ADDR_MIRROR);
            const unsigned oa = (unsigned)ya * pf, ob = (unsigned)yb * pf;
            const float t00 = ld_off(src, (oa + xa) * 4u), t10 = ld_off(src, (oa + xb) * 4u);
            const float t01 = ld_off(src, (ob + xa) * 4u), t11 = ld_off(src, (ob + xb) * 4u);
            const float a = wa[k], b = wb[k], oma = 1.0f - a, omb = 1.0f - b;
            val[k] = oma * omb * t00 + a * omb * t10 + oma * b * t01 + a * b * t11;
            tv[k] = ld_off(tgt, ((unsigned)gys[k] * pf + (unsigned)gxs[k]) * 4u);
        }
    }
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        const int i = tid + k * NT;
        if (i < WR * WC) {
            const int ly = i / WC, lx = i - ly * WC;
            Lw[ly][lx] = val[k];
            Lt[ly][lx] = tv[k];
        }
    }
    __syncthreads();
    // B. derivatives
    const bool mixed = sm.mixed != 0, ident = !mixed && sm.ident != 0;
    for (int i = tid; i < DH * DW; i += NT) {
        int ly = i / DW, lx = i - ly * DW;
        int gx = x0 - HW + lx, gy = y0 - HW + ly;
        if (gx < 0 || gx >= w || gy < 0 || gy >= h) {
            Ld[0][ly][lx] = 0.0f;
            Ld[1][ly][lx] = 0.0f;
            Ld[2][ly][lx] = 0.0f;
            continue;
        }
        if (ident) {
            // every tap is the neighbouring texel itself: tap k of derivative pixel (ly, lx) is
            // warped-tile row ly + 1 + k (column lx + 1 + k)
            const int r2 = ly + 3, c2 = lx + 3;
            auto px5 = [&](const float (*L)[WC]) {
                float t = L[r2][c2 + 2];
                t -= L[r2][c2 + 1] * 8.0f;
                t += L[r2][c2 - 1] * 8.0f;
                t -= L[r2][c2 - 2];
                t = div_n(t, 1.0 / 12.0);
                return t;
            };
            auto py5 = [&](const float (*L)[WC]) {
                float t = L[r2 + 2][c2];
                t -= L[r2 + 1][c2] * 8.0f;
                t += L[r2 - 1][c2] * 8.0f;
                t -= L[r2 - 2][c2];
                t = div_n(t, 1.0 / 12.0);
                return t;
            };
            float t0 = px5(Lw), t1 = px5(Lt);
            Ld[0][ly][lx] = (t0 + t1) * 0.5f;
            Ld[2][ly][lx] = Lw[r2][c2] - Lt[r2][c2];
            t0 = py5(Lw);
            t1 = py5(Lt);
            Ld[1][ly][lx] = (t0 + t1) * 0.5f;
            continue;
        }
        const LTap* c = sm.u.ab.Tc[lx];
        const LTap* r = sm.u.ab.Tr[ly];
        if (!mixed) {
            // every tap lands on one texel: the stencils are plain texel differences
            const int r2 = r[2].p, c2 = c[2].p;
            auto px5 = [&](const float (*L)[WC]) {
                float t = L[r2][c[4].p];
                t -= L[r2][c[3].p] * 8.0f;
                t += L[r2][c[1].p] * 8.0f;
                t -= L[r2][c[0].p];
                t = div_n(t, 1.0 / 12.0);
                return t;
            };
            auto py5 = [&](const float (*L)[WC]) {
                float t = L[r[4].p][c2];
                t -= L[r[3].p][c2] * 8.0f;
                t += L[r[1].p][c2] * 8.0f;
                t -= L[r[0].p][c2];
                t = div_n(t, 1.0 / 12.0);
                return t;
            };
            float t0 = px5(Lw), t1 = px5(Lt);
            Ld[0][ly][lx] = (t0 + t1) * 0.5f;
            Ld[2][ly][lx] = Lw[r2][c2] - Lt[r2][c2];
            t0 = py5(Lw);
            t1 = py5(Lt);
            Ld[1][ly][lx] = (t0 + t1) * 0.5f;
            continue;
        }
        auto sx = [&](const float (*L)[WC]) {
            float t = lds_fetch_w<WC>(L, c[4], r[2]);
            t -= lds_fetch_w<WC>(L, c[3], r[2]) * 8.0f;
            t += lds_fetch_w<WC>(L, c[1], r[2]) * 8.0f;
            t -= lds_fetch_w<WC>(L, c[0], r[2]);
            t = div_n(t, 1.0 / 12.0);
            return t;
        };
        auto sy = [&](const float (*L)[WC]) {
            float t = lds_fetch_w<WC>(L, c[2], r[4]);
            t -= lds_fetch_w<WC>(L, c[2], r[3]) * 8.0f;
            t += lds_fetch_w<WC>(L, c[2], r[1]) * 8.0f;
            t -= lds_fetch_w<WC>(L, c[2], r[0]);
            t = div_n(t, 1.0 / 12.0);
            return t;
        };
        float t0 = sx(Lw), t1 = sx(Lt);
        Ld[0][ly][lx] = (t0 + t1) * 0.5f;
        Ld[2][ly][lx] = lds_fetch_w<WC>(Lw, c[2], r[2]) - lds_fetch_w<WC>(Lt, c[2], r[2]);
        t0 = sy(Lw);
        t1 = sy(Lt);
        Ld[1][ly][lx] = (t0 + t1) * 0.5f;
    }
    __syncthreads();
    // C1. horizontal window sums of the five products for every region row: a task covers LKF_SEG
    // output columns of one row (the derivative window of the task stays in registers). The sums
    // overwrite the LDS of phases A and B, so they are stored after a barrier.
    constexpr int NSEG = LKF_BX / LKF_SEG, NW = LKF_SEG + 2 * HW;
    static_assert(DH * NSEG <= NT, "one horizontal task per thread");
    static_assert(sizeof(sm.u.H) <= sizeof(sm.u.ab), "the window sums fit in the phase A/B storage");
    const bool has_task = tid < DH * NSEG;
    const int ry = tid % DH, xs = (tid / DH) * LKF_SEG;
    float hs[5][LKF_SEG];
    if (has_task) {
        float pr[5][NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const float ax = Ld[0][ry][xs + k], ay = Ld[1][ry][xs + k], at = Ld[2][ry][xs + k];
            pr[0][k] = ax * ax;
            pr[1][k] = ax * ay;
            pr[2][k] = ay * ay;
            pr[3][k] = ax * at;
            pr[4][k] = ay * at;
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) window_sums<LKF_SEG, 2 * HW + 1>(pr[q], hs[q]);
    }
    __syncthreads();
    auto& H = sm.u.H;
    if (has_task)
#pragma unroll
        for (int q = 0; q < 5; ++q)
#pragma unroll
            for (int i = 0; i < LKF_SEG; ++i) H[q][ry][xs + i] = hs[q][i];
    __syncthreads();
    // C2. vertical window sums for LKF_RPT output rows of one column
    const int lx = tid % LKF_BX, oy0 = (tid / LKF_BX) * LKF_RPT;
    float V[LKF_RPT][5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        float col[LKF_RPT + 2 * HW], vs[LKF_RPT];
#pragma unroll
        for (int r = 0; r < LKF_RPT + 2 * HW; ++r) col[r] = H[q][oy0 + r][lx];
        window_sums<LKF_RPT, 2 * HW + 1>(col, vs);
#pragma unroll
        for (int o = 0; o < LKF_RPT; ++o) V[o][q] = vs[o];
    }
    // D. solve and update
    const int px = x0 + lx;
    if (px >= w) return;
#pragma unroll
    for (int o = 0; o < LKF_RPT; ++o) {
        const int py = y0 + oy0 + o;
        if (py >= h) break;
        float f0 = fpre[o].x, f1 = fpre[o].y;
        float inv[4];
        if (px >= HW && px < w - HW && py >= HW && py < h - HW) {
            if (ill_ratio > 0.0) {
                float u0, u1;
                const bool ok = lk_ill(V[o][0], V[o][1], V[o][2], ill_ratio)
                                    ? lk_solve_fp64(V[o][0], V[o][1], V[o][2], V[o][3], V[o][4], min_det, u0, u1)
                                    : lk_solve_wc(V[o][0], V[o][1], V[o][2], V[o][3], V[o][4], min_det, u0, u1);
                if (ok) {
                    f0 += u0;
                    f1 += u1;
                }
            } else if (lk_pinv<false>(V[o][0], V[o][1], V[o][2], min_det, inv)) {
                float u0 = inv[0] * V[o][3] + inv[1] * V[o][4];
                float u1 = inv[2] * V[o][3] + inv[3] * V[o][4];
                u0 = isnan(u0) ? 0.0f : u0;
                u1 = isnan(u1) ? 0.0f : u1;
                f0 += u0;
                f1 += u1;
            }
        }
        float* fo = row_ptr(flow_out, flow_pitch, py) + px * 2;
        fo[0] = f0;
        fo[1] = f1;
    }
}

}  // namespace

namespace isa {

int launch_flow_from_tiles(float* flow, int flow_pitch, const float* tiles, int tiles_pitch, int tw, int th, int w,
                           int h, float bsx, float bsy, float rot, int quant, hipStream_t s) {
    const float c = (float)std::cos((double)rot), sn = (float)std::sin((double)rot);
    k_flow_from_tiles<<<dim3(div_up(w, 4 * 64), div_up(h, 4)), dim3(64, 4), 0, s>>>(flow, flow_pitch, tiles,
                                                                                     tiles_pitch, tw, th, w, h, bsx,
                                                                                     bsy, c, sn, quant);
    return check_launch("flow_from_tiles");
}

int launch_warp(const float* src, int src_pitch, float* out, int out_pitch, const float* flow, int flow_pitch, int w,
                int h, int quant, hipStream_t s) {
    k_warp<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(src, src_pitch, out, out_pitch, flow, flow_pitch,
                                                                      w, h, quant);
    return check_launch("warp");
}

// taps: tables from build_tap_tables for (w, h, quant), or null to evaluate the taps in place.
int launch_derivatives(const float* src, const float* tgt, int img_pitch, float* ix, float* iy, float* iz, int pitch,
                       int w, int h, int quant, const void* taps, hipStream_t s) {
    const Tap* col = (const Tap*)taps;
    const dim3 grid(div_up(w, DV_BX), div_up(h, DV_BY)), block(DV_BX, DV_BY);
    if (col)
        k_derivatives_tab<true><<<grid, block, 0, s>>>(src, tgt, img_pitch, ix, iy, iz, pitch, w, h, col,
                                                       col + (size_t)w * 5, quant);
    else
        k_derivatives_tab<false><<<grid, block, 0, s>>>(src, tgt, img_pitch, ix, iy, iz, pitch, w, h, nullptr, nullptr,
                                                        quant);
    return check_launch("derivatives");
}

int launch_derivatives2(const float* img, int img_pitch, float* ix, float* iy, int pitch, int w, int h, int quant,
                        hipStream_t s) {
    k_derivatives2<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(img, img_pitch, ix, iy, pitch, w, h,
                                                                              quant);
    return check_launch("derivatives2");
}

int launch_lk_step(float* flow, int flow_pitch, const float* ix, const float* iy, const float* it, int pitch, int w,
                   int h, int window, float min_det, bool box, float ill_ratio, hipStream_t s) {
    int hw = window / 2;
    if (box && hw <= LK_HMAX) {
        k_lk_step_box<<<dim3(div_up(w, LK_BX), div_up(h, LK_BY)), dim3(LK_BX, LK_BY), 0, s>>>(
            flow, flow_pitch, ix, iy, it, pitch, w, h, hw, min_det, (double)ill_ratio);
    } else {
        k_lk_step<<<dim3(div_up(w, 32), div_up(h, 16)), dim3(32, 16), 0, s>>>(flow, flow_pitch, ix, iy, it, pitch, w,
                                                                              h, hw, min_det);
    }
    return check_launch("lucas_kanade_step");
}

// OpticalFlow.LucasKanade with an explicit work area (tmp, Ix, Iy, Iz: 4 w*h floats, pitch w*4).
size_t tap_table_bytes(int w, int h) { return (size_t)(w + h) * 5 * sizeof(Tap); }

int build_tap_tables(void* tables, int w, int h, int quant, hipStream_t s) {
    Tap* col = (Tap*)tables;
    Tap* row = col + (size_t)w * 5;
    k_tap_tables<<<div_up(w + h, 256), 256, 0, s>>>(col, row, w, h, quant);
    return check_launch("tap_tables");
}

// taps: tables from build_tap_tables for (w, h, quant).
int lucas_kanade(float* src, const float* tgt, int img_pitch, const float* tiles, int tiles_pitch, int tw, int th,
                 float* flow, int flow_pitch, float* work, int w, int h, int iterations, float bsx, float bsy,
                 float rot, float min_det, int window, int quant, bool box, hipStream_t s, const void* taps,
                 float ill_ratio, hipEvent_t* iter_ev) {
    const size_t np = (size_t)w * h;
    float* tmp = work;
    float* ix = work + np;
    float* iy = work + 2 * np;
    float* iz = work + 3 * np;
    const int wp = w * (int)sizeof(float);
    const int hw = window / 2;
    // the fused kernel addresses the images and the flow with 32-bit byte offsets
    const bool off32 = img_pitch % 4 == 0 && flow_pitch % 4 == 0 && (size_t)img_pitch * h < ((size_t)1 << 32) &&
                       (size_t)flow_pitch * h < ((size_t)1 << 32);
    if (box && hw >= 1 && hw <= 8 && iterations > 0 && flow_pitch == 2 * wp && off32) {
        // fused iterations, ping-ponging between `flow` and the work area so the last lands in `flow`
        float* bufs[2] = {flow, work};
        int cur = iterations % 2;  // buffer holding the initial field
        ISA_TRY(launch_flow_from_tiles(bufs[cur], flow_pitch, tiles, tiles_pitch, tw, th, w, h, bsx, bsy, rot, quant,
                                       s));
        if (iter_ev) (void)hipEventRecord(iter_ev[0], s);
        const Tap* col = (const Tap*)taps;
        const dim3 grid(div_up(w, LKF_BX), div_up(h, LKF_BY)), block(LKF_BX * (LKF_BY / LKF_RPT));
        for (int it = 0; it < iterations; ++it) {
            const float* fi = bufs[cur];
            float* fo = bufs[cur ^ 1];
#define ISA_LKF(H)                                                                                                   \
    case H:                                                                                                        \
        k_lk_fused<H><<<grid, block, 0, s>>>(src, tgt, img_pitch, fi, fo, flow_pitch, w, h, min_det, quant, col,   \
                                             col + (size_t)w * 5, (double)ill_ratio);                              \
        break;
            switch (hw) { ISA_LKF(1) ISA_LKF(2) ISA_LKF(3) ISA_LKF(4) ISA_LKF(5) ISA_LKF(6) ISA_LKF(7) ISA_LKF(8) }
#undef ISA_LKF
            ISA_TRY(check_launch("lucas_kanade_fused"));
            cur ^= 1;
        }
        if (iter_ev) (void)hipEventRecord(iter_ev[1], s);
        return ISA_OK;
    }
    ISA_TRY(launch_flow_from_tiles(flow, flow_pitch, tiles, tiles_pitch, tw, th, w, h, bsx, bsy, rot, quant, s));
    if (iter_ev) (void)hipEventRecord(iter_ev[0], s);
    for (int it = 0; it < iterations; ++it) {
        ISA_TRY(launch_warp(src, img_pitch, tmp, wp, flow, flow_pitch, w, h, quant, s));
        ISA_TRY(launch_derivatives(tmp, tgt, wp, ix, iy, iz, wp, w, h, quant, taps, s));
        ISA_TRY(launch_lk_step(flow, flow_pitch, ix, iy, iz, wp, w, h, window, min_det, box, ill_ratio, s));
    }
    if (iter_ev) (void)hipEventRecord(iter_ev[1], s);
    return ISA_OK;
}

}  // namespace isa

extern "C" {

int isa_create_flow_field_from_tiles(float* flow2_dev, int flow_pitch, const float* tiles2_dev, int tiles_pitch,
                                     int tex_w, int tex_h, int width, int height, const float base_shift[2],
                                     float base_rotation, int tex_quant, isa_stream_t stream) {
    ISA_REQUIRE(flow2_dev && tiles2_dev && tex_w > 0 && tex_h > 0 && width > 0 && height > 0,
                "isa_create_flow_field_from_tiles: bad arguments");
    return isa::launch_flow_from_tiles(flow2_dev, flow_pitch, tiles2_dev, tiles_pitch, tex_w, tex_h, width, height,
                                       base_shift[0], base_shift[1], base_rotation, tex_quant, isa::to_stream(stream));
}

int isa_warp(const float* src_dev, int src_pitch, float* out_dev, int out_pitch, const float* flow2_dev,
             int flow_pitch, int width, int height, int tex_quant, isa_stream_t stream) {
    ISA_REQUIRE(src_dev && out_dev && flow2_dev && width > 0 && height > 0, "isa_warp: bad arguments");
    return isa::launch_warp(src_dev, src_pitch, out_dev, out_pitch, flow2_dev, flow_pitch, width, height, tex_quant,
                            isa::to_stream(stream));
}

int isa_compute_derivatives(const float* src_dev, const float* tgt_dev, int img_pitch, float* ix_dev, float* iy_dev,
                            float* iz_dev, int pitch, int width, int height, int tex_quant, isa_stream_t stream) {
    ISA_REQUIRE(src_dev && tgt_dev && ix_dev && iy_dev && iz_dev && width > 0 && height > 0,
                "isa_compute_derivatives: bad arguments");
    return isa::launch_derivatives(src_dev, tgt_dev, img_pitch, ix_dev, iy_dev, iz_dev, pitch, width, height,
                                   tex_quant, nullptr, isa::to_stream(stream));
}

int isa_compute_derivatives2(const float* img_dev, int img_pitch, float* ix_dev, float* iy_dev, int pitch, int width,
                             int height, int tex_quant, isa_stream_t stream) {
    ISA_REQUIRE(img_dev && ix_dev && iy_dev && width > 0 && height > 0, "isa_compute_derivatives2: bad arguments");
    return isa::launch_derivatives2(img_dev, img_pitch, ix_dev, iy_dev, pitch, width, height, tex_quant,
                                    isa::to_stream(stream));
}

int isa_lucas_kanade_optim(float* flow2_dev, int flow_pitch, const float* ix_dev, const float* iy_dev,
                           const float* it_dev, int img_pitch, int width, int height, int window_size, float min_det,
                           isa_stream_t stream) {
    ISA_REQUIRE(flow2_dev && ix_dev && iy_dev && it_dev && width > 0 && height > 0 && window_size >= 1,
                "isa_lucas_kanade_optim: bad arguments");
    return isa::launch_lk_step(flow2_dev, flow_pitch, ix_dev, iy_dev, it_dev, img_pitch, width, height, window_size,
                               min_det, false, 0.0f, isa::to_stream(stream));
}

int isa_lucas_kanade(float* src_dev, const float* tgt_dev, int img_pitch, const float* tiles2_dev, int tiles_pitch,
                     int tex_w, int tex_h, float* flow2_dev, int flow_pitch, float* work_dev, int width, int height,
                     int iterations, const float base_shift[2], float base_rotation, float min_det, int window_size,
                     int tex_quant, isa_stream_t stream) {
    ISA_REQUIRE(src_dev && tgt_dev && tiles2_dev && flow2_dev && work_dev && width > 0 && height > 0 &&
                    iterations >= 0 && window_size >= 1,
                "isa_lucas_kanade: bad arguments");
    hipStream_t s = isa::to_stream(stream);
    ISA_TRY(isa::lucas_kanade(src_dev, tgt_dev, img_pitch, tiles2_dev, tiles_pitch, tex_w, tex_h, flow2_dev, flow_pitch,
                              work_dev, width, height, iterations, base_shift[0], base_shift[1], base_rotation, min_det,
                              window_size, tex_quant, false, s, nullptr, 0.0f, nullptr));
    // final warp, copied over the source image (R:PEFStudioDX/OpticalFlow.cs:162-163)
    ISA_TRY(isa::launch_warp(src_dev, img_pitch, work_dev, width * (int)sizeof(float), flow2_dev, flow_pitch, width,
                             height, tex_quant, s));
    return isa::check_hip(hipMemcpy2DAsync(src_dev, img_pitch, work_dev, width * sizeof(float), width * sizeof(float),
                                           height, hipMemcpyDeviceToDevice, s),
                          "isa_lucas_kanade: copy");
}

}  // extern "C"
