// Optical flow: flow field from tile shifts, warping, derivatives and the
// Lucas-Kanade step (R:Kernels/opticalFlow.cu; driver R:PEFStudioDX/OpticalFlow.cs:147-164).
#include "isa_device.h"
#include "isa_host.h"

namespace {

// CreateFlowFieldFromTiles (R:Kernels/opticalFlow.cu:47-93).
__global__ void k_flow_from_tiles(float* flow, int flow_pitch, const float* tiles, int tiles_pitch, int tw, int th,
                                  int w, int h, float bsx, float bsy, float rot, int quant) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    float c = cos_r(rot), s = sin_r(rot);
    float fx = c * -bsx - s * -bsy;
    float fy = s * -bsx + c * -bsy;
    float pcx = (float)(x - w / 2), pcy = (float)(y - h / 2);
    fx += c * pcx - s * pcy - pcx;
    fy += s * pcx + c * pcy - pcy;
    float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
    float t[2];
    tex_linear_all<2>(tiles, tiles_pitch, tw, th, u, v, ISA_ADDR_CLAMP, quant, t);
    float* o = row_ptr(flow, flow_pitch, y) + x * 2;
    o[0] = fx + t[0];
    o[1] = fy + t[1];
}

// WarpingKernel (R:Kernels/opticalFlow.cu:27-44): flow texture point/clamp, image linear/mirror.
__global__ void k_warp(const float* src, int src_pitch, float* out, int out_pitch, const float* flow, int flow_pitch,
                       int w, int h, int quant) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= w || iy >= h) return;
    float u = ((float)ix + 0.5f) / (float)w, v = ((float)iy + 0.5f) / (float)h;
    int fxi = tex_addr((int)floorf(u * (float)w), w, ISA_ADDR_CLAMP);
    int fyi = tex_addr((int)floorf(v * (float)h), h, ISA_ADDR_CLAMP);
    const float* f = row_ptr(flow, flow_pitch, fyi) + fxi * 2;
    float x = ((float)ix + 0.5f + f[0]) / (float)w;
    float y = ((float)iy + 0.5f + f[1]) / (float)h;
    row_ptr(out, out_pitch, iy)[ix] = tex_linear<1>(src, src_pitch, w, h, 0, x, y, ISA_ADDR_MIRROR, quant);
}

__device__ __forceinline__ float stencil5(const float* img, int pitch, int w, int h, float x, float y, float dx,
                                          float dy, int quant) {
    float t = tex_linear<1>(img, pitch, w, h, 0, x + 2.0f * dx, y + 2.0f * dy, ISA_ADDR_MIRROR, quant);
    t -= tex_linear<1>(img, pitch, w, h, 0, x + 1.0f * dx, y + 1.0f * dy, ISA_ADDR_MIRROR, quant) * 8.0f;
    t += tex_linear<1>(img, pitch, w, h, 0, x - 1.0f * dx, y - 1.0f * dy, ISA_ADDR_MIRROR, quant) * 8.0f;
    t -= tex_linear<1>(img, pitch, w, h, 0, x - 2.0f * dx, y - 2.0f * dy, ISA_ADDR_MIRROR, quant);
    t /= 12.0f;
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

__global__ void k_tap_tables(Tap* col, Tap* row, int w, int h, int quant) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool is_col = i < w;
    const int n = is_col ? w : h;
    const int ix = is_col ? i : i - w;
    if (ix >= n || i >= w + h) return;
    float d = 1.0f / (float)n;
    float x = ((float)ix + 0.5f) * d;
    Tap* out = (is_col ? col : row) + (size_t)ix * 5;
#pragma unroll
    for (int k = -2; k <= 2; ++k) {
        float u = k == 0 ? x : (k > 0 ? x + (float)k * d : x - (float)(-k) * d);
        float xf = u * (float)n - 0.5f;
        float fl = floorf(xf);
        Tap t;
        t.a = tex_frac(xf - fl, quant);
        t.i0 = tex_addr((int)fl, n, ISA_ADDR_MIRROR);
        t.i1 = tex_addr((int)fl + 1, n, ISA_ADDR_MIRROR);
        out[k + 2] = t;
    }
}

__device__ __forceinline__ float tap_fetch(const float* img, int pitch, const Tap& c, const Tap& r) {
    const float* ra = row_ptr(img, pitch, r.i0);
    const float a = c.a, b = r.a;
    // weights at the corners reduce the bilinear sum to one texel exactly (finite data)
    if (b == 0.0f) {
        if (a == 0.0f) return ra[c.i0];
        if (a == 1.0f) return ra[c.i1];
    } else if (b == 1.0f) {
        const float* rb = row_ptr(img, pitch, r.i1);
        if (a == 0.0f) return rb[c.i0];
        if (a == 1.0f) return rb[c.i1];
    }
    const float* rb = row_ptr(img, pitch, r.i1);
    float t00 = ra[c.i0], t10 = ra[c.i1], t01 = rb[c.i0], t11 = rb[c.i1];
    float oma = 1.0f - a, omb = 1.0f - b;
    return oma * omb * t00 + a * omb * t10 + oma * b * t01 + a * b * t11;
}

// ComputeDerivativesKernel (R:Kernels/opticalFlow.cu:96-147): 5-point stencils of both images along x
// and y, averaged, and the temporal difference, with the stencil taps read from the tables.
__global__ void k_derivatives_tab(const float* src, const float* tgt, int img_pitch, float* Ix, float* Iy, float* Iz,
                                  int pitch, int w, int h, const Tap* __restrict__ col, const Tap* __restrict__ row) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= w || iy >= h) return;
    Tap c[5], r[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        c[k] = col[(size_t)ix * 5 + k];
        r[k] = row[(size_t)iy * 5 + k];
    }
    auto sx = [&](const float* I) {
        float t = tap_fetch(I, img_pitch, c[4], r[2]);
        t -= tap_fetch(I, img_pitch, c[3], r[2]) * 8.0f;
        t += tap_fetch(I, img_pitch, c[1], r[2]) * 8.0f;
        t -= tap_fetch(I, img_pitch, c[0], r[2]);
        t /= 12.0f;
        return t;
    };
    auto sy = [&](const float* I) {
        float t = tap_fetch(I, img_pitch, c[2], r[4]);
        t -= tap_fetch(I, img_pitch, c[2], r[3]) * 8.0f;
        t += tap_fetch(I, img_pitch, c[2], r[1]) * 8.0f;
        t -= tap_fetch(I, img_pitch, c[2], r[0]);
        t /= 12.0f;
        return t;
    };
    float t0 = sx(src), t1 = sx(tgt);
    row_ptr(Ix, pitch, iy)[ix] = (t0 + t1) * 0.5f;
    row_ptr(Iz, pitch, iy)[ix] = tap_fetch(src, img_pitch, c[2], r[2]) - tap_fetch(tgt, img_pitch, c[2], r[2]);
    t0 = sy(src);
    t1 = sy(tgt);
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
template <bool EXACT>
__device__ __forceinline__ bool lk_pinv(float a, float b, float d, float min_det, float inv[4]) {
    auto at2 = [](float y, float x) { return EXACT ? atan2_r(y, x) : atan2f(y, x); };
    auto cs = [](float x) { return EXACT ? cos_r(x) : cosf(x); };
    auto sn = [](float x) { return EXACT ? sin_r(x) : sinf(x); };
    float c = b;
    float theta = 0.5f * at2(2.0f * a * c + 2.0f * b * d, a * a + b * b - c * c - d * d);
    float ct = cs(theta), st = sn(theta);
    float UT0 = ct, UT2 = -st, UT1 = st, UT3 = ct;
    float S1 = a * a + b * b + c * c + d * d;
    float S2 = sqrtf((a * a + b * b - c * c - d * d) * (a * a + b * b - c * c - d * d) +
                     4.0f * (a * c + b * d) * (a * c + b * d));
    float sigma1 = sqrtf((S1 + S2) / 2.0f);
    float sigma2 = sqrtf((S1 - S2) / 2.0f);
    float smin = fminf(sigma1, sigma1);
    if (smin < min_det) return false;
    sigma1 = sigma1 != 0 ? 1.0f / sigma1 : 0.0f;
    sigma2 = sigma2 != 0 ? 1.0f / sigma2 : 0.0f;
    float S0 = sigma1, Sb = 0.0f, Sc = 0.0f, S3 = sigma2;
    float eps = 0.5f * at2(2.0f * a * b + 2.0f * c * d, a * a - b * b + c * c - d * d);
    float ce = cs(eps), se = sn(eps);
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
// box sums in LDS instead of 2 (2h+1)^2 loads per pixel. Each horizontal / vertical pass starts a
// segment of LK_SEG outputs with a plain (2h+1)-term sum and continues it as a running sum. The
// solve is the same closed form; the update is pinv . (sum Ix It, sum Iy It), which equals the
// reference's sum of pinv . (Ix, Iy) It.
constexpr int LK_BX = 32, LK_BY = 16, LK_HMAX = 10, LK_SEG = 8;
__global__ __launch_bounds__(LK_BX* LK_BY) void k_lk_step_box(float* flow, int flow_pitch, const float* Fx,
                                                               const float* Fy, const float* Ft, int pitch, int w,
                                                               int h, int hw, float min_det) {
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
        float acc = 0.0f;
        for (int k = 0; k < win; ++k) acc += s_p[ly][xs + k][q];
        s_h[ly][xs][q] = acc;
        for (int i = 1; i < LK_SEG; ++i) {
            acc += s_p[ly][xs + i + win - 1][q] - s_p[ly][xs + i - 1][q];
            s_h[ly][xs + i][q] = acc;
        }
    }
    __syncthreads();
    // vertical: task (column, segment, quantity)
    constexpr int VSEG = LK_BY / LK_SEG;
    for (int task = tid; task < LK_BX * VSEG * 5; task += nt) {
        int q = task % 5, cs = task / 5, seg = cs % VSEG, lx = cs / VSEG;
        int ys = seg * LK_SEG;
        float acc = 0.0f;
        for (int k = 0; k < win; ++k) acc += s_h[ys + k][lx][q];
        s_v[ys][lx][q] = acc;
        for (int i = 1; i < LK_SEG; ++i) {
            acc += s_h[ys + i + win - 1][lx][q] - s_h[ys + i - 1][lx][q];
            s_v[ys + i][lx][q] = acc;
        }
    }
    __syncthreads();
    const int px = x0 + threadIdx.x, py = y0 + threadIdx.y;
    if (px < hw || px >= w - hw || py < hw || py >= h - hw) return;
    float g[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) g[q] = s_v[threadIdx.y][threadIdx.x][q];
    float inv[4];
    if (!lk_pinv<false>(g[0], g[1], g[2], min_det, inv)) return;
    float u0 = inv[0] * g[3] + inv[1] * g[4];
    float u1 = inv[2] * g[3] + inv[3] * g[4];
    u0 = isnan(u0) ? 0.0f : u0;
    u1 = isnan(u1) ? 0.0f : u1;
    float* f = row_ptr(flow, flow_pitch, py) + px * 2;
    f[0] += u0;
    f[1] += u1;
}

}  // namespace

namespace isa {

int launch_flow_from_tiles(float* flow, int flow_pitch, const float* tiles, int tiles_pitch, int tw, int th, int w,
                           int h, float bsx, float bsy, float rot, int quant, hipStream_t s) {
    k_flow_from_tiles<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(flow, flow_pitch, tiles, tiles_pitch,
                                                                                 tw, th, w, h, bsx, bsy, rot, quant);
    return check_launch("flow_from_tiles");
}

int launch_warp(const float* src, int src_pitch, float* out, int out_pitch, const float* flow, int flow_pitch, int w,
                int h, int quant, hipStream_t s) {
    k_warp<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(src, src_pitch, out, out_pitch, flow, flow_pitch,
                                                                      w, h, quant);
    return check_launch("warp");
}

int launch_derivatives(const float* src, const float* tgt, int img_pitch, float* ix, float* iy, float* iz, int pitch,
                       int w, int h, const void* taps, hipStream_t s) {
    const Tap* col = (const Tap*)taps;
    k_derivatives_tab<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(src, tgt, img_pitch, ix, iy, iz, pitch,
                                                                                 w, h, col, col + (size_t)w * 5);
    return check_launch("derivatives");
}

int launch_derivatives2(const float* img, int img_pitch, float* ix, float* iy, int pitch, int w, int h, int quant,
                        hipStream_t s) {
    k_derivatives2<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(img, img_pitch, ix, iy, pitch, w, h,
                                                                              quant);
    return check_launch("derivatives2");
}

int launch_lk_step(float* flow, int flow_pitch, const float* ix, const float* iy, const float* it, int pitch, int w,
                   int h, int window, float min_det, bool box, hipStream_t s) {
    int hw = window / 2;
    if (box && hw <= LK_HMAX) {
        k_lk_step_box<<<dim3(div_up(w, LK_BX), div_up(h, LK_BY)), dim3(LK_BX, LK_BY), 0, s>>>(
            flow, flow_pitch, ix, iy, it, pitch, w, h, hw, min_det);
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
                 float rot, float min_det, int window, int quant, bool box, hipStream_t s, const void* taps) {
    const size_t np = (size_t)w * h;
    float* tmp = work;
    float* ix = work + np;
    float* iy = work + 2 * np;
    float* iz = work + 3 * np;
    const int wp = w * (int)sizeof(float);
    ISA_TRY(launch_flow_from_tiles(flow, flow_pitch, tiles, tiles_pitch, tw, th, w, h, bsx, bsy, rot, quant, s));
    for (int it = 0; it < iterations; ++it) {
        ISA_TRY(launch_warp(src, img_pitch, tmp, wp, flow, flow_pitch, w, h, quant, s));
        ISA_TRY(launch_derivatives(tmp, tgt, wp, ix, iy, iz, wp, w, h, taps, s));
        ISA_TRY(launch_lk_step(flow, flow_pitch, ix, iy, iz, wp, w, h, window, min_det, box, s));
    }
    return ISA_OK;
}

// Stream-ordered tap tables for a one-off call of the public entry points (the merger keeps its own).
struct ScopedTaps {
    void* p = nullptr;
    hipStream_t s;
    int rc = ISA_OK;
    ScopedTaps(int w, int h, int quant, hipStream_t st) : s(st) {
        rc = check_hip(hipMallocAsync(&p, tap_table_bytes(w, h), s), "tap tables");
        if (rc == ISA_OK) rc = build_tap_tables(p, w, h, quant, s);
    }
    ~ScopedTaps() {
        if (p) (void)hipFreeAsync(p, s);
    }
};

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
    isa::ScopedTaps taps(width, height, tex_quant, isa::to_stream(stream));
    ISA_TRY(taps.rc);
    return isa::launch_derivatives(src_dev, tgt_dev, img_pitch, ix_dev, iy_dev, iz_dev, pitch, width, height, taps.p,
                                   isa::to_stream(stream));
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
                               min_det, false, isa::to_stream(stream));
}

int isa_lucas_kanade(float* src_dev, const float* tgt_dev, int img_pitch, const float* tiles2_dev, int tiles_pitch,
                     int tex_w, int tex_h, float* flow2_dev, int flow_pitch, float* work_dev, int width, int height,
                     int iterations, const float base_shift[2], float base_rotation, float min_det, int window_size,
                     int tex_quant, isa_stream_t stream) {
    ISA_REQUIRE(src_dev && tgt_dev && tiles2_dev && flow2_dev && work_dev && width > 0 && height > 0 &&
                    iterations >= 0 && window_size >= 1,
                "isa_lucas_kanade: bad arguments");
    hipStream_t s = isa::to_stream(stream);
    isa::ScopedTaps taps(width, height, tex_quant, s);
    ISA_TRY(taps.rc);
    ISA_TRY(isa::lucas_kanade(src_dev, tgt_dev, img_pitch, tiles2_dev, tiles_pitch, tex_w, tex_h, flow2_dev, flow_pitch,
                              work_dev, width, height, iterations, base_shift[0], base_shift[1], base_rotation, min_det,
                              window_size, tex_quant, false, s, taps.p));
    // final warp, copied over the source image (R:PEFStudioDX/OpticalFlow.cs:162-163)
    ISA_TRY(isa::launch_warp(src_dev, img_pitch, work_dev, width * (int)sizeof(float), flow2_dev, flow_pitch, width,
                             height, tex_quant, s));
    return isa::check_hip(hipMemcpy2DAsync(src_dev, img_pitch, work_dev, width * sizeof(float), width * sizeof(float),
                                           height, hipMemcpyDeviceToDevice, s),
                          "isa_lucas_kanade: copy");
}

}  // extern "C"
