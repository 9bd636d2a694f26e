// Merge kernels: half-resolution debayer, robustness mask, kernel accumulation.
#include <algorithm>
#include <climits>
#include <cmath>

#include <cstdint>

#include "isa_device.h"
#include "isa_host.h"

namespace {

// deBayersSubSample3 (R:Kernels/DeBayerKernels.cu:242-283); raw unpitched, 2*dim_x wide.
__global__ void k_subsample3(const uint16_t* __restrict__ raw, float* out, float max_val, int hw, int hh,
                             int out_pitch, Cfa cfa) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= hw || y >= hh) return;
    float p[3] = {0.0f, 0.0f, 0.0f};
    float factor = 1.0f / max_val;
#pragma unroll
    for (int ix = 0; ix < 2; ++ix)
#pragma unroll
        for (int iy = 0; iy < 2; ++iy) {
            int site = cfa.c[iy * 2 + ix];
            float v = (float)raw[(size_t)(2 * y + iy) * hw * 2 + 2 * x + ix];
            if (site == ISA_CFA_GREEN) p[1] += v * factor * 0.5f;
            else if (site == ISA_CFA_RED) p[0] = v * factor;
            else if (site == ISA_CFA_BLUE) p[2] = v * factor;
        }
    float* o = row_ptr(out, out_pitch, y) + x * 3;
    o[0] = p[0];
    o[1] = p[1];
    o[2] = p[2];
}

// The same for four output pixels per thread: one 16-byte load per raw row and three 16-byte stores.
// Needs hw % 4 == 0, 16-byte aligned raw and output rows.
__global__ void k_subsample3_x4(const uint16_t* __restrict__ raw, float* out, float max_val, int hw, int hh,
                                int out_pitch, Cfa cfa) {
    const int x0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x), y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x0 >= hw || y >= hh) return;
    const float factor = 1.0f / max_val;
    uint4 r[2];
#pragma unroll
    for (int iy = 0; iy < 2; ++iy) r[iy] = *reinterpret_cast<const uint4*>(raw + (size_t)(2 * y + iy) * hw * 2 + 2 * x0);
    float o[12];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float p[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ix = 0; ix < 2; ++ix)
#pragma unroll
            for (int iy = 0; iy < 2; ++iy) {  // the visiting order of k_subsample3
                const int site = cfa.c[iy * 2 + ix];
                const unsigned wd = k == 0 ? r[iy].x : (k == 1 ? r[iy].y : (k == 2 ? r[iy].z : r[iy].w));
                const float v = (float)(ix ? (wd >> 16) : (wd & 0xffffu));
                if (site == ISA_CFA_GREEN) p[1] += v * factor * 0.5f;
                else if (site == ISA_CFA_RED) p[0] = v * factor;
                else if (site == ISA_CFA_BLUE) p[2] = v * factor;
            }
        o[3 * k] = p[0];
        o[3 * k + 1] = p[1];
        o[3 * k + 2] = p[2];
    }
    float4* d = reinterpret_cast<float4*>(row_ptr(out, out_pitch, y) + x0 * 3);
    d[0] = make_float4(o[0], o[1], o[2], o[3]);
    d[1] = make_float4(o[4], o[5], o[6], o[7]);
    d[2] = make_float4(o[8], o[9], o[10], o[11]);
}

// ComputeRobustnessMask (R:Kernels/RobustnessModell.cu:28-158). Grid over the half-resolution
// image (w x h); flow: full-resolution texture fw x fh (linear, mirror). Border not written.
__global__ void k_robustness(const float* ref, const float* mov, int img_pitch, float* mask, int mask_pitch,
                             const float* flow, int flow_pitch, int fw, int fh, int w, int h, float alpha, float beta,
                             float thr_m, int quant) {
    int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y;
    if (px >= w - 1 || py >= h - 1 || px < 1 || py < 1) return;
    const double iw = 1.0 / (double)w, ih = 1.0 / (double)h;
    float u = div_n((float)px + 0.5f, iw), v = div_n((float)py + 0.5f, ih);
    float f[2];
    tex_linear_all<2>(flow, flow_pitch, fw, fh, u, v, ISA_ADDR_MIRROR, quant, f);
    // The reference's 5x5 min/max loop overwrites its result each pass (R:Kernels/RobustnessModell.cu:62-72),
    // so only the (+2, +2) sample determines M.
    float su = div_n((float)px + 2.0f + 0.5f, iw), sv = div_n((float)py + 2.0f + 0.5f, ih);
    float sm[2];
    tex_linear_all<2>(flow, flow_pitch, fw, fh, su, sv, ISA_ADDR_MIRROR, quant, sm);
    float maxx = fmaxf(sm[0], f[0]), maxy = fmaxf(sm[1], f[1]);
    float minx = fminf(sm[0], f[0]), miny = fminf(sm[1], f[1]);
    int shx = (int)roundf(f[0] * 0.5f), shy = (int)roundf(f[1] * 0.5f);
    float mr[3] = {0, 0, 0}, mm[3] = {0, 0, 0};
    float pr[9][3];
#pragma unroll
    for (int y = -1; y <= 1; ++y)
#pragma unroll
        for (int x = -1; x <= 1; ++x) {
            const float* p = row_ptr(ref, img_pitch, py + y) + (px + x) * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pr[(y + 1) * 3 + x + 1][c] = p[c];
                mr[c] += p[c];
            }
            int yy = clampi(py + shy + y, 0, h - 1), xx = clampi(px + shx + x, 0, w - 1);
            const float* q = row_ptr(mov, img_pitch, yy) + xx * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) mm[c] += q[c];
        }
    // divisions by 9 and 3 through div_n: the IEEE quotients, without the division sequence
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        mr[c] = div_n(mr[c], 1.0 / 9.0);
        mm[c] = div_n(mm[c], 1.0 / 9.0);
    }
    float meandist = fabsf(mr[0] - mm[0]) + fabsf(mr[1] - mm[1]) + fabsf(mr[2] - mm[2]);
    meandist = div_n(meandist, 1.0 / 3.0);
    maxx *= 0.5f * meandist;
    maxy *= 0.5f * meandist;
    minx *= 0.5f * meandist;
    miny *= 0.5f * meandist;
    float Mv = sqrtf((maxx - minx) * (maxx - minx) + (maxy - miny) * (maxy - miny));
    float sr[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) sr[c] += (pr[i][c] - mr[c]) * (pr[i][c] - mr[c]);
    float s = 1.5f;
    if (Mv > thr_m) s = 0.0f;
    float out[4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float sd = sqrtf(div_n(sr[c], 1.0 / 9.0));
        float smd = sqrtf(alpha * mr[c] + beta);
        if (c == 1) smd = smd / sqrtf(2.0f);
        float dist = fabsf(mr[c] - mm[c]);
        float sig = fmaxf(smd, sd);
        dist = dist * (sd * sd / (sd * sd + smd * smd));
        out[c] = fmaxf(fminf(s * exp_r(-dist * dist / (sig * sig)) - 0.12f, 1.0f), 0.0f);
    }
    out[3] = Mv;
    float* o = row_ptr(mask, mask_pitch, py) + px * 4;
    o[0] = out[0];
    o[1] = out[1];
    o[2] = out[2];
    o[3] = out[3];
}

struct AccArgs {
    Cfa cfa;
    float white[3], black[3];
    double inv_white[3];
    int white_int;  // every white level is an integer in [1, 2^24]: divide through div_n
};

__device__ __forceinline__ int site_channel(int site) { return (site >= 0 && site <= 2) ? site : -1; }

// (v - black) / white; DIVN when every white level is an integer (div_n is then exact)
template <bool DIVN> __device__ __forceinline__ float norm_raw(float v, int c, const AccArgs& a) {
    float d = v - pick3(c, a.black[0], a.black[1], a.black[2]);
    return DIVN ? div_n(d, pick3(c, a.inv_white[0], a.inv_white[1], a.inv_white[2]))
                : d / pick3(c, a.white[0], a.white[1], a.white[2]);
}

// One 5x5 gather of one frame at output pixel (x, y): integer sample shift (sx, sy), weights w[25].
// Summation order: the taps are visited class by class, class = ((py & 1) << 1) | (px & 1) for
// class 0..3, in raster order within a class; each tap adds to the sums of its (clamped, shifted)
// site's channel. All taps of a class share the site parity, so away from the image edge a class
// has one channel and its taps run on one accumulator pair with no per-tap channel selection.
template <int CLS, bool DIVN>
__device__ __forceinline__ void gather_class_inside(const uint16_t* __restrict__ rs, const float* __restrict__ mask,
                                                    int mask_pitch, int x, int y, int w, const float* wt,
                                                    const AccArgs& a, int ch, float acc[3], float ws[3]) {
    if (ch < 0) return;
    typedef const __attribute__((address_space(1))) uint16_t* graw;
    typedef const __attribute__((address_space(1))) float* gmask;
    float A = ch == 0 ? acc[0] : (ch == 1 ? acc[1] : acc[2]);
    float W = ch == 0 ? ws[0] : (ch == 1 ? ws[1] : ws[2]);
#pragma unroll
    for (int py = -2; py <= 2; ++py) {
        if (((py & 1) << 1) != (CLS & 2)) continue;
        graw rrow = (graw)(rs + (ptrdiff_t)py * w);
        gmask mrow = (gmask)row_ptr(mask, mask_pitch, (y + py) >> 1);
#pragma unroll
        for (int px = -2; px <= 2; ++px) {
            if ((px & 1) != (CLS & 1)) continue;
            const float wgt = wt[(py + 2) * 5 + px + 2];
            const float v = norm_raw<DIVN>((float)rrow[px], ch, a);
            float cert = mrow[((x + px) >> 1) * 4 + ch];
            if (!isfinite(cert)) cert = 0.0f;
            A += v * wgt * cert;
            W += wgt * cert;
        }
    }
    if (ch == 0) { acc[0] = A; ws[0] = W; }
    else if (ch == 1) { acc[1] = A; ws[1] = W; }
    else { acc[2] = A; ws[2] = W; }
}

template <int CLS, bool DIVN>
__device__ __forceinline__ void gather_class_edge(const uint16_t* __restrict__ raw, const float* __restrict__ mask,
                                                  int mask_pitch, int x, int y, int sx, int sy, int w, int h,
                                                  const float* wt, const AccArgs& a, float acc[3], float ws[3]) {
#pragma unroll
    for (int py = -2; py <= 2; ++py) {
        if (((py & 1) << 1) != (CLS & 2)) continue;
#pragma unroll
        for (int px = -2; px <= 2; ++px) {
            if ((px & 1) != (CLS & 1)) continue;
            int ppsx = clampi(x + px + sx, 0, w - 1), ppsy = clampi(y + py + sy, 0, h - 1);
            int ppx = clampi(x + px, 0, w - 1), ppy = clampi(y + py, 0, h - 1);
            int c = site_channel(a.cfa.at(ppsy, ppsx));
            if (c < 0) continue;
            float wgt = wt[(py + 2) * 5 + px + 2];
            float v = norm_raw<DIVN>((float)raw[(size_t)ppsy * w + ppsx], c, a);
            float cert = row_ptr(mask, mask_pitch, ppy / 2)[(ppx / 2) * 4 + c];
            if (!isfinite(cert)) cert = 0.0f;
            const float t = v * wgt * cert, u = wgt * cert;
            // constant indices only: the sums stay in registers
            if (c == 0) { acc[0] += t; ws[0] += u; }
            else if (c == 1) { acc[1] += t; ws[1] += u; }
            else { acc[2] += t; ws[2] += u; }
        }
    }
}

template <bool DIVN>
__device__ __forceinline__ void gather1x(const uint16_t* __restrict__ raw, const float* __restrict__ mask,
                                         int mask_pitch, int x, int y, int sx, int sy, int w, int h, const float* wt,
                                         const AccArgs& a, float acc[3], float ws[3]) {
    const int qx = x + sx, qy = y + sy;
    if (qx >= 2 && qx <= w - 3 && qy >= 2 && qy <= h - 3 && x >= 2 && x <= w - 3 && y >= 2 && y <= h - 3) {
        const uint16_t* rs = raw + (size_t)qy * w + qx;
        // channel of each class: the site of its tap nearest the centre (px, py in {0, 1})
        const int c0 = site_channel(a.cfa.at(qy, qx)), c1 = site_channel(a.cfa.at(qy, qx + 1));
        const int c2 = site_channel(a.cfa.at(qy + 1, qx)), c3 = site_channel(a.cfa.at(qy + 1, qx + 1));
        gather_class_inside<0, DIVN>(rs, mask, mask_pitch, x, y, w, wt, a, c0, acc, ws);
        gather_class_inside<1, DIVN>(rs, mask, mask_pitch, x, y, w, wt, a, c1, acc, ws);
        gather_class_inside<2, DIVN>(rs, mask, mask_pitch, x, y, w, wt, a, c2, acc, ws);
        gather_class_inside<3, DIVN>(rs, mask, mask_pitch, x, y, w, wt, a, c3, acc, ws);
    } else {
        gather_class_edge<0, DIVN>(raw, mask, mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
        gather_class_edge<1, DIVN>(raw, mask, mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
        gather_class_edge<2, DIVN>(raw, mask, mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
        gather_class_edge<3, DIVN>(raw, mask, mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
    }
}

// accumulateImages (R:Kernels/DeBayerKernels.cu:288-376).
template <bool DIVN>
__global__ void k_accumulate(const uint16_t* __restrict__ raw, float* sum, float* wsum, const float* mask,
                             const float* kern, const float* flow, int w, int h, int out_pitch, int mask_pitch,
                             int flow_pitch, AccArgs a) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x < 1 || y < 1 || x >= w - 1 || y >= h - 1) return;
    float* ps = row_ptr(sum, out_pitch, y) + x * 3;
    float* pw = row_ptr(wsum, out_pitch, y) + x * 3;
    const float* k = row_ptr(kern, out_pitch, y) + x * 3;
    const float* f = row_ptr(flow, flow_pitch, y) + x * 2;
    float acc[3] = {ps[0], ps[1], ps[2]}, ws[3] = {pw[0], pw[1], pw[2]};
    float wt[25];
#pragma unroll
    for (int py = -2; py <= 2; ++py)
#pragma unroll
        for (int px = -2; px <= 2; ++px) wt[(py + 2) * 5 + px + 2] = kernel_weight(px, py, k[0], k[1], k[2]);
    int sx = (int)roundf(f[0]), sy = (int)roundf(f[1]);
    gather1x<DIVN>(raw, mask, mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ps[c] = acc[c];
        pw[c] = ws[c];
    }
}

// Frame-fused accumulation: the 25 weights are computed once per pixel and the running
// sums stay in registers across frames; frames are visited in the given order, so the
// result equals running accumulateImages once per frame in that order.
constexpr int ACC_MAXF = 32;  // frames per launch (passed by value in the kernel arguments)
struct FramesArgs {
    const uint16_t* raw[ACC_MAXF];
    const float* mask[ACC_MAXF];
    const float* flow[ACC_MAXF];  // null entry = zero flow
    int nframes;
};

template <bool DIVN>
__global__ void k_accumulate_frames(float* sum, float* wsum, const float* kern, int w, int h, int out_pitch,
                                    int mask_pitch, int flow_pitch, FramesArgs fa, AccArgs a) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x < 1 || y < 1 || x >= w - 1 || y >= h - 1) return;
    float* ps = row_ptr(sum, out_pitch, y) + x * 3;
    float* pw = row_ptr(wsum, out_pitch, y) + x * 3;
    const float* k = row_ptr(kern, out_pitch, y) + x * 3;
    float acc[3] = {ps[0], ps[1], ps[2]}, ws[3] = {pw[0], pw[1], pw[2]};
    float wt[25];
#pragma unroll
    for (int py = -2; py <= 2; ++py)
#pragma unroll
        for (int px = -2; px <= 2; ++px) wt[(py + 2) * 5 + px + 2] = kernel_weight(px, py, k[0], k[1], k[2]);
    for (int f = 0; f < fa.nframes; ++f) {
        int sx = 0, sy = 0;
        if (fa.flow[f]) {
            const float* fl = row_ptr(fa.flow[f], flow_pitch, y) + x * 2;
            sx = (int)roundf(fl[0]);
            sy = (int)roundf(fl[1]);
        }
        gather1x<DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ps[c] = acc[c];
        pw[c] = ws[c];
    }
}

// The same accumulation with each wave on pixels of one (x, y) parity: the half-resolution mask
// pixels under a 5 x 5 window are then the same 3 x 3 block for every lane, loaded as nine float4,
// and each window row of raw samples is one aligned 12-byte load. Needs 4-byte aligned raw frames
// of even width and 16-byte aligned mask rows. Taps, classes and every sum as in gather1x.
constexpr int ACCW_X = 64, ACCW_Y = 4;  // block: 4 waves, one parity each, of 32 x 2 pixels

// raw samples at columns qx-2 .. qx+2 of a row, packed two per word: the three 32-bit words of the
// pair-aligned run qx-2-o .. qx+3-o (o = qx & 1), shifted by o samples
struct RawRow {
    unsigned w0, w1, w2;
    __device__ __forceinline__ float at(int j) const {  // sample at column qx - 2 + j
        const unsigned v = j < 2 ? w0 : (j < 4 ? w1 : w2);
        return (float)((j & 1) ? (v >> 16) : (v & 0xffffu));
    }
};

__device__ __forceinline__ RawRow raw_row5(const uint16_t* rowp, int qx) {
    typedef const __attribute__((address_space(1))) unsigned* gword;
    const int o = qx & 1;
    const gword p = (gword)(rowp + (qx - 2 - o));
    const unsigned a = p[0], b = p[1], c = p[2], sh = 16u * (unsigned)o;
    return RawRow{__builtin_amdgcn_alignbit(b, a, sh), __builtin_amdgcn_alignbit(c, b, sh), c >> sh};
}

__device__ __forceinline__ float chan(const float4& v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

// class CLS of the window for a pixel of parity (PX, PY): r the 5 raw rows, m the 3 x 3 mask pixels
// from ((x >> 1) - 1, (y >> 1) - 1)
template <int CLS, int PX, int PY, bool DIVN>
__device__ __forceinline__ void gather_class_wide(const RawRow (&r)[5], const float4 (&m)[3][3], const float* wt,
                                                  const AccArgs& a, int ch, float acc[3], float ws[3]) {
    if (ch < 0) return;
    float A = ch == 0 ? acc[0] : (ch == 1 ? acc[1] : acc[2]);
    float W = ch == 0 ? ws[0] : (ch == 1 ? ws[1] : ws[2]);
#pragma unroll
    for (int py = -2; py <= 2; ++py) {
        if (((py & 1) << 1) != (CLS & 2)) continue;
#pragma unroll
        for (int px = -2; px <= 2; ++px) {
            if ((px & 1) != (CLS & 1)) continue;
            const float wgt = wt[(py + 2) * 5 + px + 2];
            const float v = norm_raw<DIVN>(r[py + 2].at(px + 2), ch, a);
            float cert = chan(m[((PY + py) >> 1) + 1][((PX + px) >> 1) + 1], ch);
            if (!isfinite(cert)) cert = 0.0f;
            A += v * wgt * cert;
            W += wgt * cert;
        }
    }
    if (ch == 0) { acc[0] = A; ws[0] = W; }
    else if (ch == 1) { acc[1] = A; ws[1] = W; }
    else { acc[2] = A; ws[2] = W; }
}

// one frame, window and sample shift inside the image, pixel of parity (PX, PY)
template <int PX, int PY, bool DIVN>
__device__ __forceinline__ void frame_wide(const uint16_t* raw, const float* mask, int mask_pitch, int x, int y, int qx,
                                           int qy, int w, const float* wt, const AccArgs& a, float acc[3], float ws[3]) {
    typedef const __attribute__((address_space(1))) float* gmask;
    float4 m[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const gmask mr = (gmask)row_ptr(mask, mask_pitch, (y >> 1) - 1 + i) + ((x >> 1) - 1) * 4;
#pragma unroll
        for (int j = 0; j < 3; ++j) m[i][j] = make_float4(mr[4 * j], mr[4 * j + 1], mr[4 * j + 2], 0.0f);
    }
    RawRow r[5];
    const uint16_t* rs = raw + (size_t)qy * w;
#pragma unroll
    for (int py = -2; py <= 2; ++py) r[py + 2] = raw_row5(rs + (ptrdiff_t)py * w, qx);
    const int c0 = site_channel(a.cfa.at(qy, qx)), c1 = site_channel(a.cfa.at(qy, qx + 1));
    const int c2 = site_channel(a.cfa.at(qy + 1, qx)), c3 = site_channel(a.cfa.at(qy + 1, qx + 1));
    gather_class_wide<0, PX, PY, DIVN>(r, m, wt, a, c0, acc, ws);
    gather_class_wide<1, PX, PY, DIVN>(r, m, wt, a, c1, acc, ws);
    gather_class_wide<2, PX, PY, DIVN>(r, m, wt, a, c2, acc, ws);
    gather_class_wide<3, PX, PY, DIVN>(r, m, wt, a, c3, acc, ws);
}

template <bool DIVN>
__global__ __launch_bounds__(256) void k_accumulate_frames_wide(float* sum, float* wsum, const float* kern, int w,
                                                                int h, int out_pitch, int mask_pitch, int flow_pitch,
                                                                FramesArgs fa, AccArgs a) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int x = blockIdx.x * ACCW_X + 2 * (lane & 31) + (wave & 1), y = blockIdx.y * ACCW_Y + 2 * (lane >> 5) + (wave >> 1);
    if (x < 1 || y < 1 || x >= w - 1 || y >= h - 1) return;
    float* ps = row_ptr(sum, out_pitch, y) + x * 3;
    float* pw = row_ptr(wsum, out_pitch, y) + x * 3;
    const float* k = row_ptr(kern, out_pitch, y) + x * 3;
    float acc[3] = {ps[0], ps[1], ps[2]}, ws[3] = {pw[0], pw[1], pw[2]};
    float wt[25];
#pragma unroll
    for (int py = -2; py <= 2; ++py)
#pragma unroll
        for (int px = -2; px <= 2; ++px) wt[(py + 2) * 5 + px + 2] = kernel_weight(px, py, k[0], k[1], k[2]);
    const bool in_xy = x >= 2 && x <= w - 3 && y >= 2 && y <= h - 3;
    // each frame's flow at the pixel is loaded one frame ahead, so the sample rows it addresses do not wait
    // for a load issued in the same iteration
    float2 fnext = make_float2(0.0f, 0.0f);
    if (fa.nframes > 0 && fa.flow[0]) fnext = *(const float2*)(row_ptr(fa.flow[0], flow_pitch, y) + x * 2);
    for (int f = 0; f < fa.nframes; ++f) {
        int sx = 0, sy = 0;
        const float2 fcur = fnext;
        if (f + 1 < fa.nframes && fa.flow[f + 1])
            fnext = *(const float2*)(row_ptr(fa.flow[f + 1], flow_pitch, y) + x * 2);
        if (fa.flow[f]) {
            sx = (int)roundf(fcur.x);
            sy = (int)roundf(fcur.y);
        }
        const int qx = x + sx, qy = y + sy;
        if (in_xy && qx >= 2 && qx <= w - 3 && qy >= 2 && qy <= h - 3) {
            switch (wave) {  // the parity of every pixel of the wave
                case 0: frame_wide<0, 0, DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, qx, qy, w, wt, a, acc, ws); break;
                case 1: frame_wide<1, 0, DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, qx, qy, w, wt, a, acc, ws); break;
                case 2: frame_wide<0, 1, DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, qx, qy, w, wt, a, acc, ws); break;
                default: frame_wide<1, 1, DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, qx, qy, w, wt, a, acc, ws); break;
            }
        } else {
            gather_class_edge<0, DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
            gather_class_edge<1, DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
            gather_class_edge<2, DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
            gather_class_edge<3, DIVN>(fa.raw[f], fa.mask[f], mask_pitch, x, y, sx, sy, w, h, wt, a, acc, ws);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ps[c] = acc[c];
        pw[c] = ws[c];
    }
}

// accumulateImagesSuperRes (R:Kernels/DeBayerKernels.cu:378-468) over several frames in one pass, for an output grid
// of g.ow x g.oh pixels at half the raw pixel spacing whose pixel 0 lies g.ox, g.oy output pixels from the raw
// frame's origin. The reference's form is ow x oh = w x h with offset (w/2, h/2), the central half of the frame;
// the full-frame 2x merge is 2w x 2h with offset 0. Per output pixel the kernel parameters (texture at posX,
// clamp) and the 25 weights are frame independent and computed once; per frame the flow (texture, mirror) gives
// the sample shift round(2 flow), and the 25 taps are visited in the reference's raster order. Raw and
// certainty coordinates are clamped to the rows and columns the output grid covers
// ([ox/2, ox/2 + ow/2 - 1] x [oy/2, oy/2 + oh/2 - 1]). A null flow is the zero flow.
struct SrGeom {
    int ow, oh, ox, oy;
};

// v0, v1 or v2 for i = 0, 1, 2 as two unconditional selects (nested conditionals may become branches)
template <class T> __device__ __forceinline__ T sel012(int i, T v0, T v1, T v2) {
    T r = i >= 1 ? v1 : v0;
    return i >= 2 ? v2 : r;
}

template <bool DIVN>
__global__ __launch_bounds__(256) void k_accumulate_sr_frames(float* sum, float* wsum, const float* kern4, int w,
                                                              int h, int out_pitch, int mask_pitch, int kern_pitch,
                                                              int flow_pitch, FramesArgs fa, SrGeom g, AccArgs a,
                                                              int quant) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x < 1 || y < 1 || x >= g.ow - 1 || y >= g.oh - 1) return;
    float* ps = row_ptr(sum, out_pitch, y) + x * 3;
    float* pw = row_ptr(wsum, out_pitch, y) + x * 3;
    float acc[3] = {ps[0], ps[1], ps[2]}, ws[3] = {pw[0], pw[1], pw[2]};
    const float posx = div_n(((float)x + 0.5f + (float)g.ox) / 2.0f, 1.0 / (double)w);
    const float posy = div_n(((float)y + 0.5f + (float)g.oy) / 2.0f, 1.0 / (double)h);
    float k[4];
    tex_linear_all<4>(kern4, kern_pitch, w, h, posx, posy, ISA_ADDR_CLAMP, quant, k);
    float wt[25];
#pragma unroll
    for (int py = -2; py <= 2; ++py)
#pragma unroll
        for (int px = -2; px <= 2; ++px) wt[(py + 2) * 5 + px + 2] = kernel_weight(px, py, k[0], k[1], k[2]);
    const int cx0 = g.ox / 2, cx1 = g.ox / 2 + g.ow / 2 - 1, cy0 = g.oy / 2, cy1 = g.oy / 2 + g.oh / 2 - 1;
    // certainty rows and columns do not move with the flow: byte offsets of the 5 x 5 window, once per pixel
    size_t mrow[5];
    int mcol[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        mrow[i] = (size_t)(clampi((y + i - 2 + g.oy) / 2, cy0, cy1) / 2) * (size_t)mask_pitch;
        mcol[i] = (clampi((x + i - 2 + g.ox) / 2, cx0, cx1) / 2) * 16;
    }
    for (int f = 0; f < fa.nframes; ++f) {
        int sx = 0, sy = 0;
        if (fa.flow[f]) {
            float fl[2];
            tex_linear_all<2>(fa.flow[f], flow_pitch, w, h, posx, posy, ISA_ADDR_MIRROR, quant, fl);
            sx = (int)roundf(fl[0] * 2.0f);
            sy = (int)roundf(fl[1] * 2.0f);
        }
        typedef const __attribute__((address_space(1))) uint16_t* graw;
        typedef const __attribute__((address_space(1))) char* gbytes;
        const graw raw = (graw)fa.raw[f];
        const gbytes mask = (gbytes)fa.mask[f];
        // The 5 x 5 taps read at most 3 x 3 distinct raw samples: tap column px reads column rx[px], one of
        // bx[0..2] (rx is non-decreasing in px and grows by at most one per step), and likewise for rows.
        // The 9 samples are loaded and normalised once; a tap selects its own.
        int rx[5], ry[5], kx[5], ky[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            rx[i] = clampi((x + i - 2 + sx + g.ox) / 2, cx0, cx1);
            ry[i] = clampi((y + i - 2 + sy + g.oy) / 2, cy0, cy1);
        }
        const int bx[3] = {rx[0], min(rx[0] + 1, cx1), min(rx[0] + 2, cx1)};
        const int by[3] = {ry[0], min(ry[0] + 1, cy1), min(ry[0] + 2, cy1)};
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            kx[i] = rx[i] - rx[0];
            ky[i] = ry[i] - ry[0];
        }
        float sv[3][3];
        int sc[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const graw rrow = raw + (size_t)by[r] * w;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int c = site_channel(a.cfa.at(by[r], bx[q]));
                sc[r][q] = c;
                sv[r][q] = norm_raw<DIVN>((float)rrow[bx[q]], c, a);
            }
        }
        // certainties of the 25 taps (a site outside the three channels reads channel 0 and adds nothing)
        float cv[25];
        int ch[25];
        float rowv[5][3];
#pragma unroll
        for (int py = -2; py <= 2; ++py) {
            const int k = ky[py + 2];
            int rowc[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                rowv[py + 2][q] = sel012(k, sv[0][q], sv[1][q], sv[2][q]);
                rowc[q] = sel012(k, sc[0][q], sc[1][q], sc[2][q]);
            }
            const gbytes mr = mask + mrow[py + 2];
#pragma unroll
            for (int px = -2; px <= 2; ++px) {
                const int i = (py + 2) * 5 + px + 2, j = kx[px + 2];
                const int c = sel012(j, rowc[0], rowc[1], rowc[2]);
                ch[i] = c;
                cv[i] = *(const __attribute__((address_space(1))) float*)(mr + mcol[px + 2] + 4 * (c < 0 ? 0 : c));
            }
        }
        // taps in the reference's raster order; each adds to the sums of its site's channel only
#pragma unroll
        for (int py = -2; py <= 2; ++py) {
#pragma unroll
            for (int px = -2; px <= 2; ++px) {
                const int i = (py + 2) * 5 + px + 2, c = ch[i], j = kx[px + 2];
                const float wgt = wt[i];
                const float v = sel012(j, rowv[py + 2][0], rowv[py + 2][1], rowv[py + 2][2]);
                float cert = cv[i];
                if (!isfinite(cert)) cert = 0.0f;
                const float t = v * wgt * cert, u = wgt * cert;
                acc[0] = c == 0 ? acc[0] + t : acc[0];
                ws[0] = c == 0 ? ws[0] + u : ws[0];
                acc[1] = c == 1 ? acc[1] + t : acc[1];
                ws[1] = c == 1 ? ws[1] + u : ws[1];
                acc[2] = c == 2 ? acc[2] + t : acc[2];
                ws[2] = c == 2 ? ws[2] + u : ws[2];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ps[c] = acc[c];
        pw[c] = ws[c];
    }
}

AccArgs make_acc(const float white[3], const float black[3], const int cfa[4]) {
    AccArgs a;
    for (int i = 0; i < 4; ++i) a.cfa.c[i] = cfa[i];
    a.white_int = 1;
    for (int c = 0; c < 3; ++c) {
        a.white[c] = white[c];
        a.black[c] = black[c];
        a.inv_white[c] = 1.0 / (double)white[c];
        if (!(white[c] >= 1.0f && white[c] <= 16777216.0f && white[c] == std::floor(white[c]))) a.white_int = 0;
    }
    return a;
}

}  // namespace

namespace isa {

int launch_subsample3(const uint16_t* raw, float* out, float max_val, int hw, int hh, int out_pitch, const int cfa[4],
                      hipStream_t s) {
    Cfa c;
    for (int i = 0; i < 4; ++i) c.c[i] = cfa[i];
    if (hw % 4 == 0 && out_pitch % 16 == 0 && reinterpret_cast<uintptr_t>(raw) % 16 == 0 &&
        reinterpret_cast<uintptr_t>(out) % 16 == 0)
        k_subsample3_x4<<<dim3(div_up(hw, 4 * 64), div_up(hh, 4)), dim3(64, 4), 0, s>>>(raw, out, max_val, hw, hh,
                                                                                        out_pitch, c);
    else
        k_subsample3<<<dim3(div_up(hw, 16), div_up(hh, 16)), dim3(16, 16), 0, s>>>(raw, out, max_val, hw, hh,
                                                                                     out_pitch, c);
    return check_launch("subsample3");
}

int launch_robustness(const float* ref, const float* mov, int img_pitch, float* mask, int mask_pitch, const float* flow,
                      int flow_pitch, int fw, int fh, int w, int h, float alpha, float beta, float thr, int quant,
                      hipStream_t s) {
    k_robustness<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(ref, mov, img_pitch, mask, mask_pitch, flow,
                                                                            flow_pitch, fw, fh, w, h, alpha, beta, thr,
                                                                            quant);
    return check_launch("robustness_mask");
}

int launch_accumulate(const uint16_t* raw, float* sum, float* wsum, const float* mask, const float* kern,
                      const float* flow, const float white[3], const float black[3], int w, int h, int out_pitch,
                      int mask_pitch, int flow_pitch, const int cfa[4], hipStream_t s) {
    const AccArgs a = make_acc(white, black, cfa);
    auto kern_fn = a.white_int ? k_accumulate<true> : k_accumulate<false>;
    kern_fn<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(raw, sum, wsum, mask, kern, flow, w, h, out_pitch,
                                                                      mask_pitch, flow_pitch, a);
    return check_launch("accumulate");
}

int launch_accumulate_frames(float* sum, float* wsum, const float* kern, int w, int h, int out_pitch, int mask_pitch,
                             int flow_pitch, const uint16_t* const* raw, const float* const* mask,
                             const float* const* flow, int nframes, const float white[3], const float black[3],
                             const int cfa[4], hipStream_t s) {
    // raw, mask, flow: host arrays of device pointers, frames in accumulation order; launches of up to
    // ACC_MAXF frames run in order on the stream, each continuing the sums of the previous one
    const AccArgs a = make_acc(white, black, cfa);
    // wide form: raw frames 4-byte aligned with an even width, mask rows 16-byte aligned, flow pixels (float2)
    // 8-byte aligned
    bool wide = w % 2 == 0 && mask_pitch % 16 == 0 && flow_pitch % 8 == 0;
    for (int f = 0; f < nframes && wide; ++f)
        wide = reinterpret_cast<uintptr_t>(raw[f]) % 4 == 0 && reinterpret_cast<uintptr_t>(mask[f]) % 16 == 0 &&
               reinterpret_cast<uintptr_t>(flow[f]) % 8 == 0;
    auto kern_fn = a.white_int ? k_accumulate_frames<true> : k_accumulate_frames<false>;
    auto wide_fn = a.white_int ? k_accumulate_frames_wide<true> : k_accumulate_frames_wide<false>;
    for (int f0 = 0; f0 < nframes; f0 += ACC_MAXF) {
        FramesArgs fa{};
        fa.nframes = std::min(ACC_MAXF, nframes - f0);
        for (int i = 0; i < fa.nframes; ++i) {
            fa.raw[i] = raw[f0 + i];
            fa.mask[i] = mask[f0 + i];
            fa.flow[i] = flow[f0 + i];
        }
        if (wide)
            wide_fn<<<dim3(div_up(w, ACCW_X), div_up(h, ACCW_Y)), 256, 0, s>>>(sum, wsum, kern, w, h, out_pitch,
                                                                            mask_pitch, flow_pitch, fa, a);
        else
            kern_fn<<<dim3(div_up(w, 32), div_up(h, 8)), dim3(32, 8), 0, s>>>(sum, wsum, kern, w, h, out_pitch,
                                                                              mask_pitch, flow_pitch, fa, a);
        ISA_TRY(check_launch("accumulate_frames"));
    }
    return ISA_OK;
}

// Super-resolution accumulation of `nframes` frames (host arrays of device pointers, in accumulation order) onto
// an out_w x out_h grid with offset (ox, oy); launches of up to ACC_MAXF frames run in order on the stream.
int launch_accumulate_sr_frames(float* sum, float* wsum, const float* kern4, int w, int h, int out_pitch,
                                int mask_pitch, int kern_pitch, int flow_pitch, const uint16_t* const* raw,
                                const float* const* mask, const float* const* flow, int nframes, int out_w, int out_h,
                                int ox, int oy, const float white[3], const float black[3], const int cfa[4], int quant,
                                hipStream_t s) {
    const AccArgs a = make_acc(white, black, cfa);
    auto kern_fn = a.white_int ? k_accumulate_sr_frames<true> : k_accumulate_sr_frames<false>;
    const SrGeom g{out_w, out_h, ox, oy};
    for (int f0 = 0; f0 < nframes; f0 += ACC_MAXF) {
        FramesArgs fa{};
        fa.nframes = std::min(ACC_MAXF, nframes - f0);
        for (int i = 0; i < fa.nframes; ++i) {
            fa.raw[i] = raw[f0 + i];
            fa.mask[i] = mask[f0 + i];
            fa.flow[i] = flow[f0 + i];
        }
        kern_fn<<<dim3(div_up(out_w, 64), div_up(out_h, 4)), dim3(64, 4), 0, s>>>(
            sum, wsum, kern4, w, h, out_pitch, mask_pitch, kern_pitch, flow_pitch, fa, g, a, quant);
        ISA_TRY(check_launch("accumulate_superres"));
    }
    return ISA_OK;
}

int launch_accumulate_sr(const uint16_t* raw, float* sum, float* wsum, const float* mask, const float* kern4,
                         const float* flow, const float white[3], const float black[3], int w, int h, int out_pitch,
                         int mask_pitch, int kern_pitch, int flow_pitch, const int cfa[4], int quant, hipStream_t s) {
    return launch_accumulate_sr_frames(sum, wsum, kern4, w, h, out_pitch, mask_pitch, kern_pitch, flow_pitch, &raw,
                                       &mask, &flow, 1, w, h, w / 2, h / 2, white, black, cfa, quant, s);
}

}  // namespace isa

static bool cfa_valid(const int cfa[4]) {
    for (int i = 0; i < 4; ++i)
        if (cfa[i] < 0 || cfa[i] > 2) return false;
    return true;
}

extern "C" {

int isa_debayer_subsample3(const uint16_t* raw_dev, float* out3_dev, float max_val, int dim_x, int dim_y, int out_pitch,
                           const int cfa[4], isa_stream_t stream) {
    ISA_REQUIRE(raw_dev && out3_dev && dim_x > 0 && dim_y > 0 && cfa_valid(cfa), "isa_debayer_subsample3: bad arguments");
    return isa::launch_subsample3(raw_dev, out3_dev, max_val, dim_x, dim_y, out_pitch, cfa, isa::to_stream(stream));
}

int isa_compute_robustness_mask(const float* ref3_dev, const float* moved3_dev, int img_pitch, float* mask4_dev,
                                int mask_pitch, const float* flow2_dev, int flow_pitch, int flow_w, int flow_h,
                                int width, int height, float alpha, float beta, float threshold_m, int tex_quant,
                                isa_stream_t stream) {
    ISA_REQUIRE(ref3_dev && moved3_dev && mask4_dev && flow2_dev && width > 0 && height > 0 && flow_w > 0 &&
                    flow_h > 0,
                "isa_compute_robustness_mask: bad arguments");
    return isa::launch_robustness(ref3_dev, moved3_dev, img_pitch, mask4_dev, mask_pitch, flow2_dev, flow_pitch, flow_w,
                                  flow_h, width, height, alpha, beta, threshold_m, tex_quant, isa::to_stream(stream));
}

int isa_accumulate_images(const uint16_t* raw_dev, float* sum3_dev, float* weight3_dev, const float* mask4_dev,
                          const float* kernel3_dev, const float* flow2_dev, const float white[3], const float black[3],
                          int dim_x, int dim_y, int out_pitch, int mask_pitch, int flow_pitch, const int cfa[4],
                          isa_stream_t stream) {
    ISA_REQUIRE(raw_dev && sum3_dev && weight3_dev && mask4_dev && kernel3_dev && flow2_dev && dim_x > 0 && dim_y > 0 &&
                    cfa_valid(cfa),
                "isa_accumulate_images: bad arguments");
    return isa::launch_accumulate(raw_dev, sum3_dev, weight3_dev, mask4_dev, kernel3_dev, flow2_dev, white, black,
                                  dim_x, dim_y, out_pitch, mask_pitch, flow_pitch, cfa, isa::to_stream(stream));
}

int isa_accumulate_frames(const uint16_t* const* raw_devs, const float* const* mask4_devs,
                          const float* const* flow2_devs, int frames, float* sum3_dev, float* weight3_dev,
                          const float* kernel3_dev, const float white[3], const float black[3], int dim_x, int dim_y,
                          int out_pitch, int mask_pitch, int flow_pitch, const int cfa[4], isa_stream_t stream) {
    ISA_REQUIRE(raw_devs && mask4_devs && flow2_devs && frames >= 0 && sum3_dev && weight3_dev && kernel3_dev &&
                    dim_x > 0 && dim_y > 0 && cfa_valid(cfa),
                "isa_accumulate_frames: bad arguments");
    for (int f = 0; f < frames; ++f)
        ISA_REQUIRE(raw_devs[f] && mask4_devs[f], "isa_accumulate_frames: null frame pointer");
    return isa::launch_accumulate_frames(sum3_dev, weight3_dev, kernel3_dev, dim_x, dim_y, out_pitch, mask_pitch,
                                         flow_pitch, raw_devs, mask4_devs, flow2_devs, frames, white, black, cfa,
                                         isa::to_stream(stream));
}

int isa_accumulate_frames_superres(const uint16_t* const* raw_devs, const float* const* mask4_devs,
                                   const float* const* flow2_devs, int frames, float* sum3_dev, float* weight3_dev,
                                   const float* kernel4_dev, const float white[3], const float black[3], int dim_x,
                                   int dim_y, int out_w, int out_h, int offset_x, int offset_y, int out_pitch,
                                   int mask_pitch, int kernel_pitch, int flow_pitch, const int cfa[4], int tex_quant,
                                   isa_stream_t stream) {
    ISA_REQUIRE(raw_devs && mask4_devs && flow2_devs && frames >= 0 && sum3_dev && weight3_dev && kernel4_dev &&
                    dim_x > 0 && dim_y > 0 && out_w > 2 && out_h > 2 && offset_x >= 0 && offset_y >= 0 &&
                    cfa_valid(cfa),
                "isa_accumulate_frames_superres: bad arguments");
    // every raw and certainty coordinate stays inside the frame
    ISA_REQUIRE(offset_x / 2 + out_w / 2 <= dim_x && offset_y / 2 + out_h / 2 <= dim_y,
                "isa_accumulate_frames_superres: the output grid must lie inside the frame");
    ISA_REQUIRE(out_pitch >= out_w * 12, "isa_accumulate_frames_superres: bad output pitch");
    for (int f = 0; f < frames; ++f)
        ISA_REQUIRE(raw_devs[f] && mask4_devs[f], "isa_accumulate_frames_superres: null frame pointer");
    return isa::launch_accumulate_sr_frames(sum3_dev, weight3_dev, kernel4_dev, dim_x, dim_y, out_pitch, mask_pitch,
                                            kernel_pitch, flow_pitch, raw_devs, mask4_devs, flow2_devs, frames, out_w,
                                            out_h, offset_x, offset_y, white, black, cfa, tex_quant,
                                            isa::to_stream(stream));
}

int isa_accumulate_images_superres(const uint16_t* raw_dev, float* sum3_dev, float* weight3_dev, const float* mask4_dev,
                                   const float* kernel4_dev, const float* flow2_dev, const float white[3],
                                   const float black[3], int dim_x, int dim_y, int out_pitch, int mask_pitch,
                                   int kernel_pitch, int flow_pitch, const int cfa[4], int tex_quant,
                                   isa_stream_t stream) {
    ISA_REQUIRE(raw_dev && sum3_dev && weight3_dev && mask4_dev && kernel4_dev && flow2_dev && dim_x > 0 && dim_y > 0 &&
                    cfa_valid(cfa),
                "isa_accumulate_images_superres: bad arguments");
    return isa::launch_accumulate_sr(raw_dev, sum3_dev, weight3_dev, mask4_dev, kernel4_dev, flow2_dev, white, black,
                                     dim_x, dim_y, out_pitch, mask_pitch, kernel_pitch, flow_pitch, cfa, tex_quant,
                                     isa::to_stream(stream));
}

}  // extern "C"
