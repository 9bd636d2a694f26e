// Differentiable merge: a continuous relaxation of robustness + kernel accumulation with its adjoint.
//
// The reference path rounds the flow to integer sample shifts before gathering
// (R:Kernels/DeBayerKernels.cu:310-313, 403-406; R:Kernels/RobustnessModell.cu:76-77), so its output
// is piecewise constant in the flow. The relaxed merge keeps every other step and makes the flow
// enter continuously:
//  * output grid: 2W x 2H over the whole frame; output pixel y sits at raw position p = y/2 - 1/4,
//    its base raw pixel is b = floor(y/2);
//  * per frame f the 5x5 taps around the fixed integer centre b + anchor_f(b) (the anchor is the
//    rounded initial flow, held fixed), tap s = b + anchor + q at distance d = s - (p + flow_f(b))
//    from the warped sample position; as in the reference the sample is read at s clamped to the
//    frame while d uses the unclamped s;
//  * weight w = exp(-0.5 d^T Omega(b) d) with Omega from ComputeKernelParam on the reference's
//    smoothed structure tensor at b and the parameters theta = (k_detail, k_denoise, D_th, D_tr);
//    a non-finite weight falls back to the cross kernel as in the reference;
//  * certainty: the robustness mask of frame f at half-res pixel h = floor(b/2), where the moved
//    frame's 3x3 mean is sampled bilinearly at h + flow_f(2h)/2 instead of at the rounded shift
//    (no erosion, no M test);
//  * S_c = sum_f sum_taps m_c w v, W_c = sum m_c w over the taps of colour c; out_c = S_c / W_c.
// Loss: L = sum over the output interior (4-pixel border excluded) of (out - gt)^2.
//
// The backward pass recomputes every weight (nothing per frame is stored by the forward pass but
// the half-res masks) and is a pure gather: the thread of raw pixel b collects d L / d flow_f(b)
// from its four output pixels and, for even b, the mask term from the sixteen output pixels that
// read mask_f(b/2) (combined through LDS with its three neighbours). d L / d theta is reduced per
// block and summed in a fixed order.
#include <cmath>
#include <vector>

#include "isa_device.h"
#include "isa_host.h"

namespace {

// forward-mode dual number over the four kernel parameters
struct D4 {
    float v, d[4];
};
__device__ __forceinline__ D4 dconst(float v) { return D4{v, {0, 0, 0, 0}}; }
__device__ __forceinline__ D4 dvar(float v, int i) {
    D4 r = dconst(v);
    r.d[i] = 1.0f;
    return r;
}
__device__ __forceinline__ D4 operator+(D4 a, D4 b) {
    D4 r{a.v + b.v, {}};
    for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
__device__ __forceinline__ D4 operator-(D4 a, D4 b) {
    D4 r{a.v - b.v, {}};
    for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
__device__ __forceinline__ D4 operator*(D4 a, D4 b) {
    D4 r{a.v * b.v, {}};
    for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
__device__ __forceinline__ D4 operator/(D4 a, D4 b) {
    D4 r{a.v / b.v, {}};
    for (int i = 0; i < 4; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
    return r;
}
__device__ __forceinline__ D4 operator*(float s, D4 a) { return dconst(s) * a; }
__device__ __forceinline__ D4 dclamp01(D4 a) {
    if (a.v <= 0.0f) return dconst(0.0f);
    if (a.v >= 1.0f) return dconst(1.0f);
    return a;
}

// ComputeKernelParam (R:Kernels/kernel.cu:627-700) on one smoothed tensor (a11, a22, a12), with
// theta = (k_detail, k_denoise, D_th, D_tr) as duals. Returns Omega = (xx, yy, xy).
__device__ void omega_of(float a11, float a22, float a12, const float th[4], float kstretch, float kshrink,
                         D4 om[3]) {
    const D4 kDetail = dvar(th[0], 0), kDenoise = dvar(th[1], 1), Dth = dvar(th[2], 2), Dtr = dvar(th[3], 3);
    float help = sqrtf((a22 - a11) * (a22 - a11) + 4.0f * a12 * a12);
    float c = 2.0f * a12, s = a22 - a11 + help;
    float norm = sqrtf(c * c + s * s);
    if (norm > 0) {
        c /= norm;
        s /= norm;
    } else {
        c = 1.0f;
        s = 0.0f;
    }
    float lam1 = (a11 + a22 + help) / 2.0f, lam2 = (a11 + a22 - help) / 2.0f;
    float A = 1.0f + sqrtf((lam1 - lam2) * (lam1 - lam2) / ((lam1 + lam2) * (lam1 + lam2)));
    D4 D = dclamp01(dconst(1.0f) - dconst(sqrtf(lam1)) / Dtr + Dth);
    D4 k1h = (kstretch * A) * kDetail, k2h = (A / kshrink) * kDetail;
    D4 dn = kDetail * kDenoise;
    D4 one_d = dconst(1.0f) - D;
    D4 k1 = one_d * k1h + D * dn, k2 = one_d * k2h + D * dn;
    k1 = k1 * k1;
    k2 = k2 * k2;
    float x2 = c, y2 = s, x1 = s, y1 = -c;
    D4 b11 = (x1 * x1) * k1 + (x2 * x2) * k2;
    D4 b12 = (x1 * y1) * k1 + (x2 * y2) * k2;
    D4 b22 = (y1 * y1) * k1 + (y2 * y2) * k2;
    D4 det = b11 * b22 - b12 * b12 + dconst(1e-10f);
    om[0] = b22 / det;
    om[1] = b11 / det;
    om[2] = dconst(0.0f) - b12 / det;
}

__device__ __forceinline__ void omega_val(float a11, float a22, float a12, const float th[4], float ks, float kh,
                                          float om[3]) {
    D4 o[3];
    omega_of(a11, a22, a12, th, ks, kh, o);
    om[0] = o[0].v;
    om[1] = o[1].v;
    om[2] = o[2].v;
}

struct DiffArgs {
    int w, h, n, ref;
    const uint16_t* raw;   // n x h x w
    const float* half;     // n x (h/2) x (w/2) x 3
    const float* tensor;   // h x w x 3
    const float* anchor;   // n x h x w x 2 (integral values)
    const float* flow;     // n x h x w x 2
    float black[3], white[3];
    Cfa cfa;
    float alpha, beta, kstretch, kshrink;
    float th[4];
};

__device__ __forceinline__ float tap_weight(const float om[3], float dx, float dy, int qx, int qy) {
    float w = expf(-0.5f * (om[0] * dx * dx + 2.0f * om[2] * dx * dy + om[1] * dy * dy));
    if (!isfinite(w)) w = (qx * qy == 0) ? 1.0f : 0.0f;
    return w;
}

// ---- relaxed robustness mask
// mean and (population) variance over the 3x3 neighbourhood of half-res pixel (x, y), clamped
__global__ void k_mean3(const float* half, float* mean, float* var, int hw, int hh, int n) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, f = blockIdx.z;
    if (x >= hw || y >= hh || f >= n) return;
    const float* img = half + (size_t)f * hw * hh * 3;
    float m[3] = {0, 0, 0};
    for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) {
            const float* p = img + ((size_t)clampi(y + j, 0, hh - 1) * hw + clampi(x + i, 0, hw - 1)) * 3;
            for (int c = 0; c < 3; ++c) m[c] += p[c];
        }
    for (int c = 0; c < 3; ++c) m[c] /= 9.0f;
    float* o = mean + ((size_t)f * hw * hh + (size_t)y * hw + x) * 3;
    for (int c = 0; c < 3; ++c) o[c] = m[c];
    if (var) {
        float v[3] = {0, 0, 0};
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                const float* p = img + ((size_t)clampi(y + j, 0, hh - 1) * hw + clampi(x + i, 0, hw - 1)) * 3;
                for (int c = 0; c < 3; ++c) v[c] += (p[c] - m[c]) * (p[c] - m[c]);
            }
        float* ov = var + ((size_t)f * hw * hh + (size_t)y * hw + x) * 3;
        for (int c = 0; c < 3; ++c) ov[c] = v[c] / 9.0f;
    }
}

// bilinear sample (clamped) of a 3-channel half-res image and its gradient along x and y
__device__ __forceinline__ void bilin3(const float* img, int hw, int hh, float x, float y, float v[3], float gx[3],
                                       float gy[3]) {
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int x0 = clampi((int)fx, 0, hw - 1), x1 = clampi((int)fx + 1, 0, hw - 1);
    int y0 = clampi((int)fy, 0, hh - 1), y1 = clampi((int)fy + 1, 0, hh - 1);
    const float* p00 = img + ((size_t)y0 * hw + x0) * 3;
    const float* p10 = img + ((size_t)y0 * hw + x1) * 3;
    const float* p01 = img + ((size_t)y1 * hw + x0) * 3;
    const float* p11 = img + ((size_t)y1 * hw + x1) * 3;
    for (int c = 0; c < 3; ++c) {
        v[c] = (1 - a) * (1 - b) * p00[c] + a * (1 - b) * p10[c] + (1 - a) * b * p01[c] + a * b * p11[c];
        gx[c] = (1 - b) * (p10[c] - p00[c]) + b * (p11[c] - p01[c]);
        gy[c] = (1 - a) * (p01[c] - p00[c]) + a * (p11[c] - p10[c]);
    }
}

// mask value of channel c and d mask / d mu_m (0 where clamped)
__device__ __forceinline__ float mask_of(float mr, float vr, float mm, float alpha, float beta, int c, float* dmask) {
    float sd2 = vr, sd = sqrtf(vr);
    float smd = sqrtf(alpha * mr + beta);
    if (c == 1) smd = smd / sqrtf(2.0f);
    float k = sd2 / (sd2 + smd * smd);
    float diff = mr - mm;
    float dist = fabsf(diff) * k;
    float sig = fmaxf(smd, sd);
    float e = expf(-dist * dist / (sig * sig));
    float raw = 1.5f * e - 0.12f;
    float m = fmaxf(fminf(raw, 1.0f), 0.0f);
    if (dmask) {
        // d dist / d mm = -sign(diff) k ; d raw / d dist = 1.5 e (-2 dist / sig^2)
        float dd = diff > 0 ? -k : (diff < 0 ? k : 0.0f);
        *dmask = (raw > 0.0f && raw < 1.0f) ? 1.5f * e * (-2.0f * dist / (sig * sig)) * dd : 0.0f;
    }
    return m;
}

// mask_f(h) for every frame and interior half-res pixel (border: 1)
__global__ void k_relaxed_mask(DiffArgs a, const float* mean, const float* var, float* mask) {
    const int hw = a.w / 2, hh = a.h / 2;
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, f = blockIdx.z;
    if (x >= hw || y >= hh || f >= a.n) return;
    float* o = mask + ((size_t)f * hw * hh + (size_t)y * hw + x) * 3;
    if (x < 1 || y < 1 || x >= hw - 1 || y >= hh - 1) {
        o[0] = o[1] = o[2] = 1.0f;
        return;
    }
    const size_t np = (size_t)a.w * a.h;
    const float* fl = a.flow + ((size_t)f * np + (size_t)(2 * y) * a.w + 2 * x) * 2;
    float mm[3], gx[3], gy[3];
    bilin3(mean + (size_t)f * hw * hh * 3, hw, hh, x + 0.5f * fl[0], y + 0.5f * fl[1], mm, gx, gy);
    const float* mr = mean + ((size_t)a.ref * hw * hh + (size_t)y * hw + x) * 3;
    const float* vr = var + ((size_t)y * hw + x) * 3;
    for (int c = 0; c < 3; ++c) o[c] = mask_of(mr[c], vr[c], mm[c], a.alpha, a.beta, c, nullptr);
}

__device__ __forceinline__ float norm_raw(const DiffArgs& a, float v, int c) {
    return (v - pick3(c, a.black[0], a.black[1], a.black[2])) / pick3(c, a.white[0], a.white[1], a.white[2]);
}

// ---- forward: out (2W x 2H x 3), W (same), loss partials per block (double). One thread per raw
// pixel b computes its four output pixels 2b + {0,1}^2, which share b's taps, flow and Omega.
constexpr int kBX = 16, kBY = 16;

// adds a to s[c] (c in 0..2) without a run-time array index
__device__ __forceinline__ void add3(float& s0, float& s1, float& s2, int c, float a) {
    s0 += c == 0 ? a : 0.0f;
    s1 += c == 1 ? a : 0.0f;
    s2 += c == 2 ? a : 0.0f;
}
__device__ __forceinline__ float pick3(const float v[3], int c) { return c == 0 ? v[0] : (c == 1 ? v[1] : v[2]); }

__global__ __launch_bounds__(kBX* kBY) void k_relaxed_forward(DiffArgs a, const float* mask, float* out, float* wsum,
                                                              const float* gt, double* loss_part) {
    const int ow = 2 * a.w, oh = 2 * a.h, hw = a.w / 2, hh = a.h / 2;
    const int bx = blockIdx.x * kBX + threadIdx.x, by = blockIdx.y * kBY + threadIdx.y;
    const int tid = threadIdx.y * kBX + threadIdx.x;
    __shared__ double red[kBX * kBY];
    double l = 0.0;
    if (bx < a.w && by < a.h) {
        const size_t np = (size_t)a.w * a.h;
        const float* t = a.tensor + ((size_t)by * a.w + bx) * 3;
        float om[3];
        omega_val(t[0], t[1], t[2], a.th, a.kstretch, a.kshrink, om);
        const int hx = min(bx >> 1, hw - 1), hy = min(by >> 1, hh - 1);
        float S[4][3], W[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) S[k][c] = W[k][c] = 0.0f;
        for (int f = 0; f < a.n; ++f) {
            const size_t o = (size_t)f * np + (size_t)by * a.w + bx;
            const float fx = a.flow[o * 2], fy = a.flow[o * 2 + 1];
            const int cx = bx + (int)a.anchor[o * 2], cy = by + (int)a.anchor[o * 2 + 1];
            const float* m = mask + ((size_t)f * hw * hh + (size_t)hy * hw + hx) * 3;
            const float mk[3] = {m[0], m[1], m[2]};
            const uint16_t* raw = a.raw + (size_t)f * np;
            float zx[4], zy[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                zx[k] = (0.5f * (float)(2 * bx + (k & 1)) - 0.25f) + fx;
                zy[k] = (0.5f * (float)(2 * by + (k >> 1)) - 0.25f) + fy;
            }
#pragma unroll 1
            for (int qy = -2; qy <= 2; ++qy) {
                const int sy = clampi(cy + qy, 0, a.h - 1);
#pragma unroll
                for (int qx = -2; qx <= 2; ++qx) {
                    const int sx = clampi(cx + qx, 0, a.w - 1);
                    const int c = a.cfa.at(sy, sx);
                    const float v = norm_raw(a, (float)raw[(size_t)sy * a.w + sx], c);
                    const float mc = pick3(mk, c);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const float wt = tap_weight(om, (float)(cx + qx) - zx[k], (float)(cy + qy) - zy[k], qx, qy);
                        const float mw = mc * wt;
                        add3(S[k][0], S[k][1], S[k][2], c, mw * v);
                        add3(W[k][0], W[k][1], W[k][2], c, mw);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = 2 * bx + (k & 1), y = 2 * by + (k >> 1);
            float* op = out + ((size_t)y * ow + x) * 3;
            float* wp = wsum + ((size_t)y * ow + x) * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                op[c] = W[k][c] > 0.0f ? S[k][c] / W[k][c] : 0.0f;
                wp[c] = W[k][c];
            }
            if (gt && x >= 4 && y >= 4 && x < ow - 4 && y < oh - 4) {
                const float* g = gt + ((size_t)y * ow + x) * 3;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double d = (double)op[c] - (double)g[c];
                    l += d * d;
                }
            }
        }
    }
    red[tid] = l;
    __syncthreads();
    for (int st = kBX * kBY / 2; st > 0; st >>= 1) {
        if (tid < st) red[tid] += red[tid + st];
        __syncthreads();
    }
    if (tid == 0) loss_part[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// ---- backward: per raw pixel b, d L / d flow_f(b) for every f, and the theta partials
__global__ __launch_bounds__(kBX* kBY) void k_relaxed_backward(DiffArgs a, const float* mask, const float* mean,
                                                               const float* var, const float* out, const float* wsum,
                                                               const float* gt, float* dflow, double* theta_part) {
    const int ow = 2 * a.w, oh = 2 * a.h, hw = a.w / 2, hh = a.h / 2;
    const int bx = blockIdx.x * kBX + threadIdx.x, by = blockIdx.y * kBY + threadIdx.y;
    const int tid = threadIdx.y * kBX + threadIdx.x;
    const bool live = bx < a.w && by < a.h;
    const size_t np = (size_t)a.w * a.h;
    __shared__ float s_dm[kBY][kBX][3];
    __shared__ double red[kBX * kBY][4];
    // upstream gradients of the four output pixels of b: dL/dS_c = g/W, dL/dW_c = -g out/W
    float gS[4][3], gW[4][3];
    float omv[3] = {0, 0, 0};
    const float* tb = a.tensor + ((size_t)min(by, a.h - 1) * a.w + min(bx, a.w - 1)) * 3;
    if (live) {
        omega_val(tb[0], tb[1], tb[2], a.th, a.kstretch, a.kshrink, omv);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int ox = 2 * bx + (k & 1), oy = 2 * by + (k >> 1);
            const bool in = gt && ox >= 4 && oy >= 4 && ox < ow - 4 && oy < oh - 4;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gS[k][c] = gW[k][c] = 0.0f;
                if (!in) continue;
                const size_t o = ((size_t)oy * ow + ox) * 3 + c;
                const float W = wsum[o];
                if (!(W > 0.0f)) continue;
                const float g = 2.0f * (out[o] - gt[o]);
                gS[k][c] = g / W;
                gW[k][c] = -g * out[o] / W;
            }
        }
    }
    float dOm0 = 0.0f, dOm1 = 0.0f, dOm2 = 0.0f;
    // a non-finite Omega gives the constant cross-kernel weights: no gradient through w
    const bool om_ok = live && isfinite(omv[0]) && isfinite(omv[1]) && isfinite(omv[2]);
    const int hx = min(bx >> 1, hw - 1), hy = min(by >> 1, hh - 1);
    for (int f = 0; f < a.n; ++f) {
        float gfx = 0.0f, gfy = 0.0f, dm0 = 0.0f, dm1 = 0.0f, dm2 = 0.0f;
        if (live) {
            const size_t o = (size_t)f * np + (size_t)by * a.w + bx;
            const float fx = a.flow[o * 2], fy = a.flow[o * 2 + 1];
            const int cx = bx + (int)a.anchor[o * 2], cy = by + (int)a.anchor[o * 2 + 1];
            const float* m = mask + ((size_t)f * hw * hh + (size_t)hy * hw + hx) * 3;
            const float mk[3] = {m[0], m[1], m[2]};
            const uint16_t* raw = a.raw + (size_t)f * np;
            float zx[4], zy[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                zx[k] = (0.5f * (float)(2 * bx + (k & 1)) - 0.25f) + fx;
                zy[k] = (0.5f * (float)(2 * by + (k >> 1)) - 0.25f) + fy;
            }
#pragma unroll 1
            for (int qy = -2; qy <= 2; ++qy) {
                const int sy = clampi(cy + qy, 0, a.h - 1);
#pragma unroll 1
                for (int qx = -2; qx <= 2; ++qx) {
                    const int sx = clampi(cx + qx, 0, a.w - 1);
                    const int c = a.cfa.at(sy, sx);
                    const float v = norm_raw(a, (float)raw[(size_t)sy * a.w + sx], c);
                    const float mc = pick3(mk, c);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const float dx = (float)(cx + qx) - zx[k], dy = (float)(cy + qy) - zy[k];
                        const float wt = tap_weight(omv, dx, dy, qx, qy);
                        const float coef = pick3(gS[k], c) * v + pick3(gW[k], c);  // dL / d(m w)
                        add3(dm0, dm1, dm2, c, coef * wt);
                        if (!om_ok) continue;
                        const float t = coef * mc * wt;
                        // d w / d flow = w Omega d
                        gfx += t * (omv[0] * dx + omv[2] * dy);
                        gfy += t * (omv[2] * dx + omv[1] * dy);
                        dOm0 += t * (-0.5f * dx * dx);
                        dOm1 += t * (-0.5f * dy * dy);
                        dOm2 += t * (-dx * dy);
                    }
                }
            }
        }
        // the mask of h is read by the 2 x 2 raw pixels 2h + {0,1}^2: combine their terms at 2h
        s_dm[threadIdx.y][threadIdx.x][0] = dm0;
        s_dm[threadIdx.y][threadIdx.x][1] = dm1;
        s_dm[threadIdx.y][threadIdx.x][2] = dm2;
        __syncthreads();
        if (live && ((bx | by) & 1) == 0 && hx >= 1 && hy >= 1 && hx < hw - 1 && hy < hh - 1 && (bx >> 1) == hx &&
            (by >> 1) == hy) {
            const size_t o = (size_t)f * np + (size_t)by * a.w + bx;
            const float fx = a.flow[o * 2], fy = a.flow[o * 2 + 1];
            float mm[3], gx[3], gy[3];
            bilin3(mean + (size_t)f * hw * hh * 3, hw, hh, hx + 0.5f * fx, hy + 0.5f * fy, mm, gx, gy);
            const float* mr = mean + ((size_t)a.ref * hw * hh + (size_t)hy * hw + hx) * 3;
            const float* vr = var + ((size_t)hy * hw + hx) * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float tot = (s_dm[threadIdx.y][threadIdx.x][c] + s_dm[threadIdx.y][threadIdx.x + 1][c]) +
                                  (s_dm[threadIdx.y + 1][threadIdx.x][c] + s_dm[threadIdx.y + 1][threadIdx.x + 1][c]);
                float dmask;
                mask_of(mr[c], vr[c], mm[c], a.alpha, a.beta, c, &dmask);
                gfx += tot * dmask * gx[c] * 0.5f;
                gfy += tot * dmask * gy[c] * 0.5f;
            }
        }
        __syncthreads();
        if (live) {
            float* d = dflow + ((size_t)f * np + (size_t)by * a.w + bx) * 2;
            d[0] = gfx;
            d[1] = gfy;
        }
    }
    // d L / d theta at b: dL/dOmega . dOmega/dtheta (the duals are recomputed here), then a fixed-order
    // block reduction
    double v4[4] = {0.0, 0.0, 0.0, 0.0};
    if (om_ok) {
        D4 om[3];
        omega_of(tb[0], tb[1], tb[2], a.th, a.kstretch, a.kshrink, om);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            v4[i] = (double)dOm0 * (double)om[0].d[i] + (double)dOm1 * (double)om[1].d[i] +
                    (double)dOm2 * (double)om[2].d[i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) red[tid][i] = v4[i];
    __syncthreads();
    for (int st = kBX * kBY / 2; st > 0; st >>= 1) {
        if (tid < st)
            for (int i = 0; i < 4; ++i) red[tid][i] += red[tid + st][i];
        __syncthreads();
    }
    if (tid == 0)
        for (int i = 0; i < 4; ++i) theta_part[(size_t)(blockIdx.y * gridDim.x + blockIdx.x) * 4 + i] = red[0][i];
}

struct Ws {
    float *mean, *var, *mask, *out, *wsum;
    double* part;
    size_t bytes;
};

Ws ws_layout(const isa_merge_diff* d, char* base) {
    auto al = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t hn = (size_t)(d->width / 2) * (d->height / 2) * 3, on = (size_t)4 * d->width * d->height * 3;
    const size_t blocks = 2 * (size_t)isa::div_up(d->width, kBX) * isa::div_up(d->height, kBY);
    Ws w{};
    size_t o = 0;
    w.mean = (float*)(base + o);
    o += al(hn * d->frames * 4);
    w.var = (float*)(base + o);
    o += al(hn * 4);
    w.mask = (float*)(base + o);
    o += al(hn * d->frames * 4);
    w.out = (float*)(base + o);
    o += al(on * 4);
    w.wsum = (float*)(base + o);
    o += al(on * 4);
    w.part = (double*)(base + o);
    o += al(blocks * 4 * 8);
    w.bytes = o;
    return w;
}

int check_problem(const isa_merge_diff* d) {
    ISA_REQUIRE(d && d->raw_dev && d->half_dev && d->tensor_dev && d->anchor_dev, "isa_merge: null problem field");
    ISA_REQUIRE(d->width >= 16 && d->height >= 16 && d->width % 2 == 0 && d->height % 2 == 0 && d->frames >= 1,
                "isa_merge: bad sizes (even, >= 16)");
    ISA_REQUIRE(d->reference >= 0 && d->reference < d->frames, "isa_merge: reference out of range");
    for (int i = 0; i < 4; ++i) ISA_REQUIRE(d->cfa[i] >= 0 && d->cfa[i] <= 2, "isa_merge: CFA entries must be 0..2");
    return ISA_OK;
}

DiffArgs make_args(const isa_merge_diff* d, const float* flow, const float theta[4]) {
    DiffArgs a;
    a.w = d->width;
    a.h = d->height;
    a.n = d->frames;
    a.ref = d->reference;
    a.raw = d->raw_dev;
    a.half = d->half_dev;
    a.tensor = d->tensor_dev;
    a.anchor = d->anchor_dev;
    a.flow = flow;
    for (int c = 0; c < 3; ++c) {
        a.black[c] = d->black[c];
        a.white[c] = d->white[c];
    }
    for (int i = 0; i < 4; ++i) a.cfa.c[i] = d->cfa[i];
    a.alpha = d->noise_alpha;
    a.beta = d->noise_beta;
    a.kstretch = d->k_stretch;
    a.kshrink = d->k_shrink;
    for (int i = 0; i < 4; ++i) a.th[i] = theta[i];
    return a;
}

int masks(const DiffArgs& a, const Ws& w, hipStream_t s) {
    const int hw = a.w / 2, hh = a.h / 2;
    const dim3 g(isa::div_up(hw, 16), isa::div_up(hh, 16), a.n), b(16, 16);
    k_mean3<<<g, b, 0, s>>>(a.half, w.mean, nullptr, hw, hh, a.n);
    ISA_TRY(isa::check_launch("merge_diff: mean3"));
    k_mean3<<<dim3(g.x, g.y, 1), b, 0, s>>>(a.half + (size_t)a.ref * hw * hh * 3, w.var, w.var, hw, hh, 1);
    // (the reference frame's mean was written above; this pass writes its variance into w.var)
    ISA_TRY(isa::check_launch("merge_diff: var3"));
    k_relaxed_mask<<<g, b, 0, s>>>(a, w.mean, w.var, w.mask);
    return isa::check_launch("merge_diff: mask");
}

// Candidate of one refinement step: flow - clip(step g, +-max_step), kept within 0.5 px of the anchor (the
// fixed tap centre) so the 5 x 5 tap window stays centred on the sample position.
__global__ void k_refine_candidate(const float* flow, const float* g, const float* anchor, float* cand, size_t n,
                                   float step, float max_step) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = fminf(fmaxf(step * g[i], -max_step), max_step);
    const float a = anchor[i];
    cand[i] = fminf(fmaxf(flow[i] - d, a - 0.5f), a + 0.5f);
}

}  // namespace

extern "C" {

size_t isa_merge_refine_workspace(const isa_merge_diff* d) {
    if (check_problem(d) != ISA_OK) return 0;
    const size_t nf = (size_t)d->frames * d->width * d->height * 2 * sizeof(float);
    return ((ws_layout(d, nullptr).bytes + 255) & ~(size_t)255) + 2 * ((nf + 255) & ~(size_t)255);
}

int isa_merge_refine_step(const isa_merge_diff* d, float* flow_dev, float theta[4], const float* gt3_dev,
                          float lr_flow, const float lr_theta[4], float max_step, float* scale, double* loss,
                          int* accepted, void* workspace, isa_stream_t stream) {
    ISA_TRY(check_problem(d));
    ISA_REQUIRE(flow_dev && theta && gt3_dev && scale && loss && workspace && lr_flow > 0.0f && max_step > 0.0f &&
                    *scale > 0.0f,
                "isa_merge_refine_step: bad arguments");
    hipStream_t s = isa::to_stream(stream);
    const size_t nf = (size_t)d->frames * d->width * d->height * 2;
    char* base = static_cast<char*>(workspace);
    const size_t wsb = (ws_layout(d, nullptr).bytes + 255) & ~(size_t)255;
    float* g = (float*)(base + wsb);
    float* cand = (float*)(base + wsb + ((nf * sizeof(float) + 255) & ~(size_t)255));
    float gth[4];
    ISA_TRY(isa_merge_backward(d, flow_dev, theta, gt3_dev, g, gth, workspace, stream));
    if (accepted) *accepted = 0;
    for (int attempt = 0; attempt < 9; ++attempt) {
        k_refine_candidate<<<(unsigned)((nf + 255) / 256), 256, 0, s>>>(flow_dev, g, d->anchor_dev, cand, nf,
                                                                      *scale * lr_flow, max_step);
        ISA_TRY(isa::check_launch("isa_merge_refine_step: candidate"));
        float cth[4];
        for (int k = 0; k < 4; ++k) cth[k] = lr_theta ? theta[k] - *scale * lr_theta[k] * gth[k] : theta[k];
        double l = 0.0;
        ISA_TRY(isa_merge_forward(d, cand, cth, nullptr, gt3_dev, &l, workspace, stream));
        if (l <= *loss) {
            ISA_TRY(isa::check_hip(hipMemcpyAsync(flow_dev, cand, nf * sizeof(float), hipMemcpyDeviceToDevice, s),
                                   "isa_merge_refine_step: accept"));
            for (int k = 0; k < 4; ++k) theta[k] = cth[k];
            *loss = l;
            if (accepted) *accepted = 1;
            return ISA_OK;
        }
        *scale *= 0.5f;
    }
    return ISA_OK;
}

size_t isa_merge_diff_workspace(const isa_merge_diff* d) {
    if (check_problem(d) != ISA_OK) return 0;
    return ws_layout(d, nullptr).bytes;
}

int isa_merge_forward(const isa_merge_diff* d, const float* flow_dev, const float theta[4], float* out3_dev,
                      const float* gt3_dev, double* loss, void* workspace, isa_stream_t stream) {
    ISA_TRY(check_problem(d));
    ISA_REQUIRE(flow_dev && theta && workspace, "isa_merge_forward: null argument");
    hipStream_t s = isa::to_stream(stream);
    const DiffArgs a = make_args(d, flow_dev, theta);
    const Ws w = ws_layout(d, (char*)workspace);
    ISA_TRY(masks(a, w, s));
    const dim3 g(isa::div_up(a.w, kBX), isa::div_up(a.h, kBY)), b(kBX, kBY);
    k_relaxed_forward<<<g, b, 0, s>>>(a, w.mask, w.out, w.wsum, gt3_dev, w.part);
    ISA_TRY(isa::check_launch("isa_merge_forward"));
    if (out3_dev)
        ISA_TRY(isa::check_hip(hipMemcpyAsync(out3_dev, w.out, (size_t)4 * a.w * a.h * 3 * sizeof(float),
                                              hipMemcpyDeviceToDevice, s),
                               "isa_merge_forward: copy"));
    if (loss) {
        std::vector<double> part((size_t)g.x * g.y);
        ISA_TRY(isa::check_hip(hipMemcpyAsync(part.data(), w.part, part.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                               "isa_merge_forward: loss"));
        ISA_TRY(isa::check_hip(hipStreamSynchronize(s), "isa_merge_forward: sync"));
        double t = 0.0;
        for (double v : part) t += v;
        *loss = t;
    }
    return ISA_OK;
}

int isa_merge_backward(const isa_merge_diff* d, const float* flow_dev, const float theta[4], const float* gt3_dev,
                       float* dflow_dev, float dtheta[4], void* workspace, isa_stream_t stream) {
    ISA_TRY(check_problem(d));
    ISA_REQUIRE(flow_dev && theta && gt3_dev && dflow_dev && workspace, "isa_merge_backward: null argument");
    hipStream_t s = isa::to_stream(stream);
    const DiffArgs a = make_args(d, flow_dev, theta);
    const Ws w = ws_layout(d, (char*)workspace);
    // the forward state this gradient is taken at (masks, normalised output, weight sums)
    ISA_TRY(masks(a, w, s));
    const dim3 gb(isa::div_up(a.w, kBX), isa::div_up(a.h, kBY)), b(kBX, kBY);
    k_relaxed_forward<<<gb, b, 0, s>>>(a, w.mask, w.out, w.wsum, nullptr, w.part);
    ISA_TRY(isa::check_launch("isa_merge_backward: forward"));
    double* tpart = w.part + (size_t)gb.x * gb.y;
    k_relaxed_backward<<<gb, b, 0, s>>>(a, w.mask, w.mean, w.var, w.out, w.wsum, gt3_dev, dflow_dev, tpart);
    ISA_TRY(isa::check_launch("isa_merge_backward"));
    if (dtheta) {
        std::vector<double> part((size_t)gb.x * gb.y * 4);
        ISA_TRY(isa::check_hip(hipMemcpyAsync(part.data(), tpart, part.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                               "isa_merge_backward: theta"));
        ISA_TRY(isa::check_hip(hipStreamSynchronize(s), "isa_merge_backward: sync"));
        double t[4] = {0, 0, 0, 0};
        for (size_t i = 0; i < part.size(); i += 4)
            for (int k = 0; k < 4; ++k) t[k] += part[i + k];
        for (int k = 0; k < 4; ++k) dtheta[k] = (float)t[k];
    }
    return ISA_OK;
}

}  // extern "C"
