// Full-frame Fourier high-pass (PreAlignment.FourierFilter, R:PEFStudioDX/PreAlignment.cs:268-281;
// transfer function R:Kernels/kernel.cu:702-779 with lp = lps = 1).
//
// The transfer function H equals 1 for every bin with dist >= hp + hps, so
//   C2R(H X)/(W H) = x - C2R((1 - H) X)/(W H)
// and (1 - H) is supported on a block of ~ (hp W) x (2 hp H) bins. The low-pass
// term is computed as a partial DFT:
//   1. row DFT at kx < kxn        U = img . Tw          (GEMM, K = W)
//   2. column DFT at the band ky  Y = (1-H) c_kx/(WH) . sum_y U e^{-i..}
//   3. inverse column DFT         V = sum_j Y e^{+i..}
//   4. inverse row DFT, epilogue  img = clamp(img - V . Tw^T, 0, 1)   (GEMM, K = 2 kxn)
// clearAxis c > 0 also zeroes every bin with kx < c or |ky| < c (the reference's float test on |ky|). Those
// bins take weight 0 in the band above and are removed by three FP64 axis terms computed from the image:
//   L_A  = inverse of the row DFT at kx < c (all ky)                     per row:    U_A(y, kx)
//   L_B  = inverse of the column DFT at the cleared ky (all kx)          per column: Z(x, ky)
//   L_AB = their common bins, subtracted from U_A before its inverse
// img -= L_A + L_B - L_AB runs after the band's row DFT has read the image and before its clamped epilogue.
// The two GEMMs run on the matrix cores when the image rows are 16-byte aligned, else as a tiled
// FP32 SIMT GEMM. On the matrix cores they use FP32 products (v_mfma_f32_16x16x4_f32), or, when the
// caller bounds the image values, three fp16 products of two-term fp16 splits of both operands
// (v_mfma_f32_16x16x32_f16: hi.hi + hi.lo + lo.hi), each within about 2^-21 of the FP32 product,
// at a fifth of the matrix-core time.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "isa_device.h"
#include "isa_host.h"

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

struct Band {
    int kxn, kyn, ny;
};

Band band_of(int w, int h, float hp, float hps) {
    double t = (double)hp + (double)hps;
    Band b;
    b.kxn = std::min(w / 2 + 1, (int)std::ceil(t * w) + 1);
    b.kyn = std::min(h / 2 + 1, (int)std::ceil(t * h) + 1);
    int lo = std::min(b.kyn, h);
    int hi_start = std::max(b.kyn, h - b.kyn + 1);
    b.ny = lo + std::max(0, h - hi_start);
    return b;
}

__device__ __forceinline__ int band_ky(int j, int kyn, int h) {
    int lo = kyn < h ? kyn : h;
    if (j < lo) return j;
    int hi_start = kyn > h - kyn + 1 ? kyn : h - kyn + 1;
    return hi_start + (j - lo);
}

typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));

// Two-term fp16 split v = hi + lo, hi = RN16(v), lo = RN16(v - hi) (v - hi is exact in float): for
// |v| < 2^15 the remainder |v - hi - lo| is at most max(2^-22 |v|, 2^-25).
struct Split16 {
    _Float16 hi, lo;
};
__device__ __forceinline__ Split16 split16(float v) {
    const _Float16 hi = (_Float16)v;
    return Split16{hi, (_Float16)(v - (float)hi)};
}

// tw[x][2k] = cos(2 pi k x / W), tw[x][2k+1] = sin(2 pi k x / W) (row pitch n2p), exact argument
// reduction; twT is the transpose (row pitch twp). twh / twl and twTh / twTl hold the fp16 splits of
// the same floats. Padding columns / rows stay 0.
__global__ void k_row_twiddles(float* tw, int n2p, float* twT, int twp, _Float16* twh, _Float16* twl, _Float16* twTh,
                               _Float16* twTl, int w, int kxn) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (x >= w) return;
    double a = kTwoPi * (double)(((long long)k * x) % w) / (double)w;
    double s, c;
    sincos(a, &s, &c);
    const float cf = (float)c, sf = (float)s;
    const Split16 cs = split16(cf), ss = split16(sf);
    const size_t r = (size_t)x * n2p + 2 * k, tc = (size_t)(2 * k) * twp + x, ts = (size_t)(2 * k + 1) * twp + x;
    tw[r] = cf;
    tw[r + 1] = sf;
    twT[tc] = cf;
    twT[ts] = sf;
    twh[r] = cs.hi;
    twl[r] = cs.lo;
    twh[r + 1] = ss.hi;
    twl[r + 1] = ss.lo;
    twTh[tc] = cs.hi;
    twTl[tc] = cs.lo;
    twTh[ts] = ss.hi;
    twTl[ts] = ss.lo;
}

// Transfer function of R:Kernels/kernel.cu:702-779 (float, same operation order).
__device__ float filter_value(int kx, int ky, int w, int h, float lp, float hp, float lps, float hps) {
    float mx = (float)kx, my = (float)ky;
    if (my > (float)h * 0.5f) my = ((float)h - my) * -1.0f;
    mx /= (float)w;
    my /= (float)h;
    float dist = sqrtf(mx * mx + my * my);
    float fil = 0;
    lp = lp - lps;
    hp = hp + hps;
    if (lp > 0) { if (dist <= lp) fil = 1; }
    else { if (dist <= 1.0f) fil = 1; }
    if (lps > 0) {
        float fil2 = (-fil + 1.0f) * exp_r(-((dist - lp) * (dist - lp) / (2.0f * lps * lps)));
        if (fil2 > 0.001f) fil = fil2;
    }
    if (lps > 0 && lp == 0 && hp == 0 && hps == 0) fil = exp_r(-((dist - lp) * (dist - lp) / (2.0f * lps * lps)));
    if (hp > 0) {
        float fil2 = 0;
        if (dist >= hp) fil2 = 1;
        fil *= fil2;
        if (hps > 0) {
            float fil3 = (-fil2 + 1.0f) * exp_r(-((dist - hp) * (dist - hp) / (2.0f * hps * hps)));
            if (fil3 > 0.001f) fil = fil3;
        }
    }
    return fil;
}

// The reference's test on the row index (R:Kernels/kernel.cu:773): |ky / h| * h < c in float.
__device__ __forceinline__ bool axis_cleared_ky(int ky, int h, int c) {
    float my = (float)ky;
    if (my > (float)h * 0.5f) my = ((float)h - my) * -1.0f;
    my /= (float)h;
    return fabsf(my) * (float)h < (float)c;
}

// Tiled FP32 GEMM, 64 x 64 output tile per 256 threads (4 x 4 per thread), K step 16.
//   C(m, n) = sum_k A(m, k) B(k, n); A(m, k) = A[m*lda + k] (lda in floats, via row pitch bytes)
//   B(k, n) = BT ? B[n*ldb + k] : B[k*ldb + n]
// EPI 0: C[m*ldc + n] = acc.   EPI 1: img(m, n) = clamp(img(m, n) - acc, 0, 1) in place (C = img, ldc = pitch).
constexpr int TM = 64, TN = 64, TK = 16;
// Split K: blockIdx.z covers k in [z kc, min(K, (z+1) kc)); EPI 0 writes its partial to C + z * c_zstride.
template <bool BT, int EPI>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, const float* __restrict__ A, int a_pitch,
                                               const float* __restrict__ B, int ldb, float* C, int c_pitch, int kc,
                                               size_t c_zstride) {
    __shared__ float As[TK][TM + 4];
    __shared__ float Bs[TK][TN + 4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
    C += (size_t)blockIdx.z * c_zstride;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        for (int i = threadIdx.x; i < TM * TK; i += 256) {
            int mm = i / TK, kk = i - mm * TK;
            int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < kend) ? row_ptr(A, a_pitch, gm)[gk] : 0.0f;
        }
        for (int i = threadIdx.x; i < TN * TK; i += 256) {
            int nn, kk;
            if (BT) { nn = i / TK; kk = i - nn * TK; }
            else { kk = i / TN; nn = i - kk * TN; }
            int gn = n0 + nn, gk = k0 + kk;
            float v = 0.0f;
            if (gn < N && gk < kend) v = BT ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int gm = m0 + ty + 16 * i;
        if (gm >= M) continue;
        float* crow = row_ptr(C, c_pitch, gm);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int gn = n0 + tx + 16 * j;
            if (gn >= N) continue;
            if (EPI == 0) {
                crow[gn] = acc[i][j];
            } else {
                float v = crow[gn] - acc[i][j];
                crow[gn] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            }
        }
    }
}

// Sum of the split-K partials in a fixed order: out[i] = sum_z part[z][i].
__global__ void k_sum_partials(const float* __restrict__ part, float* __restrict__ out, size_t n, int nz) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < nz; ++z) s += part[(size_t)z * n + i];
    out[i] = s;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Operand feed of both MFMA GEMMs: lane (j, q) = (lane & 15, lane >> 4) loads A[r + j][k + 4q .. 4q+3]
// and B[k + 4q .. 4q+3][c + j] as float4 runs and issues four 16x16x4 products, e = 0..3, with
// A[.][k + 4q + e] and B[k + 4q + e][.] in slot q of product e: one permutation of the K index
// shared by both operands, so every product A[r][k] B[k][c] enters once. The result tile holds
// D[4q + i][j] in element i.
__device__ __forceinline__ f32x4 mfma4(const float4& a, const float4& b, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b.z, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b.w, acc, 0, 0, 0);
}

typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));

// One 16 x 16 x 32 step on fp16 splits (v_mfma_f32_16x16x32_f16: lane (j, q) holds row or column j
// at K indices 8q .. 8q+7 of both operands; the result tile as above): hi.hi + hi.lo + lo.hi,
// accumulated in FP32. Each fp16 product is exact in FP32; the dropped lo.lo term and the split
// remainders leave each product within about 2^-21 |a b| + 2^-25 (|a| + |b|) of a b.
__device__ __forceinline__ f32x4 mfma3(const h16x8& ah, const h16x8& al, const h16x8& bh, const h16x8& bl, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc, 0, 0, 0);
}

__device__ __forceinline__ void split8(const float (&v)[8], h16x8& hi, h16x8& lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const Split16 p = split16(v[e]);
        hi[e] = p.hi;
        lo[e] = p.lo;
    }
}

constexpr int HP_WAVES = 4;  // waves per block of both MFMA GEMMs
constexpr int HP_NTG = 10;   // 16-column tiles per wave in the row DFT

// Row DFT at the low band: part[z][y][c] = sum over x in K chunk z of img[y][x] twT[c][x]
// (c < n2p, row pitch n2p). A block takes 32 rows, 16 HP_NTG columns and one K chunk of kc; each
// wave sums a quarter of the chunk, and the four wave sums are added through LDS in the fixed order
// ((w3 + w2) + w1) + w0.
// SPLIT: the products on the fp16 splits of img (formed here) and of twT (twTh, twTl).
template <bool SPLIT>
__global__ __launch_bounds__(HP_WAVES * 64, 2) void k_hp_row_dft(const float* __restrict__ img, int pitch, int h, int w,
                                                              const float* __restrict__ twT,
                                                              const _Float16* __restrict__ twTh,
                                                              const _Float16* __restrict__ twTl, int twp, int n2p,
                                                              float* __restrict__ part, int kc) {
    __shared__ f32x4 red[2 * HP_NTG][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, q = lane >> 4;
    const int y0 = blockIdx.x * 32;
    const int t0 = blockIdx.y * HP_NTG, nt = min(HP_NTG, n2p / 16 - t0);
    const int kq = kc / HP_WAVES;
    const int kbeg = blockIdx.z * kc + wave * kq, kend = min(w, kbeg + kq);
    f32x4 acc[2][HP_NTG];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int t = 0; t < HP_NTG; ++t) acc[m][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const float* ar[2];
    bool aok[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int y = y0 + 16 * m + j;
        aok[m] = y < h;
        ar[m] = row_ptr(img, pitch, aok[m] ? y : 0);
    }
    const size_t boff = (size_t)(16 * t0 + j) * twp;
    const float* br = twT + boff;
    // K steps of 32 (SPLIT) or 16; kbeg is a multiple of 32 and twT rows are padded to 32 with zeros
    for (int k = kbeg; k < kend; k += SPLIT ? 32 : 16) {
        if (SPLIT) {
            const int kk = k + 8 * q;
            h16x8 ah[2], al[2], bh[HP_NTG], bl[HP_NTG];
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                float a[8];
                if (aok[m] && kk + 8 <= kend) {
                    const float4 u = *reinterpret_cast<const float4*>(ar[m] + kk);
                    const float4 v = *reinterpret_cast<const float4*>(ar[m] + kk + 4);
                    a[0] = u.x, a[1] = u.y, a[2] = u.z, a[3] = u.w, a[4] = v.x, a[5] = v.y, a[6] = v.z, a[7] = v.w;
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) a[e] = aok[m] && kk + e < kend ? ar[m][kk + e] : 0.0f;
                }
                split8(a, ah[m], al[m]);
            }
#pragma unroll
            for (int t = 0; t < HP_NTG; ++t)
                if (t < nt) {
                    const size_t o = boff + (size_t)16 * t * twp + kk;
                    bh[t] = *reinterpret_cast<const h16x8*>(twTh + o);
                    bl[t] = *reinterpret_cast<const h16x8*>(twTl + o);
                }
#pragma unroll
            for (int t = 0; t < HP_NTG; ++t)
                if (t < nt) {
                    acc[0][t] = mfma3(ah[0], al[0], bh[t], bl[t], acc[0][t]);
                    acc[1][t] = mfma3(ah[1], al[1], bh[t], bl[t], acc[1][t]);
                }
        } else {
            const int kk = k + 4 * q;
            float4 a[2];
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                if (aok[m] && kk + 4 <= kend) {
                    a[m] = *reinterpret_cast<const float4*>(ar[m] + kk);
                } else {
                    a[m].x = aok[m] && kk < kend ? ar[m][kk] : 0.0f;
                    a[m].y = aok[m] && kk + 1 < kend ? ar[m][kk + 1] : 0.0f;
                    a[m].z = aok[m] && kk + 2 < kend ? ar[m][kk + 2] : 0.0f;
                    a[m].w = 0.0f;
                }
            }
            float4 b[HP_NTG];
#pragma unroll
            for (int t = 0; t < HP_NTG; ++t)
                if (t < nt) b[t] = *reinterpret_cast<const float4*>(br + (size_t)16 * t * twp + kk);
#pragma unroll
            for (int t = 0; t < HP_NTG; ++t)
                if (t < nt) {
                    acc[0][t] = mfma4(a[0], b[t], acc[0][t]);
                    acc[1][t] = mfma4(a[1], b[t], acc[1][t]);
                }
        }
    }
    for (int src = HP_WAVES - 1; src >= 1; --src) {
        if (wave == src) {
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int t = 0; t < HP_NTG; ++t) {
                    f32x4 v = acc[m][t];
                    if (src != HP_WAVES - 1) v = red[m * HP_NTG + t][lane] + v;
                    red[m * HP_NTG + t][lane] = v;
                }
        }
        __syncthreads();
    }
    if (wave != 0) return;
    float* out = part + (size_t)blockIdx.z * h * n2p;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int t = 0; t < HP_NTG; ++t) {
            if (t >= nt) continue;
            const f32x4 v = red[m * HP_NTG + t][lane] + acc[m][t];
            const int c = 16 * (t0 + t) + j;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int y = y0 + 16 * m + 4 * q + i;
                if (y < h) out[(size_t)y * n2p + c] = v[i];
            }
        }
}

// U[y][c] (pitch n2) = sum over z of part[z][y][c] (pitch n2p), in z order.
__global__ void k_hp_sum_parts(const float* __restrict__ part, float* __restrict__ U, int h, int n2, int n2p, int nz) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (c >= n2) return;
    float s = 0.0f;
    for (int z = 0; z < nz; ++z) s += part[((size_t)z * h + y) * n2p + c];
    U[(size_t)y * n2 + c] = s;
}

// Inverse row DFT and epilogue: img[y][x] = clamp(img[y][x] - sum_c V[y][c] tw[x][c], 0, 1),
// K = n2p (V and tw zero-padded, row pitch n2p). A wave takes 64 rows x 64 columns (4 x 4 tiles);
// a block 2 x 2 waves.
// SPLIT: the products on the fp16 splits of V (Vh, Vl) and of tw (twh, twl).
template <bool SPLIT>
__global__ __launch_bounds__(HP_WAVES * 64, 2) void k_hp_row_idft(float* img, int pitch, int h, int w,
                                                               const float* __restrict__ V,
                                                               const float* __restrict__ tw,
                                                               const _Float16* __restrict__ Vh,
                                                               const _Float16* __restrict__ Vl,
                                                               const _Float16* __restrict__ twh,
                                                               const _Float16* __restrict__ twl, int n2p) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, q = lane >> 4;
    const int y0 = blockIdx.y * 128 + (wave >> 1) * 64, x0 = blockIdx.x * 128 + (wave & 1) * 64;
    if (y0 >= h || x0 >= w) return;
    f32x4 acc[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[m][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    size_t ao[4], bo[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) ao[m] = (size_t)min(y0 + 16 * m + j, h - 1) * n2p;
#pragma unroll
    for (int t = 0; t < 4; ++t) bo[t] = (size_t)min(x0 + 16 * t + j, w - 1) * n2p;
    for (int k = 0; k < n2p; k += SPLIT ? 32 : 16) {  // n2p is a multiple of 32
        if (SPLIT) {
            // operands swapped: the tile is C[x][y], so lane (j, q) ends with row y = .. + j at the four
            // consecutive columns x = .. + 4q + i
            const int kk = k + 8 * q;
            h16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                ah[m] = *reinterpret_cast<const h16x8*>(Vh + ao[m] + kk);
                al[m] = *reinterpret_cast<const h16x8*>(Vl + ao[m] + kk);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bh[t] = *reinterpret_cast<const h16x8*>(twh + bo[t] + kk);
                bl[t] = *reinterpret_cast<const h16x8*>(twl + bo[t] + kk);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[m][t] = mfma3(bh[t], bl[t], ah[m], al[m], acc[m][t]);
        } else {
            const int kk = k + 4 * q;
            float4 a[4], b[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) a[m] = *reinterpret_cast<const float4*>(V + ao[m] + kk);
#pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = *reinterpret_cast<const float4*>(tw + bo[t] + kk);
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[m][t] = mfma4(a[m], b[t], acc[m][t]);
        }
    }
    auto clamp01 = [](float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); };
    if (SPLIT) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int y = y0 + 16 * m + j;
            if (y >= h) continue;
            float* row = row_ptr(img, pitch, y);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int x = x0 + 16 * t + 4 * q;
                if (x + 4 <= w) {
                    float4 v = *reinterpret_cast<const float4*>(row + x);
                    v.x = clamp01(v.x - acc[m][t][0]);
                    v.y = clamp01(v.y - acc[m][t][1]);
                    v.z = clamp01(v.z - acc[m][t][2]);
                    v.w = clamp01(v.w - acc[m][t][3]);
                    *reinterpret_cast<float4*>(row + x) = v;
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (x + i < w) row[x + i] = clamp01(row[x + i] - acc[m][t][i]);
                }
            }
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int y = y0 + 16 * m + 4 * q + i;
            if (y >= h) continue;
            float* row = row_ptr(img, pitch, y);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int x = x0 + 16 * t + j;
                if (x >= w) continue;
                row[x] = clamp01(row[x] - acc[m][t][i]);
            }
        }
}

constexpr int kRowSplit = 8;

// Column twiddles for the band rows: ctab[y][j] = (cos, sin)(2 pi ky_j y / H), double.
__global__ void k_col_twiddles(double* ctab, int h, int kyn, int ny) {
    int y = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (y >= h) return;
    int ky = band_ky(j, kyn, h);
    double a = kTwoPi * (double)(((long long)ky * y) % h) / (double)h;
    double s, c;
    sincos(a, &s, &c);
    ctab[((size_t)y * ny + j) * 2] = c;
    ctab[((size_t)y * ny + j) * 2 + 1] = s;
}

constexpr int kColChunks = 64;

// Column DFT at the band rows, partial over a chunk of rows. U: [h][2 kxn] with
// U[.][2k] = sum x cos, U[.][2k+1] = sum x sin (= -Im of the row DFT).
// P: [chunk][ny][kxn] complex (double).
__global__ void k_col_dft_partial(const float* __restrict__ U, const double* __restrict__ ctab, double* __restrict__ P,
                                  int h, int kxn, int ny) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y, ch = blockIdx.z;
    if (k >= kxn) return;
    int y0 = (int)((long long)h * ch / kColChunks), y1 = (int)((long long)h * (ch + 1) / kColChunks);
    double sr = 0.0, si = 0.0;
#pragma unroll 4
    for (int y = y0; y < y1; ++y) {
        double c = ctab[((size_t)y * ny + j) * 2], s = ctab[((size_t)y * ny + j) * 2 + 1];
        double ur = U[(size_t)y * 2 * kxn + 2 * k], ui = -(double)U[(size_t)y * 2 * kxn + 2 * k + 1];
        sr += ur * c + ui * s;
        si += ui * c - ur * s;
    }
    P[(((size_t)ch * ny + j) * kxn + k) * 2] = sr;
    P[(((size_t)ch * ny + j) * kxn + k) * 2 + 1] = si;
}

// Sum of the chunks (fixed order), times (1 - H) c_kx / (W H). Y: [ny][kxn] complex.
__global__ void k_col_dft_finish(const double* __restrict__ P, double* __restrict__ Y, int w, int h, int kxn, int kyn,
                                 int ny, int clear_axis, float hp, float hps) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (k >= kxn) return;
    double sr = 0.0, si = 0.0;
    for (int ch = 0; ch < kColChunks; ++ch) {
        sr += P[(((size_t)ch * ny + j) * kxn + k) * 2];
        si += P[(((size_t)ch * ny + j) * kxn + k) * 2 + 1];
    }
    int ky = band_ky(j, kyn, h);
    double fil = (double)filter_value(k, ky, w, h, 1.0f, hp, 1.0f, hps);
    double ck = (k == 0 || (w % 2 == 0 && k == w / 2)) ? 1.0 : 2.0;
    double g = (1.0 - fil) * ck / ((double)w * (double)h);
    if (clear_axis > 0 && (k < clear_axis || axis_cleared_ky(ky, h, clear_axis))) g = 0.0;  // in the axis terms
    Y[((size_t)j * kxn + k) * 2] = sr * g;
    Y[((size_t)j * kxn + k) * 2 + 1] = si * g;
}

// Inverse column DFT: V'[y][2k] = Re, V'[y][2k+1] = -Im (so that the inverse row pass is a plain GEMM
// with Tw^T); row pitch n2p.
// Vh / Vl (when not null): the fp16 splits of the same floats.
__global__ void k_col_idft(const double* __restrict__ Y, const double* __restrict__ ctab, float* __restrict__ V,
                           _Float16* __restrict__ Vh, _Float16* __restrict__ Vl, int h, int kxn, int ny, int n2p) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (k >= kxn || y >= h) return;
    double sr = 0.0, si = 0.0;
#pragma unroll 4
    for (int j = 0; j < ny; ++j) {
        double c = ctab[((size_t)y * ny + j) * 2], s = ctab[((size_t)y * ny + j) * 2 + 1];
        double yr = Y[((size_t)j * kxn + k) * 2], yi = Y[((size_t)j * kxn + k) * 2 + 1];
        sr += yr * c - yi * s;
        si += yr * s + yi * c;
    }
    const size_t o = (size_t)y * n2p + 2 * k;
    const float re = (float)sr, im = (float)(-si);
    V[o] = re;
    V[o + 1] = im;
    if (Vh) {
        const Split16 a = split16(re), b = split16(im);
        Vh[o] = a.hi;
        Vl[o] = a.lo;
        Vh[o + 1] = b.hi;
        Vl[o + 1] = b.lo;
    }
}

// ---------------------------------------------------------------------------------------------------------
// clearAxis terms (FP64). Axis rows: ky_j = band_ky(j, cyn, h) for j < nyc covers |ky| <= c; which of them
// the reference clears is decided per ky by its float test.

// atw[m] = (cos, sin)(2 pi m / w), m < w.
__global__ void k_axis_twiddles(double* atw, int w) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= w) return;
    double s, c;
    sincos(kTwoPi * (double)m / (double)w, &s, &c);
    atw[2 * m] = c;
    atw[2 * m + 1] = s;
}

constexpr int AX_Q = 8;  // frequencies per thread in the axis DFTs

// Sum of a double over a block of 256 threads in a fixed order: the lanes of each wave, then the 4 waves.
__device__ __forceinline__ double block_sum256(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    v = __shfl(v, 0);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// U_A(y, kx) = sum_x img(x, y) e^{-2 pi i kx x / w} for kx < cx: block (row y, kx group of AX_Q).
__global__ __launch_bounds__(256) void k_axis_row_dft(const float* __restrict__ img, int pitch, int w, int cx,
                                                      const double* __restrict__ atw, double* __restrict__ UA) {
    __shared__ double red[4];
    const int y = blockIdx.x, kx0 = blockIdx.y * AX_Q;
    double re[AX_Q], im[AX_Q];
#pragma unroll
    for (int q = 0; q < AX_Q; ++q) re[q] = im[q] = 0.0;
    const float* row = row_ptr(img, pitch, y);
    for (int x = threadIdx.x; x < w; x += 256) {
        const double v = (double)row[x];
        int idx = (int)(((long long)kx0 * x) % w);  // (kx x) mod w, then + x per kx
#pragma unroll
        for (int q = 0; q < AX_Q; ++q) {
            re[q] = fma(v, atw[2 * idx], re[q]);
            im[q] = fma(-v, atw[2 * idx + 1], im[q]);
            idx += x;
            if (idx >= w) idx -= w;
        }
    }
#pragma unroll
    for (int q = 0; q < AX_Q; ++q) {
        const double r = block_sum256(re[q], red), i = block_sum256(im[q], red);
        if (threadIdx.x == 0 && kx0 + q < cx) {
            UA[((size_t)y * cx + kx0 + q) * 2] = r;
            UA[((size_t)y * cx + kx0 + q) * 2 + 1] = i;
        }
    }
}

// Partial Z(x, ky_j) = sum_y img(x, y) e^{-2 pi i ky_j y / h} over the rows of chunk blockIdx.z; actab holds
// (cos, sin)(2 pi ky_j y / h) at [y][j]. Zp: [chunk][j][x] complex.
__global__ __launch_bounds__(256) void k_axis_col_dft(const float* __restrict__ img, int pitch, int w, int h, int nyc,
                                                      const double* __restrict__ actab, int rows,
                                                      double* __restrict__ Zp) {
    const int x = blockIdx.x * 256 + threadIdx.x, j0 = blockIdx.y * AX_Q, ch = blockIdx.z;
    if (x >= w) return;
    const int nq = min(AX_Q, nyc - j0);
    const int y0 = ch * rows, y1 = min(h, y0 + rows);
    double re[AX_Q], im[AX_Q];
#pragma unroll
    for (int q = 0; q < AX_Q; ++q) re[q] = im[q] = 0.0;
    for (int y = y0; y < y1; ++y) {
        const double v = (double)row_ptr(img, pitch, y)[x];
        const double* ct = actab + ((size_t)y * nyc + j0) * 2;
#pragma unroll
        for (int q = 0; q < AX_Q; ++q)
            if (q < nq) {
                re[q] = fma(v, ct[2 * q], re[q]);
                im[q] = fma(-v, ct[2 * q + 1], im[q]);
            }
    }
#pragma unroll
    for (int q = 0; q < AX_Q; ++q)
        if (q < nq) {
            const size_t o = (((size_t)ch * nyc + j0 + q) * w + x) * 2;
            Zp[o] = re[q];
            Zp[o + 1] = im[q];
        }
}

// Z(x, ky_j) / h summed over the chunks in order, zero where the reference keeps ky_j. Z: [j][x] complex.
__global__ void k_axis_col_finish(const double* __restrict__ Zp, double* __restrict__ Z, int w, int h, int c, int cyn,
                                  int nyc, int chunks) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (x >= w) return;
    double re = 0.0, im = 0.0;
    if (axis_cleared_ky(band_ky(j, cyn, h), h, c))
        for (int ch = 0; ch < chunks; ++ch) {
            const size_t o = (((size_t)ch * nyc + j) * w + x) * 2;
            re += Zp[o];
            im += Zp[o + 1];
        }
    Z[((size_t)j * w + x) * 2] = re / (double)h;
    Z[((size_t)j * w + x) * 2 + 1] = im / (double)h;
}

// X(kx, ky_j) = sum_y U_A(y, kx) e^{-2 pi i ky_j y / h} for the cleared ky_j (else 0): block (kx, j).
__global__ __launch_bounds__(256) void k_axis_xab(const double* __restrict__ UA, const double* __restrict__ actab, int h,
                                                  int cx, int c, int cyn, int nyc, double* __restrict__ XAB) {
    __shared__ double red[4];
    const int kx = blockIdx.x, j = blockIdx.y;
    double re = 0.0, im = 0.0;
    const bool cleared = axis_cleared_ky(band_ky(j, cyn, h), h, c);
    if (cleared)
        for (int y = threadIdx.x; y < h; y += 256) {
            const double ur = UA[((size_t)y * cx + kx) * 2], ui = UA[((size_t)y * cx + kx) * 2 + 1];
            const double co = actab[((size_t)y * nyc + j) * 2], si = actab[((size_t)y * nyc + j) * 2 + 1];
            re += ur * co + ui * si;
            im += ui * co - ur * si;
        }
    re = block_sum256(re, red);
    im = block_sum256(im, red);
    if (threadIdx.x == 0) {
        XAB[((size_t)kx * nyc + j) * 2] = re;
        XAB[((size_t)kx * nyc + j) * 2 + 1] = im;
    }
}

// U'(y, kx) = U_A(y, kx) - (1/h) sum_j X(kx, ky_j) e^{+2 pi i ky_j y / h}: the row spectrum without the bins
// L_B already removes.
__global__ void k_axis_uprime(double* __restrict__ UA, const double* __restrict__ XAB, const double* __restrict__ actab,
                              int h, int cx, int nyc) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x, kx = blockIdx.y;
    if (y >= h) return;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < nyc; ++j) {
        const double xr = XAB[((size_t)kx * nyc + j) * 2], xi = XAB[((size_t)kx * nyc + j) * 2 + 1];
        const double co = actab[((size_t)y * nyc + j) * 2], si = actab[((size_t)y * nyc + j) * 2 + 1];
        re += xr * co - xi * si;
        im += xr * si + xi * co;
    }
    UA[((size_t)y * cx + kx) * 2] -= re / (double)h;
    UA[((size_t)y * cx + kx) * 2 + 1] -= im / (double)h;
}

// img(x, y) -= (1/w) sum_{kx < cx} c_kx Re(U'(y, kx) e^{+2 pi i kx x / w}) + sum_j Re(Z(x, ky_j) e^{+2 pi i ky_j y / h}),
// in FP64, rounded once to float; c_kx = 1 at kx = 0 and kx = w/2 (w even), else 2.
__global__ void k_axis_apply(float* img, int pitch, int w, int h, int cx, int nyc, const double* __restrict__ UA,
                             const double* __restrict__ Z, const double* __restrict__ atw,
                             const double* __restrict__ actab) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    double la = 0.0;
    int idx = 0;
    for (int kx = 0; kx < cx; ++kx) {
        const double ck = (kx == 0 || (w % 2 == 0 && kx == w / 2)) ? 1.0 : 2.0;
        const double ur = UA[((size_t)y * cx + kx) * 2], ui = UA[((size_t)y * cx + kx) * 2 + 1];
        la += ck * (ur * atw[2 * idx] - ui * atw[2 * idx + 1]);
        idx += x;
        if (idx >= w) idx -= w;
    }
    double lb = 0.0;
    for (int j = 0; j < nyc; ++j) {
        const double zr = Z[((size_t)j * w + x) * 2], zi = Z[((size_t)j * w + x) * 2 + 1];
        const double co = actab[((size_t)y * nyc + j) * 2], si = actab[((size_t)y * nyc + j) * 2 + 1];
        lb += zr * co - zi * si;
    }
    float* p = row_ptr(img, pitch, y) + x;
    *p = (float)((double)*p - (la / (double)w + lb));
}

struct HpPlan {
    Band b;
    int n2, n2p, twp;  // band columns 2 kxn, padded to 32; row pitch of twT (w padded to 32)
    int kc, nz;        // K chunk and chunk count of the MFMA row DFT
    // clearAxis c (0: none): kx < cx (cx = min(c, w/2 + 1)); axis rows |ky| <= c: nyc of them (band_ky with
    // cyn = min(c + 1, h/2 + 1)); the column DFT in ych chunks of ych_rows rows
    int c, cx, cyn, nyc, ych, ych_rows;
};

HpPlan plan_of(int w, int h, int c, float hp, float hps) {
    HpPlan P;
    P.c = c > 0 ? c : 0;
    P.cx = P.c > 0 ? std::min(P.c, w / 2 + 1) : 0;
    P.cyn = P.c > 0 ? std::min(P.c + 1, h / 2 + 1) : 0;
    P.nyc = 0;
    if (P.c > 0) {
        const int lo = std::min(P.cyn, h), hi_start = std::max(P.cyn, h - P.cyn + 1);
        P.nyc = lo + std::max(0, h - hi_start);
    }
    P.ych_rows = isa::div_up(h, std::min(8, h));
    P.ych = isa::div_up(h, P.ych_rows);
    P.b = band_of(w, h, hp, hps);
    P.n2 = 2 * P.b.kxn;
    P.n2p = isa::div_up(P.n2, 32) * 32;
    P.twp = isa::div_up(w, 32) * 32;
    // about 4096 waves in the row DFT
    const int groups = isa::div_up(h, 32) * isa::div_up(P.n2p / 16, HP_NTG);
    const int nz = std::max(1, std::min(isa::div_up(4096, HP_WAVES * groups), isa::div_up(w, 64)));
    P.kc = isa::div_up(isa::div_up(w, nz), 128) * 128;  // wave quarters of 32-multiples
    P.nz = isa::div_up(w, P.kc);
    return P;
}

struct HpLayout {
    float* part;
    float* tw;
    float* twT;
    double* ctab;
    float* U;
    float* V;
    double* P;
    double* Y;
    _Float16 *twh, *twl, *twTh, *twTl, *Vh, *Vl;  // fp16 splits of tw, twT and V (same pitches)
    // clearAxis terms: atw [w] and actab [h][nyc] twiddles, U_A [h][cx], X_AB [cx][nyc], Z partials
    // [ych][nyc][w] and Z [nyc][w], all complex double
    double *atw, *actab, *UA, *XAB, *Zp, *Z;
    size_t bytes;
};

HpLayout hp_layout(void* ws, int w, int h, const HpPlan& pl) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const Band& b = pl.b;
    char* base = (char*)ws;
    size_t off = 0;
    HpLayout L;
    L.part = (float*)(base + off);
    off += al(std::max((size_t)kRowSplit * h * pl.n2, (size_t)pl.nz * h * pl.n2p) * sizeof(float));
    L.tw = (float*)(base + off);
    off += al((size_t)w * pl.n2p * sizeof(float));
    L.twT = (float*)(base + off);
    off += al((size_t)pl.n2p * pl.twp * sizeof(float));
    L.ctab = (double*)(base + off);
    off += al((size_t)h * b.ny * 2 * sizeof(double));
    L.U = (float*)(base + off);
    off += al((size_t)h * pl.n2 * sizeof(float));
    L.V = (float*)(base + off);
    off += al((size_t)h * pl.n2p * sizeof(float));
    L.P = (double*)(base + off);
    off += al((size_t)kColChunks * b.ny * b.kxn * 2 * sizeof(double));
    L.Y = (double*)(base + off);
    off += al((size_t)b.ny * b.kxn * 2 * sizeof(double));
    _Float16** halves[6] = {&L.twh, &L.twl, &L.twTh, &L.twTl, &L.Vh, &L.Vl};
    const size_t counts[6] = {(size_t)w * pl.n2p, (size_t)w * pl.n2p, (size_t)pl.n2p * pl.twp,
                              (size_t)pl.n2p * pl.twp, (size_t)h * pl.n2p, (size_t)h * pl.n2p};
    for (int i = 0; i < 6; ++i) {
        *halves[i] = (_Float16*)(base + off);
        off += al(counts[i] * sizeof(_Float16));
    }
    double** axis[6] = {&L.atw, &L.actab, &L.UA, &L.XAB, &L.Zp, &L.Z};
    const size_t acounts[6] = {pl.c ? (size_t)w : 0, (size_t)h * pl.nyc, (size_t)h * pl.cx, (size_t)pl.cx * pl.nyc,
                               (size_t)pl.ych * pl.nyc * w, (size_t)pl.nyc * w};
    for (int i = 0; i < 6; ++i) {
        *axis[i] = (double*)(base + off);
        off += al(acounts[i] * 2 * sizeof(double));
    }
    L.bytes = off;
    return L;
}

__global__ void k_hp_zero(float* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace

namespace isa {

size_t highpass_workspace(int w, int h, int clear_axis, float hp, float hps) {
    return hp_layout(nullptr, w, h, plan_of(w, h, clear_axis, hp, hps)).bytes;
}

// If tables_ready, the twiddle tables (and the zero padding of V) in the workspace are reused.
// value_bound > 0 promises |img| <= value_bound: the two GEMMs then run on fp16 splits of their
// operands (three fp16 products per product) when every operand stays well inside the fp16 range:
// |img| <= B, |twiddles| <= 1 and |V| <= 4 ny B (|U| <= sqrt(2) w B, |Y| <= 2 sqrt(2) B).
int highpass(float* img, int pitch, int w, int h, int clear_axis, float hp, float hps, void* ws, bool tables_ready,
             float value_bound, hipStream_t s) {
    const HpPlan pl = plan_of(w, h, clear_axis, hp, hps);
    const Band& b = pl.b;
    HpLayout L = hp_layout(ws, w, h, pl);
    const int n2 = pl.n2, n2p = pl.n2p;
    if (!tables_ready) {
        // the tables and V are contiguous in the workspace only up to alignment: clear each (the
        // fp16 arrays hold an even element count: n2p is a multiple of 16)
        const size_t ntw = (size_t)w * n2p, ntwt = (size_t)n2p * pl.twp, nv = (size_t)h * n2p;
        float* const zf[9] = {L.tw, L.twT, L.V, (float*)L.twh, (float*)L.twl, (float*)L.twTh, (float*)L.twTl,
                              (float*)L.Vh, (float*)L.Vl};
        const size_t zn[9] = {ntw, ntwt, nv, ntw / 2, ntw / 2, ntwt / 2, ntwt / 2, nv / 2, nv / 2};
        for (int i = 0; i < 9; ++i) k_hp_zero<<<(unsigned)((zn[i] + 255) / 256), 256, 0, s>>>(zf[i], zn[i]);
        k_row_twiddles<<<dim3(div_up(w, 256), b.kxn), 256, 0, s>>>(L.tw, n2p, L.twT, pl.twp, L.twh, L.twl, L.twTh,
                                                                    L.twTl, w, b.kxn);
        k_col_twiddles<<<dim3(div_up(h, 256), b.ny), 256, 0, s>>>(L.ctab, h, b.kyn, b.ny);
        if (pl.c > 0) {
            k_axis_twiddles<<<div_up(w, 256), 256, 0, s>>>(L.atw, w);
            k_col_twiddles<<<dim3(div_up(h, 256), pl.nyc), 256, 0, s>>>(L.actab, h, pl.cyn, pl.nyc);
        }
    }
    const bool mfma = pitch % 16 == 0 && reinterpret_cast<uintptr_t>(img) % 16 == 0;
    const bool split = mfma && value_bound > 0.0f && 4.0 * b.ny * (double)value_bound <= 16384.0;
    // 1. U = img . Tw  (M = h, N = 2 kxn, K = w)
    if (mfma) {
        auto row_dft = split ? k_hp_row_dft<true> : k_hp_row_dft<false>;
        row_dft<<<dim3(div_up(h, 32), div_up(n2p / 16, HP_NTG), pl.nz), HP_WAVES * 64, 0, s>>>(
            img, pitch, h, w, L.twT, L.twTh, L.twTl, pl.twp, n2p, L.part, pl.kc);
        k_hp_sum_parts<<<dim3(div_up(n2, 64), h), 64, 0, s>>>(L.part, L.U, h, n2, n2p, pl.nz);
    } else {
        const int kc = div_up(div_up(w, kRowSplit), TK) * TK;
        k_sgemm<false, 0><<<dim3(div_up(n2, TN), div_up(h, TM), div_up(w, kc)), 256, 0, s>>>(
            h, n2, w, img, pitch, L.tw, n2p, L.part, n2 * (int)sizeof(float), kc, (size_t)h * n2);
        const size_t nu = (size_t)h * n2;
        k_sum_partials<<<(unsigned)((nu + 255) / 256), 256, 0, s>>>(L.part, L.U, nu, div_up(w, kc));
    }
    // clearAxis: the axis terms from the image, subtracted from it (the row DFT above has read it)
    if (pl.c > 0) {
        k_axis_row_dft<<<dim3(h, div_up(pl.cx, AX_Q)), 256, 0, s>>>(img, pitch, w, pl.cx, L.atw, L.UA);
        k_axis_col_dft<<<dim3(div_up(w, 256), div_up(pl.nyc, AX_Q), pl.ych), 256, 0, s>>>(img, pitch, w, h, pl.nyc,
                                                                                          L.actab, pl.ych_rows, L.Zp);
        k_axis_col_finish<<<dim3(div_up(w, 256), pl.nyc), 256, 0, s>>>(L.Zp, L.Z, w, h, pl.c, pl.cyn, pl.nyc, pl.ych);
        k_axis_xab<<<dim3(pl.cx, pl.nyc), 256, 0, s>>>(L.UA, L.actab, h, pl.cx, pl.c, pl.cyn, pl.nyc, L.XAB);
        k_axis_uprime<<<dim3(div_up(h, 256), pl.cx), 256, 0, s>>>(L.UA, L.XAB, L.actab, h, pl.cx, pl.nyc);
        k_axis_apply<<<dim3(div_up(w, 64), div_up(h, 4)), dim3(64, 4), 0, s>>>(img, pitch, w, h, pl.cx, pl.nyc, L.UA,
                                                                              L.Z, L.atw, L.actab);
        ISA_TRY(check_launch("fourier_highpass: clear axis"));
    }
    // 2. band column DFT with the transfer function (weight 0 on the cleared bins)
    k_col_dft_partial<<<dim3(div_up(b.kxn, 64), b.ny, kColChunks), 64, 0, s>>>(L.U, L.ctab, L.P, h, b.kxn, b.ny);
    k_col_dft_finish<<<dim3(div_up(b.kxn, 64), b.ny), 64, 0, s>>>(L.P, L.Y, w, h, b.kxn, b.kyn, b.ny, pl.c, hp, hps);
    // 3. inverse column DFT
    k_col_idft<<<dim3(div_up(b.kxn, 64), h), 64, 0, s>>>(L.Y, L.ctab, L.V, split ? L.Vh : nullptr,
                                                         split ? L.Vl : nullptr, h, b.kxn, b.ny, n2p);
    // 4. img = clamp(img - V' . Tw^T)  (M = h, N = w, K = 2 kxn)
    if (mfma) {
        auto row_idft = split ? k_hp_row_idft<true> : k_hp_row_idft<false>;
        row_idft<<<dim3(div_up(w, 128), div_up(h, 128)), HP_WAVES * 64, 0, s>>>(img, pitch, h, w, L.V, L.tw, L.Vh,
                                                                                L.Vl, L.twh, L.twl, n2p);
    } else
        k_sgemm<true, 1><<<dim3(div_up(w, TN), div_up(h, TM), 1), 256, 0, s>>>(
            h, w, n2, L.V, n2p * (int)sizeof(float), L.tw, n2p, img, pitch, n2, 0);
    return check_launch("fourier_highpass");
}

}  // namespace isa

extern "C" {

size_t isa_fourier_highpass_workspace(int width, int height, int clear_axis, float high_pass, float high_pass_sigma) {
    if (width <= 0 || height <= 0 || clear_axis < 0) return 0;
    return isa::highpass_workspace(width, height, clear_axis, high_pass, high_pass_sigma);
}

int isa_fourier_highpass(float* img_dev, int pitch, int width, int height, int clear_axis, float high_pass,
                         float high_pass_sigma, void* workspace_dev, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && img_dev && workspace_dev, "isa_fourier_highpass: bad arguments");
    ISA_REQUIRE(high_pass >= 0 && high_pass_sigma >= 0 && (double)high_pass + high_pass_sigma < 0.25,
                "isa_fourier_highpass: high_pass + high_pass_sigma must lie in [0, 0.25)");
    ISA_REQUIRE(clear_axis >= 0, "isa_fourier_highpass: clear_axis must be >= 0");
    return isa::highpass(img_dev, pitch, width, height, clear_axis, high_pass, high_pass_sigma, workspace_dev, false,
                         0.0f, isa::to_stream(stream));
}

int isa_fourier_highpass_bounded(float* img_dev, int pitch, int width, int height, int clear_axis, float high_pass,
                                 float high_pass_sigma, float value_bound, void* workspace_dev, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && img_dev && workspace_dev, "isa_fourier_highpass_bounded: bad arguments");
    ISA_REQUIRE(high_pass >= 0 && high_pass_sigma >= 0 && (double)high_pass + high_pass_sigma < 0.25,
                "isa_fourier_highpass_bounded: high_pass + high_pass_sigma must lie in [0, 0.25)");
    ISA_REQUIRE(!(value_bound != value_bound), "isa_fourier_highpass_bounded: value_bound is NaN");
    ISA_REQUIRE(clear_axis >= 0, "isa_fourier_highpass_bounded: clear_axis must be >= 0");
    return isa::highpass(img_dev, pitch, width, height, clear_axis, high_pass, high_pass_sigma, workspace_dev, false,
                         value_bound, isa::to_stream(stream));
}

}  // extern "C"
