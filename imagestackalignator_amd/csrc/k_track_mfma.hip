// Tile search with the cross terms on the matrix cores (the mixed-precision alignment path).
//
// For a tile and a reference frame a, every pair (a, b) of the level shares the reference interior
// r (T x T). The cross terms of all those pairs are one matrix product
//     C[s][j] = sum_u A[s][u] * Bm[u][j],   A[s][u] = r(u - s) (0 outside the interior),
//                                            Bm[u][j] = t_j(u),  u in the (T+2M)^2 track block,
// with s = (iy, ix) over the (2M+1)^2 lags and j over the pairs (their track blocks sit at each
// pair's own pre-shift). The product runs on 16x16 MFMA tiles: lags on the rows, up to 16 pairs on
// the columns, the track-block pixels as the reduction dimension. Operands:
//   ISA_XC_BF16X3: x = hi + lo with hi = bf16(x), lo = bf16(x - hi); hi.hi + hi.lo + lo.hi.
//   ISA_XC_FP16X3: the same split in fp16; the reference interior is scaled by a power of two that
//                  brings its maximum into [0.5, 1) (per tile), the track images are split as is.
//   ISA_XC_FP32:   fp32 operands (16x16x4 MFMA).
// All accumulate in fp32. The track images are split once per level into packed (hi | lo << 16)
// planes, so the product loop only loads and repacks. The window energies sum r^2 (FP64) and
// sum t(. + s)^2 (column sums in fp32, totals in FP64) are formed on the vector units, and
// D(s) = sum r^2 + sum t(. + s)^2 - 2 C(s) feeds findMinimum as in the exact search.
#include <algorithm>
#include <cmath>

#include "isa_device.h"
#include "isa_host.h"
#include "tile_search.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));  // 16 bytes at a 4-byte aligned address

constexpr int kPairsPerChunk = 16;  // MFMA columns
constexpr int kWaves = 4;           // waves per block; wave w reduces over track rows uy = w mod 4
constexpr int kMaxNB = 5;           // track-block width up to 80 (16-column blocks)

__device__ __forceinline__ unsigned bf16_rn(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_to_f(unsigned h) { return __uint_as_float(h << 16); }

// x = hi + lo packed as (hi | lo << 16)
template <int PREC> __device__ __forceinline__ unsigned split_pack(float x) {
    if (PREC == ISA_XC_BF16X3) {
        unsigned hi = bf16_rn(x);
        unsigned lo = bf16_rn(x - bf16_to_f(hi));
        return hi | (lo << 16);
    } else {
        _Float16 h = (_Float16)x;
        _Float16 l = (_Float16)(x - (float)h);
        return (unsigned)__builtin_bit_cast(unsigned short, h) | ((unsigned)__builtin_bit_cast(unsigned short, l) << 16);
    }
}

// four packed values -> (hi x4, lo x4) as two 64-bit operand halves; one byte permute per dword
// (v_perm_b32: bytes 0-3 of the 64-bit source are the second operand, 4-7 the first)
__device__ __forceinline__ void unpack4(unsigned x0, unsigned x1, unsigned x2, unsigned x3, u32x2& hi, u32x2& lo) {
    hi = u32x2{__builtin_amdgcn_perm(x1, x0, 0x05040100u), __builtin_amdgcn_perm(x3, x2, 0x05040100u)};
    lo = u32x2{__builtin_amdgcn_perm(x1, x0, 0x07060302u), __builtin_amdgcn_perm(x3, x2, 0x07060302u)};
}

template <int PREC>
__device__ __forceinline__ f32x4 mfma16(u32x2 a, u32x2 b, f32x4 c) {
    if (PREC == ISA_XC_BF16X3)
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(bf16x4, a), __builtin_bit_cast(bf16x4, b), c,
                                                        0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(f16x4, a), __builtin_bit_cast(f16x4, b), c, 0,
                                                     0, 0);
}

// four elements per thread (16-byte accesses when both arrays are 16-byte aligned, else one by one)
template <int PREC> __global__ void k_split_planes(const float* src, unsigned* dst, size_t n, bool vec) {
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (vec && i + 4 <= n) {
        const float4 v = *(const float4*)(src + i);
        *(uint4*)(dst + i) = make_uint4(split_pack<PREC>(v.x), split_pack<PREC>(v.y), split_pack<PREC>(v.z),
                                        split_pack<PREC>(v.w));
        return;
    }
    for (size_t k = i; k < i + 4 && k < n; ++k) dst[k] = split_pack<PREC>(src[k]);
}

// Vertical window energies V(y0, x) = sum_{y = y0}^{y0 + T - 1} src(y, x)^2 for y0 in [0, h - T]:
// one thread per column and run of kVRun window starts (a T-term sum, then slid in FP64). runmax gets
// the largest energy of each (run, column), rounded up: the squares are exact in FP64, so a window's
// error before its rounding to float is at most (T + 2 kVRun) 2^-53 times that maximum.
constexpr int kVRun = 64;
__global__ void k_vbox_sq(const float* src, int pitch, int w, int h, int T, float* dst, float* runmax) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y0 = blockIdx.y * kVRun, nv = h - T + 1;
    if (x >= w || y0 >= nv) return;
    auto at = [&](int y) { return (double)row_ptr(src, pitch, min(y, h - 1))[x]; };
    double e = 0.0, emax = 0.0;
    for (int yb = y0; yb < y0 + T; yb += 8) {  // loads of a batch issued together
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = yb + k < y0 + T ? at(yb + k) : 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) e += v[k] * v[k];
    }
    const int y1 = min(y0 + kVRun, nv);
    for (int yb = y0; yb < y1; yb += 8) {
        double a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a[k] = at(yb + k);
            b[k] = at(yb + k + T);  // rows past the image are clamped and never used
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int y = yb + k;
            if (y >= y1) break;
            dst[(size_t)y * w + x] = (float)e;
            emax = fmax(emax, e);
            if (y + T < h) e = (e - a[k] * a[k]) + b[k] * b[k];
        }
    }
    runmax[(size_t)blockIdx.y * w + x] = (float)(emax * (1.0 + 0x1p-20));
}

// LDS: the reference plane is live until the product loop ends; the partial products, C, the
// column energies and D reuse its bytes afterwards. misc (R2, scale, wave sums of r^2) is separate.
struct MfmaLayout {
    int RW, RH;        // padded reference plane: RH rows x RW columns (+ one zero row)
    int Bp;            // track-block width rounded up to 16
    size_t off_plane;  // reference plane, 4 bytes per element (packed hi|lo, or fp32)
    size_t off_red;    // running sum of the waves' partial products [MB][256] (after the product loop)
    size_t off_c;      // C [chunk][S*S]                                  (after the product loop)
    size_t off_col;    // column energies [chunk][B][S] (float); aliases red (after C is formed)
    size_t off_d;      // D [chunk][S*S]                                  (after the product loop)
    size_t off_w;      // certificate half-widths W [chunk][S*S] (float)   (after the product loop)
    size_t off_misc;   // R2 (double), scale, per-wave sums of r^2 (double), per-pair energy maxima
    size_t bytes;
};

__host__ __device__ inline MfmaLayout mfma_layout(int T, int M) {
    MfmaLayout L;
    const int S = 2 * M + 1, B = T + 2 * M, MB = (S * S + 15) / 16;
    L.Bp = (B + 15) / 16 * 16;
    L.RH = B + S - 1;
    L.RW = L.Bp + S - 1;  // reads reach column Bp - 1 + S - 1
    auto al = [](size_t v) { return (v + 15) / 16 * 16; };
    L.off_plane = 0;
    const size_t plane = al((size_t)(L.RH + 1) * L.RW * 4);
    size_t o = 0;
    // the running sum of the waves' partial products and, after it, the column energies share bytes
    L.off_red = o;
    L.off_col = o;
    o += std::max(al((size_t)MB * 256 * 4), al((size_t)kPairsPerChunk * B * S * 4));
    L.off_c = o;
    o += al((size_t)kPairsPerChunk * S * S * 4);
    L.off_d = o;
    o += al((size_t)kPairsPerChunk * S * S * 4);
    L.off_w = o;
    o += al((size_t)kPairsPerChunk * S * S * 4);
    L.off_misc = o > plane ? o : plane;
    L.bytes = L.off_misc + 128;
    return L;
}

struct XPlanes {
    const unsigned* const* frames;  // per frame: packed split plane, same pixel pitch as the level image
    const float* const* vbox;       // per frame: vertical T-window energies, (h - T + 1) x w
    const float* const* vrun;       // per frame: run maxima of vbox, ceil((h - T + 1) / kVRun) x w
    int* queue;                     // items left to the exact search: (pair, tile) int pairs
    int* count;                     // number of queued items
};

// chunk c of the level: up to 16 pair indices sharing one reference frame (-1 = empty slot)
// waves per SIMD the register allocation must allow: 4 for the 25-lag search, 3 above
template <int M> struct MfmaWaves { static constexpr int v = M <= 2 ? 4 : 3; };

template <int PREC, int M>
__global__ __launch_bounds__(kWaves * 64, MfmaWaves<M>::v) void k_track_mfma(TrackArgs a, XPlanes xp, const int* __restrict__ chunk_pairs,
                                                            int n_chunks, int nblk) {
    constexpr int S = 2 * M + 1, SS = S * S, MB = (SS + 15) / 16;
    constexpr int NT = kWaves * 64;
    extern __shared__ __align__(16) char smem[];
    const int T = a.tile, B = T + 2 * M;
    const MfmaLayout L = mfma_layout(T, M);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // XCD-aware order (as in the exact search): each XCD gets a contiguous run of (tile, chunk)
    const int per = (int)(gridDim.x / 8);
    const int blk = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (blk >= nblk) return;
    if (threadIdx.x < kPairsPerChunk) ((unsigned*)(smem + mfma_layout(a.tile, M).off_misc + 64))[threadIdx.x] = 0u;
    const int tile_idx = blk / n_chunks, chunk = blk - tile_idx * n_chunks;
    const int ty = tile_idx / a.cx, tx = tile_idx - ty * a.cx;
    const int* cp = chunk_pairs + (size_t)chunk * kPairsPerChunk;
    const int fa = a.pa[cp[0]];
    int rox, roy;
    tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fa * 4], a.base[fa * 4 + 1], a.base[fa * 4 + 2], a.base[fa * 4 + 3],
                   0.0f, 0.0f, rox, roy);
    const int rx0 = tx * T + M + rox, ry0 = ty * T + M + roy;
    const float* ia = a.frames[fa];
    float* misc = (float*)(smem + L.off_misc);
    double* R2s = (double*)misc;
    float* scale_s = misc + 2;
    float* red = (float*)(smem + L.off_red);
    auto rpix = [&](int y, int x) {
        return row_ptr(ia, a.img_pitch, clampi(ry0 + y, 0, a.h - 1))[clampi(rx0 + x, 0, a.w - 1)];
    };
    // ---- 0. power-of-two scale of the reference interior (fp16 only)
    float sc = 1.0f;
    if (PREC == ISA_XC_FP16X3) {
        float rmax = 0.0f;
        for (int yb = wave; yb < T; yb += kWaves * 8) {
            float v[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int y = yb + (k >> 1) * kWaves, x = lane + 64 * (k & 1);
                v[k] = (y < T && x < T) ? rpix(y, x) : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) rmax = fmaxf(rmax, fabsf(v[k]));
        }
        red[tid] = rmax;
        __syncthreads();
        if (tid == 0) {
            float mx = 0.0f;
            for (int i = 0; i < NT; ++i) mx = fmaxf(mx, red[i]);
            float f = 1.0f;
            if (mx > 0.0f && isfinite(mx)) {
                int e;
                frexpf(mx, &e);  // mx = m * 2^e, m in [0.5, 1)
                f = ldexpf(1.0f, -e);
            }
            *scale_s = f;
        }
        __syncthreads();
        sc = *scale_s;
    }
    // ---- 1. zero-padded reference plane P[yy][xx] = r(yy - (S-1), xx - (S-1)), and sum r^2
    const int RW = L.RW, RH = L.RH;
    unsigned* pu = (unsigned*)(smem + L.off_plane);
    float* pf = (float*)(smem + L.off_plane);
    double r2 = 0.0;
    constexpr int kFill = 8;
    for (int yb = wave; yb < RH + 1; yb += kWaves * kFill) {
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int xx = lane + 64 * half, x = xx - (S - 1);
            float v[kFill];
#pragma unroll
            for (int k = 0; k < kFill; ++k) {
                const int yy = yb + k * kWaves, y = yy - (S - 1);
                v[k] = (xx < RW && yy < RH && y >= 0 && y < T && x >= 0 && x < T) ? rpix(y, x) : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < kFill; ++k) {
                const int yy = yb + k * kWaves;
                if (xx >= RW || yy >= RH + 1) continue;
                if (PREC == ISA_XC_FP32) pf[yy * RW + xx] = v[k];
                else pu[yy * RW + xx] = split_pack<PREC>(v[k] * sc);
                r2 = fma((double)v[k], (double)v[k], r2);  // v = 0 outside the interior
            }
        }
    }
    // block sum of r^2: a butterfly within each wave, then the four wave sums in order
    for (int off = 32; off >= 1; off >>= 1) r2 += __shfl_xor(r2, off);
    double* r2buf = (double*)(smem + L.off_misc + 16);
    if (lane == 0) r2buf[wave] = r2;
    __syncthreads();
    if (tid == 0) *R2s = ((r2buf[0] + r2buf[1]) + r2buf[2]) + r2buf[3];
    // ---- 2. cross terms on MFMA: lane (column j = lane % 16, quad q = lane / 16)
    const int j = lane & 15, q = lane >> 4;
    const int pj = cp[j];
    int tox = 0, toy = 0;
    if (pj >= 0) {
        const int fb = a.pb[pj];
        const float* sh = a.shifts + (size_t)a.pk[pj] * a.meas_stride + (size_t)tile_idx * 2;
        tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fb * 4], a.base[fb * 4 + 1], a.base[fb * 4 + 2],
                       a.base[fb * 4 + 3], sh[0], sh[1], tox, toy);
    }
    const int fbj = pj >= 0 ? a.pb[pj] : fa;
    const int tx0 = tx * T + tox, ty0 = ty * T + toy;
    const size_t ipitch = (size_t)a.img_pitch / sizeof(float);  // the packed planes share the pixel pitch
    const float* bimg = a.frames[fbj];
    const unsigned* bx = PREC == ISA_XC_FP32 ? nullptr : xp.frames[fbj];
    // LDS offset of this lane's A operand for each row block: row s = mb*16 + (lane % 16)
    int aoff[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int s = mb * 16 + j;
        aoff[mb] = s < SS ? (S - 1 - s / S) * RW + (S - 1 - (s - (s / S) * S)) : -1;  // -1: padding row (zeros)
    }
    const int zero_row = RH * RW;
    f32x4 acc[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) acc[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int NB = L.Bp / 16;  // <= kMaxNB (checked by the launcher)
    const bool cols_in = tx0 >= 0 && tx0 + L.Bp <= a.w;
    int colx[kMaxNB][4];  // image column of each operand element of this lane
#pragma unroll
    for (int nb = 0; nb < kMaxNB; ++nb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ux = 16 * nb + 4 * q + e;  // a lane's four operand words: consecutive columns
            colx[nb][e] = cols_in ? tx0 + ux : clampi(tx0 + ux, 0, a.w - 1);
        }
    // operand words of one track row (packed split or fp32 bits); the wave's next row is fetched
    // while the current one is multiplied
    unsigned cur[kMaxNB][4], nxt[kMaxNB][4];
    // (block inside the image: a lane's four words are consecutive columns, one 16-byte load at a
    // 4-byte aligned address)
    const unsigned* bw = PREC == ISA_XC_FP32 ? (const unsigned*)bimg : bx;
    auto fetch_row = [&](int uy) {
        const size_t ro = (size_t)clampi(ty0 + min(uy, B - 1), 0, a.h - 1) * ipitch;
#pragma unroll
        for (int nb = 0; nb < kMaxNB; ++nb) {
            if (cols_in) {
                u32x4a4 v = {0u, 0u, 0u, 0u};
                if (pj >= 0 && nb < NB) v = *(const u32x4a4*)(bw + ro + colx[nb][0]);
                nxt[nb][0] = v.x;
                nxt[nb][1] = v.y;
                nxt[nb][2] = v.z;
                nxt[nb][3] = v.w;
                continue;
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                unsigned v = 0;
                if (pj >= 0 && nb < NB)
                    v = PREC == ISA_XC_FP32 ? __float_as_uint(bimg[ro + colx[nb][e]]) : bx[ro + colx[nb][e]];
                nxt[nb][e] = v;
            }
        }
    };
    fetch_row(wave);
    for (int uy = wave; uy < B; uy += kWaves) {
#pragma unroll
        for (int nb = 0; nb < kMaxNB; ++nb)
#pragma unroll
            for (int e = 0; e < 4; ++e) cur[nb][e] = nxt[nb][e];
        if (uy + kWaves < B) fetch_row(uy + kWaves);
#pragma unroll
        for (int nb = 0; nb < kMaxNB; ++nb) {
            if (nb >= NB) break;
            const int ux0 = 16 * nb;
            if (PREC == ISA_XC_FP32) {
                // product kk takes column ux0 + 4 q + kk of every lane quad q (the K slices of the four
                // products cover the 16 columns once)
                u32x4a4 av[MB];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
                    av[mb] = *(const u32x4a4*)(pu + (aoff[mb] >= 0 ? aoff[mb] + uy * RW : zero_row) + ux0 + 4 * q);
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const float bv = __uint_as_float(cur[nb][kk]);
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
                        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(av[mb][kk]), bv, acc[mb], 0, 0, 0);
                }
            } else {
                u32x2 bh, bl;
                unpack4(cur[nb][0], cur[nb][1], cur[nb][2], cur[nb][3], bh, bl);
                // all A words of this column block first, so the LDS reads overlap the products
                u32x4a4 aw[MB];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
                    aw[mb] = *(const u32x4a4*)(pu + (aoff[mb] >= 0 ? aoff[mb] + uy * RW : zero_row) + ux0 + 4 * q);
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    u32x2 ah, al;
                    unpack4(aw[mb][0], aw[mb][1], aw[mb][2], aw[mb][3], ah, al);
                    acc[mb] = mfma16<PREC>(al, bh, acc[mb]);
                    acc[mb] = mfma16<PREC>(ah, bl, acc[mb]);
                    acc[mb] = mfma16<PREC>(ah, bh, acc[mb]);
                }
            }
        }
    }
    // ---- 3. combine the waves' partial products in the fixed order ((w0 + w1) + w2) + w3: wave w adds
    // its own to the running sum in LDS, one wave after the other; lane holds C[4q+i][j]
    float* Cm = (float*)(smem + L.off_c);
    const float inv_sc = 1.0f / sc;
    __syncthreads();  // every wave is done with the reference plane, whose bytes are reused below
    for (int w = 0; w < kWaves; ++w) {
        if (wave == w) {
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int e = mb * 256 + (4 * q + i) * 16 + j, s = mb * 16 + 4 * q + i;
                    const float v = w == 0 ? acc[mb][i] : red[e] + acc[mb][i];
                    if (w + 1 < kWaves) red[e] = v;
                    else if (s < SS) Cm[j * SS + s] = v * inv_sc;
                }
        }
        __syncthreads();  // the last one also ends the reads of red, whose bytes the column energies reuse
    }
    // ---- 4. window energies of each pair's track block: per column, the S row windows (a T-row
    // sum, then slid one row at a time), all loads of a column issued in batches of 8. rmax[jj] gets
    // the largest window energy any entry of pair jj was slid through (for the certificate).
    float* colE = (float*)(smem + L.off_col);
    unsigned* rmax = (unsigned*)(smem + L.off_misc + 64);
    for (int task = tid; task < kPairsPerChunk * B; task += NT) {
        const int jj = task / B, c = task - jj * B;
        const int pp = cp[jj];
        if (pp < 0) continue;
        const int fb = a.pb[pp];
        const float* sh = a.shifts + (size_t)a.pk[pp] * a.meas_stride + (size_t)tile_idx * 2;
        int ox, oy;
        tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fb * 4], a.base[fb * 4 + 1], a.base[fb * 4 + 2],
                       a.base[fb * 4 + 3], sh[0], sh[1], ox, oy);
        const int bx0 = tx * T + ox, by0 = ty * T + oy;
        float* ce = colE + ((size_t)jj * B + c) * S;
        float rm;
        if (bx0 >= 0 && by0 >= 0 && bx0 + B <= a.w && by0 + B <= a.h) {
            // block inside the image: the S row windows of this column are entries of the
            // vertical window-energy image
            const float* vb = xp.vbox[fb] + (size_t)by0 * a.w + bx0 + c;
#pragma unroll
            for (int k = 0; k < S; ++k) ce[k] = vb[(size_t)k * a.w];
            const float* vr = xp.vrun[fb] + bx0 + c;
            rm = fmaxf(vr[(size_t)(by0 / kVRun) * a.w], vr[(size_t)((by0 + S - 1) / kVRun) * a.w]);
        } else {
            // the exact search's column energies (FP64 chain, then slid), rows and columns clamped
            const float* img = a.frames[fb];
            const int gx = clampi(bx0 + c, 0, a.w - 1), gy0 = by0;
            auto px = [&](int y) { return (double)row_ptr(img, a.img_pitch, clampi(gy0 + y, 0, a.h - 1))[gx]; };
            double e = 0.0, eb = 0.0;
            for (int y0 = 0; y0 < T; y0 += 8) {
                double v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = y0 + k < T ? px(y0 + k) : 0.0;
#pragma unroll
                for (int k = 0; k < 8; ++k) e = fma(v[k], v[k], e);
            }
            ce[0] = (float)e;
            eb = e;
            double u[S - 1], v[S - 1];
#pragma unroll
            for (int k = 0; k + 1 < S; ++k) {
                u[k] = px(k);
                v[k] = px(k + T);
            }
#pragma unroll
            for (int k = 0; k + 1 < S; ++k) {
                e = (e - u[k] * u[k]) + v[k] * v[k];
                eb += v[k] * v[k];  // energy of the whole column block bounds every window
                ce[k + 1] = (float)e;
            }
            rm = (float)(eb * (1.0 + 0x1p-20));
        }
        atomicMax(&rmax[jj], __float_as_uint(rm));  // non-negative floats order as their bits
    }
    __syncthreads();
    // ---- 5. D(s) = sum r^2 + sum_x colE[x + ix][iy] - 2 C(s), and the half-width W(s) of an interval
    // holding both D(s) and the exact search's value of it:
    //  - cross terms: each product of the split operands differs from r t by at most c_split |r t|
    //    (plus, in fp16, terms from the subnormal spacing); a wave's accumulator receives hgt - 3
    //    products (or fewer, whatever order and grouping the matrix cores use inside an
    //    instruction) and the four waves' sums are added in fp32, so any summation tree has height
    //    <= hgt and the fp32 error is at most gamma(hgt) = hgt u / (1 - hgt u) (u = 2^-23: every
    //    rounding faithful) times the sum of |products|; sum |r(u - s) t(u)| <= sqrt(R2 TT(s));
    //  - window energies: rounding to float (2^-24), the FP64 slide (T + 2 kVRun) 2^-53 rmax per
    //    column and the FP64 sum over the T columns;
    //  - the FP64 evaluation of R2 + TT - 2C on both sides and the rounding of D to float on both.
    float* Dm = (float*)(smem + L.off_d);
    float* Wm = (float*)(smem + L.off_w);
    const double R2 = *R2s;
    constexpr int kProd = PREC == ISA_XC_FP32 ? 4 : 16;   // products per matrix instruction
    constexpr int kPerBlk = PREC == ISA_XC_FP32 ? 4 : 3;  // instructions per 16-column block and row
    const int hgt = kProd * ((B + kWaves - 1) / kWaves) * NB * kPerBlk + (kWaves - 1) + (PREC == ISA_XC_FP32 ? 1 : 0);
    const float gam = (float)(hgt * 0x1p-23 / (1.0 - hgt * 0x1p-23));
    const float csplit = PREC == ISA_XC_BF16X3 ? 3.1f * 0x1p-16f : (PREC == ISA_XC_FP16X3 ? 3.1f * 0x1p-22f : 0.0f);
    const float eta = 0x1p-24f;  // fp16: bound on the absolute error of a split from the subnormal spacing
    // W is evaluated in fp32: every quantity below is rounded up by a relative 2^-20 or more, which
    // covers the roundings of the fp32 evaluation (a few 2^-24 each)
    const float sr2 = fmaxf((float)(sqrt(R2) * (1.0 + 0x1p-20)), R2 > 0.0 ? 1.17549435e-38f : 0.0f);
    const float r2f = (float)(R2 * (1.0 + 0x1p-20));
    const float kslide = (float)((double)T * (T + 2 * kVRun) * 0x1p-53 * 1.01);
    const float crel = (csplit + 1.03f * gam) * (1.0f + 0x1p-18f);
    for (int task = tid; task < kPairsPerChunk * SS; task += NT) {
        const int jj = task / SS, s = task - jj * SS;
        if (cp[jj] < 0) continue;
        const int iy = s / S, ix = s - iy * S;
        const float* ce = colE + ((size_t)jj * B + ix) * S + iy;
        double tt = 0.0;
        for (int x = 0; x < T; ++x) tt += (double)ce[(size_t)x * S];
        const double dd = R2 + tt - 2.0 * (double)Cm[jj * SS + s];
        Dm[jj * SS + s] = (float)dd;
        const float ttf = (float)tt * (1.0f + 0x1p-20f), rm = __uint_as_float(rmax[jj]);
        const float ett = 0x1p-23f * ttf + kslide * rm + (float)T * 0x1p-52f * ttf;
        const float ttu = ttf + ett;
        const float P = sr2 * sqrtf(ttu) * (1.0f + 0x1p-20f);
        float ec = crel * P;
        if (PREC == ISA_XC_FP16X3) ec += (1.1f * eta * T * (sc * sr2 + sqrtf(ttu)) + 3.0f * eta * eta * T * T) / sc;
        const float wd = 1.06f * (2.0f * ec + ett + 0x1p-44f * (r2f + ttf + 2.0f * P) + 0x1p-22f * fabsf((float)dd)) + 1e-30f;
        Wm[jj * SS + s] = wd;
    }
    __syncthreads();
    // ---- 6. findMinimum per pair, when the intervals decide the integer minimum and the threshold
    // test as the exact search would; otherwise the item is queued for the exact search. Sixteen lanes
    // per pair (aligned groups within a wave) scan the lags and combine by shuffles.
    // Undecided items are collected per block and appended with one atomic on the global count (one per
    // item serialised the appends on that address).
    __shared__ int q_n, q_base, q_pair[kPairsPerChunk];
    if (tid == 0) q_n = 0;
    __syncthreads();
    {
        static_assert(kPairsPerChunk * 16 == NT, "sixteen lanes per pair");
        const int jj = tid >> 4, l = tid & 15;
        const int pp = cp[jj];
        const float* D = Dm + jj * SS;
        const float* W = Wm + jj * SS;
        unsigned long long key = ~0ull;
        float mx = -3.402823466e+38f;
        int fin = 1;
        if (pp >= 0)
            for (int s = l; s < SS; s += 16) {
                const float v = D[s];
                key = min(key, argmin_key(v, s));  // first index of the smallest value, as find_min_core
                mx = fmaxf(mx, v);
                fin &= isfinite(v) && isfinite(W[s]);
            }
        for (int o = 8; o >= 1; o >>= 1) {
            key = min(key, (unsigned long long)__shfl_xor((long long)key, o, 16));
            mx = fmaxf(mx, __shfl_xor(mx, o, 16));
            fin &= __shfl_xor(fin, o, 16);
        }
        const int mi = key == ~0ull ? -1 : (int)(unsigned)key;
        const float mn = mi >= 0 ? D[mi] : 3.402823466e+38f;
        const double hi = mi >= 0 ? (double)mn + W[mi] : 0.0, lo = mi >= 0 ? (double)mn - W[mi] : 0.0;
        // every other lag lies strictly above the minimum's interval
        double lother = 1e300, lmax = -1e300, hmax = -1e300;
        if (pp >= 0 && mi >= 0)
            for (int s = l; s < SS; s += 16) {
                const double lv = (double)D[s] - W[s];
                lmax = fmax(lmax, lv);
                hmax = fmax(hmax, (double)D[s] + W[s]);
                if (s != mi) lother = fmin(lother, lv);
            }
        for (int o = 8; o >= 1; o >>= 1) {
            lother = fmin(lother, __shfl_xor(lother, o, 16));
            lmax = fmax(lmax, __shfl_xor(lmax, o, 16));
            hmax = fmax(hmax, __shfl_xor(hmax, o, 16));
        }
        if (pp >= 0 && l == 0) {
            float* sh = a.shifts + (size_t)a.pk[pp] * a.meas_stride + (size_t)tile_idx * 2;
            if (a.shift_images)
#pragma unroll 1
                for (int s = 0; s < SS; ++s) a.shift_images[((size_t)pp * a.cx * a.cy + tile_idx) * SS + s] = D[s];
            // threshold + min > max decided the same way for every value in the intervals
            const double thr = a.threshold;
            const bool zero_sure = (thr + lo) - hmax > 0x1p-20 * (fabs(thr) + fabs(lo) + fabs(hmax));
            const bool keep_sure = lmax - (thr + hi) > 0x1p-20 * (fabs(thr) + fabs(hi) + fabs(lmax));
            const bool ok = fin && mi >= 0 && lother > hi && (zero_sure || keep_sure);
            if (ok) {
                float cx, cy;
                const int dec = find_min_fit(D, M, mi, mn, mx, a.threshold, cx, cy);
                if (a.decisions) a.decisions[(size_t)pp * a.cx * a.cy + tile_idx] = dec;
                const float psx = sh[0], psy = sh[1];
                sh[0] = psx + cx;
                sh[1] = psy + cy;
            } else {
                q_pair[atomicAdd(&q_n, 1)] = pp;
            }
        }
    }
    __syncthreads();
    // the queue holds one slot per (pair, tile) of the level, so every slot stays in range
    if (tid == 0 && q_n > 0) q_base = atomicAdd(xp.count, q_n);
    __syncthreads();
    if (tid < q_n) {
        xp.queue[2 * (q_base + tid)] = q_pair[tid];
        xp.queue[2 * (q_base + tid) + 1] = tile_idx;
    }
}

}  // namespace

namespace isa {

// Pair chunks of one level: pairs grouped by reference frame, up to 16 per chunk, -1 padded.
int track_chunks(const int* pa, int m, int n, int* out_chunks) {
    int nc = 0;
    for (int f = 0; f < n; ++f) {
        int k = 0;
        for (int p = 0; p < m; ++p) {
            if (pa[p] != f) continue;
            if (k == 0) {
                for (int i = 0; i < kPairsPerChunk; ++i) out_chunks[nc * kPairsPerChunk + i] = -1;
            }
            out_chunks[nc * kPairsPerChunk + k] = p;
            if (++k == kPairsPerChunk) {
                ++nc;
                k = 0;
            }
        }
        if (k) ++nc;
    }
    return nc;
}

int track_chunks_capacity(int m, int n) { return m / kPairsPerChunk + n + 1; }

template <int PREC, int M>
static int launch_mfma_pm(const TrackArgs& a, XPlanes xp, const int* chunks, int n_chunks, hipStream_t s) {
    const MfmaLayout L = mfma_layout(a.tile, M);
    const int blocks = a.cx * a.cy * n_chunks;
    k_track_mfma<PREC, M><<<8 * div_up(blocks, 8), kWaves * 64, L.bytes, s>>>(a, xp, chunks, n_chunks, blocks);
    return check_launch("track_mfma");
}

template <int PREC>
static int launch_mfma_p(const TrackArgs& a, XPlanes xp, int M, const int* chunks, int n_chunks, hipStream_t s) {
    switch (M) {
        case 1: return launch_mfma_pm<PREC, 1>(a, xp, chunks, n_chunks, s);
        case 2: return launch_mfma_pm<PREC, 2>(a, xp, chunks, n_chunks, s);
        case 3: return launch_mfma_pm<PREC, 3>(a, xp, chunks, n_chunks, s);
        case 4: return launch_mfma_pm<PREC, 4>(a, xp, chunks, n_chunks, s);
        case 5: return launch_mfma_pm<PREC, 5>(a, xp, chunks, n_chunks, s);
        case 6: return launch_mfma_pm<PREC, 6>(a, xp, chunks, n_chunks, s);
        case 7: return launch_mfma_pm<PREC, 7>(a, xp, chunks, n_chunks, s);
        case 8: return launch_mfma_pm<PREC, 8>(a, xp, chunks, n_chunks, s);
    }
    set_error("track (mfma): max shift %d outside [1, 8]", M);
    return ISA_ERR_UNSUPPORTED;
}

// Packed (hi | lo << 16) split planes of count pixels (nothing to do for ISA_XC_FP32).
int split_planes(int precision, const float* src, unsigned* dst, size_t count, hipStream_t s) {
    const unsigned blocks = (unsigned)((count + 1023) / 1024);  // four elements per thread
    const bool vec = ((size_t)src & 15) == 0 && ((size_t)dst & 15) == 0;
    if (precision == ISA_XC_BF16X3) k_split_planes<ISA_XC_BF16X3><<<blocks, 256, 0, s>>>(src, dst, count, vec);
    else if (precision == ISA_XC_FP16X3) k_split_planes<ISA_XC_FP16X3><<<blocks, 256, 0, s>>>(src, dst, count, vec);
    else return ISA_OK;
    return check_launch("split_planes");
}

// Vertical T-window energies of a level image (count of window starts: h - T + 1 rows of w).
// runmax: vbox_run_rows(h, T) x w floats.
int vbox_energies(const float* src, int pitch, int w, int h, int T, float* dst, float* runmax, hipStream_t s) {
    if (h < T) return ISA_OK;
    k_vbox_sq<<<dim3(div_up(w, 256), div_up(h - T + 1, kVRun)), 256, 0, s>>>(src, pitch, w, h, T, dst, runmax);
    return check_launch("vbox_energies");
}

int vbox_run_rows(int h, int T) { return h < T ? 0 : div_up(h - T + 1, kVRun); }

bool track_mfma_supported(int tile, int M) {
    const MfmaLayout L = mfma_layout(tile, M);
    return M >= 1 && M <= 8 && L.Bp / 16 <= kMaxNB && L.bytes <= 160 * 1024;
}

// xplanes: per frame, the packed split plane of the level image (same pixel pitch; unused for fp32);
// vbox, vrun: per frame, vbox_energies of the level image (dense, row length w) and its run maxima.
// Items whose integer minimum and threshold test the error bounds cannot decide are appended to
// queue (2 ints each; room for every (pair, tile) of the level) and counted in *count, which the
// caller clears; launch_track_queue then searches them exactly.
int launch_track_mfma(int precision, const float* const* frames, const unsigned* const* xplanes,
                      const float* const* vbox, const float* const* vrun, int img_pitch, int w, int h, int tile, int M,
                      int cx, int cy, const int* pa, const int* pb, const int* pk, const int* chunks, int n_chunks,
                      const float* base, float* shifts, size_t meas_stride, float threshold, float* shift_images,
                      int* decisions, int* queue, int* count, hipStream_t s) {
    if (!track_mfma_supported(tile, M)) {
        set_error("track (mfma): tile %d with max shift %d is not supported (block wider than 80 pixels)", tile, M);
        return ISA_ERR_UNSUPPORTED;
    }
    TrackArgs a{frames, img_pitch, w, h, tile, cx, cy, pa, pb, pk, base, shifts, meas_stride, threshold,
                shift_images, decisions};
    XPlanes xp{xplanes, vbox, vrun, queue, count};
    switch (precision) {
        case ISA_XC_BF16X3: return launch_mfma_p<ISA_XC_BF16X3>(a, xp, M, chunks, n_chunks, s);
        case ISA_XC_FP16X3: return launch_mfma_p<ISA_XC_FP16X3>(a, xp, M, chunks, n_chunks, s);
        case ISA_XC_FP32: return launch_mfma_p<ISA_XC_FP32>(a, xp, M, chunks, n_chunks, s);
    }
    set_error("track (mfma): unknown precision %d", precision);
    return ISA_ERR_ARG;
}

}  // namespace isa
