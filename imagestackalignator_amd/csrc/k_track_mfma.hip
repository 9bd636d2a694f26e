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
//   ISA_XC_FP16X3: the same split in fp16 after scaling the tile by a power of two (the reference
//                  interior's maximum is brought into [0.5, 1); the track blocks use the same scale,
//                  and their squares are scaled once, not twice, so they stay far from fp16's range).
//   ISA_XC_FP32:   fp32 operands (16x16x4 MFMA).
// All accumulate in fp32. The window energies of the track blocks, E[s][j] = sum_u W[s][u] t_j(u)^2
// with W[s][u] = 1 where u - s lies in the interior, are a second product on the same fragments
// (W is exact in every format; t^2 is split like the operands, two products). The reference
// energy sum r^2 is a block sum in FP64. D(s) = sum r^2 + E(s) - 2 C(s), then findMinimum as in
// the exact search.
#include <algorithm>
#include <cmath>

#include "isa_device.h"
#include "isa_host.h"
#include "tile_search.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kPairsPerChunk = 16;  // MFMA columns
constexpr int kWaves = 4;           // waves per block; wave w reduces over track rows uy = w mod 4
constexpr int kMaxNB = 5;           // track-block width up to 80 (16-column blocks)

__device__ __forceinline__ unsigned short bf16_rn(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// hi / lo halves of x as 16-bit patterns
template <int PREC> __device__ __forceinline__ void split16(float x, unsigned short& hi, unsigned short& lo) {
    if (PREC == ISA_XC_BF16X3) {
        hi = bf16_rn(x);
        lo = bf16_rn(x - bf16_to_f(hi));
    } else {
        _Float16 h = (_Float16)x;
        _Float16 l = (_Float16)(x - (float)h);
        hi = __builtin_bit_cast(unsigned short, h);
        lo = __builtin_bit_cast(unsigned short, l);
    }
}

struct MfmaLayout {
    int RW, RH;        // padded reference plane: RH rows x RW columns
    int Bp;            // track-block width rounded up to 16
    size_t off_plane;  // reference planes (hi, lo) or one fp32 plane, plus one zero row
    size_t off_red;    // per-wave partial products [2][kWaves][MB][256] (C, E)
    size_t off_c;      // C, E [2][chunk][S*S]
    size_t off_col;    // block partial sums of r^2 (double)
    size_t off_d;      // D [chunk][S*S]
    size_t off_misc;   // R2 (double), scale
    size_t bytes;
};

__host__ __device__ inline MfmaLayout mfma_layout(int T, int M, int prec) {
    MfmaLayout L;
    const int S = 2 * M + 1, B = T + 2 * M, MB = (S * S + 15) / 16;
    L.Bp = (B + 15) / 16 * 16;
    L.RH = B + S - 1;
    L.RW = L.Bp + S - 1 + 4;  // reads reach column Bp - 1 + S - 1; +4 keeps rows 8-byte aligned
    L.RW = (L.RW + 3) / 4 * 4;
    (void)prec;
    const size_t elem = 4;  // fp32: one 4-byte plane; 16-bit precisions: two 2-byte planes
    auto al = [](size_t v) { return (v + 15) / 16 * 16; };
    size_t o = 0;
    L.off_plane = o;
    o += al((size_t)(L.RH + 1) * L.RW * elem);
    L.off_red = o;
    o += al((size_t)kWaves * MB * 256 * 4 * 2);  // C and E partials
    L.off_c = o;
    o += al((size_t)kPairsPerChunk * S * S * 4 * 2);  // C and E
    L.off_col = o;  // the block's 256 double partial sums of r^2
    o += al((size_t)kWaves * 64 * 8);
    L.off_d = o;
    o += al((size_t)kPairsPerChunk * S * S * 4);
    L.off_misc = o;
    o += 64;
    L.bytes = o;
    return L;
}

// chunk c of the level: up to 16 pair indices sharing one reference frame (-1 = empty slot)
template <int PREC, int M>
__global__ __launch_bounds__(kWaves * 64) void k_track_mfma(TrackArgs a, const int* __restrict__ chunk_pairs,
                                                            int n_chunks, int nblk) {
    constexpr int S = 2 * M + 1, SS = S * S, MB = (SS + 15) / 16;
    extern __shared__ __align__(16) char smem[];
    const int T = a.tile, B = T + 2 * M;
    const MfmaLayout L = mfma_layout(T, M, PREC);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // XCD-aware order (as in the exact search): each XCD gets a contiguous run of (tile, chunk)
    const int per = (int)(gridDim.x / 8);
    const int blk = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (blk >= nblk) return;
    const int tile_idx = blk / n_chunks, chunk = blk - tile_idx * n_chunks;
    const int ty = tile_idx / a.cx, tx = tile_idx - ty * a.cx;
    const int* cp = chunk_pairs + (size_t)chunk * kPairsPerChunk;
    const int p0 = cp[0];
    const int fa = a.pa[p0];
    int rox, roy;
    tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fa * 4], a.base[fa * 4 + 1], a.base[fa * 4 + 2], a.base[fa * 4 + 3],
                   0.0f, 0.0f, rox, roy);
    const int rx0 = tx * T + M + rox, ry0 = ty * T + M + roy;
    const float* ia = a.frames[fa];
    float* misc = (float*)(smem + L.off_misc);
    double* R2s = (double*)misc;
    float* scale_s = misc + 2;
    // ---- 0. power-of-two scale of the tile (fp16 only) and the reference energy sum r^2
    float rmax = 0.0f;
    double r2 = 0.0;
    if (PREC == ISA_XC_FP16X3)
        for (int yb = wave; yb < T; yb += kWaves * 8) {
            float v[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int y = yb + (k >> 1) * kWaves, x = lane + 64 * (k & 1);
                v[k] = (y < T && x < T)
                           ? row_ptr(ia, a.img_pitch, clampi(ry0 + y, 0, a.h - 1))[clampi(rx0 + x, 0, a.w - 1)]
                           : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) rmax = fmaxf(rmax, fabsf(v[k]));
        }
    float* red = (float*)(smem + L.off_red);
    red[tid] = rmax;
    __syncthreads();
    if (tid == 0) {
        float mx = 0.0f;
        for (int i = 0; i < kWaves * 64; ++i) mx = fmaxf(mx, red[i]);
        float sc = 1.0f;
        if (PREC == ISA_XC_FP16X3 && mx > 0.0f && isfinite(mx)) {
            int e;
            frexpf(mx, &e);  // mx = f * 2^e, f in [0.5, 1)
            sc = ldexpf(1.0f, -e);
        }
        *scale_s = sc;
    }
    __syncthreads();
    const float sc = *scale_s;
    // ---- 1. zero-padded reference plane: P[yy][xx] = r(yy - (S-1), xx - (S-1)) inside the interior
    const int RW = L.RW, RH = L.RH;
    unsigned short* ph = (unsigned short*)(smem + L.off_plane);  // hi plane (16-bit precisions)
    unsigned short* pl = ph + (size_t)(RH + 1) * RW;             // lo plane
    float* pf = (float*)(smem + L.off_plane);                    // fp32 plane
    // rows in batches of 8 per wave, all loads of a batch issued before any is used
    constexpr int kFill = 8;
    for (int yb = wave; yb < RH + 1; yb += kWaves * kFill) {
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int xx = lane + 64 * half;
            const int x = xx - (S - 1);
            float v[kFill];
#pragma unroll
            for (int k = 0; k < kFill; ++k) {
                const int yy = yb + k * kWaves, y = yy - (S - 1);
                const bool in = xx < RW && yy < RH && y >= 0 && y < T && x >= 0 && x < T;
                v[k] = in ? row_ptr(ia, a.img_pitch, clampi(ry0 + y, 0, a.h - 1))[clampi(rx0 + x, 0, a.w - 1)] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < kFill; ++k) {
                const int yy = yb + k * kWaves;
                if (xx >= RW || yy >= RH + 1) continue;
                const int i = yy * RW + xx;
                if (PREC == ISA_XC_FP32) {
                    pf[i] = v[k];
                } else {
                    unsigned short hi, lo;
                    split16<PREC>(v[k] * sc, hi, lo);
                    ph[i] = hi;
                    pl[i] = lo;
                }
                r2 = fma((double)v[k], (double)v[k], r2);  // v = 0 outside the interior
            }
        }
    }
    __syncthreads();
    // ---- 2. cross terms on MFMA: lane (col j = lane % 16, quad q = lane / 16)
    const int j = lane & 15, q = lane >> 4;
    const int pj = cp[j];
    int tox = 0, toy = 0;
    if (pj >= 0) {
        const int fb = a.pb[pj];
        const float* sh = a.shifts + (size_t)a.pk[pj] * a.meas_stride + (size_t)tile_idx * 2;
        tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fb * 4], a.base[fb * 4 + 1], a.base[fb * 4 + 2],
                       a.base[fb * 4 + 3], sh[0], sh[1], tox, toy);
    }
    const float* ib = pj >= 0 ? a.frames[a.pb[pj]] : ia;
    const int tx0 = tx * T + tox, ty0 = ty * T + toy;
    // LDS offset of this lane's A operand for each row block: row s = mb*16 + (lane % 16)
    int aoff[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int s = mb * 16 + j;
        if (s < SS) {
            const int iy = s / S, ix = s - iy * S;
            aoff[mb] = (S - 1 - iy) * RW + (S - 1 - ix);
        } else {
            aoff[mb] = -1;  // padding row: reads the zero row
        }
    }
    const int zero_row = RH * RW;
    // interior window of each row's lag, for the energy product: u - s inside [0, T)^2
    int lag_iy[MB], lag_ix[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int s = mb * 16 + j;
        lag_iy[mb] = s < SS ? s / S : -1000000;  // padding rows: never inside
        lag_ix[mb] = s < SS ? s - (s / S) * S : -1000000;
    }
    f32x4 acc[MB], accE[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        acc[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        accE[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    const bool cols_in = tx0 >= 0 && tx0 + L.Bp <= a.w;
    const int NB = L.Bp / 16;  // <= kMaxNB (checked by the launcher)
    // B values of one track row: lane (j, q) needs t_j(uy, 16 nb + 4q + e) (16-bit forms) or
    // t_j(uy, 16 nb + 4 kk + q) (fp32); the next row of the wave is fetched while this one computes
    float cur[kMaxNB][4], nxt[kMaxNB][4];
    auto fetch_row = [&](int uy, float (&dst)[kMaxNB][4]) {
        const float* trow = row_ptr(ib, a.img_pitch, clampi(ty0 + min(uy, B - 1), 0, a.h - 1));
#pragma unroll
        for (int nb = 0; nb < kMaxNB; ++nb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int ux = 16 * nb + (PREC == ISA_XC_FP32 ? 4 * e + q : 4 * q + e);
                const int gx = cols_in ? tx0 + ux : clampi(tx0 + ux, 0, a.w - 1);
                dst[nb][e] = (pj >= 0 && nb < NB) ? trow[cols_in && nb < NB ? gx : clampi(gx, 0, a.w - 1)] : 0.0f;
            }
    };
    fetch_row(wave, nxt);
    for (int uy = wave; uy < B; uy += kWaves) {
#pragma unroll
        for (int nb = 0; nb < kMaxNB; ++nb)
#pragma unroll
            for (int e = 0; e < 4; ++e) cur[nb][e] = nxt[nb][e];
        if (uy + kWaves < B) fetch_row(uy + kWaves, nxt);
#pragma unroll
        for (int nb = 0; nb < kMaxNB; ++nb) {
            if (nb >= NB) break;
            const int ux0 = 16 * nb;
            if (PREC == ISA_XC_FP32) {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int ux = ux0 + 4 * kk + q;
                    const float bv = cur[nb][kk];
                    const float bq = bv * bv;
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb) {
                        const float av = aoff[mb] >= 0 ? pf[aoff[mb] + uy * RW + ux] : pf[zero_row + ux];
                        const int ry = uy - lag_iy[mb], rx = ux - lag_ix[mb];
                        const float wv = (ry >= 0 && ry < T && rx >= 0 && rx < T) ? 1.0f : 0.0f;
                        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[mb], 0, 0, 0);
                        accE[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(wv, bq, accE[mb], 0, 0, 0);
                    }
                }
            } else {
                unsigned short bh[4], bl[4], qh[4], ql[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float bv = cur[nb][e] * sc;
                    split16<PREC>(bv, bh[e], bl[e]);
                    split16<PREC>(cur[nb][e] * bv, qh[e], ql[e]);  // t^2 scaled once: no fp16 overflow
                }
                constexpr unsigned short kOne = PREC == ISA_XC_BF16X3 ? 0x3F80 : 0x3C00;
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    const int base = (aoff[mb] >= 0 ? aoff[mb] + uy * RW : zero_row) + ux0 + 4 * q;
                    const int ry = uy - lag_iy[mb];
                    const bool row_in = ry >= 0 && ry < T;
                    unsigned short ah[4], al[4], wv[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        ah[e] = ph[base + e];
                        al[e] = pl[base + e];
                        const int rx = ux0 + 4 * q + e - lag_ix[mb];
                        wv[e] = (row_in && rx >= 0 && rx < T) ? kOne : (unsigned short)0;
                    }
                    if (PREC == ISA_XC_BF16X3) {
                        bf16x4 Ah = {(short)ah[0], (short)ah[1], (short)ah[2], (short)ah[3]};
                        bf16x4 Al = {(short)al[0], (short)al[1], (short)al[2], (short)al[3]};
                        bf16x4 Bh = {(short)bh[0], (short)bh[1], (short)bh[2], (short)bh[3]};
                        bf16x4 Bl = {(short)bl[0], (short)bl[1], (short)bl[2], (short)bl[3]};
                        bf16x4 W = {(short)wv[0], (short)wv[1], (short)wv[2], (short)wv[3]};
                        bf16x4 Qh = {(short)qh[0], (short)qh[1], (short)qh[2], (short)qh[3]};
                        bf16x4 Ql = {(short)ql[0], (short)ql[1], (short)ql[2], (short)ql[3]};
                        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Al, Bh, acc[mb], 0, 0, 0);
                        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bl, acc[mb], 0, 0, 0);
                        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bh, acc[mb], 0, 0, 0);
                        accE[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W, Ql, accE[mb], 0, 0, 0);
                        accE[mb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W, Qh, accE[mb], 0, 0, 0);
                    } else {
                        f16x4 Ah, Al, Bh, Bl, W, Qh, Ql;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            Ah[e] = __builtin_bit_cast(_Float16, ah[e]);
                            Al[e] = __builtin_bit_cast(_Float16, al[e]);
                            Bh[e] = __builtin_bit_cast(_Float16, bh[e]);
                            Bl[e] = __builtin_bit_cast(_Float16, bl[e]);
                            W[e] = __builtin_bit_cast(_Float16, wv[e]);
                            Qh[e] = __builtin_bit_cast(_Float16, qh[e]);
                            Ql[e] = __builtin_bit_cast(_Float16, ql[e]);
                        }
                        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16f16(Al, Bh, acc[mb], 0, 0, 0);
                        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16f16(Ah, Bl, acc[mb], 0, 0, 0);
                        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x16f16(Ah, Bh, acc[mb], 0, 0, 0);
                        accE[mb] = __builtin_amdgcn_mfma_f32_16x16x16f16(W, Ql, accE[mb], 0, 0, 0);
                        accE[mb] = __builtin_amdgcn_mfma_f32_16x16x16f16(W, Qh, accE[mb], 0, 0, 0);
                    }
                }
            }
        }
    }
    // ---- 3. combine the waves' partial products (fixed order w = 0..3); lane holds D[4q+i][j]
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            red[((size_t)wave * MB + mb) * 256 + (4 * q + i) * 16 + j] = acc[mb][i];
            red[((size_t)(kWaves + wave) * MB + mb) * 256 + (4 * q + i) * 16 + j] = accE[mb][i];
        }
    // reference energy: block sum in a fixed order
    double* r2buf = (double*)(smem + L.off_col);
    r2buf[tid] = r2;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int i = 0; i < kWaves * 64; ++i) t += r2buf[i];
        *R2s = t;
    }
    float* Cm = (float*)(smem + L.off_c);
    float* Em = Cm + kPairsPerChunk * SS;
    const float inv_sc = 1.0f / sc, inv_sc2 = inv_sc * inv_sc;
    for (int i = tid; i < 2 * MB * 256; i += kWaves * 64) {
        const int which = i / (MB * 256), r = i - which * MB * 256;
        const int mb = r >> 8, e = r & 255, row = e >> 4, col = e & 15, s = mb * 16 + row;
        if (s >= SS) continue;
        const size_t w0 = (size_t)which * kWaves;
        float v = red[(w0 * MB + mb) * 256 + e];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) v += red[((w0 + w) * MB + mb) * 256 + e];
        if (which == 0) Cm[col * SS + s] = v * inv_sc2;
        else Em[col * SS + s] = v * inv_sc;
    }
    __syncthreads();
    // ---- 4. D(s) = sum r^2 + E(s) - 2 C(s)
    float* Dm = (float*)(smem + L.off_d);
    const double R2 = *R2s;
    for (int task = tid; task < kPairsPerChunk * SS; task += kWaves * 64) {
        const int jj = task / SS, s = task - jj * SS;
        if (cp[jj] < 0) continue;
        Dm[jj * SS + s] = (float)(R2 + (double)Em[jj * SS + s] - 2.0 * (double)Cm[jj * SS + s]);
    }
    __syncthreads();
    // ---- 5. findMinimum per pair
    if (tid < kPairsPerChunk) {
        const int pp = cp[tid];
        if (pp >= 0) {
            float* sh = a.shifts + (size_t)a.pk[pp] * a.meas_stride + (size_t)tile_idx * 2;
            if (a.shift_images)
                for (int s = 0; s < SS; ++s)
                    a.shift_images[((size_t)pp * a.cx * a.cy + tile_idx) * SS + s] = Dm[tid * SS + s];
            float cx, cy;
            find_min_core(Dm + tid * SS, M, a.threshold, cx, cy);
            const float psx = sh[0], psy = sh[1];
            sh[0] = psx + cx;
            sh[1] = psy + cy;
        }
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
static int launch_mfma_pm(const TrackArgs& a, const int* chunks, int n_chunks, hipStream_t s) {
    const MfmaLayout L = mfma_layout(a.tile, M, PREC);
    const int blocks = a.cx * a.cy * n_chunks;
    k_track_mfma<PREC, M><<<8 * div_up(blocks, 8), kWaves * 64, L.bytes, s>>>(a, chunks, n_chunks, blocks);
    return check_launch("track_mfma");
}

template <int PREC>
static int launch_mfma_p(const TrackArgs& a, int M, const int* chunks, int n_chunks, hipStream_t s) {
    switch (M) {
        case 1: return launch_mfma_pm<PREC, 1>(a, chunks, n_chunks, s);
        case 2: return launch_mfma_pm<PREC, 2>(a, chunks, n_chunks, s);
        case 3: return launch_mfma_pm<PREC, 3>(a, chunks, n_chunks, s);
        case 4: return launch_mfma_pm<PREC, 4>(a, chunks, n_chunks, s);
        case 5: return launch_mfma_pm<PREC, 5>(a, chunks, n_chunks, s);
        case 6: return launch_mfma_pm<PREC, 6>(a, chunks, n_chunks, s);
        case 7: return launch_mfma_pm<PREC, 7>(a, chunks, n_chunks, s);
        case 8: return launch_mfma_pm<PREC, 8>(a, chunks, n_chunks, s);
    }
    set_error("track (mfma): max shift %d outside [1, 8]", M);
    return ISA_ERR_UNSUPPORTED;
}

int launch_track_mfma(int precision, const float* const* frames, int img_pitch, int w, int h, int tile, int M, int cx,
                      int cy, const int* pa, const int* pb, const int* pk, const int* chunks, int n_chunks,
                      const float* base, float* shifts, size_t meas_stride, float threshold, float* shift_images,
                      hipStream_t s) {
    if (mfma_layout(tile, M, precision).Bp / 16 > kMaxNB) {
        set_error("track (mfma): tile %d with max shift %d is wider than 80 pixels", tile, M);
        return ISA_ERR_UNSUPPORTED;
    }
    if (mfma_layout(tile, M, precision).bytes > 160 * 1024) {
        set_error("track (mfma): tile %d with max shift %d needs too much LDS", tile, M);
        return ISA_ERR_UNSUPPORTED;
    }
    TrackArgs a{frames, img_pitch, w, h, tile, cx, cy, pa, pb, pk, base, shifts, meas_stride, threshold,
                shift_images};
    switch (precision) {
        case ISA_XC_BF16X3: return launch_mfma_p<ISA_XC_BF16X3>(a, M, chunks, n_chunks, s);
        case ISA_XC_FP16X3: return launch_mfma_p<ISA_XC_FP16X3>(a, M, chunks, n_chunks, s);
        case ISA_XC_FP32: return launch_mfma_p<ISA_XC_FP32>(a, M, chunks, n_chunks, s);
    }
    set_error("track (mfma): unknown precision %d", precision);
    return ISA_ERR_ARG;
}

}  // namespace isa
