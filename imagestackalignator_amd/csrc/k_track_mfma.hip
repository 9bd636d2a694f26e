// Tile search with the cross terms on the matrix cores (the mixed-precision alignment path).
//
// For a tile and a reference frame a, every pair (a, b) of the level shares the reference interior
// r (T x T). The cross terms of all those pairs are one matrix product
//     C[s][j] = sum_u A[s][u] * Bm[u][j],   A[s][u] = r(u - s) (0 outside the interior),
//                                            Bm[u][j] = t_j(u),  u in the (T+2M)^2 track block,
// with s = (iy, ix) over the (2M+1)^2 lags and j over the pairs (their track blocks sit at each
// pair's own pre-shift). The product runs on 16x16 MFMA tiles: lags on the rows, up to 16 pairs on
// the columns, the track-block pixels as the reduction dimension. Operands:
//   ISA_XC_BF16X3: x = hi + lo with hi = bf16(x), lo = bf16(x - hi); hi.hi + hi.lo + lo.hi on
//                  v_mfma_f32_16x16x32_bf16, whose 32 reduction slots take 16 columns of two track rows;
//   ISA_XC_FP16X3: the same split in fp16 (v_mfma_f32_16x16x32_f16);
//   ISA_XC_FP32:   fp32 operands (v_mfma_f32_16x16x4f32, one track row per step).
// All accumulate in fp32. The track images are split once per level into planes of 16-bit hi and lo
// parts, and the reference interior is staged in LDS as the same two planes, so the product loop only
// loads. The window energies sum r^2 (FP64, from the fp32 pixels) and sum t(. + s)^2 (the vertical
// window energies of the level image, summed and slid along each row in FP64) are formed on the vector
// units, and D(s) = sum r^2 + sum t(. + s)^2 - 2 C(s) feeds findMinimum as in the exact search.
#include <algorithm>
#include <type_traits>
#include <cmath>

#include "isa_device.h"
#include "isa_host.h"
#include "tile_search.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
// 16 and 8 bytes at 2-byte aligned addresses (global and LDS accesses need no natural alignment here)
typedef unsigned u32x4a2 __attribute__((ext_vector_type(4), aligned(2)));
typedef unsigned u32x2a2 __attribute__((ext_vector_type(2), aligned(2)));
typedef unsigned u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float f32x4a4 __attribute__((ext_vector_type(4), aligned(4)));
// global-memory views (plain loads, not flat ones: they wait on the vector-memory counter only)
typedef __attribute__((address_space(1))) const unsigned gu32;
typedef __attribute__((address_space(1))) const float gf32;
typedef __attribute__((address_space(1))) const u32x4a4 gu32x4;
typedef __attribute__((address_space(1))) const f32x4a4 gf32x4;

constexpr int kPairsPerChunk = 16;  // MFMA columns
constexpr int kMaxNB = 5;           // track-block width up to 80 (16-column blocks)
constexpr int kVRun = 64;           // window starts per vertical slide of k_vbox_sq
constexpr int kCB = 64;             // columns per block of the slide maxima

__device__ __forceinline__ unsigned bf16_rn(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_to_f(unsigned h) { return __uint_as_float(h << 16); }

// x = hi + lo (+ the remainder), each part as its 16-bit pattern
template <int PREC> __device__ __forceinline__ void split2(float x, unsigned& hi, unsigned& lo) {
    if (PREC == ISA_XC_BF16X3) {
        hi = bf16_rn(x);
        lo = bf16_rn(x - bf16_to_f(hi));
    } else {
        const _Float16 h = (_Float16)x;
        const _Float16 l = (_Float16)(x - (float)h);
        hi = __builtin_bit_cast(unsigned short, h);
        lo = __builtin_bit_cast(unsigned short, l);
    }
}

template <int PREC> __device__ __forceinline__ f32x4 mfma32(u32x4 a, u32x4 b, f32x4 c) {
    if (PREC == ISA_XC_BF16X3)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                      0, 0);
}

// Packed split planes of a level image (row pitch P floats): word y P + x holds hi | lo << 16 of pixel
// (x, y). Four pixels per thread.
template <int PREC> __global__ void k_split_rows(const float* src, int P, int w, unsigned* dst) {
    const int y = blockIdx.y, x = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (x >= w) return;
    const float* s = src + (size_t)y * P + x;
    unsigned* d = dst + (size_t)y * P + x;
    if (x + 4 <= w) {
        const f32x4a4 v = *(const f32x4a4*)s;
        u32x4a4 o;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            unsigned h, l;
            split2<PREC>(v[e], h, l);
            o[e] = h | (l << 16);
        }
        *(u32x4a4*)d = o;
        return;
    }
    for (int e = 0; x + e < w; ++e) {
        unsigned h, l;
        split2<PREC>(s[e], h, l);
        d[e] = h | (l << 16);
    }
}

// Vertical window energies V(y0, x) = sum_{y = y0}^{y0 + T - 1} src(y, x)^2 for y0 in [0, h - T].
// A block takes kCB columns and a run of kVRun window starts: it loads the kVRun + T - 1 rows the run
// reads into LDS (16-byte loads, all in flight together), then four threads per column take a quarter
// of the run each: a T-term FP64 sum, then slid one row at a time. runmax gets, per run and block of
// kCB columns (one block), the largest energy any of its columns was slid through, rounded up: the
// squares are exact in FP64, so a window's error before its rounding to float is at most
// (T + 2 kVRun) 2^-53 times that maximum. Blocks of 256 threads; T <= kVTMax.
constexpr int kVTMax = 128;
__global__ __launch_bounds__(256) void k_vbox_sq(const float* src, int pitch, int w, int h, int T, float* dst,
                                                 float* runmax) {
    static_assert(kCB == 64 && kVRun == 64, "one column per lane, four quarters of the run per column");
    extern __shared__ __align__(16) float Lds[];  // (kVRun + T - 1) rows of kCB columns
    auto L = reinterpret_cast<float(*)[kCB]>(Lds);
    __shared__ double wmax[4];
    const int tid = threadIdx.x, c = tid & 63, quarter = tid >> 6;
    const int x0 = blockIdx.x * kCB, y0 = blockIdx.y * kVRun, nv = h - T + 1;
    const int rows = min(kVRun + T - 1, h - y0);  // rows the run reads (those past the image are never used)
    const int P = pitch / (int)sizeof(float);
    // rows y0 .. y0 + rows - 1, columns x0 .. x0 + 63: sixteen 16-byte pieces per row
    const bool full = x0 + kCB <= w;
    for (int i = tid; i < rows * 16; i += 256) {
        const int r = i >> 4, k = (i & 15) * 4;
        const float* g = src + (size_t)(y0 + r) * P + x0 + k;
        f32x4a4 v = {0.0f, 0.0f, 0.0f, 0.0f};
        if (full) {
            v = *(const f32x4a4*)g;
        } else {
            for (int e = 0; e < 4; ++e)
                if (x0 + k + e < w) v[e] = g[e];
        }
        *(f32x4a4*)&L[r][k] = v;
    }
    __syncthreads();
    double emax = 0.0;
    const int s0 = y0 + quarter * (kVRun / 4), s1 = min(s0 + kVRun / 4, nv);  // the quarter's window starts
    if (x0 + c < w && s0 < s1) {
        const int l0 = s0 - y0;
        double e = 0.0;
        for (int k = 0; k < T; ++k) {
            const double v = (double)L[l0 + k][c];
            e += v * v;
        }
        for (int y = s0; y < s1; ++y) {
            dst[(size_t)y * w + x0 + c] = (float)e;
            emax = fmax(emax, e);
            if (y + T < h && y + 1 < s1) {
                const double a = (double)L[y - y0][c], b = (double)L[y - y0 + T][c];
                e = (e - a * a) + b * b;
            }
        }
    }
    for (int off = 32; off >= 1; off >>= 1) emax = fmax(emax, __shfl_xor(emax, off));
    if (c == 0) wmax[quarter] = emax;
    __syncthreads();
    if (tid == 0) {
        const double m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        runmax[(size_t)blockIdx.y * ((w + kCB - 1) / kCB) + blockIdx.x] = (float)(m * (1.0 + 0x1p-20));
    }
}

// Waves per block, and waves per SIMD the register allocation must allow (the second launch bound:
// four up to 96 lags, two above).
// eight-wave blocks need enough (row pair, column block) units per wave: tiles below 64 take four
// (measured: levels 2-8 of the full workload 8.8 -> 7.8 ms with four; level 1 7.0 -> 8.6 ms)
__host__ __device__ constexpr int mfma_waves(int T, int M) { return M <= 4 && T >= 64 ? 8 : 4; }

// LDS of one block. Region 0 holds the reference planes until the product loop ends, then the partial
// products of half the waves and, after them, C; once C is formed, TT, the row bounds, D and W take the
// bytes before C. The column energies of the track blocks (and the slide maxima they come with) have a
// region of their own: they are copied in while the products run. The split planes hold pair words:
// word (yy, xx) = element xx in its low half and element xx + 1 in its high half, so that any run of
// elements from any column is read as aligned words.
struct MfmaLayout {
    int S, B, Bp, NB, NP;  // lags per axis, track-block width, its 16-multiple, 16-column blocks, row pairs
    int RH, RW, CO;        // planes: RH + 1 rows x RW words; interior at row S - 1, column CO
    int PE;                // words of one plane
    size_t off_plane;      // hi plane, then lo plane (pair words); the fp32 plane is the first of them
    size_t off_red;        // partial products of waves W/2 .. W-1 [W/2][MB][256] floats
    size_t off_tt;         // window energies TT [chunk][S*S] doubles
    size_t off_eb;         // per (pair, lag row): bound of every window energy of the row, slide-error maximum
    size_t off_d, off_w;   // D, W [chunk][S*S] floats
    size_t off_c;          // C [chunk][S*S] floats
    size_t off_col;        // column energies [chunk][S][Bq] floats, Bq = B rounded up to 4
    size_t off_vr;         // slide maxima of the vertical window energies a lag row reads [chunk][S][4] floats
    size_t bytes;
};

__host__ __device__ inline MfmaLayout mfma_layout(int T, int M, int W) {
    MfmaLayout L;
    const int S = 2 * M + 1, B = T + 2 * M, SS = S * S, MB = (SS + 15) / 16;
    L.S = S;
    L.B = B;
    L.Bp = (B + 15) / 16 * 16;
    L.NB = L.Bp / 16;
    L.NP = (B + 1) / 2;
    L.RH = B + S - 1;
    L.CO = (S - 1 + 7) / 8 * 8;
    L.RW = L.Bp + L.CO;
    L.PE = (L.RH + 1) * L.RW;
    auto al = [](size_t v) { return (v + 15) / 16 * 16; };
    L.off_plane = 0;
    L.off_red = 0;
    size_t o = 0;
    L.off_tt = o;
    o += al((size_t)kPairsPerChunk * SS * 8);
    L.off_eb = o;
    o += al((size_t)2 * kPairsPerChunk * S * 4);
    L.off_d = o;
    o += al((size_t)kPairsPerChunk * SS * 4);
    L.off_w = o;
    o += al((size_t)kPairsPerChunk * SS * 4);
    const size_t red = al((size_t)(W / 2) * MB * 256 * 4);
    L.off_c = o > red ? o : red;
    o = L.off_c + al((size_t)kPairsPerChunk * SS * 4);
    const size_t planes = 2 * al((size_t)L.PE * 4);
    o = o > planes ? o : planes;
    L.off_col = o;
    o += al((size_t)kPairsPerChunk * S * ((B + 3) & ~3) * 4);
    L.off_vr = o;
    o += al((size_t)kPairsPerChunk * S * 4 * 4);
    L.bytes = o;
    return L;
}

struct XPlanes {
    const unsigned* const* frames;  // per frame: packed split planes (k_split_rows), same pitch as the level image
    const float* const* vbox;             // per frame: vertical T-window energies, (h - T + 1) x w
    const float* const* vrun;             // per frame: their slide maxima, runs x ceil(w / kCB)
    int* queue;                           // items left to the exact search: (pair, tile) int pairs
    int* count;                           // number of queued items
};

// hi (sel 0x05040100) or lo (0x07060302) halves of two packed words, as one word of two elements
__device__ __forceinline__ unsigned halves(unsigned w1, unsigned w0, unsigned sel) {
    return __builtin_amdgcn_perm(w1, w0, sel);
}

// chunk c of the level: up to 16 pair indices sharing one reference frame (-1 = empty slot)
template <int PREC, int M, int W>
__global__ __launch_bounds__(W * 64, (M <= 4 ? 4 : 2)) void k_track_mfma(
    TrackArgs a, XPlanes xp, const int* __restrict__ chunk_pairs, int n_chunks, int nblk) {
    constexpr int S = 2 * M + 1, SS = S * S, MB = (SS + 15) / 16, RO = S - 1;
    constexpr int NT = W * 64;
    constexpr bool SPLIT = PREC != ISA_XC_FP32;
    // fp32 summation chains of the split products, as the registers allow: up to 32 lags, a chain per
    // track row pair for the hi.hi products and one chain for the products with a lo part; up to 96
    // lags, a chain per row pair for all three; beyond, one chain for all
    constexpr bool kFresh = MB <= 6;
    constexpr bool kSepLo = MB <= 2;
    constexpr int kRing = 4;  // units (row pair, column block) of track operands in flight per wave
    extern __shared__ __align__(16) char smem[];
    __shared__ double s_r2w[W];
    __shared__ double s_R2;
    __shared__ unsigned s_rmax_edge[kPairsPerChunk];
    __shared__ int s_bx[kPairsPerChunk], s_by[kPairsPerChunk], s_in[kPairsPerChunk];
    __shared__ const float* s_vb[kPairsPerChunk];  // vertical window energies of the pair's block, first row
    __shared__ const float* s_vr[kPairsPerChunk];  // slide maxima of its frame
    __shared__ float* s_sh[kPairsPerChunk];        // the pair's shift of this tile, and its value
    __shared__ float s_ps[kPairsPerChunk][2];
    __shared__ int s_any_edge;                     // some pair's block leaves the image
    __shared__ int q_n, q_base, q_pair[kPairsPerChunk];
    const int T = a.tile;
    const MfmaLayout L = mfma_layout(T, M, W);
    const int B = L.B, RW = L.RW, CO = L.CO, NB = L.NB, NP = L.NP, Bq = (B + 3) & ~3;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // XCD-aware order (as in the exact search): each XCD gets a contiguous run of (tile, chunk)
    const int per = (int)(gridDim.x / 8);
    const int blk = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (blk >= nblk) return;
    const int tile_idx = blk / n_chunks, chunk = blk - tile_idx * n_chunks;
    const int ty = tile_idx / a.cx, tx = tile_idx - ty * a.cx;
    const int* cp = chunk_pairs + (size_t)chunk * kPairsPerChunk;
    const int fa = a.pa[cp[0]];
    int rox, roy;
    tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fa * 4], a.base[fa * 4 + 1], a.base[fa * 4 + 2], a.base[fa * 4 + 3],
                   0.0f, 0.0f, rox, roy);
    const int rx0 = tx * T + M + rox, ry0 = ty * T + M + roy;
    const float* ia = a.frames[fa];
    const size_t P = (size_t)a.img_pitch / sizeof(float);  // pixel pitch, shared by the split planes
    // per pair: origin of its track block and whether the block lies inside the image
    if (tid == 0) s_any_edge = 0;
    __syncthreads();
    if (tid < kPairsPerChunk) {
        const int pp = cp[tid];
        int ox = 0, oy = 0;
        float* sh = nullptr;
        float psx = 0.0f, psy = 0.0f;
        if (pp >= 0) {
            const int fb = a.pb[pp];
            sh = a.shifts + (size_t)a.pk[pp] * a.meas_stride + (size_t)tile_idx * 2;
            psx = sh[0];
            psy = sh[1];
            tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fb * 4], a.base[fb * 4 + 1], a.base[fb * 4 + 2],
                           a.base[fb * 4 + 3], psx, psy, ox, oy);
        }
        const int bx0 = tx * T + ox, by0 = ty * T + oy;
        const bool in = pp >= 0 && bx0 >= 0 && by0 >= 0 && bx0 + B <= a.w && by0 + B <= a.h;
        s_bx[tid] = bx0;
        s_by[tid] = by0;
        s_in[tid] = in;
        s_vb[tid] = in ? xp.vbox[a.pb[pp]] + (size_t)by0 * a.w + bx0 : nullptr;
        s_vr[tid] = in ? xp.vrun[a.pb[pp]] : nullptr;
        s_sh[tid] = sh;
        s_ps[tid][0] = psx;
        s_ps[tid][1] = psy;
        s_rmax_edge[tid] = 0u;
        if (pp >= 0 && !in) s_any_edge = 1;
    }
    if (tid == 0) q_n = 0;
    __syncthreads();
    // ---- 0. column energies of the track blocks inside the image: the S rows of B vertical window
    // energies each, copied into LDS by direct loads (one wave per lag row), and the slide maxima of
    // the column blocks each row reads. They complete while the products run.
    float* colE = (float*)(smem + L.off_col);
    float* vr_s = (float*)(smem + L.off_vr);
    const int ncb = (a.w + kCB - 1) / kCB;
    typedef __attribute__((address_space(3))) void lds_void;
    typedef __attribute__((address_space(1))) const void glob_void;
    {
        // lane l < 16 holds pair l's values; a lag row reads them with lane reads
        const int lp = lane & (kPairsPerChunk - 1);
        const int l_in = s_in[lp], l_bx = s_bx[lp], l_by = s_by[lp];
        const unsigned long long l_vb = (unsigned long long)s_vb[lp], l_vr = (unsigned long long)s_vr[lp];
        for (int r = wave; r < kPairsPerChunk * S; r += W) {
            const int jj = r / S, iy = r - (r / S) * S;
            if (!__builtin_amdgcn_readlane(l_in, jj)) continue;
            const unsigned long long vb = ((unsigned long long)__builtin_amdgcn_readlane((unsigned)(l_vb >> 32), jj) << 32) |
                                          (unsigned)__builtin_amdgcn_readlane((unsigned)l_vb, jj);
            gf32* vr = (gf32*)vb + (size_t)iy * a.w;
            // B / 4 lanes copy 16 bytes each, the last B % 4 columns one by one
            if (lane < B / 4)
                __builtin_amdgcn_global_load_lds((glob_void*)(vr + 4 * lane), (lds_void*)(colE + r * Bq), 16, 0, 0);
            if (lane < (B & 3))
                __builtin_amdgcn_global_load_lds((glob_void*)(vr + (B & ~3) + lane), (lds_void*)(colE + r * Bq + (B & ~3)),
                                                 4, 0, 0);
            const int bx0 = __builtin_amdgcn_readlane(l_bx, jj), y = __builtin_amdgcn_readlane(l_by, jj) + iy;
            const int cb0 = bx0 / kCB, ncr = (bx0 + B - 1) / kCB - cb0 + 1;  // at most 3 (B <= 80)
            const unsigned long long vrr = ((unsigned long long)__builtin_amdgcn_readlane((unsigned)(l_vr >> 32), jj) << 32) |
                                           (unsigned)__builtin_amdgcn_readlane((unsigned)l_vr, jj);
            if (lane < ncr)
                __builtin_amdgcn_global_load_lds((glob_void*)((gf32*)vrr + (size_t)(y / kVRun) * ncb + cb0 + lane),
                                                 (lds_void*)(vr_s + r * 4), 4, 0, 0);
        }
    }
    // ---- 1. zero-padded reference planes P[yy][xx] = r(yy - (S-1), xx - CO), and sum r^2 (called by
    // the product loops once their first track operands are in flight)
    auto fill_planes = [&]() {
    unsigned* hiW = (unsigned*)(smem + L.off_plane);
    unsigned* loW = hiW + L.PE;
    float* pf = (float*)(smem + L.off_plane);
    const bool ref_in = rx0 >= 0 && ry0 >= 0 && rx0 + T <= a.w && ry0 + T <= a.h;
    double r2 = 0.0;
    if (SPLIT && ref_in && (T & 7) == 0) {
        // eight words per item, from elements x .. x + 8 (x = xx - CO): CO and T are multiples of 8, so
        // elements x .. x + 7 lie wholly inside the interior or wholly outside; inside, the split parts
        // come from the frame's packed planes
        gu32* xr = (gu32*)xp.frames[fa];
        const int G = RW / 8, items = (L.RH + 1) * G;
        for (int c = tid; c < items; c += NT) {
            const int yy = c / G, xx = (c - yy * G) * 8, y = yy - RO, x = xx - CO;
            unsigned pk[9];  // packed elements x .. x + 8
#pragma unroll
            for (int k = 0; k < 9; ++k) pk[k] = 0u;
            if (y >= 0 && y < T) {
                gu32* row = xr + (size_t)(ry0 + y) * P + rx0;
                if (x >= 0 && x < T) {
                    const u32x4 u0 = *(gu32x4*)(row + x), u1 = *(gu32x4*)(row + x + 4);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        pk[k] = u0[k];
                        pk[4 + k] = u1[k];
                    }
                    gf32* fr = (gf32*)ia + (size_t)(ry0 + y) * P + rx0 + x;
                    const f32x4 u = *(gf32x4*)fr, v = *(gf32x4*)(fr + 4);
#pragma unroll
                    for (int k = 0; k < 4; ++k) r2 = fma((double)u[k], (double)u[k], r2);
#pragma unroll
                    for (int k = 0; k < 4; ++k) r2 = fma((double)v[k], (double)v[k], r2);
                }
                if (x + 8 >= 0 && x + 8 < T) pk[8] = row[x + 8];
            }
            u32x4 h0, h1, l0, l1;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                h0[k] = __builtin_amdgcn_perm(pk[k + 1], pk[k], 0x05040100u);
                h1[k] = __builtin_amdgcn_perm(pk[k + 5], pk[k + 4], 0x05040100u);
                l0[k] = __builtin_amdgcn_perm(pk[k + 1], pk[k], 0x07060302u);
                l1[k] = __builtin_amdgcn_perm(pk[k + 5], pk[k + 4], 0x07060302u);
            }
            *(u32x4*)(hiW + yy * RW + xx) = h0;
            *(u32x4*)(hiW + yy * RW + xx + 4) = h1;
            *(u32x4*)(loW + yy * RW + xx) = l0;
            *(u32x4*)(loW + yy * RW + xx + 4) = l1;
        }
    } else if (!SPLIT && ref_in && (T & 3) == 0) {
        const int G = RW / 4, items = (L.RH + 1) * G;
        for (int c = tid; c < items; c += NT) {
            const int yy = c / G, xx = (c - yy * G) * 4, y = yy - RO, x = xx - CO;
            f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (y >= 0 && y < T && x >= 0 && x < T) {
                v = *(gf32x4*)((gf32*)ia + (size_t)(ry0 + y) * P + rx0 + x);
#pragma unroll
                for (int k = 0; k < 4; ++k) r2 = fma((double)v[k], (double)v[k], r2);
            }
            *(f32x4*)(pf + yy * RW + xx) = v;
        }
    } else {
        // element by element, the reference's rows and columns clamped to the image
        for (int yy = wave; yy <= L.RH; yy += W) {
            const int y = yy - RO;
            const float* row = row_ptr(ia, a.img_pitch, clampi(ry0 + y, 0, a.h - 1));
            auto val = [&](int xx) {
                const int x = xx - CO;
                return (y >= 0 && y < T && x >= 0 && x < T) ? row[clampi(rx0 + x, 0, a.w - 1)] : 0.0f;
            };
            for (int xx = lane; xx < RW; xx += 64) {
                const float v = val(xx);
                if (SPLIT) {
                    unsigned h0, l0, h1, l1;
                    split2<PREC>(v, h0, l0);
                    split2<PREC>(val(xx + 1), h1, l1);
                    hiW[yy * RW + xx] = h0 | (h1 << 16);
                    loW[yy * RW + xx] = l0 | (l1 << 16);
                } else {
                    pf[yy * RW + xx] = v;
                }
                r2 = fma((double)v, (double)v, r2);  // v = 0 outside the interior
            }
        }
    }
    // block sum of r^2: a butterfly within each wave, then the wave sums in order
    for (int off = 32; off >= 1; off >>= 1) r2 += __shfl_xor(r2, off);
    if (lane == 0) s_r2w[wave] = r2;
    __syncthreads();
    if (tid == 0) {
        double t = s_r2w[0];
        for (int w = 1; w < W; ++w) t += s_r2w[w];
        s_R2 = t;
    }
    };
    const unsigned* hiW = (const unsigned*)(smem + L.off_plane);
    const float* pf = (const float*)(smem + L.off_plane);
    // ---- 2. cross terms on MFMA: lane (column j = lane % 16, quad q = lane / 16)
    const int j = lane & 15, q = lane >> 4;
    const int pj = cp[j];
    const int fbj = pj >= 0 ? a.pb[pj] : fa;
    const int tx0 = s_bx[j], ty0 = s_by[j];
    const bool cols_in = tx0 >= 0 && tx0 + L.Bp <= a.w;
    f32x4 tot[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) tot[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    if (SPLIT) {
        // unit (row pair p, column block nb): reduction slot 8 q + 4 k + e <-> track pixel (row 2 p + k,
        // column 16 nb + 4 q + e), k in {0, 1}, e in 0..3; lane (j, q) holds B[slot][pair j] and
        // A[lag 16 mb + j][slot], so that the four lanes of a pair load 64 contiguous bytes of a packed
        // row per instruction. Padding lags (16 mb + j >= S*S) read lag 0's operands: their rows of C are
        // never used. Wave w takes a contiguous run of the NP * NB units.
        int aoff[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const int s = mb * 16 + j < SS ? mb * 16 + j : 0, iy = s / S, ix = s - (s / S) * S;
            aoff[mb] = (RO - iy) * RW + CO - ix + 4 * q;
        }
        gu32* bpl = (gu32*)xp.frames[fbj];
        const int cx0 = tx0 + 4 * q;
        const int NU = NP * NB, UPW = (NU + W - 1) / W;
        const int u0 = min(wave * UPW, NU), u1 = min(u0 + UPW, NU);
        f32x4 xs[MB], hh[MB];  // products with a lo part; hi.hi products (a chain per row pair)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            xs[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            hh[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        // the loop in two versions: every pair's block columns inside the image (16-byte loads), or
        // some at the edge (clamped columns, element by element); the choice is the same for all waves
        auto products = [&](auto edge_t) {
            constexpr bool EDGE = decltype(edge_t)::value;
            u32x4 r0[kRing], r1[kRing];  // packed words of the unit's two rows
            // operands of unit (p, nb) (zero for an empty pair slot); rows past the block repeat its last
            // row and meet zero rows of A, columns past it meet zero columns of A
            auto fetch = [&](int p, int nb, int slot) {
                u32x4 w0 = {0u, 0u, 0u, 0u}, w1 = {0u, 0u, 0u, 0u};
                if (pj >= 0) {
                    gu32* row0 = bpl + (size_t)clampi(ty0 + min(2 * p, B - 1), 0, a.h - 1) * P;
                    gu32* row1 = bpl + (size_t)clampi(ty0 + min(2 * p + 1, B - 1), 0, a.h - 1) * P;
                    if (!EDGE) {
                        w0 = *(gu32x4*)(row0 + cx0 + 16 * nb);
                        w1 = *(gu32x4*)(row1 + cx0 + 16 * nb);
                    } else {
                        int x0 = cx0 + 16 * nb;
                        asm volatile("" : "+v"(x0));  // keeps the clamped columns from being hoisted and held
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int x = clampi(x0 + e, 0, a.w - 1);
                            w0[e] = row0[x];
                            w1[e] = row1[x];
                        }
                    }
                }
                r0[slot] = w0;
                r1[slot] = w1;
            };
            // (p, nb) of the unit being multiplied and of the next one to fetch
            int fp = u0 / NB, fnb = u0 - fp * NB;
#pragma unroll
            for (int k = 0; k < kRing; ++k)
                if (u0 + k < u1) {
                    fetch(fp, fnb, k);
                    if (++fnb == NB) fnb = 0, ++fp;
                }
            fill_planes();
            int cpr = u0 / NB, cnb = u0 - cpr * NB;
            for (int u = u0; u < u1; u += kRing) {
#pragma unroll
                for (int k = 0; k < kRing; ++k) {
                    if (u + k >= u1) break;
                    const u32x4 w0 = r0[k], w1 = r1[k];
                    const u32x4 ch = {halves(w0.y, w0.x, 0x05040100u), halves(w0.w, w0.z, 0x05040100u),
                                      halves(w1.y, w1.x, 0x05040100u), halves(w1.w, w1.z, 0x05040100u)};
                    const u32x4 cl = {halves(w0.y, w0.x, 0x07060302u), halves(w0.w, w0.z, 0x07060302u),
                                      halves(w1.y, w1.x, 0x07060302u), halves(w1.w, w1.z, 0x07060302u)};
                    if (u + k + kRing < u1) {
                        fetch(fp, fnb, k);
                        if (++fnb == NB) fnb = 0, ++fp;
                    }
                    const unsigned* arow = hiW + 2 * cpr * RW + 16 * cnb;
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb) {
                        const unsigned* ap = arow + aoff[mb];
                        const u32x4 ah = {ap[0], ap[2], ap[RW], ap[RW + 2]};
                        const u32x4 al = {ap[L.PE], ap[L.PE + 2], ap[L.PE + RW], ap[L.PE + RW + 2]};
                        f32x4& hacc = kFresh ? hh[mb] : tot[mb];
                        f32x4& lacc = kSepLo ? xs[mb] : hacc;
                        hacc = mfma32<PREC>(ah, ch, hacc);
                        lacc = mfma32<PREC>(al, ch, lacc);
                        lacc = mfma32<PREC>(ah, cl, lacc);
                    }
                    if (++cnb == NB) {
                        cnb = 0;
                        ++cpr;
                        if (kFresh)
#pragma unroll
                            for (int mb = 0; mb < MB; ++mb) {
                                tot[mb] += hh[mb];
                                hh[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                            }
                    }
                }
            }
        };
        if (__all(cols_in)) products(std::false_type{});
        else products(std::true_type{});
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            if (kFresh) tot[mb] += hh[mb];  // the run's last row pair, when it ends inside one
            if (kSepLo) tot[mb] += xs[mb];
        }
    } else {
        // fp32: one track row per step; product kk takes column 16 nb + 4 q + kk in slot q, of A and B
        int aoff[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const int s = mb * 16 + j < SS ? mb * 16 + j : 0, iy = s / S, ix = s - (s / S) * S;
            aoff[mb] = (RO - iy) * RW + CO - ix + 4 * q;
        }
        gf32* bimg = (gf32*)a.frames[fbj];
        float cur[kMaxNB][4], nxt[kMaxNB][4];
        auto fetch_row = [&](int uy) {
            gf32* row = bimg + (size_t)clampi(ty0 + min(uy, B - 1), 0, a.h - 1) * P;
#pragma unroll
            for (int nb = 0; nb < kMaxNB; ++nb) {
                f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
                if (pj >= 0 && nb < NB) {
                    if (cols_in) {
                        v = *(gf32x4*)(row + tx0 + 16 * nb + 4 * q);
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e) v[e] = row[clampi(tx0 + 16 * nb + 4 * q + e, 0, a.w - 1)];
                    }
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) nxt[nb][e] = v[e];
            }
        };
        if (wave < B) fetch_row(wave);
        fill_planes();
        for (int uy = wave; uy < B; uy += W) {
#pragma unroll
            for (int nb = 0; nb < kMaxNB; ++nb)
#pragma unroll
                for (int e = 0; e < 4; ++e) cur[nb][e] = nxt[nb][e];
            if (uy + W < B) fetch_row(uy + W);
            f32x4 racc[MB];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) racc[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int nb = 0; nb < kMaxNB; ++nb) {
                if (nb >= NB) break;
                f32x4 av[MB];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) av[mb] = *(const f32x4a4*)(pf + aoff[mb] + uy * RW + 16 * nb);
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
                        racc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[mb][kk], cur[nb][kk], racc[mb], 0, 0, 0);
            }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) tot[mb] += racc[mb];
        }
    }
    // ---- 3. combine the waves' partial products in a fixed tree: waves W/2 .. W-1 park theirs in LDS,
    // wave w < W/2 adds wave w + W/2's to its own, and C = ((c0 + c1) + c2) + ... over those W/2 sums;
    // lane (j, q) holds C[4q+i][j] of each row block
    constexpr int HW = W / 2;
    float* red = (float*)(smem + L.off_red);
    float* Cm = (float*)(smem + L.off_c);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's direct loads into LDS have landed
    __syncthreads();  // every wave is done with the reference planes, whose bytes are reused below
    if (wave >= HW)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[((wave - HW) * MB + mb) * 256 + (4 * q + i) * 16 + j] = tot[mb][i];
    __syncthreads();
    if (wave < HW)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float* e = red + (wave * MB + mb) * 256 + (4 * q + i) * 16 + j;
                *e = tot[mb][i] + *e;
            }
    __syncthreads();
    for (int e = tid; e < MB * 256; e += NT) {
        const int s = (e >> 8) * 16 + ((e >> 4) & 15), jc = e & 15;
        float v = red[e];
        for (int w = 1; w < HW; ++w) v += red[w * MB * 256 + e];
        if (s < SS) Cm[jc * SS + s] = v;
    }
    __syncthreads();
    // ---- 4. column energies of the track blocks that leave the image (rows and columns clamped, as
    // the exact search reads them): per column, the S row windows (a T-row FP64 sum, then slid one row
    // at a time), loads issued in batches of 8; s_rmax_edge gets the block's largest column energy.
    if (s_any_edge) {
        for (int task = tid; task < kPairsPerChunk * B; task += NT) {
            const int jj = task / B, c = task - jj * B;
            const int pp = cp[jj];
            if (pp < 0 || s_in[jj]) continue;
            const float* img = a.frames[a.pb[pp]];
            const int gx = clampi(s_bx[jj] + c, 0, a.w - 1), gy0 = s_by[jj];
            auto px = [&](int y) { return (double)row_ptr(img, a.img_pitch, clampi(gy0 + y, 0, a.h - 1))[gx]; };
            double e = 0.0;
            for (int y0 = 0; y0 < T; y0 += 8) {
                double v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = y0 + k < T ? px(y0 + k) : 0.0;
#pragma unroll
                for (int k = 0; k < 8; ++k) e = fma(v[k], v[k], e);
            }
            float* ce = colE + (size_t)jj * S * Bq + c;
            ce[0] = (float)e;
            double eb = e;
            double u[S - 1], v[S - 1];
#pragma unroll
            for (int k = 0; k + 1 < S; ++k) {
                u[k] = px(k);
                v[k] = px(k + T);
            }
#pragma unroll
            for (int k = 0; k + 1 < S; ++k) {
                e = (e - u[k] * u[k]) + v[k] * v[k];
                eb += v[k] * v[k];  // energy of the whole column of the block bounds every window
                ce[(size_t)(k + 1) * Bq] = (float)e;
            }
            atomicMax(&s_rmax_edge[jj], __float_as_uint((float)(eb * (1.0 + 0x1p-20))));  // non-negative floats order as their bits
        }
        __syncthreads();
    }
    // ---- 5. window energies TT(iy, ix) = sum over the T columns ix .. ix + T - 1 of the column energies
    // of lag row iy: four lanes per row sum a quarter of the first window each (FP64), combined as
    // (q0 + q1) + (q2 + q3), then lane 0 slides the window one column at a time.
    double* TT = (double*)(smem + L.off_tt);
    float* EB = (float*)(smem + L.off_eb);  // [pair][iy]: sum of the row's B column energies, rounded up
    float* RM = EB + kPairsPerChunk * S;    // [pair][iy]: slide maximum of its column energies
    {
        const int sub = tid & 3, Q = (T + 3) / 4;
        const int c_beg = min(sub * Q, T), c_end = min(c_beg + Q, T);
        for (int r0 = 0; r0 < kPairsPerChunk * S; r0 += NT / 4) {  // the same trip count in every lane
            const int r = r0 + (tid >> 2);
            const int jj = r / S, iy = r - (r / S) * S;
            const bool on = r < kPairsPerChunk * S && cp[jj] >= 0;
            const float* cr = colE + (size_t)r * Bq;
            double part = 0.0;
            if (on)
                for (int c = c_beg; c < c_end; ++c) part += (double)cr[c];
            part += __shfl_xor(part, 1);  // q0 + q1 (q2 + q3): the same sum in both lanes
            part += __shfl_xor(part, 2);
            if (on && sub == 0) {
                double* tt = TT + jj * SS + iy * S;
                double t = part, rest = 0.0;
                tt[0] = t;
                for (int c = T; c < B; ++c) rest += (double)cr[c];
                for (int ix = 1; ix < S; ++ix) {
                    t = (t - (double)cr[ix - 1]) + (double)cr[ix - 1 + T];
                    tt[ix] = t;
                }
                EB[jj * S + iy] = (float)((part + rest) * (1.0 + 0x1p-20));
                float rm = __uint_as_float(s_rmax_edge[jj]);  // blocks at the edge
                if (s_in[jj]) {
                    const int bx0 = s_bx[jj], ncr = (bx0 + B - 1) / kCB - bx0 / kCB + 1;
                    for (int k = 0; k < ncr; ++k) rm = fmaxf(rm, vr_s[r * 4 + k]);
                }
                RM[jj * S + iy] = rm;
            }
        }
    }
    __syncthreads();
    // ---- 6. D(s) = sum r^2 + TT(s) - 2 C(s), and the half-width W(s) of an interval holding both D(s)
    // and the exact search's value of it:
    //  - cross terms: each product of the split operands differs from r t by at most c_split |r t|
    //    (plus, in fp16, terms from the subnormal spacing). Whatever order and grouping the matrix cores
    //    use inside an instruction, a product passes through at most K roundings there (K products per
    //    instruction) and through every later instruction of its accumulator's chain; the chains of the
    //    hi.hi products restart at every track row pair (row, in fp32) and are added into the wave's
    //    total in fp32, the products with a lo part (at most 2^-8 |r t| in bf16, 2^-11 in fp16) keep one
    //    chain, and the wave totals are added in fp32 (a tree of depth W/2). With h the longest path, the fp32 error is
    //    at most gamma(h) = h u / (1 - h u) (u = 2^-23: every rounding faithful) times the sum of
    //    |products| of the chain; sum |r(u - s) t(u)| <= sqrt(R2 TT(s));
    //  - window energies: rounding of the column energies to float, their FP64 slide, bounded by
    //    (T + 2 kVRun) 2^-53 times the slide maxima, and the FP64 sum and slide along the row, bounded
    //    by (T + 2 S) 2^-53 times the sum of the row's column energies;
    //  - sum r^2: the FP64 sums of the T^2 squares in two orders;
    //  - the FP64 evaluation of R2 + TT - 2C on both sides and the rounding of D to float on both.
    float* Dm = (float*)(smem + L.off_d);
    float* Wm = (float*)(smem + L.off_w);
    const double R2 = s_R2;
    // longest paths of a product through the fp32 roundings: h_main for the hi.hi products (all
    // products when kSepLo is false), h_lo for the products with a lo part in their own chain. A wave's
    // run of UPW units spans at most ceil(UPW / NB) + 1 row pairs; its total then receives the lo chain
    // and passes through at most W/2 additions of the wave combine.
    const int UPW = (NP * NB + W - 1) / W, NRw = (B + W - 1) / W, NRP = (UPW + NB - 1) / NB + 1;
    const int h_main = !SPLIT ? 16 * NB + NRw + HW + 2
                              : (kFresh ? (kSepLo ? 32 : 96) * NB + NRP + HW + 1 : 96 * UPW + HW);
    const int h_lo = 64 * UPW + HW + 1;
    auto gam = [](int hgt) { return (float)(hgt * 0x1p-23 / (1.0 - hgt * 0x1p-23)); };
    const float lo_rel = PREC == ISA_XC_BF16X3 ? 0x1p-8f : 0x1p-11f;
    const float csplit = PREC == ISA_XC_BF16X3 ? 3.1f * 0x1p-16f : (PREC == ISA_XC_FP16X3 ? 3.1f * 0x1p-22f : 0.0f);
    const float eta = 0x1p-24f;  // fp16: bound on the absolute error of a split from the subnormal spacing
    // W is evaluated in fp32: every quantity below is rounded up by a relative 2^-20 or more, which
    // covers the roundings of the fp32 evaluation (a few 2^-24 each)
    const float sr2 = fmaxf((float)(sqrt(R2) * (1.0 + 0x1p-20)), R2 > 0.0 ? 1.17549435e-38f : 0.0f);
    const float r2f = (float)(R2 * (1.0 + 0x1p-20));
    const float kvT = (float)((double)T * (T + 2 * kVRun) * 0x1p-53 * 1.01);
    const float kh = (float)((T + 2 * S) * 0x1p-52);
    const float er2 = (float)((double)T * T * 0x1p-52) * r2f;
    const float crel = (csplit + (SPLIT && !kSepLo ? 1.03f + 2.06f * lo_rel : 1.03f) * gam(h_main) +
                        (SPLIT && kSepLo ? 2.06f * lo_rel * gam(h_lo) : 0.0f)) *
                       (1.0f + 0x1p-18f);
    for (int task = tid; task < kPairsPerChunk * SS; task += NT) {
        const int jj = task / SS, s = task - jj * SS;
        if (cp[jj] < 0) continue;
        const int iy = s / S;
        const double tt = TT[jj * SS + s];
        const double dd = R2 + tt - 2.0 * (double)Cm[jj * SS + s];
        Dm[jj * SS + s] = (float)dd;
        const float ttf = (float)tt * (1.0f + 0x1p-20f), ebf = EB[jj * S + iy], rm = RM[jj * S + iy];
        const float ett = 0x1p-23f * ttf + kvT * rm + kh * ebf;
        const float ttu = ttf + ett;
        const float Pq = sr2 * sqrtf(ttu) * (1.0f + 0x1p-20f);
        float ec = crel * Pq;
        if (PREC == ISA_XC_FP16X3) ec += 1.1f * eta * T * (sr2 + sqrtf(ttu)) + 3.0f * eta * eta * T * T;
        const float wd =
            1.06f * (2.0f * ec + ett + er2 + 0x1p-44f * (r2f + ttf + 2.0f * Pq) + 0x1p-22f * fabsf((float)dd)) + 1e-30f;
        Wm[jj * SS + s] = wd;
    }
    __syncthreads();
    // ---- 7. findMinimum per pair, when the intervals decide the integer minimum and the threshold
    // test as the exact search would; otherwise the item is queued for the exact search. Sixteen lanes
    // per pair (aligned groups within a wave) scan the lags and combine by shuffles.
    // Undecided items are collected per block and appended with one atomic on the global count (one per
    // item serialised the appends on that address).
    {
        static_assert(kPairsPerChunk * 16 <= NT, "sixteen lanes per pair");
        const int jj = (tid >> 4) & (kPairsPerChunk - 1), l = tid & 15;
        const int pp = tid < kPairsPerChunk * 16 ? cp[jj] : -1;
        const float* D = Dm + jj * SS;
        const float* Wd = Wm + jj * SS;
        unsigned long long key = ~0ull;
        float mx = -3.402823466e+38f;
        int fin = 1;
        if (pp >= 0)
            for (int s = l; s < SS; s += 16) {
                const float v = D[s];
                key = min(key, argmin_key(v, s));  // first index of the smallest value, as find_min_core
                mx = fmaxf(mx, v);
                fin &= isfinite(v) && isfinite(Wd[s]);
            }
        for (int o = 8; o >= 1; o >>= 1) {
            key = min(key, (unsigned long long)__shfl_xor((long long)key, o, 16));
            mx = fmaxf(mx, __shfl_xor(mx, o, 16));
            fin &= __shfl_xor(fin, o, 16);
        }
        const int mi = key == ~0ull ? -1 : (int)(unsigned)key;
        const float mn = mi >= 0 ? D[mi] : 3.402823466e+38f;
        const double hi = mi >= 0 ? (double)mn + Wd[mi] : 0.0, lo = mi >= 0 ? (double)mn - Wd[mi] : 0.0;
        // every other lag lies strictly above the minimum's interval
        double lother = 1e300, lmax = -1e300, hmax = -1e300;
        if (pp >= 0 && mi >= 0)
            for (int s = l; s < SS; s += 16) {
                const double lv = (double)D[s] - Wd[s];
                lmax = fmax(lmax, lv);
                hmax = fmax(hmax, (double)D[s] + Wd[s]);
                if (s != mi) lother = fmin(lother, lv);
            }
        for (int o = 8; o >= 1; o >>= 1) {
            lother = fmin(lother, __shfl_xor(lother, o, 16));
            lmax = fmax(lmax, __shfl_xor(lmax, o, 16));
            hmax = fmax(hmax, __shfl_xor(hmax, o, 16));
        }
        if (pp >= 0 && l == 0) {
            float* sh = s_sh[jj];
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
                sh[0] = s_ps[jj][0] + cx;
                sh[1] = s_ps[jj][1] + cy;
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

// LDS the kernel declares statically, next to its dynamic layout
constexpr size_t kMfmaStaticLds = 1024;

template <int PREC, int M, int W>
static int launch_mfma_pmw(const TrackArgs& a, XPlanes xp, const int* chunks, int n_chunks, hipStream_t s) {
    const MfmaLayout L = mfma_layout(a.tile, M, W);
    const int blocks = a.cx * a.cy * n_chunks;
    k_track_mfma<PREC, M, W><<<8 * div_up(blocks, 8), W * 64, L.bytes, s>>>(a, xp, chunks, n_chunks, blocks);
    return check_launch("track_mfma");
}

template <int PREC, int M>
static int launch_mfma_pm(const TrackArgs& a, XPlanes xp, const int* chunks, int n_chunks, hipStream_t s) {
    if constexpr (M <= 4) {
        if (mfma_waves(a.tile, M) == 8) return launch_mfma_pmw<PREC, M, 8>(a, xp, chunks, n_chunks, s);
    }
    return launch_mfma_pmw<PREC, M, 4>(a, xp, chunks, n_chunks, s);
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

// Packed split planes of a w x h level image of row pitch P floats into dst (h x P 32-bit words, word
// y P + x = hi | lo << 16 of pixel (x, y)); nothing to do for ISA_XC_FP32.
int split_planes(int precision, const float* src, int P, int w, int h, unsigned* dst, hipStream_t s) {
    if (w <= 0 || h <= 0) return ISA_OK;
    const dim3 grid((unsigned)div_up(w, 1024), (unsigned)h);  // four pixels per thread
    if (precision == ISA_XC_BF16X3) k_split_rows<ISA_XC_BF16X3><<<grid, 256, 0, s>>>(src, P, w, dst);
    else if (precision == ISA_XC_FP16X3) k_split_rows<ISA_XC_FP16X3><<<grid, 256, 0, s>>>(src, P, w, dst);
    else return ISA_OK;
    return check_launch("split_planes");
}

// Vertical T-window energies of a level image (count of window starts: h - T + 1 rows of w) and their
// slide maxima (vbox_run_rows(h, T) x ceil(w / 64) floats, fewer than the same rows of w).
int vbox_energies(const float* src, int pitch, int w, int h, int T, float* dst, float* runmax, hipStream_t s) {
    if (h < T) return ISA_OK;
    if (T > kVTMax) {
        set_error("track (mfma): tile %d above %d", T, kVTMax);
        return ISA_ERR_UNSUPPORTED;
    }
    const size_t lds = (size_t)(kVRun + T - 1) * kCB * sizeof(float);
    k_vbox_sq<<<dim3(div_up(w, kCB), div_up(h - T + 1, kVRun)), 256, lds, s>>>(src, pitch, w, h, T, dst, runmax);
    return check_launch("vbox_energies");
}

int vbox_run_rows(int h, int T) { return h < T ? 0 : div_up(h - T + 1, kVRun); }

bool track_mfma_supported(int tile, int M) {
    const MfmaLayout L = mfma_layout(tile, M, mfma_waves(tile, M));
    return M >= 1 && M <= 8 && L.Bp / 16 <= kMaxNB && L.bytes + kMfmaStaticLds <= 160 * 1024;
}

// xplanes: per frame, the split planes of the level image (split_planes; unused for fp32);
// vbox, vrun: per frame, vbox_energies of the level image (dense, row length w) and its slide maxima.
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
