// Tile alignment: tile extraction, SSD search, findMinimum, shift up-sampling,
// the fused per-level tracking launch, and the FP64 ShiftMinimizer.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <vector>

#include "isa_device.h"
#include "isa_host.h"
#include "tile_search.h"

namespace {

constexpr int kMaxShift = 8;  // largest supported MaxShift (search window (2M+1)^2 <= 289)

// Integer offset of a tile (R:Kernels/kernel.cu:183-202 and :238-261): round(preshift + R(theta)(-b) + R(theta)c - c).
__device__ __forceinline__ void tile_offset(int tx, int ty, int tile, int w, int h, float bsx, float bsy, float rot,
                                            float psx, float psy, int& ox, int& oy) {
    float sf = sin_r(rot), cf = cos_r(rot);
    float sx = psx + (cf * -bsx - sf * -bsy);
    float sy = psy + (sf * -bsx + cf * -bsy);
    float pcx = (float)(tx * tile + tile / 2 - w / 2);
    float pcy = (float)(ty * tile + tile / 2 - h / 2);
    sx += cf * pcx - sf * pcy - pcx;
    sy += sf * pcx + cf * pcy - pcy;
    ox = (int)roundf(sx);
    oy = (int)roundf(sy);
}

// SSD core. r: T x T reference interior (LDS, row stride RS), t: B x B track block (row stride TS),
// B = T + 2M, S = 2M + 1, lag index i = s + M. Canonical column-wise FP64 order:
//   R2[x]     = sum_y r(y,x)^2                                (fma chain in increasing y)
//   TT[c][0]  = sum_{y<T} t[y][c]^2                          (fma chain in increasing y)
//   TT[c][k+1] = (TT[c][k] - t[k][c]^2) + t[k+T][c]^2         (squares exact in double)
//   C[x][i]   = sum_y r(y,x) t[y+iy][x+ix]                    (fma chain in increasing y)
//   Q[x][i]   = fma(-2, C[x][i], TT[x+ix][iy])
//   D(s)      = (float)(tree_x R2 + tree_x Q[.][i])
// with the pairwise tree: for stride 1, 2, 4, ...: v[x] += v[x + stride] for x a multiple of
// 2 stride. LDS: Q T*S*S doubles, TT B*S doubles, R2 T doubles, D S*S floats.
__device__ __forceinline__ int ssd_rs(int T) { return T + 1; }
__device__ __forceinline__ int ssd_ts(int T, int M) { return T + 2 * M + 1; }

template <int M>
__device__ void ssd_core(const float* r, const float* t, int T, double* Q, double* TT, double* R2, float* D, int tid,
                         int nthreads) {
    constexpr int S = 2 * M + 1;
    const int B = T + 2 * M, RS = ssd_rs(T), TS = ssd_ts(T, M);
    // 1. column energies of the reference, window energies of the track columns
    for (int task = tid; task < B + T; task += nthreads) {
        double acc = 0.0;
        if (task < B) {
            const int c = task;
            for (int y = 0; y < T; ++y) {
                double v = (double)t[y * TS + c];
                acc = fma(v, v, acc);
            }
            TT[c * S] = acc;
            for (int k = 0; k + 1 < S; ++k) {
                double a = (double)t[k * TS + c], b = (double)t[(k + T) * TS + c];
                acc = (acc - a * a) + b * b;
                TT[c * S + k + 1] = acc;
            }
        } else {
            const int x = task - B;
            for (int y = 0; y < T; ++y) {
                double v = (double)r[y * RS + x];
                acc = fma(v, v, acc);
            }
            R2[x] = acc;
        }
    }
    __syncthreads();
    // 2. cross terms, task (x, ix). The S-tall window of track column x+ix stays in registers: row
    // q lives in slot q % S and the y loop is unrolled by S, so the window never moves.
    for (int task = tid; task < T * S; task += nthreads) {
        const int x = task / S, ix = task - x * S;
        const float* tcol = t + x + ix;
        const float* rcol = r + x;
        double acc[S], win[S];
#pragma unroll
        for (int i = 0; i < S; ++i) acc[i] = 0.0;
#pragma unroll
        for (int i = 0; i < S - 1; ++i) win[i] = (double)tcol[i * TS];
        int y = 0;
        for (; y + S <= T; y += S) {
#pragma unroll
            for (int k = 0; k < S; ++k) {
                win[(k + S - 1) % S] = (double)tcol[(y + k + S - 1) * TS];
                const double rv = (double)rcol[(y + k) * RS];
#pragma unroll
                for (int i = 0; i < S; ++i) acc[i] = fma(rv, win[(k + i) % S], acc[i]);
            }
        }
        const int rem = T - y;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (k < rem) {
                win[(k + S - 1) % S] = (double)tcol[(y + k + S - 1) * TS];
                const double rv = (double)rcol[(y + k) * RS];
#pragma unroll
                for (int i = 0; i < S; ++i) acc[i] = fma(rv, win[(k + i) % S], acc[i]);
            }
        }
#pragma unroll
        for (int iy = 0; iy < S; ++iy) Q[(x * S + iy) * S + ix] = fma(-2.0, acc[iy], TT[(x + ix) * S + iy]);
    }
    __syncthreads();
    // 3. pairwise trees over x (Q for every lag, R2 once)
    for (int stride = 1; stride < T; stride *= 2) {
        const int pairs = (T + 2 * stride - 1) / (2 * stride);
        for (int task = tid; task < pairs * (S * S + 1); task += nthreads) {
            int pr = task / (S * S + 1), s = task - pr * (S * S + 1);
            int x = pr * 2 * stride;
            if (x + stride >= T) continue;
            if (s < S * S) Q[x * S * S + s] += Q[(x + stride) * S * S + s];
            else R2[x] += R2[x + stride];
        }
        __syncthreads();
    }
    for (int s = tid; s < S * S; s += nthreads) D[s] = (float)(R2[0] + Q[s]);
    __syncthreads();
}

size_t ssd_smem(int T, int M) {
    int S = 2 * M + 1, B = T + 2 * M;
    size_t f = (size_t)T * (T + 1) + (size_t)B * (B + 1) + (size_t)S * S;  // floats
    size_t d = (size_t)T * S * S + (size_t)B * S + (size_t)T;             // doubles
    return ((f * 4 + 7) & ~(size_t)7) + d * 8;
}

struct SsdLds {
    float *r, *t, *D;
    double *Q, *TT, *R2;
};

__device__ __forceinline__ SsdLds ssd_lds(char* smem, int T, int M) {
    int S = 2 * M + 1, B = T + 2 * M;
    SsdLds L;
    L.r = (float*)smem;
    L.t = L.r + T * (T + 1);
    L.D = L.t + B * (B + 1);
    size_t f = (size_t)T * (T + 1) + (size_t)B * (B + 1) + (size_t)S * S;
    L.Q = (double*)(smem + ((f * 4 + 7) & ~(size_t)7));
    L.TT = L.Q + T * S * S;
    L.R2 = L.TT + B * S;
    return L;
}

// ---------------------------------------------------------------------------
// Fused PatchTracker.Track for every pair of a level at once.
// Level images: frames[f] (device array of pointers), row pitch img_pitch bytes.
// shifts: [meas][tiles] float2 (meas_stride floats per measurement); pair p uses measurement index pk[p].
// One (pair, tile) item searched by a whole block (LDS from ssd_lds).
template <int M>
__device__ void track_item(const TrackArgs& a, int p, int tile_idx, char* smem) {
    const int T = a.tile, B = T + 2 * M;
    constexpr int S = 2 * M + 1;
    SsdLds L = ssd_lds(smem, T, M);
    const int ty = tile_idx / a.cx, tx = tile_idx - ty * a.cx;
    const int fa = a.pa[p], fb = a.pb[p], k = a.pk[p];
    float* sh = a.shifts + (size_t)k * a.meas_stride + (size_t)tile_idx * 2;
    const float psx = sh[0], psy = sh[1];
    int rox, roy, tox, toy;
    tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fa * 4], a.base[fa * 4 + 1], a.base[fa * 4 + 2], a.base[fa * 4 + 3],
                   0.0f, 0.0f, rox, roy);
    tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fb * 4], a.base[fb * 4 + 1], a.base[fb * 4 + 2], a.base[fb * 4 + 3],
                   psx, psy, tox, toy);
    const float* ia = a.frames[fa];
    const float* ib = a.frames[fb];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < T * T; i += nt) {
        int py = i / T, px = i - py * T;
        int gx = clampi(tx * T + px + M + rox, 0, a.w - 1), gy = clampi(ty * T + py + M + roy, 0, a.h - 1);
        L.r[py * (T + 1) + px] = row_ptr(ia, a.img_pitch, gy)[gx];
    }
    for (int i = tid; i < B * B; i += nt) {
        int py = i / B, px = i - py * B;
        int gx = clampi(tx * T + px + tox, 0, a.w - 1), gy = clampi(ty * T + py + toy, 0, a.h - 1);
        L.t[py * (B + 1) + px] = row_ptr(ib, a.img_pitch, gy)[gx];
    }
    __syncthreads();
    ssd_core<M>(L.r, L.t, T, L.Q, L.TT, L.R2, L.D, tid, nt);
    if (a.shift_images)
        for (int s = tid; s < S * S; s += nt) a.shift_images[((size_t)p * a.cx * a.cy + tile_idx) * S * S + s] = L.D[s];
    if (tid == 0) {
        float cx, cy;
        const int dec = find_min_core(L.D, M, a.threshold, cx, cy);
        if (a.decisions) a.decisions[(size_t)p * a.cx * a.cy + tile_idx] = dec;
        sh[0] = psx + cx;
        sh[1] = psy + cy;
    }
}

// grid (pairs, tiles): the pairs of one tile run together and share the tile regions of the frames
// through L2
template <int M>
__global__ void k_track_level(TrackArgs a) {
    extern __shared__ __align__(16) char smem[];
    track_item<M>(a, blockIdx.x, blockIdx.y, smem);
}

// The (pair, tile) items a matrix-core search could not certify (queue[2 i] = pair, queue[2 i + 1] =
// tile, *count items), searched exactly; each block takes items count-strided.
template <int M>
__global__ void k_track_queue(TrackArgs a, const int* __restrict__ queue, const int* __restrict__ count) {
    extern __shared__ __align__(16) char smem[];
    const int n = *count;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        track_item<M>(a, queue[2 * i], queue[2 * i + 1], smem);
        __syncthreads();  // the next item reuses the LDS
    }
}

// Column-parallel form of the same search for tiles with 64 % T == 0 (T >= 4) and M in {1, 2, 4}.
// Lane (g, x) owns reference column x and the lags ix = g*RPG .. g*RPG+RPG-1 (all iy): it keeps
// RPG x S cross-term accumulators and, per ix, the S-tall register window of track column x+ix,
// so one load and one conversion feed S FMAs. Every operand is read straight from the level
// image (a row of the tile is contiguous across lanes), the window energies TT go through LDS,
// and the trees over x run on lane shuffles (a group's T lanes lie within one wave). The order of
// every sum is that of ssd_core, so D is bit-identical to the block-parallel form.
// One block holds PPB (pair, tile) items; items are spread so that each XCD works on a
// contiguous range of tiles.
template <int M> struct ColsPerGroup;
template <> struct ColsPerGroup<1> { static constexpr int v = 3; };
template <> struct ColsPerGroup<2> { static constexpr int v = 5; };
template <> struct ColsPerGroup<4> { static constexpr int v = 3; };

struct ColsShape {
    int lanes_per_item, items_per_block, threads;
};

ColsShape cols_shape(int T, int M) {
    const int S = 2 * M + 1, rpg = M == 2 ? 5 : 3;
    ColsShape c;
    c.lanes_per_item = (S / rpg) * T;
    int g = std::gcd(c.lanes_per_item, 64);
    c.items_per_block = 64 / g;
    c.threads = c.items_per_block * c.lanes_per_item;
    return c;
}

// per item: TT (B S doubles), the reference energy (1 double), D (S S floats, padded to 8 bytes),
// the argmin key (8 bytes) and the max key (4 bytes, padded to 8), then the tree buffer: the T/2
// partial sums over x of every lag (T/2 x S S doubles)
__host__ __device__ inline size_t cols_item_bytes(int T, int M) {
    const int S = 2 * M + 1, B = T + 2 * M;
    return (size_t)B * S * 8 + 8 + (((size_t)S * S * 4 + 7) & ~(size_t)7) + 16 + (size_t)(T / 2) * S * S * 8;
}

size_t cols_smem(int T, int M, int items) { return (size_t)items * cols_item_bytes(T, M); }

// One step of the pairwise tree over x: lane x adds lane x + ST (ST = 1, 2, 4, ... < T). Only the
// lanes with x % (2 ST) == 0 keep a meaningful sum, and they only ever read lanes that kept one at
// every earlier step, so every lane adds (no branch) and the kept sums are those of the branchy
// form. Strides below 16 stay within a 16-lane row (DPP row shift), 16 swaps 32-lane halves' rows
// (ds_swizzle), 32 crosses them (bpermute).
template <int ST>
__device__ __forceinline__ double tree_step(double v) {
    const long long b = __double_as_longlong(v);
    int lo = (int)b, hi = (int)(b >> 32);
    if constexpr (ST < 16) {
        lo = __builtin_amdgcn_mov_dpp(lo, 0x100 + ST, 0xf, 0xf, false);
        hi = __builtin_amdgcn_mov_dpp(hi, 0x100 + ST, 0xf, 0xf, false);
    } else if constexpr (ST == 16) {
        lo = __builtin_amdgcn_ds_swizzle(lo, 0x401f);
        hi = __builtin_amdgcn_ds_swizzle(hi, 0x401f);
    } else {
        lo = __shfl_down(lo, ST);
        hi = __shfl_down(hi, ST);
    }
    return v + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Pairwise sum of N partial sums read from LDS at v[0], v[stride], ...: (v0 + v1) + (v2 + v3), ... by
// halves, the order of the stride-doubling tree.
template <int I0, int NN, int N>
__device__ __forceinline__ double tree_lds_sum(const double (&v)[N]) {
    if constexpr (NN == 1) return v[I0];
    else return tree_lds_sum<I0, NN / 2>(v) + tree_lds_sum<I0 + NN / 2, NN / 2>(v);
}
template <int N>
__device__ __forceinline__ double tree_lds(const double* p, int stride) {
    if constexpr (N > 8) {
        return tree_lds<N / 2>(p, stride) + tree_lds<N / 2>(p + (N / 2) * stride, stride);
    } else {
        double v[N];
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = p[i * stride];
        return tree_lds_sum<0, N>(v);
    }
}

// minimum waves per SIMD the register allocation must allow: M >= 2 runs faster at 3 waves (with a
// few spills outside the row loop) than at 2 with more registers; M = 1 fits 4 as it is
template <int M> struct TrackColsWaves { static constexpr int v = M >= 2 ? 3 : 1; };

// Group blk of the items: item i is (pair i % pairs, tile i / pairs), or, with a queue, (queue[2 i],
// queue[2 i + 1]).
template <int M, bool DEC>
__device__ __forceinline__ void track_cols_group(const TrackArgs& a, int pairs, int items_total, int ppb, int lpi,
                                                 int blk, const int* __restrict__ queue, char* smem) {
    constexpr int S = 2 * M + 1, RPG = ColsPerGroup<M>::v, G = S / RPG;
    static_assert(G * RPG == S, "groups must tile the lags");
    const int T = a.tile, B = T + 2 * M;
    const int tid = threadIdx.x;
    const size_t item_bytes = cols_item_bytes(T, M);
    // per-item lane coordinates
    const int pp = tid / lpi, lane = tid - pp * lpi;
    const int g = lane / T, x = lane - g * T;
    const int item = blk * ppb + pp;
    const bool live = item < items_total;
    const int it = live ? item : items_total - 1;
    // the item's pair and tile, and its shift (re-derived after the row loop rather than held in
    // registers across it)
    auto decode = [&](int& p, int& tile_idx) {
        if (queue) {
            p = queue[2 * it];
            tile_idx = queue[2 * it + 1];
        } else {
            tile_idx = it / pairs;
            p = it - tile_idx * pairs;
        }
    };
    auto shift_of = [&](int p, int tile_idx) {
        return a.shifts + (size_t)a.pk[p] * a.meas_stride + (size_t)tile_idx * 2;
    };
    int tile_idx, p;
    decode(p, tile_idx);
    const int ty = tile_idx / a.cx, tx = tile_idx - ty * a.cx;
    const int fa = a.pa[p], fb = a.pb[p];
    const float* sh0 = shift_of(p, tile_idx);
    int rox, roy, tox, toy;
    tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fa * 4], a.base[fa * 4 + 1], a.base[fa * 4 + 2], a.base[fa * 4 + 3],
                   0.0f, 0.0f, rox, roy);
    tile_offset_sc(tx, ty, T, a.w, a.h, a.base[fb * 4], a.base[fb * 4 + 1], a.base[fb * 4 + 2], a.base[fb * 4 + 3],
                   sh0[0], sh0[1], tox, toy);
    const float* ia = a.frames[fa];
    const float* ib = a.frames[fb];
    char* mine = smem + (size_t)pp * item_bytes;
    double* TT = (double*)mine;
    double* R2tot = TT + B * S;
    float* Dm = (float*)(R2tot + 1);
    unsigned long long* kmin = (unsigned long long*)((char*)Dm + (((size_t)S * S * 4 + 7) & ~(size_t)7));
    unsigned* kmax = (unsigned*)(kmin + 1);
    double* tbuf = (double*)(kmin + 2);
    // block origins in the level image
    const int rx0 = tx * T + M + rox, ry0 = ty * T + M + roy;  // reference interior
    const int tx0 = tx * T + tox, ty0 = ty * T + toy;          // track block
    auto colc = [&](int gx) { return clampi(gx, 0, a.w - 1); };
    // 1. cross terms, reference column energy and the track column energies, in one pass over y.
    // Lane (g, x) keeps, besides its RPG x S cross terms, the energy chains of its first window
    // column x + g*RPG and of its last one x + S - 1 (the latter is needed only from the lanes of
    // the last group with x > T - RPG, which own the track columns no first column reaches).
    const bool inside = rx0 >= 0 && rx0 + T <= a.w && ry0 >= 0 && ry0 + T <= a.h && tx0 >= 0 && tx0 + B <= a.w &&
                        ty0 >= 0 && ty0 + B <= a.h;
    // Addresses: a lane's columns are held as global pointers and a row adds y * pitch elements for a
    // uniform y (a scalar product). Waves whose items all lie inside the image use the unclamped
    // addresses (the RPG track columns are then consecutive); a wave holding a boundary item clamps
    // every column once and every row offset into [0, (h-1) pitch].
    typedef const __attribute__((address_space(1))) float* gptr;
    const int pf = a.img_pitch / (int)sizeof(float), maxoff = (a.h - 1) * pf;
    int roff0 = ry0 * pf, toff0 = ty0 * pf;
    // opaque to the optimiser: the row offset stays roff0 + y * pitch instead of being refactored
    // into a per-lane (ry0 + y) * pitch multiply
    asm("" : "+v"(roff0));
    asm("" : "+v"(toff0));
    double acc[RPG][S], win[RPG][S];
    double r2 = 0.0, tt0 = 0.0, tt1 = 0.0;
#pragma unroll
    for (int q = 0; q < RPG; ++q)
#pragma unroll
        for (int i = 0; i < S; ++i) acc[q][i] = 0.0;
    auto pass = [&](auto clamp_rows) {
        constexpr bool CL = decltype(clamp_rows)::value;
        // rv(y): reference value at row y; tv(j, q): track value at row j, column q of the lane
        gptr rcol, tcol[RPG];
        if (CL) {
            rcol = (gptr)ia + colc(rx0 + x);
#pragma unroll
            for (int q = 0; q < RPG; ++q) tcol[q] = (gptr)ib + colc(tx0 + x + g * RPG + q);
        } else {
            rcol = (gptr)ia + (rx0 + x) + roff0;
            tcol[0] = (gptr)ib + (tx0 + x + g * RPG) + toff0;
#pragma unroll
            for (int q = 1; q < RPG; ++q) tcol[q] = tcol[0] + q;
        }
        auto rv = [&](int y) -> float {
            if (CL) return rcol[(unsigned)min(max(roff0 + y * pf, 0), maxoff)];
            return rcol[y * pf];
        };
        auto trow = [&](int j) -> int {  // element offset of track row j from tcol[q]
            if (CL) return (int)(unsigned)min(max(toff0 + j * pf, 0), maxoff);
            return j * pf;
        };
        auto tv = [&](int o, int q) -> float { return CL ? tcol[q][(unsigned)o] : (tcol[0] + o)[q]; };
#pragma unroll
        for (int i = 0; i < S - 1; ++i) {
            const int o = trow(i);
#pragma unroll
            for (int q = 0; q < RPG; ++q) win[q][i] = (double)tv(o, q);
        }
        // operands of S steps are fetched as one block
        float nr[S], nt[S][RPG];
        auto fetch = [&](int y0s) {
#pragma unroll
            for (int kk = 0; kk < S; ++kk) {
                const int y = min(y0s + kk, T - 1);  // past the end: a valid address, value unused
                nr[kk] = rv(y);
                const int o = trow(y + S - 1);
#pragma unroll
                for (int q = 0; q < RPG; ++q) nt[kk][q] = tv(o, q);
            }
        };
        // (one block ahead only for M = 1; for larger M the prefetch registers would cost occupancy)
        constexpr bool kAhead = M <= 1;
        if (kAhead) fetch(0);
        for (int y = 0; y < T; y += S) {
            float cr[S], ct[S][RPG];
            if (!kAhead) fetch(y);
#pragma unroll
            for (int kk = 0; kk < S; ++kk) {
                cr[kk] = nr[kk];
#pragma unroll
                for (int q = 0; q < RPG; ++q) ct[kk][q] = nt[kk][q];
            }
            if (kAhead && y + S < T) fetch(y + S);
#pragma unroll
            for (int kk = 0; kk < S; ++kk) {
                if (y + kk < T) {
                    const double r = (double)cr[kk];
                    r2 = fma(r, r, r2);
#pragma unroll
                    for (int q = 0; q < RPG; ++q) win[q][(kk + S - 1) % S] = (double)ct[kk][q];
                    tt0 = fma(win[0][kk], win[0][kk], tt0);
                    tt1 = fma(win[RPG - 1][kk], win[RPG - 1][kk], tt1);
#pragma unroll
                    for (int q = 0; q < RPG; ++q)
#pragma unroll
                        for (int i = 0; i < S; ++i) acc[q][i] = fma(r, win[q][(kk + i) % S], acc[q][i]);
                }
            }
        }
        // window energies of the other row offsets: TT[c][k+1] = (TT[c][k] - t[k][c]^2) + t[k+T][c]^2
        const int c0 = x + g * RPG, c1 = x + S - 1;
        const bool own1 = g == G - 1 && x > T - RPG;
        TT[c0 * S] = tt0;
        if (own1) TT[c1 * S] = tt1;
        for (int q = 0; q + 1 < S; ++q) {
            const int oa = trow(q), ob = trow(q + T);
            double u0 = (double)tv(oa, 0), v0 = (double)tv(ob, 0);
            double u1 = (double)tv(oa, RPG - 1), v1 = (double)tv(ob, RPG - 1);
            tt0 = (tt0 - u0 * u0) + v0 * v0;
            tt1 = (tt1 - u1 * u1) + v1 * v1;
            TT[c0 * S + q + 1] = tt0;
            if (own1) TT[c1 * S + q + 1] = tt1;
        }
    };
    if (__all(inside))
        pass(std::false_type{});
    else
        pass(std::true_type{});
    __syncthreads();  // TT complete
#pragma unroll
    for (int q = 0; q < RPG; ++q)
#pragma unroll
        for (int iy = 0; iy < S; ++iy) acc[q][iy] = fma(-2.0, acc[q][iy], TT[(x + g * RPG + q) * S + iy]);
    // 3. pairwise trees over x (the tree of ssd_core: stride 1, 2, 4, ... < T). r2's runs on lane
    // shuffles. For the lags, the first level (stride 1) runs on lane pairs; the T/2 partial sums of each
    // lag then go through LDS, where one lane per lag adds them in the same pairwise order (partial p
    // holds x = 2p, 2p + 1), so a lag's remaining levels cost one lane instead of every lane of the item.
    auto level = [&](auto st) {
        constexpr int ST = decltype(st)::value;
        if (ST >= T) return;
        r2 = tree_step<ST>(r2);
    };
    level(std::integral_constant<int, 1>{});
    level(std::integral_constant<int, 2>{});
    level(std::integral_constant<int, 4>{});
    level(std::integral_constant<int, 8>{});
    level(std::integral_constant<int, 16>{});
    level(std::integral_constant<int, 32>{});
#pragma unroll
    for (int q = 0; q < RPG; ++q)
#pragma unroll
        for (int iy = 0; iy < S; ++iy) acc[q][iy] = tree_step<1>(acc[q][iy]);
    if (x == 0 && g == 0) {
        *R2tot = r2;
        *kmin = ~0ull;
        *kmax = ordered_bits(-3.402823466e+38f);
    }
    // partial p of lag t = iy S + ix at tbuf[p S S + t]
    if ((x & 1) == 0) {
        double* wp = tbuf + (x >> 1) * (S * S) + g * RPG;
#pragma unroll
        for (int iy = 0; iy < S; ++iy)
#pragma unroll
            for (int q = 0; q < RPG; ++q) wp[iy * S + q] = acc[q][iy];
    }
    __syncthreads();  // partials and R2tot complete
    for (int t = lane; t < S * S; t += lpi) {  // one pass unless the item has fewer lanes than lags
        double total;
        switch (T) {
            case 4: total = tree_lds<2>(tbuf + t, S * S); break;
            case 8: total = tree_lds<4>(tbuf + t, S * S); break;
            case 16: total = tree_lds<8>(tbuf + t, S * S); break;
            case 32: total = tree_lds<16>(tbuf + t, S * S); break;
            default: total = tree_lds<32>(tbuf + t, S * S); break;
        }
        Dm[t] = (float)(*R2tot + total);
    }
    __syncthreads();  // Dm complete
    // findMinimum's scan over the S x S values, spread over the item's lanes
    for (int i = lane; i < S * S; i += lpi) {
        const float v = Dm[i];
        atomicMin(kmin, argmin_key(v, i));
        if (!isnan(v)) atomicMax(kmax, ordered_bits(v));
    }
    __syncthreads();
    if (!live) return;
    asm volatile("" ::: "memory");  // the item's pair, tile and shift are loaded again, not kept from before
    decode(p, tile_idx);
    float* sh = shift_of(p, tile_idx);
    if (a.shift_images)
        for (int s = lane; s < S * S; s += lpi) a.shift_images[((size_t)p * a.cx * a.cy + tile_idx) * S * S + s] = Dm[s];
    if (lane == 0) {
        const float psx = sh[0], psy = sh[1];
        const unsigned long long km = *kmin;
        const int mi = km == ~0ull ? -1 : (int)(unsigned)km;
        const float mn = mi < 0 ? 3.402823466e+38f : Dm[mi];
        float cx, cy;
        const int dec = find_min_fit(Dm, M, mi, mn, from_ordered_bits(*kmax), a.threshold, cx, cy);
        if (DEC) a.decisions[(size_t)p * a.cx * a.cy + tile_idx] = dec;
        sh[0] = psx + cx;
        sh[1] = psy + cy;
    }
}

// DEC: write a.decisions (instantiated separately, so the plain search carries no extra registers)
template <int M, bool DEC>
__global__ __launch_bounds__(256, TrackColsWaves<M>::v) void k_track_cols(TrackArgs a, int pairs, int items_total,
                                                                         int ppb, int lpi, int nblk) {
    extern __shared__ __align__(16) char smem[];
    // XCD-aware block order: consecutive hardware block ids go round-robin over 8 XCDs; give each
    // XCD a contiguous run of items so its L2 holds the tile regions its pairs share. The grid has
    // 8 * per blocks, so (id % 8, id / 8) covers every slot of [0, 8 * per); slots >= nblk idle.
    const int per = (int)(gridDim.x / 8);
    const int blk = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (blk >= nblk) return;
    track_cols_group<M, DEC>(a, pairs, items_total, ppb, lpi, blk, nullptr, smem);
}

// The queued items (queue, *count) in groups of ppb; each block takes groups count-strided.
template <int M, bool DEC>
__global__ __launch_bounds__(256, TrackColsWaves<M>::v) void k_track_cols_queue(TrackArgs a, int ppb, int lpi,
                                                                               const int* __restrict__ queue,
                                                                               const int* __restrict__ count) {
    extern __shared__ __align__(16) char smem[];
    const int n = *count, nblk = (n + ppb - 1) / ppb;
    for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        track_cols_group<M, DEC>(a, 0, n, ppb, lpi, blk, queue, smem);
        __syncthreads();  // the next group reuses the LDS
    }
}

// Stand-alone SSD over pre-extracted (T+2M)^2 tiles (isa_tile_ssd).
template <int M>
__global__ void k_tile_ssd(const float* __restrict__ rt, const float* __restrict__ tt, float* out, int T) {
    extern __shared__ __align__(16) char smem[];
    const int B = T + 2 * M;
    constexpr int S = 2 * M + 1;
    SsdLds L = ssd_lds(smem, T, M);
    const int t = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const float* r = rt + (size_t)t * B * B;
    const float* k = tt + (size_t)t * B * B;
    for (int i = tid; i < T * T; i += nt) {
        int py = i / T, px = i - py * T;
        L.r[py * (T + 1) + px] = r[(py + M) * B + px + M];
    }
    for (int i = tid; i < B * B; i += nt) {
        int py = i / B, px = i - py * B;
        L.t[py * (B + 1) + px] = k[i];
    }
    __syncthreads();
    ssd_core<M>(L.r, L.t, T, L.Q, L.TT, L.R2, L.D, tid, nt);
    for (int s = tid; s < S * S; s += nt) out[(size_t)t * S * S + s] = L.D[s];
}

int block_threads(int T, int M) {
    int work = T * (2 * M + 1);
    int bt = ((work + 63) / 64) * 64;
    return std::max(64, std::min(bt, 512));
}

__global__ void k_tiles_border(const float* img, int pitch, float* tiles, int w, int h, int M, int T, int cx, int cy,
                               float bsx, float bsy, float rot) {
    int B = T + 2 * M;
    int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y;
    int t = blockIdx.z * blockDim.z + threadIdx.z;
    if (px >= B || py >= B || t >= cx * cy) return;
    size_t o = (size_t)t * B * B + py * B + px;
    if (px < M || py < M || px >= T + M || py >= T + M) {
        tiles[o] = 0.0f;
        return;
    }
    int ty = t / cx, tx = t - ty * cx, ox, oy;
    tile_offset(tx, ty, T, w, h, bsx, bsy, rot, 0.0f, 0.0f, ox, oy);
    tiles[o] = row_ptr(img, pitch, clampi(ty * T + py + oy, 0, h - 1))[clampi(tx * T + px + ox, 0, w - 1)];
}

__global__ void k_tiles_preshift(const float* img, int pitch, float* tiles, const float* pre, int pre_pitch, int w,
                                 int h, int M, int T, int cx, int cy, float bsx, float bsy, float rot) {
    int B = T + 2 * M;
    int px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y * blockDim.y + threadIdx.y;
    int t = blockIdx.z * blockDim.z + threadIdx.z;
    if (px >= B || py >= B || t >= cx * cy) return;
    int ty = t / cx, tx = t - ty * cx, ox, oy;
    const float* ps = row_ptr(pre, pre_pitch, ty) + tx * 2;
    tile_offset(tx, ty, T, w, h, bsx, bsy, rot, ps[0], ps[1], ox, oy);
    tiles[(size_t)t * B * B + py * B + px] =
        row_ptr(img, pitch, clampi(ty * T + py + oy, 0, h - 1))[clampi(tx * T + px + ox, 0, w - 1)];
}

__global__ void k_find_minimum(const float* simg, float* coords, int pitch, int M, int cx, int cy, float thr) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cx * cy) return;
    int S = 2 * M + 1;
    float x, y;
    find_min_core(simg + (size_t)t * S * S, M, thr, x, y);
    int ty = t / cx, tx = t - ty * cx;
    float* o = row_ptr(coords, pitch, ty) + tx * 2;
    o[0] = x;
    o[1] = y;
}

__global__ void k_add_shifts(float* pre, int pre_pitch, const float* c, int c_pitch, int cx, int cy) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= cx * 2 || y >= cy) return;
    row_ptr(pre, pre_pitch, y)[x] += row_ptr(c, c_pitch, y)[x];
}

// UpSampleShifts (R:Kernels/kernel.cu:551-598).
__global__ void k_upsample(const float* in, float* out, int in_pitch, int out_pitch, int old_level, int new_level,
                           int ocx, int ocy, int ncx, int ncy, int old_tile, int new_tile) {
    int bx = blockIdx.x * blockDim.x + threadIdx.x, by = blockIdx.y * blockDim.y + threadIdx.y;
    if (bx >= ncx || by >= ncy) return;
    float factor = (float)old_level * (float)old_tile / (float)(new_level * new_tile);
    float ox = (float)bx / factor, oy = (float)by / factor;
    int x0 = (int)floorf(ox), x1 = (int)ceilf(ox), y0 = (int)floorf(oy), y1 = (int)ceilf(oy);
    x0 = min(x0, ocx - 1); x1 = min(x1, ocx - 1);
    y0 = min(y0, ocy - 1); y1 = min(y1, ocy - 1);
    const float* r0 = row_ptr(in, in_pitch, y0);
    const float* r1 = row_ptr(in, in_pitch, y1);
    float* o = row_ptr(out, out_pitch, by) + bx * 2;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        float mm = r0[x0 * 2 + c], pm = r0[x1 * 2 + c], mp = r1[x0 * 2 + c], pp = r1[x1 * 2 + c];
        float t1 = mm + (pm - mm) * (1.0f - ((float)x1 - ox));
        float t2 = mp + (pp - mp) * (1.0f - ((float)x1 - ox));
        float v = t1 + (t2 - t1) * (1.0f - ((float)y1 - oy));
        v *= (float)old_level / (float)new_level;
        o[c] = v;
    }
}

// All measurements at once: in/out are [meas][tiles] float2 with meas_stride floats per measurement
// (dense rows of old / new count). Same arithmetic as k_upsample.
__global__ void k_upsample_all(const float* in, float* out, size_t meas_stride, int old_level, int new_level, int ocx,
                               int ocy, int ncx, int ncy, int old_tile, int new_tile) {
    int bx = blockIdx.x * blockDim.x + threadIdx.x, by = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (bx >= ncx || by >= ncy) return;
    const float* i0 = in + (size_t)k * meas_stride;
    float* o0 = out + (size_t)k * meas_stride;
    float factor = (float)old_level * (float)old_tile / (float)(new_level * new_tile);
    float ox = (float)bx / factor, oy = (float)by / factor;
    int x0 = (int)floorf(ox), x1 = (int)ceilf(ox), y0 = (int)floorf(oy), y1 = (int)ceilf(oy);
    x0 = min(x0, ocx - 1); x1 = min(x1, ocx - 1);
    y0 = min(y0, ocy - 1); y1 = min(y1, ocy - 1);
    const float* r0 = i0 + (size_t)y0 * ocx * 2;
    const float* r1 = i0 + (size_t)y1 * ocx * 2;
    float* o = o0 + ((size_t)by * ncx + bx) * 2;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        float mm = r0[x0 * 2 + c], pm = r0[x1 * 2 + c], mp = r1[x0 * 2 + c], pp = r1[x1 * 2 + c];
        float t1 = mm + (pm - mm) * (1.0f - ((float)x1 - ox));
        float t2 = mp + (pp - mp) * (1.0f - ((float)x1 - ox));
        float v = t1 + (t2 - t1) * (1.0f - ((float)y1 - oy));
        v *= (float)old_level / (float)new_level;
        o[c] = v;
    }
}

// ---------------------------------------------------------------------------
// ShiftMinimizer (R:PEFStudioDX/ShiftCollection.cs:738-822) in FP64, one 64-lane workgroup per tile.
// Measurement row k estimates x_b - x_a (x_j = position of frame j, x_0 = 0). Minimising
// sum_k |x_b - x_a - d_k|^2 over the active rows gives the same solution as the reference's
// r = (A^T A)^-1 A^T d (x = prefix sums of r); the normal matrix is the reduced graph Laplacian,
// solved by Cholesky. Singular <=> the active measurement graph is disconnected (status -1, r = 0).
constexpr int kMinFrames = 64;
constexpr int kMinThreads = 64;

struct MinArgs {
    float* meas;  // [m][tiles] float2, meas_stride floats per measurement
    size_t meas_stride;
    float* one_to_one;  // [tile][n1] float2
    int* status;
    int* iters;
    const int* rows;  // [m][2] (a, b)
    const double* lfull;  // Cholesky factor of the all-rows reduced Laplacian (n1 x n1), or null
    int tiles, n, m;
};

__global__ __launch_bounds__(kMinThreads) void k_minimize(MinArgs a) {
    extern __shared__ __align__(16) char smem[];
    const int n = a.n, n1 = n - 1, m = a.m, tid = threadIdx.x;
    double* d = (double*)smem;          // m x 2
    double* L = d + 2 * m;              // n1 x n1
    double* rhs = L + n1 * n1;          // n1 x 2
    double* x = rhs + 2 * n1;           // n x 2
    double* red_v = x + 2 * n;          // kMinThreads
    int* red_i = (int*)(red_v + kMinThreads);  // kMinThreads
    int* list = red_i + kMinThreads;    // n1 x (n - 1): rows of each reduced node
    char* active = (char*)(list + n1 * (n - 1));
    __shared__ unsigned long long s_nb[kMinFrames];
    __shared__ int s_flag;
    static_assert(kMinThreads == 64 && kMinFrames <= 64, "one wave, one lane per frame");
    const int t = blockIdx.x;
    for (int k = tid; k < m; k += kMinThreads) {
        const float* v = a.meas + (size_t)k * a.meas_stride + (size_t)t * 2;
        d[2 * k] = (double)v[0];
        d[2 * k + 1] = (double)v[1];
        active[k] = 1;
    }
    // rows of node tid + 1 in increasing k: 2 k when it is the row's second frame, 2 k + 1 when the first
    // (the rows are distinct frame pairs: at most n - 1 per node)
    int cnt = 0;
    if (tid < n1)
        for (int k = 0; k < m; ++k) {
            const int ra = a.rows[2 * k], rb = a.rows[2 * k + 1];
            const int side = rb == tid + 1 ? 0 : (ra == tid + 1 ? 1 : -1);
            if (side >= 0 && cnt < n - 1) list[tid * (n - 1) + cnt++] = 2 * k + side;
        }
    __syncthreads();
    int status = 0, it;
    for (it = 0; it < 10; ++it) {
        // every tile starts with all rows active: the graph, its Laplacian and the Laplacian's Cholesky
        // factor are the same for all tiles and come precomputed (lfull is given only for a connected
        // graph); the right-hand side is still the tile's own
        const bool pre = it == 0 && a.lfull != nullptr;
        // connectivity of the active measurement graph: per-node neighbour bit masks (LDS atomics over
        // the rows), then a breadth-first sweep from node 0 with wave-wide ORs (lane i = node i)
        if (pre) {
            if (tid == 0) s_flag = 1;
        } else {
            if (tid < kMinFrames) s_nb[tid] = 0ull;
            __syncthreads();
            for (int k = tid; k < m; k += kMinThreads) {
                if (!active[k]) continue;
                const int u = a.rows[2 * k], v = a.rows[2 * k + 1];
                atomicOr(&s_nb[u], 1ull << v);
                atomicOr(&s_nb[v], 1ull << u);
            }
            __syncthreads();
            const unsigned long long nb = tid < n ? s_nb[tid] : 0ull;
            unsigned long long reached = 1ull;
            for (int step = 0; step < n; ++step) {
                unsigned long long add = ((reached >> tid) & 1ull) ? nb : 0ull;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) add |= __shfl_xor(add, off);
                const unsigned long long nr = reached | add;
                if (nr == reached) break;
                reached = nr;
            }
            const unsigned long long all = n == 64 ? ~0ull : ((1ull << n) - 1ull);
            if (tid == 0) s_flag = reached == all;
        }
        __syncthreads();
        if (!s_flag) {
            status = -1;
            for (int i = tid; i < 2 * n; i += kMinThreads) x[i] = 0.0;
            __syncthreads();
            break;
        }
        // reduced Laplacian (nodes 1..n-1) and right-hand side; lane i owns row i. The Laplacian comes from
        // the neighbour masks of the connectivity test (an off-diagonal entry is -1 or 0, a diagonal one the
        // node's degree), the right-hand side from the node's own active rows in increasing order.
        if (pre)
            for (int i = tid; i < n1 * n1; i += kMinThreads) L[i] = a.lfull[i];
        if (tid < n1) {
            if (!pre) {
                const unsigned long long nbi = s_nb[tid + 1];
                for (int j = 0; j < n1; ++j)
                    L[tid * n1 + j] = j == tid ? (double)__popcll(nbi) : -(double)((nbi >> (j + 1)) & 1ull);
            }
            double bx = 0.0, by = 0.0;
            for (int e = 0; e < cnt; ++e) {
                const int v = list[tid * (n - 1) + e], k = v >> 1;
                if (!active[k]) continue;
                if (v & 1) {
                    bx -= d[2 * k];
                    by -= d[2 * k + 1];
                } else {
                    bx += d[2 * k];
                    by += d[2 * k + 1];
                }
            }
            rhs[2 * tid] = bx;
            rhs[2 * tid + 1] = by;
        }
        __syncthreads();
        // Cholesky, lower triangle in place (the same steps as minimize_full_factor on the host)
        for (int j = 0; j < n1 && !pre; ++j) {
            if (tid == 0) L[j * n1 + j] = sqrt(L[j * n1 + j]);
            __syncthreads();
            const double djj = L[j * n1 + j];
            for (int i = j + 1 + tid; i < n1; i += kMinThreads) L[i * n1 + j] /= djj;
            __syncthreads();
            for (int i = j + 1 + tid; i < n1; i += kMinThreads) {
                const double lij = L[i * n1 + j];
                for (int c = j + 1; c <= i; ++c) L[i * n1 + c] -= lij * L[c * n1 + j];
            }
            __syncthreads();
        }
        // forward substitution L y = rhs, one column at a time: y_i, then every later row subtracts
        // L_ri y_i (each row sees its subtractions in increasing i, as a row-by-row solve)
        for (int i = 0; i < n1; ++i) {
            if (tid < 2) rhs[2 * i + tid] = rhs[2 * i + tid] / L[i * n1 + i];
            __syncthreads();
            for (int e = tid; e < 2 * (n1 - 1 - i); e += kMinThreads) {
                const int r = i + 1 + (e >> 1), c = e & 1;
                rhs[2 * r + c] -= L[r * n1 + i] * rhs[2 * i + c];
            }
            __syncthreads();
        }
        // backward substitution L^T x = y, one column at a time from the last
        for (int i = n1 - 1; i >= 0; --i) {
            if (tid < 2) rhs[2 * i + tid] = rhs[2 * i + tid] / L[i * n1 + i];
            __syncthreads();
            for (int e = tid; e < 2 * i; e += kMinThreads) {
                const int r = e >> 1, c = e & 1;
                rhs[2 * r + c] -= L[i * n1 + r] * rhs[2 * i + c];
            }
            __syncthreads();
        }
        if (tid < 2) x[tid] = 0.0;
        for (int e = tid; e < 2 * n1; e += kMinThreads) x[2 + e] = rhs[e];
        __syncthreads();
        // largest residual above 1 px^2 (first index on ties, as the sequential scan)
        double best = 1.0;
        int bi = -1;
        for (int k = tid; k < m; k += kMinThreads) {
            if (!active[k]) continue;  // a removed row has d = 0 and a zero row of A: residual 0
            int ra = a.rows[2 * k], rb = a.rows[2 * k + 1];
            double ex = d[2 * k] - (x[2 * rb] - x[2 * ra]), ey = d[2 * k + 1] - (x[2 * rb + 1] - x[2 * ra + 1]);
            double dist = ex * ex + ey * ey;
            if (dist > best) { best = dist; bi = k; }
        }
        // wave-wide arg-max (largest residual; the smallest row index among equal residuals)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(bi, off);
            if (oi >= 0 && (bi < 0 || ob > best || (ob == best && oi < bi))) {
                best = ob;
                bi = oi;
            }
        }
        if (tid == 0) {
            if (bi >= 0) {
                d[2 * bi] = 0.0;
                d[2 * bi + 1] = 0.0;
                active[bi] = 0;
            }
            s_flag = bi;
        }
        __syncthreads();
        status = s_flag;
        __syncthreads();  // every lane has read s_flag before lane 0 reuses it
        if (status < 0) break;
    }
    // outputs: r_i = x_{i+1} - x_i; every measurement := x_b - x_a (A_safe r)
    for (int i = tid; i < n1; i += kMinThreads) {
        float* o = a.one_to_one + ((size_t)t * n1 + i) * 2;
        o[0] = (float)(x[2 * (i + 1)] - x[2 * i]);
        o[1] = (float)(x[2 * (i + 1) + 1] - x[2 * i + 1]);
    }
    for (int k = tid; k < m; k += kMinThreads) {
        int ra = a.rows[2 * k], rb = a.rows[2 * k + 1];
        float* v = a.meas + (size_t)k * a.meas_stride + (size_t)t * 2;
        v[0] = (float)(x[2 * rb] - x[2 * ra]);
        v[1] = (float)(x[2 * rb + 1] - x[2 * ra + 1]);
    }
    if (tid == 0) {
        if (a.status) a.status[t] = status;
        if (a.iters) a.iters[t] = it < 10 ? it + 1 : 10;
    }
}

// The same search with the reduced Laplacian in registers, for n - 1 <= N1MAX: lane i owns row i of
// L (Lr[j] = L[i][j]) and the part of column i below the diagonal (Lc[r] = L[r][i]); rows and columns
// are exchanged by lane broadcasts, so the factorisation and the two substitutions have no LDS round
// trips. The Laplacian comes from the neighbour bit masks of the connectivity test (the rows are
// distinct frame pairs, so an off-diagonal entry is -1 or 0 and a diagonal one the node's degree), and
// each node's right-hand side from its own rows in increasing order. Every floating-point operation
// is that of k_minimize, in the same order, so the results are identical.
__device__ __forceinline__ double bcast_lane(double v, int lane) { return __shfl(v, lane); }

template <int N1MAX>
__global__ __launch_bounds__(kMinThreads) void k_minimize_reg(MinArgs a) {
    extern __shared__ __align__(16) char smem[];
    const int n = a.n, n1 = n - 1, m = a.m, tid = threadIdx.x;
    double* d = (double*)smem;          // m x 2
    double* xs = d + 2 * m;             // n x 2
    int* rows = (int*)(xs + 2 * n);     // m x 2
    int* list = rows + 2 * m;           // n1 x (n - 1)
    char* active = (char*)(list + n1 * (n - 1));
    __shared__ unsigned long long s_nb[kMinFrames];
    __shared__ int s_flag;
    static_assert(kMinThreads == 64 && N1MAX < 64, "one wave, one lane per node");
    const int t = blockIdx.x;
    for (int k = tid; k < m; k += kMinThreads) {
        const float* v = a.meas + (size_t)k * a.meas_stride + (size_t)t * 2;
        d[2 * k] = (double)v[0];
        d[2 * k + 1] = (double)v[1];
        active[k] = 1;
        rows[2 * k] = a.rows[2 * k];
        rows[2 * k + 1] = a.rows[2 * k + 1];
    }
    __syncthreads();
    // rows of node tid + 1 in increasing k: 2 k when it is the row's second frame, 2 k + 1 when the first
    int cnt = 0;
    if (tid < n1)
        for (int k = 0; k < m; ++k) {
            const int ra = rows[2 * k], rb = rows[2 * k + 1];
            const int side = rb == tid + 1 ? 0 : (ra == tid + 1 ? 1 : -1);
            if (side >= 0 && cnt < n - 1) list[tid * (n - 1) + cnt++] = 2 * k + side;
        }
    double Lr[N1MAX], Lc[N1MAX];
    double bx = 0.0, by = 0.0;
    int status = 0, it;
    for (it = 0; it < 10; ++it) {
        const bool pre = it == 0 && a.lfull != nullptr;
        if (pre) {
            if (tid == 0) s_flag = 1;
        } else {
            if (tid < kMinFrames) s_nb[tid] = 0ull;
            __syncthreads();
            for (int k = tid; k < m; k += kMinThreads) {
                if (!active[k]) continue;
                const int u = rows[2 * k], v = rows[2 * k + 1];
                atomicOr(&s_nb[u], 1ull << v);
                atomicOr(&s_nb[v], 1ull << u);
            }
            __syncthreads();
            const unsigned long long nb = tid < n ? s_nb[tid] : 0ull;
            unsigned long long reached = 1ull;
            for (int step = 0; step < n; ++step) {
                unsigned long long add = ((reached >> tid) & 1ull) ? nb : 0ull;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) add |= __shfl_xor(add, off);
                const unsigned long long nr = reached | add;
                if (nr == reached) break;
                reached = nr;
            }
            const unsigned long long all = n == 64 ? ~0ull : ((1ull << n) - 1ull);
            if (tid == 0) s_flag = reached == all;
        }
        __syncthreads();
        if (!s_flag) {
            status = -1;
            for (int i = tid; i < 2 * n; i += kMinThreads) xs[i] = 0.0;
            __syncthreads();
            break;
        }
        // L (lane tid: reduced node tid, frame tid + 1) and the right-hand side
        if (pre) {
#pragma unroll
            for (int j = 0; j < N1MAX; ++j) {
                Lr[j] = (tid < n1 && j <= tid) ? a.lfull[tid * n1 + j] : 0.0;
                Lc[j] = (tid < n1 && j > tid && j < n1) ? a.lfull[j * n1 + tid] : 0.0;
            }
        } else {
            const unsigned long long nbi = tid < n1 ? s_nb[tid + 1] : 0ull;
#pragma unroll
            for (int j = 0; j < N1MAX; ++j)
                Lr[j] = j == tid ? (double)__popcll(nbi) : -(double)((nbi >> (j + 1)) & 1ull);
        }
        bx = 0.0;
        by = 0.0;
        for (int e = 0; e < cnt; ++e) {
            const int v = list[tid * (n - 1) + e], k = v >> 1;
            if (!active[k]) continue;
            if (v & 1) {
                bx -= d[2 * k];
                by -= d[2 * k + 1];
            } else {
                bx += d[2 * k];
                by += d[2 * k + 1];
            }
        }
        // Cholesky, lower triangle (step j: the diagonal, column j divided, then the trailing rows); the
        // loops run to N1MAX with their bodies guarded, so every register index is a constant
        if (!pre)
#pragma unroll
            for (int j = 0; j < N1MAX; ++j) {
                if (j < n1) {
                    if (tid == j) Lr[j] = sqrt(Lr[j]);
                    const double djj = bcast_lane(Lr[j], j);
                    if (tid > j && tid < n1) Lr[j] /= djj;
#pragma unroll
                    for (int c = j + 1; c < N1MAX; ++c) {
                        if (c < n1) {
                            const double lcj = bcast_lane(Lr[j], c);  // L[c][j]
                            if (tid == j) Lc[c] = lcj;
                            if (tid >= c && tid < n1) Lr[c] -= Lr[j] * lcj;
                        }
                    }
                }
            }
        // forward substitution L y = rhs
#pragma unroll
        for (int i = 0; i < N1MAX; ++i) {
            if (i < n1) {
                if (tid == i) {
                    bx = bx / Lr[i];
                    by = by / Lr[i];
                }
                const double yx = bcast_lane(bx, i), yy = bcast_lane(by, i);
                if (tid > i && tid < n1) {
                    bx -= Lr[i] * yx;
                    by -= Lr[i] * yy;
                }
            }
        }
        // backward substitution L^T x = y
#pragma unroll
        for (int i = N1MAX - 1; i >= 0; --i) {
            if (i < n1) {
                if (tid == i) {
                    bx = bx / Lr[i];
                    by = by / Lr[i];
                }
                const double xx = bcast_lane(bx, i), xy = bcast_lane(by, i);
                if (tid < i) {
                    bx -= Lc[i] * xx;
                    by -= Lc[i] * xy;
                }
            }
        }
        if (tid < 2) xs[tid] = 0.0;
        if (tid < n1) {
            xs[2 + 2 * tid] = bx;
            xs[3 + 2 * tid] = by;
        }
        __syncthreads();
        // largest residual above 1 px^2 (first index on ties, as the sequential scan)
        double best = 1.0;
        int bi = -1;
        for (int k = tid; k < m; k += kMinThreads) {
            if (!active[k]) continue;
            const int ra = rows[2 * k], rb = rows[2 * k + 1];
            double ex = d[2 * k] - (xs[2 * rb] - xs[2 * ra]), ey = d[2 * k + 1] - (xs[2 * rb + 1] - xs[2 * ra + 1]);
            double dist = ex * ex + ey * ey;
            if (dist > best) { best = dist; bi = k; }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(bi, off);
            if (oi >= 0 && (bi < 0 || ob > best || (ob == best && oi < bi))) {
                best = ob;
                bi = oi;
            }
        }
        if (tid == 0) {
            if (bi >= 0) {
                d[2 * bi] = 0.0;
                d[2 * bi + 1] = 0.0;
                active[bi] = 0;
            }
            s_flag = bi;
        }
        __syncthreads();
        status = s_flag;
        __syncthreads();
        if (status < 0) break;
    }
    for (int i = tid; i < n1; i += kMinThreads) {
        float* o = a.one_to_one + ((size_t)t * n1 + i) * 2;
        o[0] = (float)(xs[2 * (i + 1)] - xs[2 * i]);
        o[1] = (float)(xs[2 * (i + 1) + 1] - xs[2 * i + 1]);
    }
    for (int k = tid; k < m; k += kMinThreads) {
        int ra = rows[2 * k], rb = rows[2 * k + 1];
        float* v = a.meas + (size_t)k * a.meas_stride + (size_t)t * 2;
        v[0] = (float)(xs[2 * rb] - xs[2 * ra]);
        v[1] = (float)(xs[2 * rb + 1] - xs[2 * ra + 1]);
    }
    if (tid == 0) {
        if (a.status) a.status[t] = status;
        if (a.iters) a.iters[t] = it < 10 ? it + 1 : 10;
    }
}

size_t minimize_reg_smem(int n, int m) {
    return (size_t)(2 * m + 2 * n) * 8 + (size_t)(2 * m + (n - 1) * (n - 1)) * 4 + m + 16;
}

size_t minimize_smem(int n, int m) {
    int n1 = n - 1;
    return (size_t)(2 * m + n1 * n1 + 2 * n1 + 2 * n + kMinThreads) * 8 + (size_t)(kMinThreads + n1 * (n - 1)) * 4 +
           m + 16;
}

// getOptimalShifts (R:Kernels/ShiftMinimizerKernels.cu:178-218).
__global__ void k_optimal_shifts(float* out, int pitch, const float* r, int n, int cx, int cy, int ref, int trk) {
    int tx = blockIdx.x * blockDim.x + threadIdx.x, ty = blockIdx.y * blockDim.y + threadIdx.y;
    if (tx >= cx || ty >= cy) return;
    int n1 = n - 1;
    const float* rr = r + (size_t)(tx + ty * cx) * n1 * 2;
    float sx = 0.0f, sy = 0.0f;
    if (ref < trk)
        for (int i = ref; i < trk; ++i) { sx += rr[2 * i]; sy += rr[2 * i + 1]; }
    else if (trk < ref)
        for (int i = trk; i < ref; ++i) { sx -= rr[2 * i]; sy -= rr[2 * i + 1]; }
    float* o = row_ptr(out, pitch, ty) + tx * 2;
    o[0] = sx;
    o[1] = sy;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
namespace isa {
int launch_track_mfma(int precision, const float* const* frames, const unsigned* const* xplanes,
                      const float* const* vbox, const float* const* vrun, int img_pitch, int w, int h, int tile, int M,
                      int cx, int cy, const int* pa, const int* pb, const int* pk, const int* chunks, int n_chunks,
                      const float* base, float* shifts, size_t meas_stride, float threshold, float* shift_images,
                      int* decisions, int* queue, int* count, hipStream_t s);
int vbox_energies(const float* src, int pitch, int w, int h, int T, float* dst, float* runmax, hipStream_t s);
int vbox_run_rows(int h, int T);
int split_planes(int precision, const float* src, int P, int w, int h, unsigned* dst, hipStream_t s);
}  // namespace isa

namespace isa {

int shift_tables(int n, int ref, int strategy, int block_in, int* pairs, int* index, int* rows) {
    if (n < 2 || n > kMinFrames || ref < 0 || ref >= n) return -ISA_ERR_ARG;
    int block = block_in >= n ? n - 1 : block_in;
    int n1 = n - 1, m;
    switch (strategy) {
        case ISA_STRATEGY_FULL: m = (int)((n - 1) * ((float)n / 2.0f)); break;
        case ISA_STRATEGY_ONLY_ON_REFERENCE: m = n - 1; break;
        case ISA_STRATEGY_ON_REFERENCE_BLOCK: m = (int)((block - 1) * ((float)block / 2.0f)) + block * (n - block); break;
        case ISA_STRATEGY_BLOCKS: m = (int)(-0.5 * block * (block - 2 * n + 1)); break;
        default: return -ISA_ERR_ARG;
    }
    if (block < 1 && strategy >= ISA_STRATEGY_ON_REFERENCE_BLOCK) return -ISA_ERR_ARG;
    std::vector<int> pa, pb, idx((size_t)n * n, -1);
    if (strategy == ISA_STRATEGY_FULL || strategy == ISA_STRATEGY_BLOCKS) {
        int dmax = strategy == ISA_STRATEGY_FULL ? n1 : block;
        if (strategy == ISA_STRATEGY_FULL) {
            for (int a = 0; a < n; ++a)
                for (int b = a + 1; b < n; ++b) { pa.push_back(a); pb.push_back(b); }
        } else {
            for (int s = 0; s < block; ++s)
                for (int line = 0; line < n1 - s; ++line) { pa.push_back(line); pb.push_back(line + s + 1); }
        }
        int c = 0;
        for (int d = 1; d <= dmax; ++d)
            for (int f = 0; f + d < n; ++f) idx[(size_t)f * n + f + d] = c++;
    } else if (strategy == ISA_STRATEGY_ONLY_ON_REFERENCE) {
        int c = 0;
        for (int k = 0; k < n; ++k)
            if (k != ref) { pa.push_back(ref); pb.push_back(k); idx[(size_t)ref * n + k] = c++; }
    } else {
        int first, last;
        if (ref < n / 2) { first = std::max(0, ref - block / 2); last = first + block - 1; }
        else { last = std::min(n - 1, ref + block / 2); first = last - block + 1; }
        if (first < 0 || last >= n) return -ISA_ERR_ARG;
        int c = 0;
        for (int s = 0; s < block - 1; ++s)
            for (int line = 0; line < block - 1 - s; ++line) {
                pa.push_back(line + first); pb.push_back(line + first + s + 1);
                idx[(size_t)(line + first) * n + line + first + s + 1] = c++;
            }
        for (int i = 0; i < first; ++i)
            for (int tb = first; tb <= last; ++tb) { pa.push_back(tb); pb.push_back(i); idx[(size_t)tb * n + i] = c++; }
        for (int i = last + 1; i < n; ++i)
            for (int tb = first; tb <= last; ++tb) { pa.push_back(tb); pb.push_back(i); idx[(size_t)tb * n + i] = c++; }
    }
    if ((int)pa.size() != m) {
        set_error("shift tables: pair count %d does not match shift count %d", (int)pa.size(), m);
        return -ISA_ERR_ARG;
    }
    for (int p = 0; p < m; ++p) {
        int k = idx[(size_t)pa[p] * n + pb[p]];
        if (k < 0 || k >= m) return -ISA_ERR_ARG;
        if (pairs) { pairs[2 * p] = pa[p]; pairs[2 * p + 1] = pb[p]; }
        if (rows) { rows[2 * k] = pa[p]; rows[2 * k + 1] = pb[p]; }
    }
    if (index) std::copy(idx.begin(), idx.end(), index);
    return m;
}

template <int M>
static int launch_track_level_m(const TrackArgs& a, int pairs, hipStream_t s) {
    if constexpr (M == 1 || M == 2 || M == 4) {
        if (a.tile >= 4 && a.tile <= 64 && 64 % a.tile == 0) {
            const ColsShape c = cols_shape(a.tile, M);
            const int items = pairs * a.cx * a.cy;
            const int blocks = div_up(items, c.items_per_block);
            const size_t csm = cols_smem(a.tile, M, c.items_per_block);
            const unsigned grid = 8 * div_up(blocks, 8);
            if (a.decisions)
                k_track_cols<M, true><<<grid, c.threads, csm, s>>>(a, pairs, items, c.items_per_block,
                                                                   c.lanes_per_item, blocks);
            else
                k_track_cols<M, false><<<grid, c.threads, csm, s>>>(a, pairs, items, c.items_per_block,
                                                                    c.lanes_per_item, blocks);
            return check_launch("track_cols");
        }
    }
    size_t smem = ssd_smem(a.tile, M);
    k_track_level<M><<<dim3(pairs, a.cx * a.cy), block_threads(a.tile, M), smem, s>>>(a);
    return check_launch("track_level");
}

int launch_track_level(const float* const* frames, int img_pitch, int w, int h, int tile, int M, int cx, int cy,
                       const int* pa, const int* pb, const int* pk, int pairs, const float* base, float* shifts,
                       size_t meas_stride, float threshold, float* shift_images, int* decisions, hipStream_t s) {
    if (ssd_smem(tile, M) > 160 * 1024) {
        set_error("track: tile %d with max shift %d needs too much LDS", tile, M);
        return ISA_ERR_UNSUPPORTED;
    }
    TrackArgs a{frames, img_pitch, w, h, tile, cx, cy, pa, pb, pk, base, shifts, meas_stride, threshold,
                shift_images, decisions};
    switch (M) {
        case 1: return launch_track_level_m<1>(a, pairs, s);
        case 2: return launch_track_level_m<2>(a, pairs, s);
        case 3: return launch_track_level_m<3>(a, pairs, s);
        case 4: return launch_track_level_m<4>(a, pairs, s);
        case 5: return launch_track_level_m<5>(a, pairs, s);
        case 6: return launch_track_level_m<6>(a, pairs, s);
        case 7: return launch_track_level_m<7>(a, pairs, s);
        case 8: return launch_track_level_m<8>(a, pairs, s);
    }
    set_error("track: max shift %d outside [1, %d]", M, kMaxShift);
    return ISA_ERR_UNSUPPORTED;
}

// Exact search of the queued items (queue, *count: from launch_track_mfma). The grid is fixed; its
// blocks take the items count-strided, so the count stays on the device.
int launch_track_queue(const float* const* frames, int img_pitch, int w, int h, int tile, int M, int cx, int cy,
                       const int* pa, const int* pb, const int* pk, const float* base, float* shifts,
                       size_t meas_stride, float threshold, int* decisions, const int* queue, const int* count,
                       hipStream_t s) {
    const size_t smem = ssd_smem(tile, M);
    if (smem > 160 * 1024) {
        set_error("track: tile %d with max shift %d needs too much LDS", tile, M);
        return ISA_ERR_UNSUPPORTED;
    }
    TrackArgs a{frames, img_pitch, w, h, tile, cx, cy, pa, pb, pk, base, shifts, meas_stride, threshold,
                nullptr, decisions};
    if ((M == 1 || M == 2 || M == 4) && tile >= 4 && tile <= 64 && 64 % tile == 0) {
        const ColsShape c = cols_shape(tile, M);
        const size_t csm = cols_smem(tile, M, c.items_per_block);
        const int ppb = c.items_per_block, lpi = c.lanes_per_item;
#define ISA_TCQ(MM)                                                                                          \
    if (M == MM) {                                                                                         \
        if (decisions) k_track_cols_queue<MM, true><<<4096, c.threads, csm, s>>>(a, ppb, lpi, queue, count);  \
        else k_track_cols_queue<MM, false><<<4096, c.threads, csm, s>>>(a, ppb, lpi, queue, count);           \
    }
        ISA_TCQ(1) ISA_TCQ(2) ISA_TCQ(4)
#undef ISA_TCQ
        return check_launch("track_cols_queue");
    }
    const int grid = 2048, threads = block_threads(tile, M);
#define ISA_TQ(MM)                                                                   \
    case MM:                                                                       \
        k_track_queue<MM><<<grid, threads, smem, s>>>(a, queue, count);            \
        return check_launch("track_queue");
    switch (M) { ISA_TQ(1) ISA_TQ(2) ISA_TQ(3) ISA_TQ(4) ISA_TQ(5) ISA_TQ(6) ISA_TQ(7) ISA_TQ(8) }
#undef ISA_TQ
    set_error("track: max shift %d outside [1, %d]", M, kMaxShift);
    return ISA_ERR_UNSUPPORTED;
}

int launch_upsample(const float* in, float* out, int in_pitch, int out_pitch, int ol, int nl, int ocx, int ocy, int ncx,
                    int ncy, int ot, int nt, hipStream_t s) {
    k_upsample<<<dim3(div_up(ncx, 16), div_up(ncy, 16)), dim3(16, 16), 0, s>>>(in, out, in_pitch, out_pitch, ol, nl,
                                                                               ocx, ocy, ncx, ncy, ot, nt);
    return check_launch("upsample_shifts");
}

int launch_upsample_all(const float* in, float* out, size_t meas_stride, int m, int ol, int nl, int ocx, int ocy,
                        int ncx, int ncy, int ot, int nt, hipStream_t s) {
    k_upsample_all<<<dim3(div_up(ncx, 16), div_up(ncy, 16), m), dim3(16, 16), 0, s>>>(in, out, meas_stride, ol, nl, ocx,
                                                                                      ocy, ncx, ncy, ot, nt);
    return check_launch("upsample_shifts_all");
}

// The first iteration's normal matrix, factored on the host: the reduced Laplacian of all m rows
// and its Cholesky factor, with the same operations in the same order as k_minimize (so the factor
// is the one every tile would compute). Returns false when the graph is disconnected.
bool minimize_full_factor(const int* rows, int n, int m, double* L) {
    const int n1 = n - 1;
    std::vector<int> parent(n);
    std::iota(parent.begin(), parent.end(), 0);
    auto root = [&](int u) {
        while (parent[u] != u) u = parent[u];
        return u;
    };
    for (int k = 0; k < m; ++k) {
        int u = root(rows[2 * k]), v = root(rows[2 * k + 1]);
        if (u != v) parent[std::max(u, v)] = std::min(u, v);
    }
    for (int i = 0; i < n; ++i)
        if (root(i) != 0) return false;
    for (int i = 0; i < n1 * n1; ++i) L[i] = 0.0;
    for (int i = 0; i < n1; ++i) {
        const int node = i + 1;
        for (int k = 0; k < m; ++k) {
            const int ra = rows[2 * k], rb = rows[2 * k + 1];
            if (rb == node) {
                L[i * n1 + i] += 1.0;
                if (ra != 0) L[i * n1 + ra - 1] -= 1.0;
            } else if (ra == node) {
                L[i * n1 + i] += 1.0;
                if (rb != 0) L[i * n1 + rb - 1] -= 1.0;
            }
        }
    }
    for (int j = 0; j < n1; ++j) {
        L[j * n1 + j] = std::sqrt(L[j * n1 + j]);
        const double djj = L[j * n1 + j];
        for (int i = j + 1; i < n1; ++i) L[i * n1 + j] /= djj;
        for (int i = j + 1; i < n1; ++i) {
            const double lij = L[i * n1 + j];
            for (int c = j + 1; c <= i; ++c) L[i * n1 + c] -= lij * L[c * n1 + j];
        }
    }
    return true;
}

int launch_minimize(float* meas, size_t meas_stride, float* one_to_one, int* status, int* iters, const int* rows,
                    const double* lfull, int tiles, int n, int m, hipStream_t s) {
    size_t smem = minimize_smem(n, m);
    if (n > kMinFrames || smem > 160 * 1024) {
        set_error("minimize: %d frames / %d measurements exceed the supported size", n, m);
        return ISA_ERR_UNSUPPORTED;
    }
    MinArgs a{meas, meas_stride, one_to_one, status, iters, rows, lfull, tiles, n, m};
    if (n <= 8) {
        k_minimize_reg<7><<<tiles, kMinThreads, minimize_reg_smem(n, m), s>>>(a);
    } else if (n <= 16) {
        k_minimize_reg<15><<<tiles, kMinThreads, minimize_reg_smem(n, m), s>>>(a);
    } else if (n <= 24) {
        k_minimize_reg<23><<<tiles, kMinThreads, minimize_reg_smem(n, m), s>>>(a);
    } else if (n <= 32) {
        k_minimize_reg<31><<<tiles, kMinThreads, minimize_reg_smem(n, m), s>>>(a);
    } else {
        k_minimize<<<tiles, kMinThreads, smem, s>>>(a);
    }
    return check_launch("minimize_shifts");
}

int launch_optimal_shifts(float* out, int pitch, const float* r, int n, int cx, int cy, int ref, int trk,
                          hipStream_t s) {
    k_optimal_shifts<<<dim3(div_up(cx, 32), div_up(cy, 8)), dim3(32, 8), 0, s>>>(out, pitch, r, n, cx, cy, ref, trk);
    return check_launch("get_optimal_shifts");
}

}  // namespace isa

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

int isa_convert_to_tiles_border(const float* img_dev, int img_pitch, float* tiles_dev, int width, int height,
                                int max_shift, int tile_size, int tile_count_x, int tile_count_y,
                                const float base_shift[2], float base_rotation, isa_stream_t stream) {
    ISA_REQUIRE(img_dev && tiles_dev && width > 0 && height > 0 && max_shift >= 0 && tile_size > 0 &&
                    tile_count_x > 0 && tile_count_y > 0,
                "isa_convert_to_tiles_border: bad arguments");
    int B = tile_size + 2 * max_shift;
    k_tiles_border<<<dim3(isa::div_up(B, 24), B, isa::div_up(tile_count_x * tile_count_y, 8)), dim3(24, 1, 8), 0,
                     isa::to_stream(stream)>>>(img_dev, img_pitch, tiles_dev, width, height, max_shift, tile_size,
                                               tile_count_x, tile_count_y, base_shift[0], base_shift[1], base_rotation);
    return isa::check_launch("isa_convert_to_tiles_border");
}

int isa_convert_to_tiles_preshift(const float* img_dev, int img_pitch, float* tiles_dev, const float* preshift2_dev,
                                  int preshift_pitch, int width, int height, int max_shift, int tile_size,
                                  int tile_count_x, int tile_count_y, const float base_shift[2], float base_rotation,
                                  isa_stream_t stream) {
    ISA_REQUIRE(img_dev && tiles_dev && preshift2_dev && width > 0 && height > 0 && max_shift >= 0 && tile_size > 0 &&
                    tile_count_x > 0 && tile_count_y > 0,
                "isa_convert_to_tiles_preshift: bad arguments");
    int B = tile_size + 2 * max_shift;
    k_tiles_preshift<<<dim3(isa::div_up(B, 24), B, isa::div_up(tile_count_x * tile_count_y, 8)), dim3(24, 1, 8), 0,
                       isa::to_stream(stream)>>>(img_dev, img_pitch, tiles_dev, preshift2_dev, preshift_pitch, width,
                                                 height, max_shift, tile_size, tile_count_x, tile_count_y,
                                                 base_shift[0], base_shift[1], base_rotation);
    return isa::check_launch("isa_convert_to_tiles_preshift");
}

int isa_tile_ssd(const float* ref_tiles_dev, const float* track_tiles_dev, float* shift_image_dev, int max_shift,
                 int tile_size, int tile_count, isa_stream_t stream) {
    ISA_REQUIRE(ref_tiles_dev && track_tiles_dev && shift_image_dev && tile_size > 0 && tile_count > 0,
                "isa_tile_ssd: bad arguments");
    ISA_REQUIRE(max_shift >= 1 && max_shift <= kMaxShift, "isa_tile_ssd: max_shift must be in [1, 8]");
    size_t smem = ssd_smem(tile_size, max_shift);
    if (smem > 160 * 1024) {
        isa::set_error("isa_tile_ssd: tile too large for LDS");
        return ISA_ERR_UNSUPPORTED;
    }
    hipStream_t s = isa::to_stream(stream);
    int bt = block_threads(tile_size, max_shift);
    switch (max_shift) {
        case 1: k_tile_ssd<1><<<tile_count, bt, smem, s>>>(ref_tiles_dev, track_tiles_dev, shift_image_dev, tile_size); break;
        case 2: k_tile_ssd<2><<<tile_count, bt, smem, s>>>(ref_tiles_dev, track_tiles_dev, shift_image_dev, tile_size); break;
        case 3: k_tile_ssd<3><<<tile_count, bt, smem, s>>>(ref_tiles_dev, track_tiles_dev, shift_image_dev, tile_size); break;
        case 4: k_tile_ssd<4><<<tile_count, bt, smem, s>>>(ref_tiles_dev, track_tiles_dev, shift_image_dev, tile_size); break;
        case 5: k_tile_ssd<5><<<tile_count, bt, smem, s>>>(ref_tiles_dev, track_tiles_dev, shift_image_dev, tile_size); break;
        case 6: k_tile_ssd<6><<<tile_count, bt, smem, s>>>(ref_tiles_dev, track_tiles_dev, shift_image_dev, tile_size); break;
        case 7: k_tile_ssd<7><<<tile_count, bt, smem, s>>>(ref_tiles_dev, track_tiles_dev, shift_image_dev, tile_size); break;
        default: k_tile_ssd<8><<<tile_count, bt, smem, s>>>(ref_tiles_dev, track_tiles_dev, shift_image_dev, tile_size); break;
    }
    return isa::check_launch("isa_tile_ssd");
}

int isa_find_minimum(const float* shift_image_dev, float* coords2_dev, int coords_pitch, int max_shift,
                     int tile_count_x, int tile_count_y, float threshold, isa_stream_t stream) {
    ISA_REQUIRE(shift_image_dev && coords2_dev && max_shift >= 0 && tile_count_x > 0 && tile_count_y > 0,
                "isa_find_minimum: bad arguments");
    int n = tile_count_x * tile_count_y;
    k_find_minimum<<<isa::div_up(n, 128), 128, 0, isa::to_stream(stream)>>>(shift_image_dev, coords2_dev, coords_pitch,
                                                                           max_shift, tile_count_x, tile_count_y,
                                                                           threshold);
    return isa::check_launch("isa_find_minimum");
}

int isa_add_shifts(float* preshift2_dev, int preshift_pitch, const float* coords2_dev, int coords_pitch,
                   int tile_count_x, int tile_count_y, isa_stream_t stream) {
    ISA_REQUIRE(preshift2_dev && coords2_dev && tile_count_x > 0 && tile_count_y > 0, "isa_add_shifts: bad arguments");
    k_add_shifts<<<dim3(isa::div_up(2 * tile_count_x, 32), isa::div_up(tile_count_y, 8)), dim3(32, 8), 0,
                   isa::to_stream(stream)>>>(preshift2_dev, preshift_pitch, coords2_dev, coords_pitch, tile_count_x,
                                             tile_count_y);
    return isa::check_launch("isa_add_shifts");
}

int isa_upsample_shifts(const float* in2_dev, float* out2_dev, int in_pitch, int out_pitch, int old_level,
                        int new_level, int old_count_x, int old_count_y, int new_count_x, int new_count_y,
                        int old_tile_size, int new_tile_size, isa_stream_t stream) {
    ISA_REQUIRE(in2_dev && out2_dev && old_level > 0 && new_level > 0 && old_count_x > 0 && old_count_y > 0 &&
                    new_count_x > 0 && new_count_y > 0 && old_tile_size > 0 && new_tile_size > 0,
                "isa_upsample_shifts: bad arguments");
    return isa::launch_upsample(in2_dev, out2_dev, in_pitch, out_pitch, old_level, new_level, old_count_x, old_count_y,
                                new_count_x, new_count_y, old_tile_size, new_tile_size, isa::to_stream(stream));
}

}  // extern "C"

namespace {

// Workspace of the per-call tile search (one pair): the argument block the level kernels read
// (frame pointers, pair indices, base shifts, the exact-search counter), the tracked shifts in
// dense rows, and for the matrix-core forms the split planes, window energies and queue.
struct TrackHead {
    const float* frames[2];
    const unsigned* xplanes[2];
    const float* vbox[2];
    const float* vrun[2];
    int pa, pb, pk, count;
    int chunk[16];
    float base[8];
};

struct TrackLayout {
    size_t head, shifts, planes, vbox, vrun, queue, bytes;
};

size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

TrackLayout track_layout(int width, int height, int img_pitch, int tile_size, int max_shift, int precision) {
    const int cx = (width - 2 * max_shift) / tile_size, cy = (height - 2 * max_shift) / tile_size;
    const bool split = precision == ISA_XC_BF16X3 || precision == ISA_XC_FP16X3;
    const bool mfma = precision != ISA_XC_FP64;
    TrackLayout t;
    size_t o = 0;
    t.head = o; o += al256(sizeof(TrackHead));
    t.shifts = o; o += al256((size_t)cx * cy * 2 * sizeof(float));
    t.planes = o; o += split ? 2 * al256(((size_t)img_pitch / 4) * height * sizeof(unsigned)) : 0;
    t.vbox = o; o += mfma ? al256((size_t)(height - tile_size + 1) * width * sizeof(float)) : 0;
    t.vrun = o; o += mfma ? al256((size_t)isa::vbox_run_rows(height, tile_size) * width * sizeof(float)) : 0;
    t.queue = o; o += mfma ? al256((size_t)cx * cy * 2 * sizeof(int)) : 0;
    t.bytes = o;
    return t;
}

__global__ void k_put_track_head(TrackHead h, TrackHead* dst) { *dst = h; }

__global__ void k_set_int(int* dst, const int* src, int value) { *dst = src ? *src : value; }

}  // namespace

extern "C" {

size_t isa_track_workspace(int width, int height, int img_pitch, int tile_size, int max_shift, int xc_precision) {
    if (width <= 0 || height <= 0 || tile_size <= 0 || max_shift < 1 || max_shift > kMaxShift ||
        img_pitch < width * 4 || xc_precision < ISA_XC_FP64 || xc_precision > ISA_XC_FP32)
        return 0;
    if ((width - 2 * max_shift) / tile_size <= 0 || (height - 2 * max_shift) / tile_size <= 0) return 0;
    return track_layout(width, height, img_pitch, tile_size, max_shift, xc_precision).bytes;
}

int isa_track_xc_ex(const float* img_track_dev, const float* img_ref_dev, int img_pitch, float* preshift2_dev,
                    int preshift_pitch, int width, int height, int tile_size, int max_shift,
                    const float base_shift_ref[2], float base_rotation_ref, const float base_shift_track[2],
                    float base_rotation_track, float threshold, int xc_precision, int* decisions_dev,
                    int* exact_count_dev, void* workspace_dev, isa_stream_t stream) {
    // PatchTracker.Track through the fused level kernel with one pair (frame 0 = reference, frame 1 = tracked).
    ISA_REQUIRE(img_track_dev && img_ref_dev && preshift2_dev && workspace_dev && width > 0 && height > 0 &&
                    tile_size > 0,
                "isa_track: bad arguments");
    ISA_REQUIRE(reinterpret_cast<uintptr_t>(workspace_dev) % 256 == 0, "isa_track: workspace must be 256-byte aligned");
    ISA_REQUIRE(xc_precision >= ISA_XC_FP64 && xc_precision <= ISA_XC_FP32, "isa_track: unknown xc_precision");
    ISA_REQUIRE(max_shift >= 1 && max_shift <= kMaxShift, "isa_track: max_shift must be in [1, 8]");
    const int cx = (width - 2 * max_shift) / tile_size, cy = (height - 2 * max_shift) / tile_size;
    ISA_REQUIRE(cx > 0 && cy > 0, "isa_track: image smaller than one tile");
    ISA_REQUIRE(img_pitch % 4 == 0 && img_pitch >= width * 4, "isa_track: bad image pitch");
    ISA_REQUIRE(preshift_pitch >= cx * 8, "isa_track: bad shift pitch");
    hipStream_t s = isa::to_stream(stream);
    const bool split = xc_precision == ISA_XC_BF16X3 || xc_precision == ISA_XC_FP16X3;
    const bool mfma = xc_precision != ISA_XC_FP64;
    const TrackLayout tl = track_layout(width, height, img_pitch, tile_size, max_shift, xc_precision);
    char* d = static_cast<char*>(workspace_dev);
    TrackHead h{{img_ref_dev, img_track_dev}, {nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}, 0, 1, 0, 0,
                {0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
                {base_shift_ref[0], base_shift_ref[1], (float)std::sin((double)base_rotation_ref),
                 (float)std::cos((double)base_rotation_ref), base_shift_track[0], base_shift_track[1],
                 (float)std::sin((double)base_rotation_track), (float)std::cos((double)base_rotation_track)}};
    const size_t plane = split ? (tl.vbox - tl.planes) / 2 : 0;
    if (split) {
        h.xplanes[0] = (const unsigned*)(d + tl.planes);
        h.xplanes[1] = (const unsigned*)(d + tl.planes + plane);
    }
    if (mfma) {
        h.vbox[1] = (const float*)(d + tl.vbox);
        h.vrun[1] = (const float*)(d + tl.vrun);
    }
    TrackHead* sd = (TrackHead*)(d + tl.head);
    float* dn = (float*)(d + tl.shifts);
    int* queue = (int*)(d + tl.queue);
    const size_t row = (size_t)cx * 2 * sizeof(float);
    // the argument block travels in the kernel's own arguments: nothing on the host has to outlive a copy
    k_put_track_head<<<1, 1, 0, s>>>(h, sd);
    ISA_TRY(isa::check_launch("isa_track: arguments"));
    ISA_TRY(isa::check_hip(hipMemcpy2DAsync(dn, row, preshift2_dev, preshift_pitch, row, cy, hipMemcpyDeviceToDevice, s),
                           "isa_track: copy"));
    for (int f = 0; f < 2 && split; ++f)
        ISA_TRY(isa::split_planes(xc_precision, f ? img_track_dev : img_ref_dev, img_pitch / 4, width, height,
                                  (unsigned*)h.xplanes[f], s));
    if (mfma)
        ISA_TRY(isa::vbox_energies(img_track_dev, img_pitch, width, height, tile_size, (float*)h.vbox[1],
                                   (float*)h.vrun[1], s));
    if (!mfma)
        ISA_TRY(isa::launch_track_level(sd->frames, img_pitch, width, height, tile_size, max_shift, cx, cy, &sd->pa,
                                        &sd->pb, &sd->pk, 1, sd->base, dn, 0, threshold, nullptr, decisions_dev, s));
    else {
        ISA_TRY(isa::launch_track_mfma(xc_precision, sd->frames, sd->xplanes, sd->vbox, sd->vrun, img_pitch, width,
                                       height, tile_size, max_shift, cx, cy, &sd->pa, &sd->pb, &sd->pk, sd->chunk, 1,
                                       sd->base, dn, 0, threshold, nullptr, decisions_dev, queue, &sd->count, s));
        ISA_TRY(isa::launch_track_queue(sd->frames, img_pitch, width, height, tile_size, max_shift, cx, cy, &sd->pa,
                                        &sd->pb, &sd->pk, sd->base, dn, 0, threshold, decisions_dev, queue,
                                        &sd->count, s));
    }
    ISA_TRY(isa::check_hip(hipMemcpy2DAsync(preshift2_dev, preshift_pitch, dn, row, row, cy, hipMemcpyDeviceToDevice, s),
                           "isa_track: copy back"));
    if (exact_count_dev) {
        k_set_int<<<1, 1, 0, s>>>(exact_count_dev, mfma ? &sd->count : nullptr, cx * cy);
        ISA_TRY(isa::check_launch("isa_track: count"));
    }
    return ISA_OK;
}

int isa_track_xc(const float* img_track_dev, const float* img_ref_dev, int img_pitch, float* preshift2_dev,
                 int preshift_pitch, int width, int height, int tile_size, int max_shift, const float base_shift_ref[2],
                 float base_rotation_ref, const float base_shift_track[2], float base_rotation_track, float threshold,
                 int xc_precision, void* workspace_dev, isa_stream_t stream) {
    return isa_track_xc_ex(img_track_dev, img_ref_dev, img_pitch, preshift2_dev, preshift_pitch, width, height,
                           tile_size, max_shift, base_shift_ref, base_rotation_ref, base_shift_track,
                           base_rotation_track, threshold, xc_precision, nullptr, nullptr, workspace_dev, stream);
}

int isa_track(const float* img_track_dev, const float* img_ref_dev, int img_pitch, float* preshift2_dev,
              int preshift_pitch, int width, int height, int tile_size, int max_shift, const float base_shift_ref[2],
              float base_rotation_ref, const float base_shift_track[2], float base_rotation_track, float threshold,
              void* workspace_dev, isa_stream_t stream) {
    return isa_track_xc(img_track_dev, img_ref_dev, img_pitch, preshift2_dev, preshift_pitch, width, height, tile_size,
                        max_shift, base_shift_ref, base_rotation_ref, base_shift_track, base_rotation_track, threshold,
                        ISA_XC_FP64, workspace_dev, stream);
}

int isa_shift_tables(int frame_count, int reference_index, int strategy, int block_size, int* pairs_out,
                     int* index_out, int* rows_out) {
    return isa::shift_tables(frame_count, reference_index, strategy, block_size, pairs_out, index_out, rows_out);
}

}  // extern "C"

namespace {

// Workspace of isa_minimize_shifts: the measurement rows, then (for a connected graph of at most
// kMinFrames frames) the Cholesky factor of the all-rows normal equations. At most kMinFrames = 64
// frames: m <= 64 * 63 / 2 = 2016 rows (16,128 bytes) and a 63 x 63 factor (31,752 bytes).
constexpr size_t kMinimizeTablesMax = 16128 + 31752;

bool minimize_layout(int frame_count, int reference_index, int strategy, int block_size, int* m_out,
                     size_t* rows_bytes, size_t* total) {
    if (frame_count < 2 || frame_count > kMinFrames) return false;
    int m = isa::shift_tables(frame_count, reference_index, strategy, block_size, nullptr, nullptr, nullptr);
    if (m <= 0) return false;
    const int n1 = frame_count - 1;
    *m_out = m;
    *rows_bytes = ((size_t)2 * m * sizeof(int) + 15) & ~(size_t)15;
    *total = *rows_bytes + (size_t)n1 * n1 * sizeof(double);
    return *total <= kMinimizeTablesMax;
}

// The minimizer on the tables in `tables_dev` (at least the layout's total bytes, 16-byte aligned):
// the rows and factor are built in a host stack buffer and written in stream order through kernel
// arguments. No device allocation, no host synchronisation.
int minimize_with_tables(float* measured2_dev, int meas_pitch, float* one_to_one2_dev, int* status_dev,
                         int* iters_dev, int tile_count, int frame_count, int reference_index, int strategy,
                         int block_size, void* tables_dev, const char* who, hipStream_t s) {
    ISA_REQUIRE(measured2_dev && one_to_one2_dev && tile_count > 0 && meas_pitch >= tile_count * 8 &&
                    meas_pitch % 8 == 0,
                "%s: bad arguments", who);
    int m = 0;
    size_t rows_bytes = 0, total = 0;
    ISA_REQUIRE(minimize_layout(frame_count, reference_index, strategy, block_size, &m, &rows_bytes, &total),
                "%s: invalid strategy / frame count (2..%d frames)", who, kMinFrames);
    alignas(16) unsigned char host[kMinimizeTablesMax];  // 47,880 bytes of stack
    int* rows = reinterpret_cast<int*>(host);
    isa::shift_tables(frame_count, reference_index, strategy, block_size, nullptr, nullptr, rows);
    double* factor = reinterpret_cast<double*>(host + rows_bytes);
    const bool have_factor = isa::minimize_full_factor(rows, frame_count, m, factor);
    ISA_TRY(isa::upload_by_args(tables_dev, host, have_factor ? total : rows_bytes, s));
    const int* d_rows = (const int*)tables_dev;
    const double* d_factor = have_factor ? (const double*)((char*)tables_dev + rows_bytes) : nullptr;
    return isa::launch_minimize(measured2_dev, meas_pitch / sizeof(float), one_to_one2_dev, status_dev, iters_dev,
                                d_rows, d_factor, tile_count, frame_count, m, s);
}

// Device-resident tables of isa_minimize_shifts_noalloc: one copy per GPU, held until the process ends.
static __device__ __align__(256) unsigned char scratch[49152];
static_assert(sizeof(scratch) >= kMinimizeTablesMax, "scratch holds the largest tables");

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` on each GPU, looked up once

int scratch_address(void** out) {
    int dev = 0;
    ISA_TRY(isa::check_hip(hipGetDevice(&dev), "isa_minimize_shifts_noalloc: hipGetDevice"));
    ISA_REQUIRE(dev >= 0 && dev < kMaxDevices, "isa_minimize_shifts_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        ISA_TRY(isa::check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                               "isa_minimize_shifts_noalloc: scratch address"));
        ISA_REQUIRE(p != nullptr, "isa_minimize_shifts_noalloc: no scratch address");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return ISA_OK;
}

}  // namespace

extern "C" {

size_t isa_minimize_shifts_workspace(int frame_count, int reference_index, int strategy, int block_size) {
    int m = 0;
    size_t rb = 0, tot = 0;
    return minimize_layout(frame_count, reference_index, strategy, block_size, &m, &rb, &tot) ? tot : 0;
}

int isa_minimize_shifts(float* measured2_dev, int meas_pitch, float* one_to_one2_dev, int* status_dev, int* iters_dev,
                        int tile_count, int frame_count, int reference_index, int strategy, int block_size,
                        void* workspace_dev, isa_stream_t stream) {
    ISA_REQUIRE(workspace_dev && reinterpret_cast<uintptr_t>(workspace_dev) % 16 == 0,
                "isa_minimize_shifts: workspace must be a 16-byte aligned device buffer");
    return minimize_with_tables(measured2_dev, meas_pitch, one_to_one2_dev, status_dev, iters_dev, tile_count,
                                frame_count, reference_index, strategy, block_size, workspace_dev,
                                "isa_minimize_shifts", isa::to_stream(stream));
}

int isa_minimize_shifts_noalloc(float* measured2_dev, int meas_pitch, float* one_to_one2_dev, int* status_dev,
                                int* iters_dev, int tile_count, int frame_count, int reference_index, int strategy,
                                int block_size, isa_stream_t stream) {
    int m = 0;
    size_t rb = 0, total = 0;
    ISA_REQUIRE(minimize_layout(frame_count, reference_index, strategy, block_size, &m, &rb, &total),
                "isa_minimize_shifts_noalloc: tables of %d frames exceed the %zu-byte scratch (2..%d frames)",
                frame_count, sizeof(scratch), kMinFrames);
    void* tables = nullptr;
    ISA_TRY(scratch_address(&tables));
    return minimize_with_tables(measured2_dev, meas_pitch, one_to_one2_dev, status_dev, iters_dev, tile_count,
                                frame_count, reference_index, strategy, block_size, tables,
                                "isa_minimize_shifts_noalloc", isa::to_stream(stream));
}

int isa_get_optimal_shifts(float* out2_dev, int out_pitch, const float* one_to_one2_dev, int frame_count,
                           int tile_count_x, int tile_count_y, int reference_image, int image_to_track,
                           isa_stream_t stream) {
    ISA_REQUIRE(out2_dev && one_to_one2_dev && frame_count >= 2 && tile_count_x > 0 && tile_count_y > 0 &&
                    reference_image >= 0 && reference_image < frame_count && image_to_track >= 0 &&
                    image_to_track < frame_count,
                "isa_get_optimal_shifts: bad arguments");
    return isa::launch_optimal_shifts(out2_dev, out_pitch, one_to_one2_dev, frame_count, tile_count_x, tile_count_y,
                                      reference_image, image_to_track, isa::to_stream(stream));
}

}  // extern "C"
