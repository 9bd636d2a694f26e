// Global pre-alignment: the translation part of PreAlignment.ScanAngles (R:PEFStudioDX/PreAlignment.cs:177-261)
// as driven by ComputePreAlignment (R:PEFStudioDX/ImageStackAlignatorController.cs:1451-1528).
//
// The reference correlates each tracking image with frame 0 by a full-frame cuFFT: the circular
// cross-correlation c(s) = sum_x r(x) t((x + s) mod (W, H)) (conjugateComplexMulKernel on R2C spectra,
// C2R, divided by W H), takes the first maximum, wraps it into (-W/2, W/2] x (-H/2, H/2] and returns
// minus it. Here the same c(s) is evaluated directly at the lags that can hold the maximum: every lag
// of the search radius on a box-downsampled image (the first level whose radius is at most 8 lags, or
// the last whose sides stay at least 64 px), then a window around twice the previous maximum at each
// finer level: 15 x 15 lags at the level below the coarsest, then 5 x 5, searched twice above full
// resolution (the second window centred on the first's maximum: a maximum that stays put is the largest
// of its own 5 x 5 neighbourhood). Sums are FP64 per thread and reduced across blocks in a fixed order,
// so the result is deterministic; nothing is allocated and the host never waits (the window centre of
// each level is read on the device). Up to PA_BATCH tracked frames share every launch: their lag tiles
// are blocks of one grid and each frame keeps its own partial sums, maxima and window centres.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "isa_device.h"
#include "isa_host.h"

namespace isa {
int launch_resize_box(const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int l, hipStream_t s);
}

namespace {

constexpr int PA_K = 5, PA_LAGS = PA_K * PA_K;  // lags per tile: 5 x 5
constexpr int PA_THREADS = 256;                 // columns per block, one per thread
constexpr int PA_MAX_LEVELS = 6;               // levels 1, 2, 4, 8, 16, 32
constexpr int PA_COARSE_RADIUS = 8;            // lags each way at the coarsest level (before the slack)
constexpr int PA_MIN_SIDE = 64;                // smallest side of the coarsest level
constexpr int PA_MIN_ROWS = 32;                // smallest row chunk
constexpr int PA_BATCH = 8;                    // tracked frames per launch
constexpr int PA_BSTRIDE = 2 * PA_MAX_LEVELS;  // ints of one frame's maxima (x, y per level)

struct FrameSet {
    const float* p[PA_BATCH];  // per frame of the batch: its image at the level searched
};

struct LagWindow {
    int tiles;       // lag tiles per row and column (odd); lag tile (tx, ty) covers 5 x 5 lags
    int radius;      // the window is centre + [-radius, radius]^2, radius = (5 tiles - 1) / 2
};

// 2 x 2 box means of every frame of a batch (frame blockIdx.z: in.p[z] -> out + z * out_stride floats): the
// sums (rows in order, columns in order, from 0) and the division of k_resize_box's 2 x 2 box, so the
// values are identical to its.
__global__ void k_box2_batch(FrameSet in, int in_pitch, float* out, size_t out_stride, int out_pitch, int ow, int oh) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= ow || y >= oh) return;
    const float* r0 = row_ptr(in.p[blockIdx.z], in_pitch, 2 * y) + 2 * x;
    const float* r1 = row_ptr(in.p[blockIdx.z], in_pitch, 2 * y + 1) + 2 * x;
    float s = 0.0f;
    s += r0[0];
    s += r0[1];
    s += r1[0];
    s += r1[1];
    row_ptr(out + blockIdx.z * out_stride, out_pitch, y)[x] = s / 4.0f;
}

// Partial c(s) over one (column block, row chunk) of the image for the 25 lags of one lag tile of one
// frame of the batch (blockIdx.z = frame * tiles + tile): s = centre + o with o in the tile, centre =
// scale * (the frame's previous maximum, centre_in + frame * PA_BSTRIDE), or 0.
// Thread x walks the rows of its column: each b row is loaded once (5 neighbouring columns) and meets the
// 5 a rows of the tile's row lags, held in a ring of registers (zero outside the chunk), so a pixel costs
// 6 loads for 25 FP64 products (float products are exact in FP64). Lags are reduced modulo the image
// size first, so one wrap per coordinate suffices.
__global__ __launch_bounds__(PA_THREADS) void k_cc_rows(const float* __restrict__ a, FrameSet bs, int pitch, int w,
                                                        int h, const int* centre_in, int scale, LagWindow win,
                                                        int rows, double* partial, size_t part_stride) {
    const int tiles = win.tiles * win.tiles, fb = blockIdx.z / tiles;
    const int tile = blockIdx.z - fb * tiles, ty = tile / win.tiles, tx = tile - ty * win.tiles;
    const float* __restrict__ b = bs.p[fb];
    partial += fb * part_stride;
    int cx = 0, cy = 0;
    if (centre_in) {
        cx = centre_in[fb * PA_BSTRIDE] * scale;
        cy = centre_in[fb * PA_BSTRIDE + 1] * scale;
    }
    int sx0 = (cx - win.radius + PA_K * tx) % w, sy0 = (cy - win.radius + PA_K * ty) % h;
    if (sx0 < 0) sx0 += w;
    if (sy0 < 0) sy0 += h;
    const int x = blockIdx.x * PA_THREADS + threadIdx.x;
    const bool active = x < w;
    const int xa = active ? x : w - 1;  // lanes past the last column read valid pixels and add zero
    int xs[PA_K];
#pragma unroll
    for (int i = 0; i < PA_K; ++i) {
        int xx = xa + sx0 + i;  // < 2 w + 4 <= 3 w (w >= 8)
        if (xx >= w) xx -= w;
        if (xx >= w) xx -= w;
        xs[i] = xx;
    }
    const int y0 = blockIdx.y * rows;
    const int nrows = min(rows, h - y0), tend = nrows + PA_K - 1;
    int rb = y0 + sy0;  // b row of step t = 0 (y0 < h, sy0 < h)
    if (rb >= h) rb -= h;
    double acc[PA_LAGS], ar[PA_K];
#pragma unroll
    for (int l = 0; l < PA_LAGS; ++l) acc[l] = 0.0;
#pragma unroll
    for (int j = 0; j < PA_K; ++j) ar[j] = 0.0;
    // Step t pairs b row y0 + sy0 + t with a rows y0 + t - j (lag row j); a row y0 + t sits in ring slot
    // t mod 5, so with t0 a multiple of 5 every slot index below is a constant. Whole groups of 5 steps
    // issue their 30 loads before the products; the last partial group checks each step.
    int t0 = 0;
    for (; t0 + PA_K <= tend; t0 += PA_K) {
        float av[PA_K], bv[PA_K][PA_K];
#pragma unroll
        for (int u = 0; u < PA_K; ++u) {
            const int t = t0 + u;
            av[u] = row_ptr(a, pitch, y0 + min(t, nrows - 1))[xa];
            if (!active || t >= nrows) av[u] = 0.0f;
            const float* br = row_ptr(b, pitch, rb);
#pragma unroll
            for (int i = 0; i < PA_K; ++i) bv[u][i] = br[xs[i]];
            if (++rb == h) rb = 0;
        }
#pragma unroll
        for (int u = 0; u < PA_K; ++u) {
            ar[u] = (double)av[u];
#pragma unroll
            for (int i = 0; i < PA_K; ++i) {
                const double bd = (double)bv[u][i];
#pragma unroll
                for (int j = 0; j < PA_K; ++j) acc[j * PA_K + i] = fma(ar[(u - j + PA_K) % PA_K], bd, acc[j * PA_K + i]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < PA_K; ++u) {
        const int t = t0 + u;
        if (t < tend) {
            ar[u] = (active && t < nrows) ? (double)row_ptr(a, pitch, y0 + t)[xa] : 0.0;
            const float* br = row_ptr(b, pitch, rb);
#pragma unroll
            for (int i = 0; i < PA_K; ++i) {
                const double bd = (double)br[xs[i]];
#pragma unroll
                for (int j = 0; j < PA_K; ++j) acc[j * PA_K + i] = fma(ar[(u - j + PA_K) % PA_K], bd, acc[j * PA_K + i]);
            }
            if (++rb == h) rb = 0;
        }
    }
    // block sum: lanes of a wave, then the waves through LDS
    __shared__ double s_part[PA_THREADS / 64][PA_LAGS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int l = 0; l < PA_LAGS; ++l) {
        double v = acc[l];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) s_part[wave][l] = v;
    }
    __syncthreads();
    if (threadIdx.x < PA_LAGS) {
        double v = s_part[0][threadIdx.x];
        for (int k = 1; k < PA_THREADS / 64; ++k) v += s_part[k][threadIdx.x];
        const int chunks = gridDim.x * gridDim.y, chunk = blockIdx.y * gridDim.x + blockIdx.x;
        partial[((size_t)tile * PA_LAGS + threadIdx.x) * chunks + chunk] = v;
    }
}

// c(s) of every lag of the window, then the first maximum in raster order of the window (lag y, then
// lag x). Each lag is summed by a group of G lanes (a power of two up to 64, so that a lane sums a few
// chunks): lane j of the group sums chunks j, j + G, ... in order, the group is combined by a fixed
// butterfly and its first lane's sum is the value. Writes the maximum's lag to best (2 ints, absolute lag)
// and its value. partial is lag-major: [tile][lag][chunk]. One block per frame of the batch (blockIdx.x):
// its partial sums at partial + frame * part_stride, its centre, maximum and value at frame * PA_BSTRIDE
// (ints) and frame * PA_MAX_LEVELS (doubles).
constexpr int PA_ARGMAX_THREADS = 1024;
__global__ __launch_bounds__(PA_ARGMAX_THREADS) void k_cc_argmax(const double* partial, size_t part_stride,
                                                                 int chunks, int group, const int* centre_in,
                                                                 int scale, LagWindow win, int* best,
                                                                 double* best_value) {
    const int fb = blockIdx.x;
    partial += fb * part_stride;
    if (centre_in) centre_in += fb * PA_BSTRIDE;
    best += fb * PA_BSTRIDE;
    best_value += fb * PA_MAX_LEVELS;
    constexpr int waves = PA_ARGMAX_THREADS / 64;
    const int n = win.tiles * win.tiles * PA_LAGS;
    const int side = win.tiles * PA_K;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane & (group - 1), first = lane - g;
    const int per_wave = 64 / group;
    double bv = -INFINITY;
    int bi = 0x7fffffff;
    // every lane of a wave runs the same number of rounds (the shuffles need all of them)
    const int rounds = (n + waves * per_wave - 1) / (waves * per_wave);
    for (int r = 0; r < rounds; ++r) {
        const int i = (r * waves + wave) * per_wave + first / group;
        double v = 0.0;
        if (i < n) {
            const double* pl = partial + (size_t)i * chunks;
#pragma unroll 8
            for (int c = g; c < chunks; c += group) v += pl[c];
        }
        for (int o = group >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
        v = __shfl(v, first);
        if (i < n && g == 0) {
            const int tile = i / PA_LAGS, l = i - tile * PA_LAGS;
            const int ty = tile / win.tiles, tx = tile - ty * win.tiles;
            const int oy = ty * PA_K + l / PA_K, ox = tx * PA_K + l % PA_K;
            const int ri = oy * side + ox;  // raster index within the window
            if (v > bv || (v == bv && ri < bi)) {
                bv = v;
                bi = ri;
            }
        }
    }
    __shared__ double s_v[PA_ARGMAX_THREADS];
    __shared__ int s_i[PA_ARGMAX_THREADS];
    s_v[threadIdx.x] = bv;
    s_i[threadIdx.x] = bi;
    __syncthreads();
    for (int o = PA_ARGMAX_THREADS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const double v2 = s_v[threadIdx.x + o];
            const int i2 = s_i[threadIdx.x + o];
            if (v2 > s_v[threadIdx.x] || (v2 == s_v[threadIdx.x] && i2 < s_i[threadIdx.x])) {
                s_v[threadIdx.x] = v2;
                s_i[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int cx = 0, cy = 0;
        if (centre_in) {
            cx = centre_in[0] * scale;
            cy = centre_in[1] * scale;
        }
        const int ri = s_i[0];
        best[0] = cx - win.radius + ri % side;
        best[1] = cy - win.radius + ri / side;
        *best_value = s_v[0];
    }
}

// Lanes per lag in k_cc_argmax: the smallest power of two that leaves each lane at most 8 chunks.
int argmax_group(int chunks) {
    int g = 1;
    while (g < 64 && g * 8 < chunks) g <<= 1;
    return g;
}

// Final result of each frame of the batch (thread t): the full-resolution maximum s wrapped into
// (-W/2, W/2] x (-H/2, H/2] as the reference does, returned negated (R:PEFStudioDX/PreAlignment.cs:248-258),
// with angle 0 and the correlation value.
__global__ void k_cc_result(const int* best, const double* value, int w, int h, float* out4, int nb) {
    const int t = threadIdx.x;
    if (t >= nb) return;
    best += t * PA_BSTRIDE;
    value += t * PA_MAX_LEVELS;
    out4 += 4 * t;
    int x = best[0] % w, y = best[1] % h;
    if (x < 0) x += w;
    if (y < 0) y += h;
    if (x > w / 2) x -= w;
    if (y > h / 2) y -= h;
    out4[0] = (float)(-x);
    out4[1] = (float)(-y);
    out4[2] = 0.0f;
    out4[3] = (float)*value;
}

__global__ void k_zero4(float* out4) {
    if (threadIdx.x < 4) out4[threadIdx.x] = 0.0f;
}

struct Plan {
    int levels;                       // pyramid levels used: 1, 2, ..., 2^(levels-1)
    int lw[PA_MAX_LEVELS], lh[PA_MAX_LEVELS], lpitch[PA_MAX_LEVELS];
    size_t pyr_off[PA_MAX_LEVELS];    // offsets of the downsampled copies (level >= 1) in one pyramid
    size_t pyr_bytes;                 // one pyramid (levels 2, 4, ...)
    LagWindow coarse;
    int rows[PA_MAX_LEVELS], row_chunks[PA_MAX_LEVELS];  // row chunk of the correlation blocks, and their count
    size_t partial_bytes, bytes;      // workspace: the reference's and PA_BATCH tracked frames' pyramids,
                                      // partial sums, maxima (partial_bytes: one frame's partial sums)
};

size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

LagWindow window_of(const Plan& P, int i) {
    // the coarsest level searches the whole radius, the next finer one 15 x 15 lags around twice its
    // maximum, the others 5 x 5
    return i == P.levels - 1 ? P.coarse : (i == P.levels - 2 ? LagWindow{3, 7} : LagWindow{1, 2});
}

// Windows searched at level i: the 5 x 5 windows above full resolution twice (the second centred on the
// first's maximum), every other window once.
int passes_of(const Plan& P, int i) { return i < P.levels - 2 && i >= 1 ? 2 : 1; }

bool make_plan(int w, int h, int max_shift, Plan& P) {
    if (w < 8 || h < 8 || max_shift < 1) return false;
    // coarsest level: the first whose search radius is at most PA_COARSE_RADIUS lags, or the last whose
    // sides stay >= PA_MIN_SIDE
    int L = 1;
    while (L < PA_MAX_LEVELS && ((max_shift + (1 << (L - 1)) - 1) >> (L - 1)) > PA_COARSE_RADIUS &&
           (w >> L) >= PA_MIN_SIDE && (h >> L) >= PA_MIN_SIDE)
        ++L;
    P.levels = L;
    size_t off = 0;
    for (int i = 0; i < L; ++i) {
        P.lw[i] = w >> i;
        P.lh[i] = h >> i;
        P.lpitch[i] = P.lw[i] * 4;
        P.pyr_off[i] = off;
        if (i > 0) off += al256((size_t)P.lw[i] * P.lh[i] * 4);
    }
    P.pyr_bytes = off;
    const int lc = L - 1;
    const int rc = ((max_shift + (1 << lc) - 1) >> lc) + 1;  // coarse radius, one lag of slack
    const int k = std::max(0, (rc - 2 + PA_K - 1) / PA_K);
    P.coarse.tiles = 2 * k + 1;
    P.coarse.radius = (PA_K * P.coarse.tiles - 1) / 2;
    size_t part = 0;
    for (int i = 0; i < L; ++i) {
        const LagWindow win = window_of(P, i);
        const int tiles = win.tiles * win.tiles, cols = isa::div_up(P.lw[i], PA_THREADS);
        // about 1024 (tile, column block, row chunk) blocks (4 waves of 126 VGPRs per SIMD), row chunks of at
        // least PA_MIN_ROWS rows
        const int want = std::max(1, 1024 / (tiles * cols));
        const int rc_max = isa::div_up(P.lh[i], PA_MIN_ROWS);
        P.rows[i] = isa::div_up(P.lh[i], std::min(want, rc_max));
        P.row_chunks[i] = isa::div_up(P.lh[i], P.rows[i]);
        part = std::max(part, (size_t)tiles * cols * P.row_chunks[i] * PA_LAGS * sizeof(double));
    }
    P.partial_bytes = al256(part);
    P.bytes = (1 + PA_BATCH) * P.pyr_bytes + PA_BATCH * P.partial_bytes +
              al256((size_t)PA_BATCH * PA_BSTRIDE * sizeof(int)) + al256((size_t)PA_BATCH * PA_MAX_LEVELS * sizeof(double));
    return true;
}

}  // namespace

namespace isa {

size_t prealign_workspace(int width, int height, int max_shift) {
    Plan P;
    return make_plan(width, height, max_shift, P) ? P.bytes : 0;
}

// frames: host array of `n` device pointers (tracking images, pitch img_pitch bytes, 4-byte aligned).
int prealign(const float* const* frames, int img_pitch, int w, int h, int n, int max_shift, float* shifts4,
             void* workspace, hipStream_t s) {
    Plan P;
    ISA_REQUIRE(make_plan(w, h, max_shift, P), "prealign: bad size or search radius");
    ISA_REQUIRE(img_pitch % 4 == 0 && img_pitch >= w * 4, "prealign: bad image pitch");
    char* ws = static_cast<char*>(workspace);
    char* pyr_ref = ws;
    char* pyr_trk = ws + P.pyr_bytes;  // PA_BATCH pyramids
    double* partial = (double*)(ws + (1 + PA_BATCH) * P.pyr_bytes);
    const size_t part_stride = P.partial_bytes / sizeof(double);
    int* best = (int*)((char*)partial + PA_BATCH * P.partial_bytes);
    double* best_value = (double*)((char*)best + al256((size_t)PA_BATCH * PA_BSTRIDE * sizeof(int)));
    auto level_ptr = [&](char* pyr, const float* full, int i) -> const float* {
        return i == 0 ? full : (const float*)(pyr + P.pyr_off[i]);
    };
    auto build = [&](char* pyr, const float* full) -> int {
        // each level from the one above: a 2 x 2 box of 2^(i-1) boxes is the 2^i box of the same pixels
        for (int i = 1; i < P.levels; ++i)
            ISA_TRY(launch_resize_box(level_ptr(pyr, full, i - 1), i == 1 ? img_pitch : P.lpitch[i - 1],
                                      (float*)(pyr + P.pyr_off[i]), P.lpitch[i], P.lw[i - 1], P.lh[i - 1], 2, s));
        return ISA_OK;
    };
    k_zero4<<<1, 64, 0, s>>>(shifts4);
    ISA_TRY(check_launch("prealign: frame 0"));
    if (n < 2) return ISA_OK;
    ISA_TRY(build(pyr_ref, frames[0]));
    const int lc = P.levels - 1;
    for (int f0 = 1; f0 < n; f0 += PA_BATCH) {
        const int nb = std::min(PA_BATCH, n - f0);
        // the batch's pyramids, one launch per level for all its frames
        for (int i = 1; i < P.levels; ++i) {
            FrameSet src;
            for (int j = 0; j < PA_BATCH; ++j) {
                const int jj = std::min(j, nb - 1);
                src.p[j] = level_ptr(pyr_trk + jj * P.pyr_bytes, frames[f0 + jj], i - 1);
            }
            k_box2_batch<<<dim3(div_up(P.lw[i], 32), div_up(P.lh[i], 8), nb), dim3(32, 8), 0, s>>>(
                src, i == 1 ? img_pitch : P.lpitch[i - 1], (float*)(pyr_trk + P.pyr_off[i]), P.pyr_bytes / sizeof(float),
                P.lpitch[i], P.lw[i], P.lh[i]);
            ISA_TRY(check_launch("prealign: pyramid"));
        }
        for (int i = lc; i >= 0; --i) {
            const LagWindow win = window_of(P, i);
            const float* a = level_ptr(pyr_ref, frames[0], i);
            FrameSet bs;
            for (int j = 0; j < PA_BATCH; ++j)
                bs.p[j] = level_ptr(pyr_trk + std::min(j, nb - 1) * P.pyr_bytes, frames[f0 + std::min(j, nb - 1)], i);
            const int pitch = i == 0 ? img_pitch : P.lpitch[i];
            const int cols = div_up(P.lw[i], PA_THREADS), chunks = cols * P.row_chunks[i];
            for (int pass = 0; pass < passes_of(P, i); ++pass) {
                // centre: none at the coarsest level, twice the coarser maximum, then this level's own
                const int* centre = i == lc ? nullptr : (pass == 0 ? best + 2 * (i + 1) : best + 2 * i);
                const int scale = pass == 0 ? 2 : 1;
                k_cc_rows<<<dim3(cols, P.row_chunks[i], win.tiles * win.tiles * nb), PA_THREADS, 0, s>>>(
                    a, bs, pitch, P.lw[i], P.lh[i], centre, scale, win, P.rows[i], partial, part_stride);
                ISA_TRY(check_launch("prealign: correlation"));
                k_cc_argmax<<<nb, PA_ARGMAX_THREADS, 0, s>>>(partial, part_stride, chunks, argmax_group(chunks),
                                                             centre, scale, win, best + 2 * i, best_value + i);
                ISA_TRY(check_launch("prealign: maximum"));
            }
        }
        k_cc_result<<<1, 64, 0, s>>>(best, best_value, w, h, shifts4 + 4 * f0, nb);
        ISA_TRY(check_launch("prealign: result"));
    }
    return ISA_OK;
}

}  // namespace isa

extern "C" {

size_t isa_prealign_workspace(int width, int height, int max_shift) {
    return isa::prealign_workspace(width, height, max_shift);
}

int isa_prealign(const float* const* frames_dev, int img_pitch, int width, int height, int frames, int max_shift,
                 float* shifts4_dev, void* workspace_dev, isa_stream_t stream) {
    ISA_REQUIRE(frames_dev && shifts4_dev && workspace_dev && frames >= 1 && width > 0 && height > 0,
                "isa_prealign: bad arguments");
    ISA_REQUIRE(reinterpret_cast<uintptr_t>(workspace_dev) % 256 == 0, "isa_prealign: workspace must be 256-byte aligned");
    for (int f = 0; f < frames; ++f) ISA_REQUIRE(frames_dev[f], "isa_prealign: null frame pointer");
    return isa::prealign(frames_dev, img_pitch, width, height, frames, max_shift, shifts4_dev, workspace_dev,
                         isa::to_stream(stream));
}

}  // extern "C"
