// CPU oracle for the burst-merge hot path of ImageStackAlignator.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU, what the
// reference's CUDA kernels and C# orchestration compute. It is the checker
// for the HIP product path: only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it. The product library never links it.
//
// Every routine is templated on the arithmetic type T:
//   T = float  : "reference semantics" -- the reference's FP32 expressions,
//                evaluated in the same operation order (no FMA contraction;
//                compile with -ffp-contract=off), quirks kept.
//   T = double : the FP64 oracle -- the same algorithm evaluated in double.
// Two stages are defined in double for both instantiations because the
// product computes them that way: the tile SSD search (exact direct SSD,
// canonical summation order) and the ShiftMinimizer least squares.
//
// Image layout: dense row-major, channels interleaved (pitch = width).
//
// Citations R:<file>:<line> are relative to the reference checkout.
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <vector>

namespace ora {

enum { CFA_RED = 0, CFA_GREEN = 1, CFA_BLUE = 2 };
enum { ADDR_CLAMP = 0, ADDR_MIRROR = 1 };

extern int g_threads;

// Split [0, n) into contiguous chunks, one per worker thread.
inline void parallel_for(int n, const std::function<void(int, int)>& body) {
    int nt = std::max(1, std::min(g_threads, n));
    if (nt == 1) { body(0, n); return; }
    std::vector<std::thread> th;
    int chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        int lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=, &body] { body(lo, hi); });
    }
    for (auto& t : th) t.join();
}

// Transcendentals: evaluated in double and rounded once to T. For T = float this
// is the correctly rounded result except in ~2^-28 of cases; the HIP kernels do
// the same, so the two sides agree bit for bit.
inline float exp_r(float x) { return (float)std::exp((double)x); }
inline double exp_r(double x) { return std::exp(x); }
inline float pow_r(float x, float y) { return (float)std::pow((double)x, (double)y); }
inline double pow_r(double x, double y) { return std::pow(x, y); }
inline float atan2_r(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline double atan2_r(double y, double x) { return std::atan2(y, x); }
inline float sin_r(float x) { return (float)std::sin((double)x); }
inline double sin_r(double x) { return std::sin(x); }
inline float cos_r(float x) { return (float)std::cos((double)x); }
inline double cos_r(double x) { return std::cos(x); }

inline int cfa_at(const int cfa[4], int y, int x) { return cfa[(y % 2) * 2 + (x % 2)]; }

// ---------------------------------------------------------------------------
// Texture-unit emulation (normalised coordinates). Linear filtering follows the
// CUDA programming guide's "Texture Fetching" appendix: x = u*W - 0.5,
// i = floor(x), alpha = frac(x); with quant the weight keeps 8 fractional
// bits (rounded to nearest). Mirror = half-sample symmetric reflection.
inline int tex_addr(int i, int n, int mode) {
    if (mode == ADDR_CLAMP) return i < 0 ? 0 : (i >= n ? n - 1 : i);
    int p = 2 * n;
    i %= p;
    if (i < 0) i += p;
    return i < n ? i : p - 1 - i;
}

template <class T> inline T tex_frac(T a, int quant) {
    return quant ? std::nearbyint(a * (T)256) / (T)256 : a;
}

template <class T>
inline T tex_linear(const T* img, int w, int h, int ch, int c, T u, T v, int mode, int quant) {
    T x = u * (T)w - (T)0.5;
    T y = v * (T)h - (T)0.5;
    T fx = std::floor(x), fy = std::floor(y);
    T a = tex_frac(x - fx, quant), b = tex_frac(y - fy, quant);
    int x0 = (int)fx, y0 = (int)fy;
    int xa = tex_addr(x0, w, mode), xb = tex_addr(x0 + 1, w, mode);
    int ya = tex_addr(y0, h, mode), yb = tex_addr(y0 + 1, h, mode);
    T t00 = img[((size_t)ya * w + xa) * ch + c];
    T t10 = img[((size_t)ya * w + xb) * ch + c];
    T t01 = img[((size_t)yb * w + xa) * ch + c];
    T t11 = img[((size_t)yb * w + xb) * ch + c];
    T oma = (T)1 - a, omb = (T)1 - b;
    return oma * omb * t00 + a * omb * t10 + oma * b * t01 + a * b * t11;
}

template <class T>
inline T tex_point(const T* img, int w, int h, int ch, int c, T u, T v, int mode) {
    int x = tex_addr((int)std::floor(u * (T)w), w, mode);
    int y = tex_addr((int)std::floor(v * (T)h), h, mode);
    return img[((size_t)y * w + x) * ch + c];
}

// ---------------------------------------------------------------------------
// Shift-collection tables (R:PEFStudioDX/ShiftCollection.cs:113-496).
enum { STRAT_FULL = 0, STRAT_ONLY_ON_REFERENCE = 1, STRAT_ON_REFERENCE_BLOCK = 2, STRAT_BLOCKS = 3 };

struct ShiftTables {
    int n = 0, m = 0, n1 = 0, ref = 0, strategy = 0, block = 0;
    std::vector<float> A;              // m x n1, column-major (ld m), as CreateShiftMatrix
    std::vector<int> pair_ref, pair_track;  // iteration order of GetShiftPairs
    std::vector<int> index;            // n x n, -1 if the pair is not measured
    // each measurement row k measures sign_k * (x_{b_k} - x_{a_k}) where
    // x_j = sum_{i<j} r_i (x_0 = 0); derived from the columns of A
    std::vector<int> row_a, row_b;
};

int shift_count(int n, int strategy, int block);
bool build_shift_tables(int n, int ref, int strategy, int block_in, ShiftTables& t);

}  // namespace ora
