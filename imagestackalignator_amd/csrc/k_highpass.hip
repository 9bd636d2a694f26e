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
#include <cmath>

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

// tw[x][2k] = cos(2 pi k x / W), tw[x][2k+1] = sin(2 pi k x / W), exact argument reduction.
__global__ void k_row_twiddles(float* tw, int w, int kxn) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (x >= w) return;
    double a = kTwoPi * (double)(((long long)k * x) % w) / (double)w;
    double s, c;
    sincos(a, &s, &c);
    tw[(size_t)x * 2 * kxn + 2 * k] = (float)c;
    tw[(size_t)x * 2 * kxn + 2 * k + 1] = (float)s;
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

constexpr int kColChunks = 16;

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
                                 int ny, float hp, float hps) {
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
    Y[((size_t)j * kxn + k) * 2] = sr * g;
    Y[((size_t)j * kxn + k) * 2 + 1] = si * g;
}

// Inverse column DFT: V'[y][2k] = Re, V'[y][2k+1] = -Im (so that the inverse row pass is a plain GEMM with Tw^T).
__global__ void k_col_idft(const double* __restrict__ Y, const double* __restrict__ ctab, float* __restrict__ V, int h,
                           int kxn, int ny) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (k >= kxn || y >= h) return;
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < ny; ++j) {
        double c = ctab[((size_t)y * ny + j) * 2], s = ctab[((size_t)y * ny + j) * 2 + 1];
        double yr = Y[((size_t)j * kxn + k) * 2], yi = Y[((size_t)j * kxn + k) * 2 + 1];
        sr += yr * c - yi * s;
        si += yr * s + yi * c;
    }
    V[(size_t)y * 2 * kxn + 2 * k] = (float)sr;
    V[(size_t)y * 2 * kxn + 2 * k + 1] = (float)(-si);
}

struct HpLayout {
    float* part;
    float* tw;
    double* ctab;
    float* U;
    float* V;
    double* P;
    double* Y;
    size_t bytes;
};

HpLayout hp_layout(void* ws, int w, int h, const Band& b) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    char* base = (char*)ws;
    size_t off = 0;
    HpLayout L;
    L.part = (float*)(base + off);
    off += al((size_t)kRowSplit * h * 2 * b.kxn * sizeof(float));
    L.tw = (float*)(base + off);
    off += al((size_t)w * 2 * b.kxn * sizeof(float));
    L.ctab = (double*)(base + off);
    off += al((size_t)h * b.ny * 2 * sizeof(double));
    L.U = (float*)(base + off);
    off += al((size_t)h * 2 * b.kxn * sizeof(float));
    L.V = (float*)(base + off);
    off += al((size_t)h * 2 * b.kxn * sizeof(float));
    L.P = (double*)(base + off);
    off += al((size_t)kColChunks * b.ny * b.kxn * 2 * sizeof(double));
    L.Y = (double*)(base + off);
    off += al((size_t)b.ny * b.kxn * 2 * sizeof(double));
    L.bytes = off;
    return L;
}

}  // namespace

namespace isa {

size_t highpass_workspace(int w, int h, float hp, float hps) {
    Band b = band_of(w, h, hp, hps);
    return hp_layout(nullptr, w, h, b).bytes;
}

// If tables_ready, the twiddle tables at the start of the workspace are reused.
int highpass(float* img, int pitch, int w, int h, float hp, float hps, void* ws, bool tables_ready, hipStream_t s) {
    Band b = band_of(w, h, hp, hps);
    HpLayout L = hp_layout(ws, w, h, b);
    const int n2 = 2 * b.kxn;
    if (!tables_ready) {
        k_row_twiddles<<<dim3(div_up(w, 256), b.kxn), 256, 0, s>>>(L.tw, w, b.kxn);
        k_col_twiddles<<<dim3(div_up(h, 256), b.ny), 256, 0, s>>>(L.ctab, h, b.kyn, b.ny);
    }
    // 1. U = img . Tw  (M = h, N = 2 kxn, K = w)
    const int kc = div_up(div_up(w, kRowSplit), TK) * TK;
    k_sgemm<false, 0><<<dim3(div_up(n2, TN), div_up(h, TM), div_up(w, kc)), 256, 0, s>>>(
        h, n2, w, img, pitch, L.tw, n2, L.part, n2 * (int)sizeof(float), kc, (size_t)h * n2);
    const size_t nu = (size_t)h * n2;
    k_sum_partials<<<(unsigned)((nu + 255) / 256), 256, 0, s>>>(L.part, L.U, nu, div_up(w, kc));
    // 2. band column DFT with the transfer function
    k_col_dft_partial<<<dim3(div_up(b.kxn, 64), b.ny, kColChunks), 64, 0, s>>>(L.U, L.ctab, L.P, h, b.kxn, b.ny);
    k_col_dft_finish<<<dim3(div_up(b.kxn, 64), b.ny), 64, 0, s>>>(L.P, L.Y, w, h, b.kxn, b.kyn, b.ny, hp, hps);
    // 3. inverse column DFT
    k_col_idft<<<dim3(div_up(b.kxn, 64), h), 64, 0, s>>>(L.Y, L.ctab, L.V, h, b.kxn, b.ny);
    // 4. img = clamp(img - V' . Tw^T)  (M = h, N = w, K = 2 kxn)
    k_sgemm<true, 1><<<dim3(div_up(w, TN), div_up(h, TM), 1), 256, 0, s>>>(h, w, n2, L.V, n2 * (int)sizeof(float),
                                                                           L.tw, n2, img, pitch, n2, 0);
    return check_launch("fourier_highpass");
}

}  // namespace isa

extern "C" {

size_t isa_fourier_highpass_workspace(int width, int height, float high_pass, float high_pass_sigma) {
    if (width <= 0 || height <= 0) return 0;
    return isa::highpass_workspace(width, height, high_pass, high_pass_sigma);
}

int isa_fourier_highpass(float* img_dev, int pitch, int width, int height, int clear_axis, float high_pass,
                         float high_pass_sigma, void* workspace_dev, isa_stream_t stream) {
    ISA_REQUIRE(width > 0 && height > 0 && img_dev && workspace_dev, "isa_fourier_highpass: bad arguments");
    ISA_REQUIRE(high_pass >= 0 && high_pass_sigma >= 0 && (double)high_pass + high_pass_sigma < 0.25,
                "isa_fourier_highpass: high_pass + high_pass_sigma must lie in [0, 0.25)");
    if (clear_axis != 0) {
        isa::set_error("isa_fourier_highpass: clear_axis != 0 is not supported");
        return ISA_ERR_UNSUPPORTED;
    }
    return isa::highpass(img_dev, pitch, width, height, high_pass, high_pass_sigma, workspace_dev, false,
                         isa::to_stream(stream));
}

}  // extern "C"
