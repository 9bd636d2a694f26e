// Device-resident burst merge: the C++ mirror of the controller pipeline
// SkipPreAlignment -> TrackPatches -> PrepareAccumulation -> Accumulate -> ApplyWeighting
// (R:PEFStudioDX/ImageStackAlignatorController.cs:1530-1569, 1601-1746, 1795-2002) with the
// PatchTracker / ShiftCollection / OpticalFlow modules folded in. All buffers are allocated
// at creation; a run launches kernels only (no host round trips of images).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "isa_device.h"
#include "isa_host.h"

namespace isa {
// launchers defined in the kernel translation units
int launch_raw_to_gray(const uint16_t* raw, int raw_pitch, float* gray, int gray_pitch, int w, int h,
                       const float black[3], const float white[3], const float camwhite[3], const int cfa[4],
                       int green_only, hipStream_t s);
int launch_gauss1(const float* in, int in_pitch, float* out, int out_pitch, float* tmp, int w, int h,
                  const float* taps, int n, hipStream_t s);
int launch_resize_box(const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int l, hipStream_t s);
bool resize_box248_ok(const float* in, int in_pitch, int w, int h);
int launch_resize_box248(const float* in, int in_pitch, float* o2, int p2, float* o4, int p4, float* o8, int p8, int w,
                         int h, hipStream_t s);
int launch_kernel_param4(float* img, int w, int h, int pitch, float dth, float dtr, float kd, float kn, float ks,
                         float kh, hipStream_t s);
size_t highpass_workspace(int w, int h, int clear_axis, float hp, float hps);
int highpass(float* img, int pitch, int w, int h, int clear_axis, float hp, float hps, void* ws, bool tables_ready,
             float value_bound, hipStream_t s);
int launch_track_level(const float* const* frames, int img_pitch, int w, int h, int tile, int M, int cx, int cy,
                       const int* pa, const int* pb, const int* pk, int pairs, const float* base, float* shifts,
                       size_t meas_stride, float threshold, float* shift_images, int* decisions, hipStream_t s);
int launch_track_queue(const float* const* frames, int img_pitch, int w, int h, int tile, int M, int cx, int cy,
                       const int* pa, const int* pb, const int* pk, const float* base, float* shifts,
                       size_t meas_stride, float threshold, int* decisions, const int* queue, const int* count,
                       hipStream_t s);
int launch_upsample_all(const float* in, float* out, size_t meas_stride, int m, int ol, int nl, int ocx, int ocy,
                        int ncx, int ncy, int ot, int nt, hipStream_t s);
bool minimize_full_factor(const int* rows, int n, int m, double* L);
int launch_minimize(float* meas, size_t meas_stride, float* one_to_one, int* status, int* iters, const int* rows,
                    const double* lfull, int tiles, int n, int m, hipStream_t s);
int launch_optimal_shifts(float* out, int pitch, const float* r, int n, int cx, int cy, int ref, int trk,
                          hipStream_t s);
int launch_derivatives2(const float* img, int img_pitch, float* ix, float* iy, int pitch, int w, int h, int quant,
                        hipStream_t s);
int lucas_kanade(float* src, const float* tgt, int img_pitch, const float* tiles, int tiles_pitch, int tw, int th,
                 float* flow, int flow_pitch, float* work, int w, int h, int iterations, float bsx, float bsy,
                 float rot, float min_det, int window, int quant, bool box, hipStream_t s, const void* taps,
                 float ill_ratio, hipEvent_t* iter_ev);
size_t tap_table_bytes(int w, int h);
// Pyramid level i of a merge with these parameters runs the matrix-core tile search (otherwise the exact one).
inline bool xc_at(const isa_params& P, int i) {
    if (P.xc_precision == ISA_XC_FP64) return false;
    return P.xc_levels < 0 ? P.tile_size[i] == 32 : ((P.xc_levels >> i) & 1) != 0;
}
int track_chunks(const int* pa, int m, int n, int* out_chunks);
int track_chunks_capacity(int m, int n);
int launch_track_mfma(int precision, const float* const* frames, const unsigned* const* xplanes,
                      const float* const* vbox, const float* const* vrun, int img_pitch, int w, int h, int tile, int M,
                      int cx, int cy, const int* pa, const int* pb, const int* pk, const int* chunks, int n_chunks,
                      const float* base, float* shifts, size_t meas_stride, float threshold, float* shift_images,
                      int* decisions, int* queue, int* count, hipStream_t s);
int vbox_energies(const float* src, int pitch, int w, int h, int T, float* dst, float* runmax, hipStream_t s);
int vbox_run_rows(int h, int T);
int split_planes(int precision, const float* src, int P, int w, int h, unsigned* dst, hipStream_t s);
bool track_mfma_supported(int tile, int M);
int build_tap_tables(void* tables, int w, int h, int quant, hipStream_t s);
int launch_subsample3(const uint16_t* raw, float* out, float max_val, int hw, int hh, int out_pitch, const int cfa[4],
                      hipStream_t s);
size_t prealign_workspace(int width, int height, int max_shift);
int prealign(const float* const* frames, int img_pitch, int w, int h, int n, int max_shift, float* shifts4,
             void* workspace, hipStream_t s);
int launch_robustness(const float* ref, const float* mov, int img_pitch, float* mask, int mask_pitch, const float* flow,
                      int flow_pitch, int fw, int fh, int w, int h, float alpha, float beta, float thr, int quant,
                      hipStream_t s);
int launch_accumulate_frames(float* sum, float* wsum, const float* kern, int w, int h, int out_pitch, int mask_pitch,
                             int flow_pitch, const uint16_t* const* raw, const float* const* mask,
                             const float* const* flow, int nframes, const float white[3], const float black[3],
                             const int cfa[4], hipStream_t s);
int launch_accumulate_sr_frames(float* sum, float* wsum, const float* kern4, int w, int h, int out_pitch,
                                int mask_pitch, int kern_pitch, int flow_pitch, const uint16_t* const* raw,
                                const float* const* mask, const float* const* flow, int nframes, int out_w, int out_h,
                                int ox, int oy, const float white[3], const float black[3], const int cfa[4], int quant,
                                hipStream_t s);
}  // namespace isa

extern "C" int isa_compute_structure_tensor(const float*, const float*, int, float*, int, int, int, isa_stream_t);
extern "C" int isa_gauss_filter_border(const float*, int, float*, int, float*, int, int, int, const float*, int,
                                       isa_stream_t);
extern "C" int isa_compute_kernel_param(float*, int, int, int, float, float, float, float, float, float, isa_stream_t);
extern "C" int isa_erode_border4(const float*, int, float*, int, int, int, int, isa_stream_t);
extern "C" int isa_debayer_full(int, int, const uint16_t*, int, float*, int, const float*, const float*, const float*,
                                const int*, isa_stream_t);
extern "C" int isa_apply_weighting(float*, const float*, const float*, int, int, int, float, isa_stream_t);
extern "C" int isa_gaussian_taps(float, float*, int);

namespace {

__global__ void k_fill(float* p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int fill(float* p, size_t n, float v, hipStream_t s) {
    k_fill<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
    return isa::check_launch("fill");
}

enum Stage { ST_TRACKIMG = 0, ST_TRACK, ST_MINIMIZE, ST_PREPARE, ST_LK, ST_ROBUST, ST_ACCUM, ST_FINAL, ST_COUNT };

}  // namespace

struct isa_merger {
    int w = 0, h = 0, n = 0;
    int ow = 0, oh = 0;  // output size: w x h, or 2w x 2h for the full-frame 2x merge
    isa_params P{};
    int L = 0, mcx = 0, mcy = 0;
    int cx[ISA_MAX_LEVELS]{}, cy[ISA_MAX_LEVELS]{};
    int m_alloc = 0;  // capacity of the measurement tables
    // buffers
    std::vector<void*> allocs;
    float* trk = nullptr;   // n * w * h
    float* lk = nullptr;    // n * w * h
    float* gray = nullptr;  // w * h scratch
    float* gtmp = nullptr;  // w * h scratch (Gaussian fallback / structure tensor)
    float* lev[ISA_MAX_LEVELS]{};  // n * (w/l) * (h/l) for levels with l > 1
    float* shifts = nullptr;       // m * mtiles * 2
    float* shifts_tmp = nullptr;
    float* one_to_one = nullptr;  // tiles0 * n1 * 2
    float* tile_img = nullptr;    // tiles0 * 2
    float* kern = nullptr;        // w * h * (3 or 4)
    float* dx = nullptr;
    float* dy = nullptr;
    float* tensor = nullptr;  // w * h * 3
    float* ref_half = nullptr;
    float* trk_half = nullptr;
    float* mask = nullptr;   // (w/2)(h/2)*4, persistent (border stays 1)
    float* masks = nullptr;  // n * (w/2)(h/2)*4 eroded
    float* flows = nullptr;  // n * w * h * 2
    float* lk_src = nullptr;  // w * h
    float* lk_work = nullptr; // 4 * w * h
    float* sum = nullptr;
    float* wsum = nullptr;
    void* hp_ws = nullptr;
    bool hp_tables = false;
    char* taps = nullptr;  // derivative tap tables (w, h, tex_quant), built at create
    unsigned* xplanes = nullptr;  // packed split planes of every frame at every level (16-bit cross terms)
    size_t xplane_off[ISA_MAX_LEVELS]{};
    float* vbox = nullptr;  // vertical window energies of every frame at every level (MFMA cross terms)
    size_t vbox_off[ISA_MAX_LEVELS]{}, vbox_len[ISA_MAX_LEVELS]{};
    float* vrun = nullptr;  // their run maxima (error bounds of the MFMA search)
    size_t vrun_off[ISA_MAX_LEVELS]{}, vrun_len[ISA_MAX_LEVELS]{};
    int* xc_queue = nullptr;  // (pair, tile) items the MFMA search leaves to the exact search
    int* xc_count = nullptr;  // per level: number of such items
    // device tables: pa, pb, pk (m each), rows (2m), base (L * n * 4 floats), frame pointers (L * n),
    // accumulate pointer arrays (raw n, mask n, flow n), MFMA pair chunks, vbox and run-maximum pointers
    char* tab = nullptr;
    char* tab_host = nullptr;  // pinned staging copy
    size_t tab_bytes = 0;
    void* pre_ws = nullptr;      // global pre-alignment (prealign_max_shift > 0): workspace,
    float* pre_dev = nullptr;    // per-frame (x, y, angle, value) on the device
    float* pre_host = nullptr;   // and their pinned host copy
    hipEvent_t pre_ev[2]{};      // start and end of the pre-alignment
    hipEvent_t upload_done = nullptr;
    hipEvent_t ev[ST_COUNT + 1]{};
    // per level: start, track launch, minimizer launch, end, end of the main track launch (the matrix-core
    // search before its exact-search queue)
    hipEvent_t lev_ev[ISA_MAX_LEVELS][5]{};
    hipEvent_t lk_ev[64][2]{};  // per frame: start and end of the Lucas-Kanade iterations
    float stage_ms[ST_COUNT]{};
    bool have_times = false;
    int ref = 0;
    int m_last = 0;  // measurements (pairs) of the last run
    bool pre_last = false;  // the last run computed its base shifts by the pre-alignment

    ~isa_merger() {
        for (void* p : allocs) (void)hipFree(p);
        if (tab_host) (void)hipHostFree(tab_host);
        if (pre_host) (void)hipHostFree(pre_host);
        for (auto& e : pre_ev)
            if (e) (void)hipEventDestroy(e);
        if (upload_done) (void)hipEventDestroy(upload_done);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& l : lev_ev)
            for (auto& e : l)
                if (e) (void)hipEventDestroy(e);
        for (auto& l : lk_ev)
            for (auto& e : l)
                if (e) (void)hipEventDestroy(e);
    }
    template <class T> int alloc(T*& p, size_t count) {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256)) != hipSuccess) {
            isa::set_error("isa_merger_create: device allocation of %zu bytes failed", count * sizeof(T));
            return ISA_ERR_ALLOC;
        }
        allocs.push_back(q);
        p = (T*)q;
        return ISA_OK;
    }
};

namespace {

struct TabLayout {
    size_t pa, pb, pk, rows, lfull, base, frames, xframes, vframes, rframes, acc_raw, acc_mask, acc_flow, chunks,
        bytes;
};

TabLayout tab_layout(int m, int n, int L) {
    TabLayout t;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t o = 0;
    t.pa = o; o += al(m * sizeof(int));
    t.pb = o; o += al(m * sizeof(int));
    t.pk = o; o += al(m * sizeof(int));
    t.rows = o; o += al(2 * m * sizeof(int));
    t.lfull = o; o += al((size_t)(n - 1) * (n - 1) * sizeof(double));
    t.base = o; o += al((size_t)L * n * 4 * sizeof(float));
    t.frames = o; o += al((size_t)L * n * sizeof(void*));
    t.xframes = o; o += al((size_t)L * n * sizeof(void*));
    t.vframes = o; o += al((size_t)L * n * sizeof(void*));
    t.rframes = o; o += al((size_t)L * n * sizeof(void*));
    t.acc_raw = o; o += al(n * sizeof(void*));
    t.acc_mask = o; o += al(n * sizeof(void*));
    t.acc_flow = o; o += al(n * sizeof(void*));
    t.chunks = o; o += al((size_t)isa::track_chunks_capacity(m, n) * 16 * sizeof(int));
    t.bytes = o;
    return t;
}

// PreAlignmentStore (R:PEFStudioDX/PreAlignmentStore.cs:36-143): per-frame shifts relative to the
// reference frame chosen by the store's rule (the frame minimising the summed shift length).
struct Store {
    std::vector<float> sx, sy, rot;
    int ref = 0;
};

Store make_store(int n, const float* shift, const float* rot_deg, int ref_override) {
    std::vector<double> ix(n, 0), iy(n, 0), iz(n, 0), bx(n, 0), by(n, 0), bz(n, 0);
    for (int i = 0; i < n; ++i) {
        if (shift) { ix[i] = shift[2 * i]; iy[i] = shift[2 * i + 1]; }
        if (rot_deg) iz[i] = rot_deg[i];
    }
    for (int i = 1; i < n; ++i) { bx[i - 1] = ix[i] - ix[i - 1]; by[i - 1] = iy[i] - iy[i - 1]; bz[i - 1] = iz[i] - iz[i - 1]; }
    int best = 0;
    double best_len = std::numeric_limits<double>::max();
    for (int i = 0; i < n - 1; ++i) {
        // total shift when frame i is the reference: -(sum of steps before i) + (sum of steps from i on)
        double lx = 0, ly = 0;
        for (int j = 0; j < i; ++j) { lx -= bx[j]; ly -= by[j]; }
        for (int j = i; j < n - 1; ++j) { lx += bx[j]; ly += by[j]; }
        double d = std::sqrt(lx * lx + ly * ly);
        if (d < best_len) { best_len = d; best = i; }
    }
    Store s;
    s.sx.resize(n); s.sy.resize(n); s.rot.resize(n);
    for (int i = 0; i < n; ++i) {
        double ax = 0, ay = 0, az = 0;
        for (int j = i; j < best; ++j) { ax -= bx[j]; ay -= by[j]; az -= bz[j]; }
        for (int j = best; j < i; ++j) { ax += bx[j]; ay += by[j]; az += bz[j]; }
        s.sx[i] = (float)ax; s.sy[i] = (float)ay; s.rot[i] = (float)az;
    }
    s.ref = ref_override >= 0 ? ref_override : best;
    return s;
}

float store_rotation(const Store& s, int from, int to) {
    float a = s.rot[to] - s.rot[from];
    return a / 180.0f * (float)3.14159265358979323846;
}

int validate(const isa_params& P, int w, int h, int n) {
    ISA_REQUIRE(w >= 16 && h >= 16 && w % 2 == 0 && h % 2 == 0, "merger: width and height must be even and >= 16");
    ISA_REQUIRE(n >= 2 && n <= 64, "merger: 2..64 frames supported");
    ISA_REQUIRE(P.n_levels >= 1 && P.n_levels <= ISA_MAX_LEVELS, "merger: 1..8 pyramid levels");
    for (int i = 0; i < P.n_levels; ++i) {
        ISA_REQUIRE(P.resize_level[i] >= 1 && P.tile_size[i] >= 2 && P.max_shift[i] >= 1 && P.max_shift[i] <= 8,
                    "merger: level %d: resize >= 1, tile >= 2, max shift in [1, 8]", i);
        int lw = w / P.resize_level[i], lh = h / P.resize_level[i];
        ISA_REQUIRE((lw - 2 * P.max_shift[i]) / P.tile_size[i] >= 1 && (lh - 2 * P.max_shift[i]) / P.tile_size[i] >= 1,
                    "merger: level %d has no tile", i);
    }
    ISA_REQUIRE(P.iterations_lk >= 0 && P.window_size_lk >= 1 && P.erode_size >= 1 && P.erode_size <= 41,
                "merger: LK / erode parameters out of range");
    ISA_REQUIRE(P.super_resolution >= 0 && P.super_resolution <= 2, "merger: super_resolution must be 0, 1 or 2");
    ISA_REQUIRE(P.lk_ill_ratio >= 0.0f && P.lk_ill_ratio <= 0.25f, "merger: lk_ill_ratio must be in [0, 0.25]");
    ISA_REQUIRE(P.prealign_max_shift >= 0 && P.prealign_max_shift <= std::min(w, h) / 2,
                "merger: prealign_max_shift must be in [0, min(width, height) / 2]");
    for (int i = 0; i < 4; ++i) ISA_REQUIRE(P.cfa[i] >= 0 && P.cfa[i] <= 2, "merger: CFA entries must be 0, 1 or 2");
    ISA_REQUIRE(P.xc_precision >= ISA_XC_FP64 && P.xc_precision <= ISA_XC_FP32, "merger: unknown xc_precision");
    ISA_REQUIRE(P.xc_levels >= -1, "merger: xc_levels must be a bit mask of levels or -1");
    if (P.xc_precision != ISA_XC_FP64)
        for (int i = 0; i < P.n_levels; ++i)
            if (!isa::track_mfma_supported(P.tile_size[i], P.max_shift[i])) {
                isa::set_error("merger: tile %d with max shift %d is too wide for the MFMA tile search",
                               P.tile_size[i], P.max_shift[i]);
                return ISA_ERR_UNSUPPORTED;
            }
    ISA_REQUIRE(P.clear_axis >= 0, "merger: clear_axis must be >= 0");
    return ISA_OK;
}

}  // namespace

extern "C" {

int isa_merger_create(isa_merger** out, int width, int height, int frames, const isa_params* params) {
    ISA_REQUIRE(out && params, "isa_merger_create: null argument");
    *out = nullptr;
    ISA_TRY(validate(*params, width, height, frames));
    isa_merger* M = new isa_merger();
    M->w = width; M->h = height; M->n = frames; M->P = *params; M->L = params->n_levels;
    const isa_params& P = M->P;
    const size_t np = (size_t)width * height, nh = (size_t)(width / 2) * (height / 2);
    int tiles0 = 0;
    for (int i = 0; i < M->L; ++i) {
        int lw = width / P.resize_level[i], lh = height / P.resize_level[i];
        M->cx[i] = (lw - 2 * P.max_shift[i]) / P.tile_size[i];
        M->cy[i] = (lh - 2 * P.max_shift[i]) / P.tile_size[i];
        M->mcx = std::max(M->mcx, M->cx[i]);
        M->mcy = std::max(M->mcy, M->cy[i]);
    }
    tiles0 = M->cx[0] * M->cy[0];
    const size_t mtiles = (size_t)M->mcx * M->mcy;
    M->m_alloc = frames * (frames - 1) / 2;
    int rc = ISA_OK;
    auto A = [&](auto*& p, size_t cnt) { if (rc == ISA_OK) rc = M->alloc(p, cnt); };
    A(M->trk, (size_t)frames * np);
    A(M->lk, (size_t)frames * np);
    A(M->gray, np);
    A(M->gtmp, np * 3);
    for (int i = 0; i < M->L; ++i)
        if (P.resize_level[i] > 1)
            A(M->lev[i], (size_t)frames * (width / P.resize_level[i]) * (height / P.resize_level[i]));
    A(M->shifts, (size_t)M->m_alloc * mtiles * 2);
    A(M->shifts_tmp, (size_t)M->m_alloc * mtiles * 2);
    A(M->one_to_one, (size_t)tiles0 * (frames - 1) * 2);
    A(M->tile_img, (size_t)tiles0 * 2);
    A(M->kern, np * (P.super_resolution ? 4 : 3));
    A(M->dx, np);
    A(M->dy, np);
    A(M->tensor, np * 3);
    A(M->ref_half, nh * 3);
    A(M->trk_half, nh * 3);
    A(M->mask, nh * 4);
    A(M->masks, (size_t)frames * nh * 4);
    A(M->flows, (size_t)frames * np * 2);
    A(M->lk_src, np);
    A(M->lk_work, 4 * np);
    M->ow = P.super_resolution == 2 ? 2 * width : width;
    M->oh = P.super_resolution == 2 ? 2 * height : height;
    const size_t nout = (size_t)M->ow * M->oh;
    A(M->sum, nout * 3);
    A(M->wsum, nout * 3);
    A(M->taps, isa::tap_table_bytes(width, height));
    if (P.xc_precision == ISA_XC_BF16X3 || P.xc_precision == ISA_XC_FP16X3) {
        size_t tot = 0;
        for (int i = 0; i < M->L; ++i) {
            M->xplane_off[i] = tot;
            if (isa::xc_at(P, i)) tot += (size_t)frames * (width / P.resize_level[i]) * (height / P.resize_level[i]);
        }
        A(M->xplanes, tot);
    }
    if (P.xc_precision != ISA_XC_FP64) {
        size_t tot = 0;
        for (int i = 0; i < M->L; ++i) {
            const int lw = width / P.resize_level[i], lh = height / P.resize_level[i];
            M->vbox_off[i] = tot;
            M->vbox_len[i] = isa::xc_at(P, i) ? (size_t)std::max(lh - P.tile_size[i] + 1, 1) * lw : 0;
            tot += (size_t)frames * M->vbox_len[i];
        }
        A(M->vbox, tot);
        tot = 0;
        for (int i = 0; i < M->L; ++i) {
            const int lw = width / P.resize_level[i], lh = height / P.resize_level[i];
            M->vrun_off[i] = tot;
            M->vrun_len[i] = isa::xc_at(P, i) ? (size_t)std::max(isa::vbox_run_rows(lh, P.tile_size[i]), 1) * lw : 0;
            tot += (size_t)frames * M->vrun_len[i];
        }
        A(M->vrun, tot);
        A(M->xc_queue, (size_t)M->m_alloc * mtiles * 2);
        A(M->xc_count, ISA_MAX_LEVELS);
    }
    if (rc == ISA_OK && P.use_high_pass) {
        char* ws = nullptr;
        A(ws, isa::highpass_workspace(width, height, P.clear_axis, P.high_pass, P.high_pass_sigma));
        M->hp_ws = ws;
    }
    if (rc == ISA_OK && P.prealign_max_shift > 0) {
        char* pw = nullptr;
        A(pw, isa::prealign_workspace(width, height, P.prealign_max_shift));
        M->pre_ws = pw;
        A(M->pre_dev, (size_t)frames * 4);
        if (rc == ISA_OK && hipHostMalloc((void**)&M->pre_host, (size_t)frames * 4 * sizeof(float),
                                          hipHostMallocDefault) != hipSuccess) {
            isa::set_error("isa_merger_create: pinned host allocation failed");
            rc = ISA_ERR_ALLOC;
        }
        for (int j = 0; j < 2 && rc == ISA_OK; ++j) rc = isa::check_hip(hipEventCreate(&M->pre_ev[j]), "event");
    }
    TabLayout tl = tab_layout(M->m_alloc, frames, M->L);
    M->tab_bytes = tl.bytes;
    A(M->tab, tl.bytes);
    if (rc == ISA_OK && hipHostMalloc((void**)&M->tab_host, tl.bytes, hipHostMallocDefault) != hipSuccess) {
        isa::set_error("isa_merger_create: pinned host allocation failed");
        rc = ISA_ERR_ALLOC;
    }
    if (rc == ISA_OK) rc = isa::check_hip(hipEventCreateWithFlags(&M->upload_done, hipEventDisableTiming), "event");
    for (int i = 0; i <= ST_COUNT && rc == ISA_OK; ++i) rc = isa::check_hip(hipEventCreate(&M->ev[i]), "event");
    for (int i = 0; i < M->L && rc == ISA_OK; ++i)
        for (int j = 0; j < 5 && rc == ISA_OK; ++j) rc = isa::check_hip(hipEventCreate(&M->lev_ev[i][j]), "event");
    for (int f = 0; f < frames && rc == ISA_OK; ++f)
        for (int j = 0; j < 2 && rc == ISA_OK; ++j) rc = isa::check_hip(hipEventCreate(&M->lk_ev[f][j]), "event");
    // derivative tap tables depend only on (width, height, tex_quant): built once, on the null stream,
    // and complete before create returns
    if (rc == ISA_OK) rc = isa::build_tap_tables(M->taps, width, height, P.tex_quant, 0);
    if (rc == ISA_OK) rc = isa::check_hip(hipStreamSynchronize(0), "tap tables");
    if (rc == ISA_OK) rc = isa::check_hip(hipEventRecord(M->upload_done, 0), "event");
    if (rc != ISA_OK) {
        delete M;
        return rc;
    }
    *out = M;
    return ISA_OK;
}

int isa_merger_destroy(isa_merger* m) {
    if (!m) return ISA_OK;
    (void)hipDeviceSynchronize();
    delete m;
    return ISA_OK;
}

int isa_merger_run(isa_merger* M, const uint16_t* raw_dev, const float* base_shift, const float* base_rotation,
                   float* out3_dev, int out_pitch, isa_stream_t stream) {
    ISA_REQUIRE(M && raw_dev && out3_dev && out_pitch >= M->ow * 12, "isa_merger_run: bad arguments");
    hipStream_t s = isa::to_stream(stream);
    const isa_params& P = M->P;
    const int w = M->w, h = M->h, n = M->n, n1 = n - 1, hw = w / 2, hh = h / 2;
    const size_t np = (size_t)w * h, nh = (size_t)hw * hh;
    const int wp = w * (int)sizeof(float);
    float taps_track[99], taps_acc[99], taps_st[99];
    int nt_track = isa_gaussian_taps(P.sigma_debayer_tracking, taps_track, 99);
    int nt_acc = isa_gaussian_taps(P.sigma_debayer_accumulation, taps_acc, 99);
    int nt_st = isa_gaussian_taps(P.sigma_structure_tensor, taps_st, 99);
    ISA_REQUIRE(nt_track > 0 && nt_acc > 0 && nt_st > 0, "isa_merger_run: invalid Gaussian sigma");
    // ---- 1. tracking images (DeBayerBWGaussWB twice per frame, sharing the debayer)
    auto tracking_images = [&]() -> int {
        (void)hipEventRecord(M->ev[0], s);
        for (int f = 0; f < n; ++f) {
            ISA_TRY(isa::launch_raw_to_gray(raw_dev + (size_t)f * np, wp / 2, M->gray, wp, w, h, P.black, P.white,
                                            P.camera_white, P.cfa, P.green_channel_only, s));
            ISA_TRY(isa::launch_gauss1(M->gray, wp, M->lk + (size_t)f * np, wp, M->gtmp, w, h, taps_acc, nt_acc, s));
            if (P.use_high_pass) {
                // the grey image is a convex combination of clamped sRGB values: within [0, 1] up to
                // rounding, so 2 bounds it
                ISA_TRY(isa::highpass(M->gray, wp, w, h, P.clear_axis, P.high_pass, P.high_pass_sigma, M->hp_ws,
                                      M->hp_tables, 2.0f, s));
                M->hp_tables = true;
            }
            ISA_TRY(isa::launch_gauss1(M->gray, wp, M->trk + (size_t)f * np, wp, M->gtmp, w, h, taps_track, nt_track,
                                       s));
        }
        (void)hipEventRecord(M->ev[1], s);
        return ISA_OK;
    };
    // ---- 1b. ComputePreAlignment (R:...Controller.cs:1451-1528) when asked and no base shifts are given: the
    // initial shifts come from the tracking images, so they are computed first and read back once
    const bool pre = P.prealign_max_shift > 0 && base_shift == nullptr;
    std::vector<float> pre_shift, pre_rot;
    if (pre) {
        ISA_TRY(tracking_images());
        std::vector<const float*> fptr(n);
        for (int f = 0; f < n; ++f) fptr[f] = M->trk + (size_t)f * np;
        (void)hipEventRecord(M->pre_ev[0], s);
        ISA_TRY(isa::prealign(fptr.data(), wp, w, h, n, P.prealign_max_shift, M->pre_dev, M->pre_ws, s));
        (void)hipEventRecord(M->pre_ev[1], s);
        ISA_TRY(isa::check_hip(hipMemcpyAsync(M->pre_host, M->pre_dev, (size_t)n * 4 * sizeof(float),
                                              hipMemcpyDeviceToHost, s),
                               "isa_merger_run: pre-alignment read-back"));
        ISA_TRY(isa::check_hip(hipStreamSynchronize(s), "isa_merger_run: pre-alignment read-back"));
        pre_shift.resize(2 * n);
        pre_rot.resize(n);
        for (int f = 0; f < n; ++f) {
            pre_shift[2 * f] = M->pre_host[4 * f];
            pre_shift[2 * f + 1] = M->pre_host[4 * f + 1];
            pre_rot[f] = M->pre_host[4 * f + 2];
        }
        base_shift = pre_shift.data();
        base_rotation = pre_rot.data();
    }
    // ---- host tables (pre-alignment store, shift collection), uploaded through pinned memory
    Store st = make_store(n, base_shift, base_rotation, P.reference_index);
    ISA_REQUIRE(st.ref >= 0 && st.ref < n, "isa_merger_run: reference index out of range");
    M->ref = st.ref;
    const int ref = st.ref;
    std::vector<int> pairs(2 * M->m_alloc), rows(2 * M->m_alloc), index((size_t)n * n);
    int m = isa::shift_tables(n, ref, P.tracking_strategy, P.tracking_block_size, pairs.data(), index.data(),
                              rows.data());
    ISA_REQUIRE(m > 0 && m <= M->m_alloc, "isa_merger_run: invalid tracking strategy for %d frames", n);
    M->m_last = m;
    TabLayout tl = tab_layout(M->m_alloc, n, M->L);
    ISA_TRY(isa::check_hip(hipEventSynchronize(M->upload_done), "isa_merger_run: wait"));
    char* th = M->tab_host;
    int* pa = (int*)(th + tl.pa);
    int* pb = (int*)(th + tl.pb);
    int* pk = (int*)(th + tl.pk);
    for (int p = 0; p < m; ++p) {
        pa[p] = pairs[2 * p];
        pb[p] = pairs[2 * p + 1];
        pk[p] = index[(size_t)pa[p] * n + pb[p]];
    }
    std::memcpy(th + tl.rows, rows.data(), 2 * m * sizeof(int));
    // the minimizer's first iteration (all rows active) is the same for every tile: factor it here
    const bool have_factor = isa::minimize_full_factor(rows.data(), n, m, (double*)(th + tl.lfull));
    const int n_chunks = isa::track_chunks(pa, m, n, (int*)(th + tl.chunks));
    std::vector<char> tracked(n, 0);  // frames some pair tracks (the matrix-core search reads their window energies)
    for (int p = 0; p < m; ++p) tracked[pb[p]] = 1;
    float* base = (float*)(th + tl.base);
    const float** fr = (const float**)(th + tl.frames);
    const unsigned** xfr = (const unsigned**)(th + tl.xframes);
    const float** vfr = (const float**)(th + tl.vframes);
    const float** rfr = (const float**)(th + tl.rframes);
    for (int i = 0; i < M->L; ++i) {
        const int l = P.resize_level[i];
        const size_t lsz = (size_t)(w / l) * (h / l);
        for (int f = 0; f < n; ++f) {
            // GetShift(frame) / level and GetRotation(frame) (R:...Controller.cs:1705-1708)
            float gx = st.sx[f] - st.sx[ref], gy = st.sy[f] - st.sy[ref];
            const float rot = store_rotation(st, ref, f);
            base[((size_t)i * n + f) * 4 + 0] = gx / (float)l;
            base[((size_t)i * n + f) * 4 + 1] = gy / (float)l;
            base[((size_t)i * n + f) * 4 + 2] = (float)std::sin((double)rot);
            base[((size_t)i * n + f) * 4 + 3] = (float)std::cos((double)rot);
            fr[(size_t)i * n + f] = l > 1 ? M->lev[i] + (size_t)f * lsz : M->trk + (size_t)f * np;
            xfr[(size_t)i * n + f] = M->xplanes ? M->xplanes + M->xplane_off[i] + (size_t)f * lsz : nullptr;
            vfr[(size_t)i * n + f] = M->vbox ? M->vbox + M->vbox_off[i] + (size_t)f * M->vbox_len[i] : nullptr;
            rfr[(size_t)i * n + f] = M->vrun ? M->vrun + M->vrun_off[i] + (size_t)f * M->vrun_len[i] : nullptr;
        }
    }
    const uint16_t** acc_raw = (const uint16_t**)(th + tl.acc_raw);
    const float** acc_mask = (const float**)(th + tl.acc_mask);
    const float** acc_flow = (const float**)(th + tl.acc_flow);
    {
        int k = 0;
        acc_raw[k] = raw_dev + (size_t)ref * np;
        acc_mask[k] = M->masks + (size_t)ref * nh * 4;
        acc_flow[k] = nullptr;
        ++k;
        for (int f = 0; f < n; ++f) {
            if (f == ref) continue;
            acc_raw[k] = raw_dev + (size_t)f * np;
            acc_mask[k] = M->masks + (size_t)f * nh * 4;
            acc_flow[k] = M->flows + (size_t)f * np * 2;
            ++k;
        }
    }
    ISA_TRY(isa::check_hip(hipMemcpyAsync(M->tab, th, tl.bytes, hipMemcpyHostToDevice, s), "isa_merger_run: upload"));
    ISA_TRY(isa::check_hip(hipEventRecord(M->upload_done, s), "isa_merger_run: event"));
    const int* d_pa = (const int*)(M->tab + tl.pa);
    const int* d_pb = (const int*)(M->tab + tl.pb);
    const int* d_pk = (const int*)(M->tab + tl.pk);
    const int* d_rows = (const int*)(M->tab + tl.rows);
    const double* d_lfull = have_factor ? (const double*)(M->tab + tl.lfull) : nullptr;
    const float* d_base = (const float*)(M->tab + tl.base);
    const int* d_chunks = (const int*)(M->tab + tl.chunks);
    const unsigned* const* d_xframes = (const unsigned* const*)(M->tab + tl.xframes);
    const float* const* d_vframes = (const float* const*)(M->tab + tl.vframes);
    const float* const* d_rframes = (const float* const*)(M->tab + tl.rframes);
    const float* const* d_frames = (const float* const*)(M->tab + tl.frames);

    if (!pre) ISA_TRY(tracking_images());
    // ---- 2. TrackPatches: levels coarse -> fine
    const size_t mtiles = (size_t)M->mcx * M->mcy, mstride = mtiles * 2;
    ISA_TRY(isa::check_hip(hipMemsetAsync(M->shifts, 0, (size_t)m * mstride * sizeof(float), s), "memset"));
    if (M->xc_count)
        ISA_TRY(isa::check_hip(hipMemsetAsync(M->xc_count, 0, ISA_MAX_LEVELS * sizeof(int), s), "memset"));
    // pyramid levels 2, 4 and 8 of every frame in one pass over its tracking image, when they are
    // the only levels above 1 (each at most once) and the frame tiles into 8 x 8 blocks
    float* lev248[3] = {nullptr, nullptr, nullptr};
    bool pyr248 = isa::resize_box248_ok(M->trk, wp, w, h) && (np * sizeof(float)) % 16 == 0;
    for (int i = 0; i < M->L && pyr248; ++i) {
        const int l = P.resize_level[i];
        if (l == 1) continue;
        const int k = l == 2 ? 0 : (l == 4 ? 1 : (l == 8 ? 2 : -1));
        if (k < 0 || lev248[k]) pyr248 = false;
        else lev248[k] = M->lev[i];
    }
    if (pyr248 && (lev248[0] || lev248[1] || lev248[2])) {
        for (int f = 0; f < n; ++f) {
            auto at = [&](int k, int l) { return lev248[k] ? lev248[k] + (size_t)f * (w / l) * (h / l) : nullptr; };
            ISA_TRY(isa::launch_resize_box248(M->trk + (size_t)f * np, wp, at(0, 2), (w / 2) * 4, at(1, 4), (w / 4) * 4,
                                              at(2, 8), (w / 8) * 4, w, h, s));
        }
    }
    for (int i = M->L - 1; i >= 0; --i) {
        const int l = P.resize_level[i], lw = w / l, lh = h / l;
        (void)hipEventRecord(M->lev_ev[i][0], s);
        if (i != M->L - 1) {
            ISA_TRY(isa::launch_upsample_all(M->shifts, M->shifts_tmp, mstride, m, P.resize_level[i + 1], l, M->cx[i + 1],
                                             M->cy[i + 1], M->cx[i], M->cy[i], P.tile_size[i + 1], P.tile_size[i], s));
            std::swap(M->shifts, M->shifts_tmp);
        }
        if (l > 1 && !pyr248)
            for (int f = 0; f < n; ++f)
                ISA_TRY(isa::launch_resize_box(M->trk + (size_t)f * np, wp, M->lev[i] + (size_t)f * lw * lh,
                                               lw * (int)sizeof(float), w, h, l, s));
        const bool xc = isa::xc_at(P, i);
        if (M->xplanes && xc)
            for (int f = 0; f < n; ++f)
                ISA_TRY(isa::split_planes(P.xc_precision,
                                          l > 1 ? M->lev[i] + (size_t)f * lw * lh : M->trk + (size_t)f * np, lw, lw,
                                          lh, M->xplanes + M->xplane_off[i] + (size_t)f * lw * lh, s));
        if (M->vbox && xc)
            for (int f = 0; f < n; ++f) {
                if (!tracked[f]) continue;
                ISA_TRY(isa::vbox_energies(l > 1 ? M->lev[i] + (size_t)f * lw * lh : M->trk + (size_t)f * np,
                                           lw * (int)sizeof(float), lw, lh, P.tile_size[i],
                                           M->vbox + M->vbox_off[i] + (size_t)f * M->vbox_len[i],
                                           M->vrun + M->vrun_off[i] + (size_t)f * M->vrun_len[i], s));
            }
        (void)hipEventRecord(M->lev_ev[i][1], s);
        if (!xc) {
            ISA_TRY(isa::launch_track_level(d_frames + (size_t)i * n, lw * (int)sizeof(float), lw, lh, P.tile_size[i],
                                            P.max_shift[i], M->cx[i], M->cy[i], d_pa, d_pb, d_pk, m,
                                            d_base + (size_t)i * n * 4, M->shifts, mstride, P.matching_threshold,
                                            nullptr, nullptr, s));
            (void)hipEventRecord(M->lev_ev[i][4], s);
        } else {
            ISA_TRY(isa::launch_track_mfma(P.xc_precision, d_frames + (size_t)i * n, d_xframes + (size_t)i * n,
                                           d_vframes + (size_t)i * n, d_rframes + (size_t)i * n,
                                           lw * (int)sizeof(float), lw, lh, P.tile_size[i], P.max_shift[i], M->cx[i],
                                           M->cy[i], d_pa, d_pb, d_pk, d_chunks, n_chunks, d_base + (size_t)i * n * 4,
                                           M->shifts, mstride, P.matching_threshold, nullptr, nullptr, M->xc_queue,
                                           M->xc_count + i, s));
            (void)hipEventRecord(M->lev_ev[i][4], s);
            ISA_TRY(isa::launch_track_queue(d_frames + (size_t)i * n, lw * (int)sizeof(float), lw, lh, P.tile_size[i],
                                            P.max_shift[i], M->cx[i], M->cy[i], d_pa, d_pb, d_pk,
                                            d_base + (size_t)i * n * 4, M->shifts, mstride, P.matching_threshold,
                                            nullptr, M->xc_queue, M->xc_count + i, s));
        }
        (void)hipEventRecord(M->lev_ev[i][2], s);
        ISA_TRY(isa::launch_minimize(M->shifts, mstride, M->one_to_one, nullptr, nullptr, d_rows, d_lfull,
                                     M->cx[i] * M->cy[i], n, m, s));
        (void)hipEventRecord(M->lev_ev[i][3], s);
    }
    (void)hipEventRecord(M->ev[2], s);
    (void)hipEventRecord(M->ev[3], s);
    // ---- 3. PrepareAccumulation: kernel parameters of the reference frame
    const float* lk_ref = M->lk + (size_t)ref * np;
    ISA_TRY(isa::launch_derivatives2(lk_ref, wp, M->dx, M->dy, wp, w, h, P.tex_quant, s));
    ISA_TRY(isa_compute_structure_tensor(M->dx, M->dy, wp, M->tensor, 3 * wp, w, h, s));
    const int kch = P.super_resolution ? 4 : 3;
    if (!P.super_resolution) {
        ISA_TRY(isa_gauss_filter_border(M->tensor, 3 * wp, M->kern, 3 * wp, M->gtmp, w, h, 3, taps_st, nt_st, s));
        ISA_TRY(isa_compute_kernel_param(M->kern, w, h, 3 * wp, P.dth, P.dtr, P.k_detail, P.k_denoise, P.k_stretch,
                                         P.k_shrink, s));
    } else {
        // float4 copy for the SR texture path (4th channel 0)
        ISA_TRY(fill(M->kern, np * 4, 0.0f, s));
        float* k3 = M->lk_work;  // 3 * np floats of scratch
        ISA_TRY(isa_gauss_filter_border(M->tensor, 3 * wp, k3, 3 * wp, M->gtmp, w, h, 3, taps_st, nt_st, s));
        ISA_TRY(isa_compute_kernel_param(k3, w, h, 3 * wp, P.dth, P.dtr, P.k_detail, P.k_denoise, P.k_stretch,
                                         P.k_shrink, s));
        ISA_TRY(isa::check_hip(hipMemcpy2DAsync(M->kern, 16, k3, 12, 12, np, hipMemcpyDeviceToDevice, s), "kern4"));
    }
    float maxv = -std::numeric_limits<float>::max();
    for (int c = 0; c < 3; ++c) maxv = std::max(maxv, P.white[c] + P.black[c]);
    ISA_TRY(isa::launch_subsample3(raw_dev + (size_t)ref * np, M->ref_half, maxv, hw, hh, hw * 12, P.cfa, s));
    ISA_TRY(fill(M->mask, nh * 4, 1.0f, s));
    ISA_TRY(fill(M->masks + (size_t)ref * nh * 4, nh * 4, 1.0f, s));
    const size_t nout = (size_t)M->ow * M->oh;
    ISA_TRY(isa::check_hip(hipMemsetAsync(M->sum, 0, nout * 3 * sizeof(float), s), "memset"));
    ISA_TRY(isa::check_hip(hipMemsetAsync(M->wsum, 0, nout * 3 * sizeof(float), s), "memset"));
    (void)hipEventRecord(M->ev[4], s);
    // ---- 4. Accumulate: flow per frame (LK), robustness, erosion
    for (int f = 0; f < n; ++f) {
        if (f == ref) continue;
        ISA_TRY(isa::launch_optimal_shifts(M->tile_img, M->cx[0] * 8, M->one_to_one, n, M->cx[0], M->cy[0], ref, f, s));
        float bsx = st.sx[f] - st.sx[ref], bsy = st.sy[f] - st.sy[ref];
        // LucasKanade overwrites its source with the warped image; the merge does not use that image, so
        // the final warp of R:PEFStudioDX/OpticalFlow.cs:162-163 is skipped.
        ISA_TRY(isa::lucas_kanade(M->lk + (size_t)f * np, lk_ref, wp, M->tile_img, M->cx[0] * 8, M->cx[0], M->cy[0],
                                  M->flows + (size_t)f * np * 2, 2 * wp, M->lk_work, w, h, P.iterations_lk, bsx, bsy,
                                  store_rotation(st, ref, f), P.min_det_lk, P.window_size_lk, P.tex_quant, true, s,
                                  M->taps, P.lk_ill_ratio, M->lk_ev[f]));
    }
    (void)hipEventRecord(M->ev[5], s);
    for (int f = 0; f < n; ++f) {
        if (f == ref) continue;
        ISA_TRY(isa::launch_subsample3(raw_dev + (size_t)f * np, M->trk_half, maxv, hw, hh, hw * 12, P.cfa, s));
        ISA_TRY(isa::launch_robustness(M->ref_half, M->trk_half, hw * 12, M->mask, hw * 16, M->flows + (size_t)f * np * 2,
                                       2 * wp, w, h, hw, hh, P.noise_alpha, P.noise_beta, P.threshold_m, P.tex_quant,
                                       s));
        ISA_TRY(isa_erode_border4(M->mask, hw * 16, M->masks + (size_t)f * nh * 4, hw * 16, hw, hh, P.erode_size, s));
    }
    (void)hipEventRecord(M->ev[6], s);
    // ---- 5. accumulation of every frame (reference first, then index order)
    if (!P.super_resolution) {
        ISA_TRY(isa::launch_accumulate_frames(M->sum, M->wsum, M->kern, w, h, 3 * wp, hw * 16, 2 * wp,
                                              (const uint16_t* const*)(M->tab_host + tl.acc_raw),
                                              (const float* const*)(M->tab_host + tl.acc_mask),
                                              (const float* const*)(M->tab_host + tl.acc_flow), n, P.white, P.black,
                                              P.cfa, s));
    } else {
        // accumulateImagesSuperRes: the reference's central-half form (w x h grid, offset (w/2, h/2)) or the
        // full-frame 2x grid (offset 0); the reference frame's flow is the zero flow
        const bool full = P.super_resolution == 2;
        ISA_TRY(isa::launch_accumulate_sr_frames(M->sum, M->wsum, M->kern, w, h, M->ow * 12, hw * 16, kch * wp, 2 * wp,
                                                 (const uint16_t* const*)(M->tab_host + tl.acc_raw),
                                                 (const float* const*)(M->tab_host + tl.acc_mask),
                                                 (const float* const*)(M->tab_host + tl.acc_flow), n, M->ow, M->oh,
                                                 full ? 0 : w / 2, full ? 0 : h / 2, P.white, P.black, P.cfa,
                                                 P.tex_quant, s));
    }
    (void)hipEventRecord(M->ev[7], s);
    // ---- 6. final weighting (1x: fill with the reference debayer, threshold 1; SR: threshold 0)
    if (!P.super_resolution) {
        ISA_TRY(isa_debayer_full(w, h, raw_dev + (size_t)ref * np, wp / 2, out3_dev, out_pitch, P.black, P.white,
                                 P.camera_white, P.cfa, s));
        ISA_TRY(isa_apply_weighting(out3_dev, M->sum, M->wsum, w, h, out_pitch, 1.0f, s));
    } else {
        ISA_TRY(isa::check_hip(hipMemset2DAsync(out3_dev, out_pitch, 0, (size_t)M->ow * 12, M->oh, s), "memset"));
        ISA_TRY(isa_apply_weighting(out3_dev, M->sum, M->wsum, M->ow, M->oh, out_pitch, 0.0f, s));
    }
    (void)hipEventRecord(M->ev[8], s);
    M->have_times = true;
    M->pre_last = pre;
    (void)n1;
    (void)kch;
    return isa::check_launch("isa_merger_run");
}

int isa_merger_reference_index(isa_merger* m) { return m ? m->ref : -1; }

int isa_merger_stage_ms(isa_merger* M, float* ms, int capacity) {
    if (!M || !ms || !M->have_times) return 0;
    if (hipEventSynchronize(M->ev[ST_COUNT]) != hipSuccess) return 0;
    float t[ST_COUNT];
    for (int i = 0; i < ST_COUNT; ++i) {
        t[i] = 0;
        (void)hipEventElapsedTime(&t[i], M->ev[i], M->ev[i + 1]);
    }
    // split the track-patches interval into tile search and minimizer time
    float tr = 0, mn = 0;
    for (int i = 0; i < M->L; ++i) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, M->lev_ev[i][0], M->lev_ev[i][2]);
        (void)hipEventElapsedTime(&b, M->lev_ev[i][2], M->lev_ev[i][3]);
        tr += a;
        mn += b;
    }
    t[ST_TRACK] = tr;
    t[ST_MINIMIZE] = mn;
    int k = 0;
    for (; k < ST_COUNT && k < capacity; ++k) ms[k] = t[k];
    if (k < capacity) {  // stage 8: global pre-alignment (0 when the base shifts came from the caller)
        float pa = 0.0f;
        if (M->pre_last) (void)hipEventElapsedTime(&pa, M->pre_ev[0], M->pre_ev[1]);
        ms[k++] = pa;
    }
    return k;
}

int isa_merger_level_ms(isa_merger* M, int level, float ms[3]) {
    ISA_REQUIRE(M && ms && level >= 0 && level < M->L, "isa_merger_level_ms: bad arguments");
    if (!M->have_times) return ISA_ERR_ARG;
    ISA_TRY(isa::check_hip(hipEventSynchronize(M->ev[ST_COUNT]), "isa_merger_level_ms: sync"));
    for (int j = 0; j < 3; ++j) {
        ms[j] = 0;
        (void)hipEventElapsedTime(&ms[j], M->lev_ev[level][j], M->lev_ev[level][j + 1]);
    }
    return ISA_OK;
}

int isa_merger_kernel_ms(isa_merger* M, int kernel, int level, float* ms, int* launches) {
    ISA_REQUIRE(M && ms && launches && kernel >= 0 && kernel <= 3 && level >= -1 && level < M->L,
                "isa_merger_kernel_ms: bad arguments");
    if (!M->have_times) return ISA_ERR_ARG;
    ISA_TRY(isa::check_hip(hipEventSynchronize(M->ev[ST_COUNT]), "isa_merger_kernel_ms: sync"));
    float tot = 0.0f;
    int cnt = 0;
    auto el = [](hipEvent_t a, hipEvent_t b) {
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, a, b);
        return t;
    };
    if (kernel == 0) {  // Lucas-Kanade iterations of every non-reference frame
        for (int f = 0; f < M->n; ++f) {
            if (f == M->ref) continue;
            tot += el(M->lk_ev[f][0], M->lk_ev[f][1]);
            cnt += M->P.iterations_lk;
        }
    } else if (kernel == 1 || kernel == 2) {  // tile-search launch / exact-search queue per level
        for (int i = 0; i < M->L; ++i) {
            if (level >= 0 && i != level) continue;
            tot += kernel == 1 ? el(M->lev_ev[i][1], M->lev_ev[i][4]) : el(M->lev_ev[i][4], M->lev_ev[i][2]);
            cnt += kernel == 1 || isa::xc_at(M->P, i) ? 1 : 0;
        }
    } else {  // accumulation of every frame
        tot = el(M->ev[ST_ACCUM], M->ev[ST_ACCUM + 1]);
        cnt = 1;
    }
    *ms = tot;
    *launches = cnt;
    return ISA_OK;
}

int isa_merger_exact_items(isa_merger* M, int level, int* exact_items, int* items) {
    ISA_REQUIRE(M && exact_items && items && level >= 0 && level < M->L, "isa_merger_exact_items: bad arguments");
    if (!M->have_times) return ISA_ERR_ARG;
    *items = M->m_last * M->cx[level] * M->cy[level];
    if (!M->xc_count || !isa::xc_at(M->P, level)) {
        *exact_items = *items;  // the exact search ran on every item
        return ISA_OK;
    }
    ISA_TRY(isa::check_hip(hipEventSynchronize(M->ev[ST_COUNT]), "isa_merger_exact_items: sync"));
    return isa::check_hip(hipMemcpy(exact_items, M->xc_count + level, sizeof(int), hipMemcpyDeviceToHost),
                          "isa_merger_exact_items: copy");
}

int isa_merger_buffer(isa_merger* M, int which, int index, void** ptr, int* pitch, int* width, int* height) {
    ISA_REQUIRE(M && ptr, "isa_merger_buffer: bad arguments");
    const int w = M->w, h = M->h, n = M->n;
    const size_t np = (size_t)w * h, nh = (size_t)(w / 2) * (h / 2);
    auto set = [&](void* p, int pt, int ww, int hh) {
        *ptr = p;
        if (pitch) *pitch = pt;
        if (width) *width = ww;
        if (height) *height = hh;
        return ISA_OK;
    };
    bool fi = index >= 0 && index < n;
    switch (which) {
        case 0: ISA_REQUIRE(fi, "frame index"); return set(M->trk + (size_t)index * np, w * 4, w, h);
        case 1: ISA_REQUIRE(fi, "frame index"); return set(M->lk + (size_t)index * np, w * 4, w, h);
        case 2: return set(M->one_to_one, (n - 1) * 8, n - 1, M->cx[0] * M->cy[0]);
        case 3: ISA_REQUIRE(fi, "frame index"); return set(M->flows + (size_t)index * np * 2, w * 8, w, h);
        case 4: ISA_REQUIRE(fi, "frame index"); return set(M->masks + (size_t)index * nh * 4, (w / 2) * 16, w / 2, h / 2);
        case 5: return set(M->kern, w * 4 * (M->P.super_resolution ? 4 : 3), w, h);
        case 6: return set(M->sum, M->ow * 12, M->ow, M->oh);
        case 7: return set(M->wsum, M->ow * 12, M->ow, M->oh);
        case 8: {
            ISA_REQUIRE(index >= 0 && index < M->L, "level index");
            // shifts of the last run are those of the finest level; only level 0 is retained
            ISA_REQUIRE(index == 0, "only the finest level's minimised shifts are retained");
            return set(M->shifts, (int)((size_t)M->mcx * M->mcy * 8), M->cx[0] * M->cy[0], M->m_alloc);
        }
        case 9: return set(M->trk_half, (w / 2) * 12, w / 2, h / 2);
        case 10:
            ISA_REQUIRE(M->pre_dev, "isa_merger_buffer: no pre-alignment (prealign_max_shift is 0)");
            return set(M->pre_dev, 16, 1, n);
    }
    isa::set_error("isa_merger_buffer: unknown buffer %d", which);
    return ISA_ERR_ARG;
}

}  // extern "C"
