// Library-wide C ABI: version, defaults, error text, device query.
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstdint>
#include <cstring>

#include "isa_host.h"

namespace {
thread_local char g_err[512] = "";

// One piece of a host table carried to the device in the kernel's argument block.
constexpr int kPutWords = 480;  // 1920 bytes of payload per launch
struct PutChunk {
    unsigned w[kPutWords];
};

__global__ void k_put_words(PutChunk c, unsigned* dst, int nwords) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) dst[i] = c.w[i];
}
}  // namespace

int isa::upload_by_args(void* dst_dev, const void* src, size_t bytes, hipStream_t s) {
    ISA_REQUIRE(reinterpret_cast<uintptr_t>(dst_dev) % 4 == 0 && bytes % 4 == 0, "upload: unaligned table");
    const unsigned* in = static_cast<const unsigned*>(src);
    unsigned* out = static_cast<unsigned*>(dst_dev);
    const size_t words = bytes / 4;
    for (size_t o = 0; o < words; o += kPutWords) {
        const int nw = (int)std::min<size_t>(kPutWords, words - o);
        PutChunk c;
        std::memcpy(c.w, in + o, (size_t)nw * 4);
        k_put_words<<<1, 256, 0, s>>>(c, out + o, nw);
        ISA_TRY(check_launch("upload"));
    }
    return ISA_OK;
}

void isa::set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" {

const char* isa_version(void) { return "isa-mi355x 0.1.0 (gfx950)"; }

const char* isa_last_error(void) { return g_err; }

int isa_device_arch(char* buf, int len) {
    ISA_REQUIRE(buf && len > 0, "isa_device_arch: bad arguments");
    int dev = 0;
    ISA_TRY(isa::check_hip(hipGetDevice(&dev), "hipGetDevice"));
    hipDeviceProp_t prop;
    ISA_TRY(isa::check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties"));
    snprintf(buf, (size_t)len, "%s", prop.gcnArchName);
    return ISA_OK;
}

// Defaults of the controller's DependencyProperties (R:PEFStudioDX/ImageStackAlignatorController.cs:443-675)
// and PatchTrackingLevel (R:PEFStudioDX/PatchTrackingLevel.cs:33-38); raw levels for a 14-bit RGGB sensor.
void isa_default_params(isa_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->sigma_debayer_tracking = 0.5f;
    p->sigma_debayer_accumulation = 0.5f;
    p->high_pass = 0.01f;
    p->high_pass_sigma = 0.0025f;
    p->clear_axis = 0;
    p->use_high_pass = 1;
    p->matching_threshold = 1.0f;
    p->n_levels = 1;
    p->resize_level[0] = 1;
    p->tile_size[0] = 32;
    p->max_shift[0] = 2;
    p->tracking_strategy = ISA_STRATEGY_FULL;
    p->tracking_block_size = 5;
    p->sigma_structure_tensor = 1.0f;
    p->dth = 0.001f;
    p->dtr = 0.006f;
    p->k_detail = 0.25f;
    p->k_denoise = 3.0f;
    p->k_stretch = 4.0f;
    p->k_shrink = 2.0f;
    p->iterations_lk = 3;
    p->window_size_lk = 11;
    p->min_det_lk = 0.01f;
    p->erode_size = 5;
    p->threshold_m = 1e5f;
    p->super_resolution = 0;
    p->green_channel_only = 0;
    p->noise_alpha = 0.5f / 16383.0f;
    p->noise_beta = 9.0f / (16383.0f * 16383.0f);
    for (int c = 0; c < 3; ++c) {
        p->black[c] = 0.0f;
        p->white[c] = 16383.0f;
        p->camera_white[c] = 1.0f;
    }
    p->cfa[0] = 0;
    p->cfa[1] = 1;
    p->cfa[2] = 1;
    p->cfa[3] = 2;
    p->reference_index = -1;
    p->tex_quant = 1;
    p->xc_precision = ISA_XC_FP64;
    p->lk_ill_ratio = 0.01f;
    p->prealign_max_shift = 0;
    p->xc_levels = -1;
}

}  // extern "C"
