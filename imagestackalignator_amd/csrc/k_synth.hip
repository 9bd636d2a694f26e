// Seeded synthetic RGGB bursts for benchmarks and tests: an analytic scene
// (Gaussian blobs, oriented step edges, 1/f band-limited waves) sampled at
// x + g_f + flow_f(x), optional occluding patches, Poisson-Gaussian noise
// (normal approximation), 14-bit u16 output. Not part of the reference.
#include <cmath>
#include <cstdint>
#include <vector>

#include "isa_device.h"
#include "isa_host.h"

namespace {

struct Blob {
    float cx, cy, inv2s2, a[3];
};
struct Edge {
    float px, py, nx, ny, a[3];
};
struct Wave {
    float kx, ky, ph, a[3];
};

struct SceneDev {
    const Blob* blobs;
    const Edge* edges;
    const Wave* waves;
    const float* frame;  // per frame: gx, gy, flow phases (4), occlusion seed offset
    int nb, ne, nw;
    float base[3];
    float flow_max, flow_scale, occ_frac, gain, read, white;
    uint64_t seed;
    int w, h;
};

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline double u01(uint64_t z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

__device__ void scene_rgb(const SceneDev& s, float x, float y, float rgb[3]) {
    for (int c = 0; c < 3; ++c) rgb[c] = s.base[c];
    for (int i = 0; i < s.nb; ++i) {
        const Blob& b = s.blobs[i];
        float dx = x - b.cx, dy = y - b.cy;
        float e = __expf(-(dx * dx + dy * dy) * b.inv2s2);
        for (int c = 0; c < 3; ++c) rgb[c] += b.a[c] * e;
    }
    for (int i = 0; i < s.ne; ++i) {
        const Edge& e = s.edges[i];
        float d = (x - e.px) * e.nx + (y - e.py) * e.ny;
        float st = 1.0f / (1.0f + __expf(-4.0f * d));  // step of ~0.5 px width
        for (int c = 0; c < 3; ++c) rgb[c] += e.a[c] * st;
    }
    for (int i = 0; i < s.nw; ++i) {
        const Wave& v = s.waves[i];
        float sv = __sinf(v.kx * x + v.ky * y + v.ph);
        for (int c = 0; c < 3; ++c) rgb[c] += v.a[c] * sv;
    }
    for (int c = 0; c < 3; ++c) rgb[c] = fminf(fmaxf(rgb[c], 0.0f), 1.0f);
}

__device__ void frame_position(const SceneDev& s, int f, float x, float y, float& px, float& py) {
    const float* fr = s.frame + f * 8;
    float k = 6.283185307f / s.flow_scale;
    float fm = s.flow_max * fr[6];  // fr[6] = 0 for frame 0
    float fx = fm * __sinf(k * x + fr[2]) * __cosf(k * y + fr[3]);
    float fy = fm * __cosf(k * x + fr[4]) * __sinf(k * y + fr[5]);
    px = x + fr[0] + fx;
    py = y + fr[1] + fy;
}

// Frames other than 0 carry occluding patches: a fraction occ_frac of the 32 x 32 pixel tiles of the frame's
// pixel grid get a random offset in [-0.3, 0.3] on every channel.
__device__ float occluded(const SceneDev& s, int f, int x, int y, float v) {
    if (f == 0 || s.occ_frac <= 0.0f) return v;
    const int tx = x >> 5, ty = y >> 5;
    const uint64_t hsh = mix64(s.seed ^ mix64(0x0CC1ull + (uint64_t)f * 1000003ull + (uint64_t)ty * 7919ull + tx));
    if (u01(hsh) < s.occ_frac) v = fminf(fmaxf(v + (float)(u01(mix64(hsh)) - 0.5) * 0.6f, 0.0f), 1.0f);
    return v;
}

__global__ void k_synth(SceneDev s, uint16_t* raw, int f) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= s.w || y >= s.h) return;
    float px, py;
    frame_position(s, f, (float)x, (float)y, px, py);
    float rgb[3];
    scene_rgb(s, px, py, rgb);
    int c = (y & 1) == 0 ? ((x & 1) == 0 ? 0 : 1) : ((x & 1) == 0 ? 1 : 2);  // RGGB
    float v = rgb[c];
    v = occluded(s, f, x, y, v);
    float dn = v * s.white;
    uint64_t base = mix64(s.seed ^ 0xA5A5ull) + ((uint64_t)f * s.h + y) * (uint64_t)s.w + x;
    double u1 = u01(mix64(base * 2 + 1)), u2 = u01(mix64(base * 2 + 2));
    float g = (float)(sqrt(-2.0 * log(u1 + 1e-300)) * cos(6.283185307179586 * u2));
    float noisy = dn + sqrtf(s.gain * dn + s.read * s.read) * g;
    noisy = rintf(noisy);
    noisy = fminf(fmaxf(noisy, 0.0f), s.white);
    raw[((size_t)f * s.h + y) * s.w + x] = (uint16_t)noisy;
}

__global__ void k_ground_truth(SceneDev s, int f, float* out, int pitch, int ow, int oh, float scale, float ox,
                               float oy) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= ow || y >= oh) return;
    float px, py;
    const float fx = (float)x * scale + ox, fy = (float)y * scale + oy;
    frame_position(s, f, fx, fy, px, py);
    float rgb[3];
    scene_rgb(s, px, py, rgb);
    // the frame's occluders, at the pixel nearest to the sample (a merge onto this frame keeps them)
    const int qx = (int)floorf(fx + 0.5f), qy = (int)floorf(fy + 0.5f);
    for (int c = 0; c < 3; ++c) rgb[c] = occluded(s, f, qx < 0 ? 0 : qx, qy < 0 ? 0 : qy, rgb[c]);
    float* o = row_ptr(out, pitch, y) + x * 3;
    o[0] = rgb[0];
    o[1] = rgb[1];
    o[2] = rgb[2];
}

struct Host {
    std::vector<Blob> blobs;
    std::vector<Edge> edges;
    std::vector<Wave> waves;
    std::vector<float> frame;
};

Host build_scene(const isa_synth_config& c) {
    Host H;
    uint64_t st = c.seed * 0x2545F4914F6CDD1Dull + 17;
    auto next = [&]() { st = mix64(st); return u01(st); };
    const double W = c.width, Hh = c.height;
    for (int i = 0; i < c.n_blobs; ++i) {
        Blob b;
        b.cx = (float)(next() * W);
        b.cy = (float)(next() * Hh);
        double sg = 1.0 + next() * 19.0;
        b.inv2s2 = (float)(1.0 / (2.0 * sg * sg));
        double amp = (next() - 0.4) * 0.6;
        for (int k = 0; k < 3; ++k) b.a[k] = (float)(amp * (0.6 + 0.8 * next()));
        H.blobs.push_back(b);
    }
    for (int i = 0; i < c.n_edges; ++i) {
        Edge e;
        e.px = (float)(next() * W);
        e.py = (float)(next() * Hh);
        double ang = next() * 6.283185307179586;
        e.nx = (float)std::cos(ang);
        e.ny = (float)std::sin(ang);
        double amp = (next() - 0.5) * 0.16;
        for (int k = 0; k < 3; ++k) e.a[k] = (float)(amp * (0.7 + 0.6 * next()));
        H.edges.push_back(e);
    }
    // 1/f: frequencies log-uniform in [1/512, 1/3] cycles/px, amplitude ~ 1/f, total std ~ 0.04
    double norm = 0.0;
    std::vector<double> f(c.n_waves);
    for (int i = 0; i < c.n_waves; ++i) {
        f[i] = std::exp(std::log(1.0 / 512) + next() * (std::log(1.0 / 3) - std::log(1.0 / 512)));
        norm += 0.5 / (f[i] * f[i]);
    }
    double scale = c.n_waves > 0 ? 0.04 / std::sqrt(norm) : 0.0;
    for (int i = 0; i < c.n_waves; ++i) {
        Wave v;
        double ang = next() * 6.283185307179586, k = 6.283185307179586 * f[i];
        v.kx = (float)(k * std::cos(ang));
        v.ky = (float)(k * std::sin(ang));
        v.ph = (float)(next() * 6.283185307179586);
        double amp = scale / f[i];
        for (int q = 0; q < 3; ++q) v.a[q] = (float)(amp * (0.8 + 0.4 * next()));
        H.waves.push_back(v);
    }
    for (int fr = 0; fr < c.frames; ++fr) {
        float gx = fr == 0 ? 0.0f : (float)((next() * 2.0 - 1.0) * c.global_shift_max);
        float gy = fr == 0 ? 0.0f : (float)((next() * 2.0 - 1.0) * c.global_shift_max);
        float p[4];
        for (int k = 0; k < 4; ++k) p[k] = (float)(next() * 6.283185307179586);
        float fm = fr == 0 ? 0.0f : 1.0f;  // frame 0 (the reference) carries no local flow
        H.frame.insert(H.frame.end(), {gx, gy, p[0], p[1], p[2], p[3], fm, 0.0f});
    }
    return H;
}

}  // namespace

namespace {
struct DevScene {
    SceneDev s;
    void* mem = nullptr;
};

int upload_scene(const isa_synth_config& c, const Host& H, DevScene& D, hipStream_t st) {
    size_t nb = H.blobs.size() * sizeof(Blob), ne = H.edges.size() * sizeof(Edge), nw = H.waves.size() * sizeof(Wave),
           nf = H.frame.size() * sizeof(float);
    size_t total = nb + ne + nw + nf + 64;
    ISA_TRY(isa::check_hip(hipMalloc(&D.mem, total), "synth: allocation"));
    char* p = (char*)D.mem;
    int rc = ISA_OK;
    if (nb) rc = isa::check_hip(hipMemcpy(p, H.blobs.data(), nb, hipMemcpyHostToDevice), "synth: upload");
    if (rc == ISA_OK && ne) rc = isa::check_hip(hipMemcpy(p + nb, H.edges.data(), ne, hipMemcpyHostToDevice), "synth: upload");
    if (rc == ISA_OK && nw)
        rc = isa::check_hip(hipMemcpy(p + nb + ne, H.waves.data(), nw, hipMemcpyHostToDevice), "synth: upload");
    if (rc == ISA_OK)
        rc = isa::check_hip(hipMemcpy(p + nb + ne + nw, H.frame.data(), nf, hipMemcpyHostToDevice), "synth: upload");
    if (rc != ISA_OK) {
        (void)hipFree(D.mem);
        D.mem = nullptr;
        return rc;
    }
    (void)st;
    SceneDev& s = D.s;
    s.blobs = (const Blob*)p;
    s.edges = (const Edge*)(p + nb);
    s.waves = (const Wave*)(p + nb + ne);
    s.frame = (const float*)(p + nb + ne + nw);
    s.nb = (int)H.blobs.size();
    s.ne = (int)H.edges.size();
    s.nw = (int)H.waves.size();
    s.base[0] = 0.30f;
    s.base[1] = 0.35f;
    s.base[2] = 0.28f;
    s.flow_max = c.flow_max;
    s.flow_scale = c.flow_scale > 0 ? c.flow_scale : 128.0f;
    s.occ_frac = c.occluded_fraction;
    s.gain = c.gain;
    s.read = c.read_sigma;
    s.white = c.white;
    s.seed = c.seed;
    s.w = c.width;
    s.h = c.height;
    return ISA_OK;
}
}  // namespace

extern "C" {

void isa_default_synth_config(isa_synth_config* c) {
    c->seed = 1234;
    c->width = 1024;
    c->height = 1024;
    c->frames = 8;
    c->global_shift_max = 16.0f;
    c->flow_max = 1.5f;
    c->flow_scale = 128.0f;
    c->occluded_fraction = 0.0f;
    c->gain = 0.5f;
    c->read_sigma = 3.0f;
    c->n_blobs = 200;
    c->n_edges = 50;
    c->n_waves = 64;
    c->white = 16383.0f;
}

int isa_synth_burst(const isa_synth_config* c, uint16_t* raw_dev, float* frame_shift_out, isa_stream_t stream) {
    ISA_REQUIRE(c && raw_dev && c->width > 0 && c->height > 0 && c->frames > 0, "isa_synth_burst: bad arguments");
    Host H = build_scene(*c);
    DevScene D;
    hipStream_t s = isa::to_stream(stream);
    ISA_TRY(upload_scene(*c, H, D, s));
    int rc = ISA_OK;
    for (int f = 0; f < c->frames && rc == ISA_OK; ++f) {
        k_synth<<<dim3(isa::div_up(c->width, 32), isa::div_up(c->height, 8)), dim3(32, 8), 0, s>>>(D.s, raw_dev, f);
        rc = isa::check_launch("isa_synth_burst");
    }
    if (rc == ISA_OK) rc = isa::check_hip(hipStreamSynchronize(s), "isa_synth_burst: sync");
    (void)hipFree(D.mem);
    if (frame_shift_out)
        for (int f = 0; f < c->frames; ++f) {
            frame_shift_out[2 * f] = H.frame[f * 8];
            frame_shift_out[2 * f + 1] = H.frame[f * 8 + 1];
        }
    return rc;
}

int isa_synth_ground_truth(const isa_synth_config* c, int frame, float* out3_dev, int out_pitch, int out_w, int out_h,
                           float scale, float ox, float oy, isa_stream_t stream) {
    ISA_REQUIRE(c && out3_dev && frame >= 0 && frame < c->frames && out_w > 0 && out_h > 0,
                "isa_synth_ground_truth: bad arguments");
    Host H = build_scene(*c);
    DevScene D;
    hipStream_t s = isa::to_stream(stream);
    ISA_TRY(upload_scene(*c, H, D, s));
    k_ground_truth<<<dim3(isa::div_up(out_w, 32), isa::div_up(out_h, 8)), dim3(32, 8), 0, s>>>(
        D.s, frame, out3_dev, out_pitch, out_w, out_h, scale, ox, oy);
    int rc = isa::check_launch("isa_synth_ground_truth");
    if (rc == ISA_OK) rc = isa::check_hip(hipStreamSynchronize(s), "isa_synth_ground_truth: sync");
    (void)hipFree(D.mem);
    return rc;
}

}  // extern "C"
