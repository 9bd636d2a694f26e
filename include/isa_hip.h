/*
 * isa_hip.h -- C ABI of the MI355X (gfx950) burst-merge library libisa_hip.so.
 *
 * It replaces the PTX-loading kernel wrappers of ImageStackAlignator
 * (R:PEFStudioDX/BasicKernels.cs, DeBayerKernels.cs, OpticalFlowKernels.cs,
 * RobustnessModellKernel.cs, ShiftMinimizerKernels.cs), the NPP / cuFFT /
 * cuBLAS calls their host modules make on the merge path, and the host
 * modules themselves (PatchTracker, ShiftCollection, OpticalFlow and the
 * controller's pipeline methods) for callers that want the whole path.
 *
 * Conventions
 *  - Every pointer argument named *_dev is device memory owned by the caller.
 *    No per-kernel entry point allocates device memory or waits for the
 *    device: scratch comes in a caller-owned workspace (with a *_workspace()
 *    size query) and host tables travel in kernel arguments. The exceptions
 *    are the calls that return values in host memory, which synchronise the
 *    stream: isa_merge_forward (loss), isa_merge_backward (dtheta),
 *    isa_merge_refine_step, isa_merger_stage_ms / _level_ms / _exact_items,
 *    isa_merger_run with the pre-alignment (it reads the initial shifts back),
 *    isa_synth_burst and isa_synth_ground_truth (which also stage their scene
 *    tables in memory of their own). isa_merger_create allocates every buffer
 *    of the whole-path merger once, as the reference's controller did.
 *    Pitches are in BYTES. float2/3/4 images are interleaved (x, y[, z[, w]])
 *    per pixel.
 *  - `cfa` replaces the reference's module constant c_cfaPattern: four ints,
 *    row-major 2x2, values 0 = red, 1 = green, 2 = blue
 *    (R:Kernels/DeBayerKernels.cu:28-41; set via the BayerPattern setters,
 *    R:PEFStudioDX/DeBayerKernels.cs:53-64).
 *  - `stream` is a hipStream_t (0 = the null stream). Launches are
 *    asynchronous; the reference launched synchronously.
 *  - Return value: ISA_OK (0) or an error code; isa_last_error() gives the
 *    message of the last failure on the calling thread. The reference threw
 *    (ManagedCuda CudaException / ArgumentOutOfRangeException); a binding
 *    turns a non-zero code into its exception.
 *  - Texture fetches of the reference are evaluated in-kernel with the same
 *    address and filter modes (normalised coordinates; linear filtering with
 *    exact or, with tex_quant = 1, 8-bit fractional weights).
 */
#ifndef ISA_HIP_H
#define ISA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* isa_stream_t;

enum isa_status {
    ISA_OK = 0,
    ISA_ERR_ARG = 1,         /* invalid size / argument (reference: ArgumentOutOfRangeException) */
    ISA_ERR_HIP = 2,         /* HIP runtime error (reference: CudaException) */
    ISA_ERR_UNSUPPORTED = 3, /* combination not implemented by this library */
    ISA_ERR_ALLOC = 4,       /* device allocation failed */
    ISA_ERR_DATA = 5         /* malformed file or pixel stream (reference: FileLoadException, ArgumentException,
                                IndexOutOfRangeException or EndOfStreamException while reading a raw file) */
};

enum isa_tracking_strategy { /* R:PEFStudioDX/ShiftCollection.cs:37-47 */
    ISA_STRATEGY_FULL = 0,
    ISA_STRATEGY_ONLY_ON_REFERENCE = 1,
    ISA_STRATEGY_ON_REFERENCE_BLOCK = 2,
    ISA_STRATEGY_BLOCKS = 3
};

/* Arithmetic of the tile cross terms sum r(x) t(x+s). No counterpart in the reference (which uses
 * FP32 cuFFT); FP64 is the exact direct search, the others run the cross terms as a matrix product
 * on the matrix cores with FP32 accumulation. */
enum isa_xc_precision {
    ISA_XC_FP64 = 0,   /* exact direct SSD in FP64 (default) */
    ISA_XC_BF16X3 = 1, /* bf16 hi+lo split operands, three products */
    ISA_XC_FP16X3 = 2, /* fp16 hi+lo split operands with per-tile power-of-two scaling, three products */
    ISA_XC_FP32 = 3    /* fp32 operands */
};

/* Output of isa_color_pipeline. */
enum isa_color_format {
    ISA_COLOR_RGB16 = 0, /* 3 x u16 per pixel, scale 65535: SaveAs16BitTiff */
    ISA_COLOR_BGRA8 = 1  /* 4 x u8 per pixel (B, G, R, 255), scale 255: ProcessColorsAndShow's display image */
};

#define ISA_MAX_LEVELS 8

/* Parameters of the merge path: the controller's DependencyProperties with
 * their defaults (R:PEFStudioDX/ImageStackAlignatorController.cs:443-675) plus
 * the raw-file levels the controller reads from RawFile
 * (R:PefFile/RawFile.cs:124-229). Level 0 is the finest pyramid level. */
typedef struct isa_params {
    float sigma_debayer_tracking;     /* 0.5 */
    float sigma_debayer_accumulation; /* 0.5 */
    float high_pass;                  /* 0.01 */
    float high_pass_sigma;            /* 0.0025 */
    int clear_axis;                   /* 0; > 0 zeroes the high-pass spectrum at kx < clear_axis and |ky| < clear_axis */
    int use_high_pass;                /* 1: tracking images are high-passed (SkipPreAlignment/ComputePreAlignment path) */
    float matching_threshold;         /* 1.0 */
    int n_levels;                     /* 1 */
    int resize_level[ISA_MAX_LEVELS]; /* {1} */
    int tile_size[ISA_MAX_LEVELS];    /* {32} */
    int max_shift[ISA_MAX_LEVELS];    /* {2} */
    int tracking_strategy;            /* ISA_STRATEGY_FULL */
    int tracking_block_size;          /* 5 */
    float sigma_structure_tensor;     /* 1.0 */
    float dth, dtr, k_detail, k_denoise, k_stretch, k_shrink; /* 0.001 0.006 0.25 3.0 4 2 */
    int iterations_lk;                /* 3 */
    int window_size_lk;               /* 11 */
    float min_det_lk;                 /* 0.01 */
    int erode_size;                   /* 5 */
    float threshold_m;                /* 1e5 */
    int super_resolution;             /* 0: 1x merge; 1: reference 2x merge of the central half (W x H output);
                                         2: full-frame 2x merge (2W x 2H output, same per-pixel rule, offset 0) */
    int green_channel_only;           /* 0 */
    float noise_alpha, noise_beta;    /* noise model sigma^2 = alpha x + beta on the [0,1] scale */
    float black[3];                   /* black level per colour */
    float white[3];                   /* white level per colour */
    float camera_white[3];            /* as-shot white balance (1,1,1 = identity) */
    int cfa[4];                       /* Bayer pattern, row-major 2x2 */
    int reference_index;              /* -1: PreAlignmentStore rule (R:PEFStudioDX/PreAlignmentStore.cs:36-110) */
    int tex_quant;                    /* 1 (default): 8-bit texture weights as the texture unit; 0: exact weights */
    int xc_precision;                 /* tile cross terms: ISA_XC_FP64 (default, exact direct SSD) or an MFMA form */
    float lk_ill_ratio;               /* 0.01: Lucas-Kanade windows whose FP32 sums give det < ratio * trace^2
                                         (lambda2 / lambda1 below the ratio) are solved in FP64; 0: the reference's
                                         FP32 closed form everywhere. No counterpart in the reference. */
    int prealign_max_shift;           /* 0 (SkipPreAlignment): base shifts from the caller; > 0: when the caller
                                         passes none, ComputePreAlignment's translation search (isa_prealign) within
                                         this radius in pixels sets them */
    int xc_levels;                    /* with xc_precision other than ISA_XC_FP64: bit i set = pyramid level i runs
                                         the matrix-core search, the other levels the exact one; -1 (default): the
                                         levels with 32-pixel tiles, where the matrix-core search is the faster one
                                         (with 64- and 16-pixel tiles the exact search is as fast or faster). No
                                         counterpart in the reference. */
} isa_params;

void isa_default_params(isa_params* p);
const char* isa_version(void);
const char* isa_last_error(void);
/* Writes the gcnArchName of the current device (e.g. "gfx950:sramecc+:xnack-"). */
int isa_device_arch(char* buf, int len);

/* ------------------------------------------------------------------------
 * Per-kernel entry points (one per reference kernel / library call).
 * --------------------------------------------------------------------- */

/* deBayerGreenKernel, R:Kernels/DeBayerKernels.cu:53-149; wrapper R:PEFStudioDX/DeBayerKernels.cs:31-65. */
int isa_debayer_green(int width, int height, const float* in_dev, int in_pitch, float* out3_dev, int out_pitch,
                      const float black[3], const float scale[3], const int cfa[4], isa_stream_t stream);
/* deBayerRedBlueKernel, R:Kernels/DeBayerKernels.cu:151-231; wrapper R:PEFStudioDX/DeBayerKernels.cs:67-101. */
int isa_debayer_redblue(int width, int height, const float* in_dev, int in_pitch, float* out3_dev, int out_pitch,
                        const float black[3], const float scale[3], const int cfa[4], isa_stream_t stream);
/* DeBayerFullRes as one launch: u16 raw -> float3 decoded image: NPP Convert, green, red/blue, then
 * NPP Div by white/camera_white (R:...Controller.cs:1268-1311). The 2-px border is written as 0. */
int isa_debayer_full(int width, int height, const uint16_t* raw_dev, int raw_pitch, float* out3_dev, int out_pitch,
                     const float black[3], const float white[3], const float camera_white[3], const int cfa[4],
                     isa_stream_t stream);
/* deBayersSubSample3, R:Kernels/DeBayerKernels.cu:242-283; wrapper R:PEFStudioDX/DeBayerKernels.cs:103-132.
 * raw is unpitched (2*dim_x wide). */
int isa_debayer_subsample3(const uint16_t* raw_dev, float* out3_dev, float max_val, int dim_x, int dim_y, int out_pitch,
                           const int cfa[4], isa_stream_t stream);
/* accumulateImages, R:Kernels/DeBayerKernels.cu:288-376; wrapper R:PEFStudioDX/DeBayerKernels.cs:134-175. */
int isa_accumulate_images(const uint16_t* raw_dev, float* sum3_dev, float* weight3_dev, const float* mask4_dev,
                          const float* kernel3_dev, const float* flow2_dev, const float white[3], const float black[3],
                          int dim_x, int dim_y, int out_pitch, int mask_pitch, int flow_pitch, const int cfa[4],
                          isa_stream_t stream);
/* accumulateImages over several frames in one pass (the controller's calls for the reference frame,
 * R:PEFStudioDX/ImageStackAlignatorController.cs:1901, and for every other frame, :1984; kernel
 * R:Kernels/DeBayerKernels.cu:288-376):
 * raw_devs, mask4_devs, flow2_devs are host arrays of `frames` device pointers in accumulation order
 * (a NULL flow is the zero flow); the sums equal calling isa_accumulate_images once per frame in that order. */
int isa_accumulate_frames(const uint16_t* const* raw_devs, const float* const* mask4_devs,
                          const float* const* flow2_devs, int frames, float* sum3_dev, float* weight3_dev,
                          const float* kernel3_dev, const float white[3], const float black[3], int dim_x, int dim_y,
                          int out_pitch, int mask_pitch, int flow_pitch, const int cfa[4], isa_stream_t stream);
/* accumulateImagesSuperRes, R:Kernels/DeBayerKernels.cu:378-468; wrapper R:PEFStudioDX/DeBayerKernels.cs:177-230.
 * kernel4 (clamp) and flow2 (mirror) are dim_x x dim_y textures sampled bilinearly. */
int isa_accumulate_images_superres(const uint16_t* raw_dev, float* sum3_dev, float* weight3_dev, const float* mask4_dev,
                                   const float* kernel4_dev, const float* flow2_dev, const float white[3],
                                   const float black[3], int dim_x, int dim_y, int out_pitch, int mask_pitch,
                                   int kernel_pitch, int flow_pitch, const int cfa[4], int tex_quant,
                                   isa_stream_t stream);
/* accumulateImagesSuperRes (R:Kernels/DeBayerKernels.cu:378-468) over several frames in one pass and for any output
 * grid at half the raw pixel spacing: out_w x out_h pixels whose pixel 0 lies (offset_x, offset_y) output pixels from
 * the frame's origin (output pixel x sits at raw position (x + 0.5 + offset_x) / 2). The reference's form is
 * out = dim (dim_x x dim_y) with offset (dim_x/2, dim_y/2), the central half of the frame
 * (isa_accumulate_images_superres); the full-frame 2x merge is 2 dim_x x 2 dim_y with offset 0. Raw and certainty
 * coordinates are clamped to the frame rows and columns the grid covers. raw_devs, mask4_devs, flow2_devs: host
 * arrays of `frames` device pointers in accumulation order (a NULL flow is the zero flow); the sums equal one
 * single-frame accumulation per frame in that order, taps in the reference's raster order. */
int isa_accumulate_frames_superres(const uint16_t* const* raw_devs, const float* const* mask4_devs,
                                   const float* const* flow2_devs, int frames, float* sum3_dev, float* weight3_dev,
                                   const float* kernel4_dev, const float white[3], const float black[3], int dim_x,
                                   int dim_y, int out_w, int out_h, int offset_x, int offset_y, int out_pitch,
                                   int mask_pitch, int kernel_pitch, int flow_pitch, const int cfa[4], int tex_quant,
                                   isa_stream_t stream);
/* ApplyWeighting, R:Kernels/kernel.cu:316-372; wrapper R:PEFStudioDX/BasicKernels.cs:259-285. */
int isa_apply_weighting(float* inout3_dev, const float* sum3_dev, const float* weight3_dev, int width, int height,
                        int pitch, float threshold, isa_stream_t stream);
/* GammasRGB, R:Kernels/kernel.cu:271-313; wrapper R:PEFStudioDX/BasicKernels.cs:243-258. */
int isa_gamma_srgb(float* img3_dev, int width, int height, int pitch, isa_stream_t stream);
/* The OETF of GammasRGB (NaN -> 0, clamp to [0, 1], sRGB curve) on `count` floats: exact = 1 evaluates the
 * power as pow in double rounded once to float, exact = 0 through the certified fast form the debayer
 * and gamma kernels use (the same floats). */
int isa_srgb_oetf(const float* in_dev, float* out_dev, size_t count, int exact, isa_stream_t stream);
/* NPP ColorToGray (R:...Controller.cs:2404-2408): gray = c0 R + c1 G + c2 B. */
int isa_color_to_gray(const float* in3_dev, int in_pitch, float* out_dev, int out_pitch, int width, int height,
                      const float coef[3], isa_stream_t stream);
/* The colour pipeline after the merge (R:...Controller.cs:2198-2262 SaveAs16BitTiff, 2264-2322
 * ProcessColorsAndShow) in one pass: NPP ThresholdLTGT to [0, camera_white], ColorTwist camera_to_prophoto
 * (3 x 4 row-major, the host's DNGColorSpace.ProPhoto.FromPCS * CameraToPCS), ThresholdLTGT to [0, 1],
 * LUTCubic through the tone curve (lut_n levels lut_x_dev -> lut_y_dev, x increasing, 4..1024; cubic through
 * the four levels around the value), ColorTwist prophoto_to_srgb, Mul 2^exposure, GammasRGB, Mul 65535 or
 * 255, then the crop ROI converted with round-to-nearest-even and saturation into out_dev (crop_width x
 * crop_height, pitch out_pitch bytes) as isa_color_format. The NPP steps' own rounding is not documented: the
 * float operations here run in the written order. */
int isa_color_pipeline(const float* img3_dev, int img_pitch, int width, int height, const float camera_white[3],
                       const float camera_to_prophoto[12], const float* lut_x_dev, const float* lut_y_dev, int lut_n,
                       const float prophoto_to_srgb[12], float exposure, int crop_left, int crop_top, int crop_width,
                       int crop_height, int format, void* out_dev, int out_pitch, isa_stream_t stream);
/* ImageFileDirectory(width, height).SaveAsTiff (R:PefFile/ImageFileDirectory.cs:76-123): the same little-endian
 * baseline TIFF (13 entries in the reference's order, uncompressed 16-bit RGB in one strip, 72 dpi, "MK TIFF"),
 * byte for byte. isa_tiff16_encode writes it to dst (host memory, capacity bytes) and returns its size, or 0
 * when capacity is too small (dst may be NULL to query the size); isa_write_tiff16 writes it to a file.
 * rgb_host: width x height x 3 u16, row-major. */
size_t isa_tiff16_encode(void* dst, size_t capacity, const uint16_t* rgb_host, int width, int height);
int isa_write_tiff16(const char* path, const uint16_t* rgb_host, int width, int height);
/* gaussian_filter_1D (R:...Controller.cs:2351-2382): writes the taps to taps_out (host), returns their count
 * (<= 99) or a negative value when capacity is too small. */
int isa_gaussian_taps(float sigma, float* taps_out, int capacity);
/* NPP FilterGaussBorder with replicate border (R:...Controller.cs:2414, 1855): separable, horizontal then
 * vertical; channels = 1, 3 or 4; taps on the host (<= 99). tmp_dev: width*height*channels floats. */
int isa_gauss_filter_border(const float* in_dev, int in_pitch, float* out_dev, int out_pitch, float* tmp_dev, int width,
                            int height, int channels, const float* taps, int n_taps, isa_stream_t stream);
/* PreAlignment.FourierFilter (R:PEFStudioDX/PreAlignment.cs:268-281 with fourierFilter R:Kernels/kernel.cu:702-779,
 * lp = lps = 1): high-pass in place, then clamp to [0,1]. The transfer function equals 1 outside a low band, so the
 * filter subtracts the low band's part, computed as a partial DFT (row DFT at the band's columns, column DFT at its
 * rows, and their inverses) in FP32 products. clear_axis > 0 also zeroes the bins with kx < clear_axis or
 * |ky / height| * height < clear_axis (float, as the reference tests it, R:Kernels/kernel.cu:773); their part is
 * removed in FP64 from the row DFT at those kx and the column DFT at those ky (cost grows with clear_axis).
 * workspace_dev: isa_fourier_highpass_workspace() bytes (0 for bad sizes or clear_axis < 0). */
size_t isa_fourier_highpass_workspace(int width, int height, int clear_axis, float high_pass, float high_pass_sigma);
int isa_fourier_highpass(float* img_dev, int pitch, int width, int height, int clear_axis, float high_pass,
                         float high_pass_sigma, void* workspace_dev, isa_stream_t stream);
/* The same filter (R:PEFStudioDX/PreAlignment.cs:268-281) for an image the caller bounds: value_bound > 0 promises
 * |img| <= value_bound (a grey image in [0,1] passes 1). With 16-byte aligned rows the two GEMMs of the partial DFT
 * then run on the matrix cores on fp16 hi + lo splits of both operands (three fp16 products per FP32 product, each
 * within about 2^-21 of it); value_bound <= 0, or a bound that could take an operand near the fp16 range, gives
 * isa_fourier_highpass. The merger filters its tracking images this way. Same workspace. */
int isa_fourier_highpass_bounded(float* img_dev, int pitch, int width, int height, int clear_axis, float high_pass,
                                 float high_pass_sigma, float value_bound, void* workspace_dev, isa_stream_t stream);
/* NPP ResizeSqrPixel(1/level, SuperSampling) at an integer factor (R:...Controller.cs:1694-1703). */
int isa_resize_box(const float* in_dev, int in_pitch, float* out_dev, int out_pitch, int width, int height, int level,
                   isa_stream_t stream);
/* The levels 2, 4 and 8 of isa_resize_box in one pass over the input (each output may be NULL). No counterpart in
 * the reference: the merger's pyramid. width and height multiples of 8; input 16-byte aligned; level-2 rows 16-byte
 * and level-4 rows 8-byte aligned. Values identical to isa_resize_box. */
int isa_resize_box_levels(const float* in_dev, int in_pitch, int width, int height, float* out2_dev, int pitch2,
                          float* out4_dev, int pitch4, float* out8_dev, int pitch8, isa_stream_t stream);
/* convertToTilesOverlapBorder, R:Kernels/kernel.cu:171-228; wrapper R:PEFStudioDX/BasicKernels.cs:48-63. */
int isa_convert_to_tiles_border(const float* img_dev, int img_pitch, float* tiles_dev, int width, int height,
                                int max_shift, int tile_size, int tile_count_x, int tile_count_y,
                                const float base_shift[2], float base_rotation, isa_stream_t stream);
/* convertToTilesOverlapPreShift, R:Kernels/kernel.cu:230-288; wrapper R:PEFStudioDX/BasicKernels.cs:31-46. */
int isa_convert_to_tiles_preshift(const float* img_dev, int img_pitch, float* tiles_dev, const float* preshift2_dev,
                                  int preshift_pitch, int width, int height, int max_shift, int tile_size,
                                  int tile_count_x, int tile_count_y, const float base_shift[2], float base_rotation,
                                  isa_stream_t stream);
/* Tile SSD map: replaces cuFFT R2C x2 + conjugateComplexMulKernel + C2R + DivC + squaredSum + boxFilterWithBorderX/Y
 * + normalizedCC (R:PEFStudioDX/PatchTracker.cs:328-348; R:Kernels/kernel.cu:27-169, 394-411). Writes
 * D(s) = sum |r - t(.+s)|^2 over the tile interior for s in [-M,M]^2, [tile][(2M+1)^2], evaluated exactly in FP64. */
int isa_tile_ssd(const float* ref_tiles_dev, const float* track_tiles_dev, float* shift_image_dev, int max_shift,
                 int tile_size, int tile_count, isa_stream_t stream);
/* findMinimum, R:Kernels/kernel.cu:413-546; wrapper R:PEFStudioDX/BasicKernels.cs:202-217. */
int isa_find_minimum(const float* shift_image_dev, float* coords2_dev, int coords_pitch, int max_shift,
                     int tile_count_x, int tile_count_y, float threshold, isa_stream_t stream);
/* NPP Add of the float view (R:PEFStudioDX/PatchTracker.cs:355-358): preshift += coords. */
int isa_add_shifts(float* preshift2_dev, int preshift_pitch, const float* coords2_dev, int coords_pitch,
                   int tile_count_x, int tile_count_y, isa_stream_t stream);
/* UpSampleShifts, R:Kernels/kernel.cu:551-598; wrapper R:PEFStudioDX/BasicKernels.cs:219-234. */
int isa_upsample_shifts(const float* in2_dev, float* out2_dev, int in_pitch, int out_pitch, int old_level,
                        int new_level, int old_count_x, int old_count_y, int new_count_x, int new_count_y,
                        int old_tile_size, int new_tile_size, isa_stream_t stream);
/* Device workspace bytes of isa_track, isa_track_xc and isa_track_xc_ex for one level image size, pitch, tile size,
 * maximum shift and precision (0 on invalid input). The workspace is the caller's, 256-byte aligned, and may be reused
 * by the next call on the same stream. Replaces the buffers PatchTracker allocates in InitForSize / AllocateDeviceMemory
 * (R:PEFStudioDX/PatchTracker.cs:211-253). */
size_t isa_track_workspace(int width, int height, int img_pitch, int tile_size, int max_shift, int xc_precision);
/* PatchTracker.Track for ONE pair (R:PEFStudioDX/PatchTracker.cs:317-360): tile extraction of both images, SSD map,
 * findMinimum and preshift += result, fused; the images are the level images (width x height). Stream-ordered: no
 * allocation, no host synchronisation. workspace_dev: isa_track_workspace(..., ISA_XC_FP64) bytes. */
int isa_track(const float* img_track_dev, const float* img_ref_dev, int img_pitch, float* preshift2_dev,
              int preshift_pitch, int width, int height, int tile_size, int max_shift, const float base_shift_ref[2],
              float base_rotation_ref, const float base_shift_track[2], float base_rotation_track, float threshold,
              void* workspace_dev, isa_stream_t stream);
/* isa_track with the cross terms in the given isa_xc_precision (ISA_XC_FP64 is isa_track itself). No counterpart in
 * the reference: the entry point of the mixed-precision alignment for one pair. workspace_dev:
 * isa_track_workspace(..., xc_precision) bytes. */
int isa_track_xc(const float* img_track_dev, const float* img_ref_dev, int img_pitch, float* preshift2_dev,
                 int preshift_pitch, int width, int height, int tile_size, int max_shift, const float base_shift_ref[2],
                 float base_rotation_ref, const float base_shift_track[2], float base_rotation_track, float threshold,
                 int xc_precision, void* workspace_dev, isa_stream_t stream);
/* isa_track_xc with two outputs for checking the mixed-precision search. No counterpart in the reference.
 * decisions_dev (device, tiles ints, may be NULL): per tile the integer part of findMinimum, 2 * argmin + 1 when the
 * threshold test zeroes the shift, -1 without a minimum. The matrix-core forms bound the error of every lag's value;
 * where the bounds cannot decide the argmin and the threshold test as the exact search would, the tile is searched
 * exactly, so the decisions equal isa_track's. exact_count_dev (device, one int, may be NULL): tiles searched
 * exactly. */
int isa_track_xc_ex(const float* img_track_dev, const float* img_ref_dev, int img_pitch, float* preshift2_dev,
                    int preshift_pitch, int width, int height, int tile_size, int max_shift,
                    const float base_shift_ref[2], float base_rotation_ref, const float base_shift_track[2],
                    float base_rotation_track, float threshold, int xc_precision, int* decisions_dev,
                    int* exact_count_dev, void* workspace_dev, isa_stream_t stream);

/* ShiftCollection tables (R:PEFStudioDX/ShiftCollection.cs:113-496). pairs_out: 2*m ints (reference, toTrack) in
 * GetShiftPairs order; index_out: n*n ints (-1 where unmeasured); rows_out: 2*m ints (a, b): measurement row k
 * measures x_b - x_a with x_j = sum_{i<j} r_i. Any output may be NULL. Returns m (>0) or -ISA_ERR_ARG. */
int isa_shift_tables(int frame_count, int reference_index, int strategy, int block_size, int* pairs_out,
                     int* index_out, int* rows_out);
/* Device workspace bytes of isa_minimize_shifts for a frame count and tracking strategy (0 on invalid input): the
 * measurement rows and the all-rows factor, which ShiftCollection's constructor keeps on the device
 * (R:PEFStudioDX/ShiftCollection.cs:498-586). */
size_t isa_minimize_shifts_workspace(int frame_count, int reference_index, int strategy, int block_size);
/* ShiftCollection.MinimizeCUBLAS (R:PEFStudioDX/ShiftCollection.cs:738-822) in FP64, on the device, all tiles at once
 * (no host round trip per iteration). measured2_dev: [m][tile_count] float2 (measurement-major, pitch
 * meas_pitch bytes per measurement) -- read, then overwritten with A_safe r as the reference's separateShifts does;
 * one_to_one2_dev: [tile][n1] float2 (shiftsOneToOne_d); status_dev (optional): [tile] final status, iters_dev
 * (optional): iterations used. workspace_dev: isa_minimize_shifts_workspace() bytes, 16-byte aligned; its tables are
 * written in stream order on each call (no allocation, no host synchronisation). */
int isa_minimize_shifts(float* measured2_dev, int meas_pitch, float* one_to_one2_dev, int* status_dev, int* iters_dev,
                        int tile_count, int frame_count, int reference_index, int strategy, int block_size,
                        void* workspace_dev, isa_stream_t stream);
/* isa_minimize_shifts with its tables in a fixed device array of the library (48 KiB per GPU, enough for every
 * supported frame count, 2..64) instead of a caller workspace. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * It allocates no device memory and does not synchronise. Every call on one GPU rewrites that GPU's one copy of the
 * tables, so calls that overlap on one GPU (on different streams or threads) race; serialise them. */
int isa_minimize_shifts_noalloc(float* measured2_dev, int meas_pitch, float* one_to_one2_dev, int* status_dev,
                                int* iters_dev, int tile_count, int frame_count, int reference_index, int strategy,
                                int block_size, isa_stream_t stream);
/* getOptimalShifts, R:Kernels/ShiftMinimizerKernels.cu:178-218; wrapper R:PEFStudioDX/ShiftMinimizerKernels.cs:69-96. */
int isa_get_optimal_shifts(float* out2_dev, int out_pitch, const float* one_to_one2_dev, int frame_count,
                           int tile_count_x, int tile_count_y, int reference_image, int image_to_track,
                           isa_stream_t stream);

/* CreateFlowFieldFromTiles, R:Kernels/opticalFlow.cu:47-93; wrapper R:PEFStudioDX/OpticalFlowKernels.cs:58-85.
 * tiles2 is a tex_w x tex_h texture (linear, clamp). */
int isa_create_flow_field_from_tiles(float* flow2_dev, int flow_pitch, const float* tiles2_dev, int tiles_pitch,
                                     int tex_w, int tex_h, int width, int height, const float base_shift[2],
                                     float base_rotation, int tex_quant, isa_stream_t stream);
/* WarpingKernel, R:Kernels/opticalFlow.cu:27-44; wrapper R:PEFStudioDX/OpticalFlowKernels.cs:32-57. */
int isa_warp(const float* src_dev, int src_pitch, float* out_dev, int out_pitch, const float* flow2_dev,
             int flow_pitch, int width, int height, int tex_quant, isa_stream_t stream);
/* ComputeDerivativesKernel, R:Kernels/opticalFlow.cu:96-147; wrapper R:PEFStudioDX/OpticalFlowKernels.cs:86-119.
 * Ix, Iy, Iz share `pitch`; source/target share `img_pitch`. */
int isa_compute_derivatives(const float* src_dev, const float* tgt_dev, int img_pitch, float* ix_dev, float* iy_dev,
                            float* iz_dev, int pitch, int width, int height, int tex_quant, isa_stream_t stream);
/* ComputeDerivatives2Kernel, R:Kernels/opticalFlow.cu:150-185; wrapper R:PEFStudioDX/OpticalFlowKernels.cs:120-144. */
int isa_compute_derivatives2(const float* img_dev, int img_pitch, float* ix_dev, float* iy_dev, int pitch, int width,
                             int height, int tex_quant, isa_stream_t stream);
/* lucasKanadeOptim, R:Kernels/opticalFlow.cu:189-325; wrapper R:PEFStudioDX/OpticalFlowKernels.cs:146-175.
 * Window sums as in the reference (per-pixel (2h+1)^2 loops). */
int isa_lucas_kanade_optim(float* flow2_dev, int flow_pitch, const float* ix_dev, const float* iy_dev,
                           const float* it_dev, int img_pitch, int width, int height, int window_size, float min_det,
                           isa_stream_t stream);
/* OpticalFlow.LucasKanade (R:PEFStudioDX/OpticalFlow.cs:147-164): flow from tiles, `iterations` x {warp, derivatives,
 * LK step}, final warp written over src. work_dev: 4*width*height floats (tmp, Ix, Iy, Iz), the buffers OpticalFlow's
 * constructor allocates (R:PEFStudioDX/OpticalFlow.cs:115-145). No allocation, no host synchronisation. */
int isa_lucas_kanade(float* src_dev, const float* tgt_dev, int img_pitch, const float* tiles2_dev, int tiles_pitch,
                     int tex_w, int tex_h, float* flow2_dev, int flow_pitch, float* work_dev, int width, int height,
                     int iterations, const float base_shift[2], float base_rotation, float min_det, int window_size,
                     int tex_quant, isa_stream_t stream);
/* ComputeStructureTensor, R:Kernels/kernel.cu:600-625; wrapper R:PEFStudioDX/BasicKernels.cs:287-302. */
int isa_compute_structure_tensor(const float* dx_dev, const float* dy_dev, int dxdy_pitch, float* out3_dev,
                                 int out_pitch, int width, int height, isa_stream_t stream);
/* ComputeKernelParam, R:Kernels/kernel.cu:627-700; wrapper R:PEFStudioDX/BasicKernels.cs:304-319. */
int isa_compute_kernel_param(float* kernel3_dev, int width, int height, int pitch, float dth, float dtr,
                             float k_detail, float k_denoise, float k_stretch, float k_shrink, isa_stream_t stream);
/* ComputeRobustnessMask, R:Kernels/RobustnessModell.cu:28-158; wrapper R:PEFStudioDX/RobustnessModellKernel.cs:30-67.
 * width/height: half-resolution grid; flow2: flow_w x flow_h texture (linear, mirror). */
int isa_compute_robustness_mask(const float* ref3_dev, const float* moved3_dev, int img_pitch, float* mask4_dev,
                                int mask_pitch, const float* flow2_dev, int flow_pitch, int flow_w, int flow_h,
                                int width, int height, float alpha, float beta, float threshold_m, int tex_quant,
                                isa_stream_t stream);
/* NPP ErodeBorder, size x size all-set mask, anchor size/2, replicate border (R:...Controller.cs:1976). */
int isa_erode_border4(const float* in4_dev, int in_pitch, float* out4_dev, int out_pitch, int width, int height,
                      int size, isa_stream_t stream);

/* Global pre-alignment, the translation part of PreAlignment.ScanAngles (R:PEFStudioDX/PreAlignment.cs:177-261) as
 * ComputePreAlignment runs it (R:PEFStudioDX/ImageStackAlignatorController.cs:1451-1528): for every frame f >= 1 the
 * lag s maximising the circular cross-correlation sum_x img_0(x) img_f((x + s) mod (W, H)) of the tracking images
 * (the reference's cuFFT R2C, conjugateComplexMulKernel, C2R, / (W H), MaxIndex), searched within |s| <= max_shift
 * per axis, wrapped into (-W/2, W/2] x (-H/2, H/2]. shifts4_dev (device, frames x float4): (-s.x, -s.y, angle 0,
 * correlation value) per frame, frame 0 all zero: the initialShifts the PreAlignmentStore is built from
 * (isa_merger_run's base_shift). The rotation scan is not implemented (angle 0). frames_dev: host array of device
 * pointers; workspace_dev: isa_prealign_workspace() bytes, 256-byte aligned (box pyramids of the reference and of
 * the eight tracked frames searched per launch, about 12 B per pixel). No allocation, no host synchronisation. */
size_t isa_prealign_workspace(int width, int height, int max_shift);
int isa_prealign(const float* const* frames_dev, int img_pitch, int width, int height, int frames, int max_shift,
                 float* shifts4_dev, void* workspace_dev, isa_stream_t stream);

/* ------------------------------------------------------------------------
 * Whole-path entry points (device-resident; the fast path).
 * --------------------------------------------------------------------- */
typedef struct isa_merger isa_merger;

/* Allocates every buffer the path needs for `frames` raw frames of width x height. */
int isa_merger_create(isa_merger** out, int width, int height, int frames, const isa_params* params);
int isa_merger_destroy(isa_merger* m);
/* Runs SkipPreAlignment-style tracking-image preparation, TrackPatches, PrepareAccumulation, Accumulate and the
 * final ApplyWeighting (R:...Controller.cs:1530-1569, 1601-1746, 1795-2002). raw_dev: frames x height x width u16,
 * contiguous. base_shift (2*frames floats, host, may be NULL = 0) and base_rotation (frames floats, host, degrees,
 * may be NULL) are the initial shifts the PreAlignmentStore is built from. With base_shift NULL and
 * params.prealign_max_shift > 0 they are computed as ComputePreAlignment does (isa_prealign on the tracking images,
 * R:...Controller.cs:1451-1528); the run then reads them back once (one stream synchronisation).
 * out3_dev: width x height float3 (pitch out_pitch); 2 width x 2 height with super_resolution = 2. */
int isa_merger_run(isa_merger* m, const uint16_t* raw_dev, const float* base_shift, const float* base_rotation,
                   float* out3_dev, int out_pitch, isa_stream_t stream);
/* Device pointer and geometry of an internal buffer, for inspection:
 *   0 tracking image of frame `index` (high-passed)   1 LK image of frame `index`
 *   2 finest-level one-to-one shifts [tile][n1] float2 3 flow of frame `index` (float2)
 *   4 eroded mask of frame `index` (float4, half res) 5 kernel parameters (float3 / float4 in SR)
 *   6 accumulated sum (float3)                        7 accumulated weight (float3)
 *   8 per-pair shifts of level `index` after minimisation, [m][tiles] float2
 *   9 half-resolution debayer of frame `index` (float3)
 *  10 pre-alignment result of the last run, frames x float4 (x, y, angle, value); needs prealign_max_shift > 0
 * Returns ISA_ERR_ARG for an unknown buffer. */
int isa_merger_buffer(isa_merger* m, int which, int index, void** ptr, int* pitch, int* width, int* height);
/* Reference index chosen by the last run, and per-stage milliseconds of the last run (HIP events): stages are
 * 0 tracking images, 1 track patches, 2 minimizer, 3 prepare accumulation, 4 LK, 5 robustness, 6 accumulate,
 * 7 final weighting, 8 global pre-alignment. Returns the number of stages written. */
int isa_merger_reference_index(isa_merger* m);
int isa_merger_stage_ms(isa_merger* m, float* ms, int capacity);
/* Milliseconds of the last run at pyramid level `level` (0 = finest): ms[0] shift up-sampling and level
 * images, ms[1] the fused tile-search launch (all pairs), ms[2] the ShiftMinimizer launch. */
int isa_merger_level_ms(isa_merger* m, int level, float ms[3]);
/* Device time of the last run spent in one kernel family, from HIP events on the merger's stream, and its number of
 * launches: kernel 0 the Lucas-Kanade iterations (fused warp, derivatives, window sums, solve; every non-reference
 * frame), 1 the tile-search launch at pyramid level `level` (-1: all levels; the matrix-core search for a mixed
 * precision), 2 the exact-search queue after a matrix-core search, 3 the accumulation of every frame. No counterpart
 * in the reference. */
int isa_merger_kernel_ms(isa_merger* m, int kernel, int level, float* ms, int* launches);
/* (pair, tile) items of the last run at pyramid level `level`, and how many of them the exact FP64 search decided
 * (all of them for ISA_XC_FP64; with a matrix-core precision, those whose error bounds left the integer minimum or
 * the threshold test open). No counterpart in the reference. */
int isa_merger_exact_items(isa_merger* m, int level, int* exact_items, int* items);

/* ------------------------------------------------------------------------
 * Differentiable merge (no counterpart in the reference, whose merge is piecewise constant in the
 * flow). A continuous relaxation of robustness + kernel accumulation onto a 2W x 2H output of the
 * whole frame: output pixel y at raw position y/2 - 1/4, base raw pixel b = floor(y/2); per frame
 * the 5x5 taps around b + anchor_f(b) (fixed integers) weighted by exp(-d^T Omega d / 2) with
 * d = tap - (y/2 - 1/4 + flow_f(b)), Omega from ComputeKernelParam on the reference's smoothed
 * structure tensor and theta = {k_detail, k_denoise, D_th, D_tr}; certainty = the robustness mask
 * of frame f at half-res pixel floor(b/2) with the moved 3x3 mean sampled bilinearly at
 * h + flow_f(2h)/2 (no erosion). Loss: sum over the output interior (4-pixel border excluded) of
 * (out - gt)^2. The backward pass recomputes the weights (recompute-in-backward) and is a gather.
 * ------------------------------------------------------------------------ */
typedef struct isa_merge_diff {
    int width, height, frames;  /* raw frame size (even) and count */
    int reference;              /* reference frame index */
    const uint16_t* raw_dev;    /* frames x height x width */
    const float* half_dev;      /* frames x (height/2) x (width/2) x 3: isa_debayer_subsample3 of each frame */
    const float* tensor_dev;    /* height x width x 3: smoothed structure tensor of the reference (a11, a22, a12) */
    const float* anchor_dev;    /* frames x height x width x 2: integer tap centres relative to b (held fixed) */
    float black[3], white[3];   /* raw normalisation per colour */
    int cfa[4];
    float noise_alpha, noise_beta;
    float k_stretch, k_shrink;  /* fixed kernel parameters */
} isa_merge_diff;
/* Device workspace bytes for isa_merge_forward / isa_merge_backward (0 on invalid input). */
size_t isa_merge_diff_workspace(const isa_merge_diff* problem);
/* Forward: out3 (2H x 2W x 3, may be NULL), loss (host, may be NULL; needs gt3). flow: frames x H x W x 2.
 * theta: {k_detail, k_denoise, D_th, D_tr}. */
int isa_merge_forward(const isa_merge_diff* problem, const float* flow_dev, const float theta[4], float* out3_dev,
                      const float* gt3_dev, double* loss, void* workspace, isa_stream_t stream);
/* Backward of the loss: dflow (frames x H x W x 2, device) and dtheta (host, 4 floats). */
int isa_merge_backward(const isa_merge_diff* problem, const float* flow_dev, const float theta[4], const float* gt3_dev,
                       float* dflow_dev, float dtheta[4], void* workspace, isa_stream_t stream);

/* One step of joint refinement of the flow (and optionally theta) by gradient descent with backtracking, the
 * refinement the differentiable merge exists for: the gradient at (flow, theta), then candidates
 * flow - clip(scale lr_flow dflow, +-max_step), each kept within 0.5 px of its anchor, and theta - scale lr_theta
 * dtheta (lr_theta may be NULL: theta fixed), tried at most 9 times with scale halved after each one that raises
 * the loss. The first that does not replaces flow_dev / theta / loss (*accepted = 1); *scale keeps its halvings
 * for the next step. loss (in/out, host): the loss at (flow, theta) on entry. workspace:
 * isa_merge_refine_workspace() bytes (the forward/backward workspace plus the gradient and a candidate flow).
 * Synchronises the stream (losses are read on the host). No counterpart in the reference. */
size_t isa_merge_refine_workspace(const isa_merge_diff* problem);
int isa_merge_refine_step(const isa_merge_diff* problem, float* flow_dev, float theta[4], const float* gt3_dev,
                          float lr_flow, const float lr_theta[4], float max_step, float* scale, double* loss,
                          int* accepted, void* workspace, isa_stream_t stream);

/* ------------------------------------------------------------------------
 * Raw ingest: PEF and DNG files to the u16 Bayer image and its levels.
 * Replaces RawFile / PEFFile / DNGFile (R:PefFile/RawFile.cs:28-279, R:PefFile/PEFFile.cs:28-316,
 * R:PefFile/DNGFile.cs:36-950), which the controller constructs one task per file (".pef" -> PEFFile, else
 * DNGFile; R:PEFStudioDX/ImageStackAlignatorController.cs:1066-1119). The container (TIFF IFDs,
 * SubIFDs, the Exif IFD, the Pentax maker note, DNGPrivateData) is parsed on the host by isa_raw_open; the
 * pixel data (Pentax Huffman stream, lossless-JPEG tiles or strips, uncompressed strips) is decoded on the
 * device by isa_raw_decode, many files per call.
 * --------------------------------------------------------------------- */
enum isa_raw_format { ISA_RAW_AUTO = 0, ISA_RAW_PEF = 1, ISA_RAW_DNG = 2 };
enum isa_raw_compression {
    ISA_RAW_UNCOMPRESSED = 1,      /* TIFF Compression 1 */
    ISA_RAW_LOSSLESS_JPEG = 7,     /* DNG lossless JPEG (SOF3) tiles or strips */
    ISA_RAW_PENTAX_HUFFMAN = 65535 /* PEF: maker-note tag 0x0220 present */
};
/* The RawFile properties (R:PefFile/RawFile.cs:121-229), in R, G, B order where the reference says so. */
typedef struct isa_raw_info {
    int format;                        /* ISA_RAW_PEF or ISA_RAW_DNG */
    int compression;                   /* isa_raw_compression */
    int width, height;                 /* RawWidth, RawHeight */
    int bit_depth;                     /* BitsPerSample */
    int crop_left, crop_top;           /* CropLeft, CropTop */
    int cropped_width, cropped_height; /* CroppedWidth, CroppedHeight */
    int bayer_pattern[4];              /* BayerPattern: 0 red, 1 green, 2 blue, 3 cyan, 4 magenta, 5 yellow, 6 white */
    float white_level[3];              /* WhiteLevel (DNG: after the decode, less the largest black delta) */
    float black_level[3];              /* BlackLevel */
    float white_balance[3];            /* WhiteBalance */
    float color_twist[12];             /* ColorTwist, 3 x 4 row-major */
    int color_twist_is_identity;       /* ColorTwistIsIdentity */
    float roll_angle;                  /* RollAngle (degrees) */
    int roll_angle_present;            /* RollAnglePresent */
    float noise_alpha, noise_beta;     /* NoiseModelAlpha, NoiseModelBeta (0 when the file holds none) */
    int iso;                           /* ISO */
    unsigned exposure_num, exposure_den; /* ExposureTime */
    int orientation;                   /* Orientation (TIFF value, 1 = normal) */
    int max_black_delta;               /* DNG: largest BlackLevelDeltaH + BlackLevelDeltaV over the decoded pixels */
    char make[64];                     /* Make */
    char unique_model[64];             /* UniqueModelName */
    char recording_date[32];           /* RecordingDate as written in the file ("YYYY:MM:DD HH:MM:SS") */
} isa_raw_info;
typedef struct isa_raw_file isa_raw_file;
/* Parses the file held in host memory (file, size bytes) as the PEFFile / DNGFile constructors do (format
 * ISA_RAW_AUTO: DNG when IFD0 holds DNGVersion or UniqueCameraModel, else PEF). Keeps a copy of what the
 * decode needs; the caller's buffer may be freed on return. *out receives a handle for isa_raw_close. */
int isa_raw_open(const void* file, size_t size, int format, isa_raw_file** out);
int isa_raw_get_info(const isa_raw_file* f, isa_raw_info* info);
void isa_raw_close(isa_raw_file* f);
/* Workspace bytes isa_raw_decode needs for these n files. */
int isa_raw_decode_workspace(const isa_raw_file* const* files, int n, size_t* bytes);
/* Decodes n files at once: file_dev[i] is the device copy of file i's bytes (its size as given to isa_raw_open),
 * raw_dev[i] its width x height u16 RawImage (raw_pitch[i] bytes per row; pixels no tile or strip covers are
 * 0). Synchronises the stream and returns ISA_ERR_DATA, naming the file, where the reference's decode throws.
 * PEF Huffman and lossless-JPEG streams are decoded in parallel by self-synchronising chunks (one work-group
 * per stream, or per restart interval), then the predictors are undone by scans. */
int isa_raw_decode(const isa_raw_file* const* files, int n, const void* const* file_dev, void* const* raw_dev,
                   const int* raw_pitch, void* workspace_dev, size_t workspace_bytes, isa_stream_t stream);
/* Of the last isa_raw_decode on this thread: the most chain walks one stream needed before its chunks agreed
 * (1 when every guessed chunk start was right) and the number of chunk re-decodes. No counterpart in the
 * reference (its decode is sequential). */
int isa_raw_decode_stats(int* max_sync_walks, int* chunk_redecodes);

/* ------------------------------------------------------------------------
 * Synthetic bursts (benchmark / test input; not part of the reference).
 * --------------------------------------------------------------------- */
typedef struct isa_synth_config {
    uint64_t seed;
    int width, height, frames;
    float global_shift_max;   /* global shift ~ U(-g, g) px per frame (frame 0 gets 0) */
    float flow_max;           /* local flow amplitude (px) */
    float flow_scale;         /* spatial scale of the local flow (px) */
    float occluded_fraction;  /* fraction of 32x32 tiles with an occluding offset patch (frames != 0) */
    float gain, read_sigma;   /* Poisson-Gaussian noise in DN: var = gain * signal + read_sigma^2 */
    int n_blobs, n_edges, n_waves;
    float white;              /* 16383 for 14-bit */
} isa_synth_config;
void isa_default_synth_config(isa_synth_config* c);
/* Renders the burst (frames x height x width u16, RGGB) on the device. frame_shift_out (host, 2*frames, may be
 * NULL) receives each frame's global shift. */
int isa_synth_burst(const isa_synth_config* c, uint16_t* raw_dev, float* frame_shift_out, isa_stream_t stream);
/* Noise-free RGB rendering of frame `frame`'s scene geometry, with that frame's occluding patches, on an
 * out_w x out_h grid covering the frame with sample spacing `scale` (1 = pixel grid, 0.5 = 2x) and origin
 * offset (ox, oy) in frame pixels. */
int isa_synth_ground_truth(const isa_synth_config* c, int frame, float* out3_dev, int out_pitch, int out_w, int out_h,
                           float scale, float ox, float oy, isa_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* ISA_HIP_H */
