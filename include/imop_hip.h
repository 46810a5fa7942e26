/*
 * imop_hip.h -- C ABI of the MI355X (gfx950) Bayer demosaic path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int holding
 * an imop::DemosaicError value (include/imop/types.hpp:31-39 of the
 * reference): 0 = Ok, 1 = NullInput, 2 = InvalidDimensions,
 * 3 = InvalidBitDepth, 4 = InvalidPattern, 5 = ImageTooSmall,
 * 6 = InternalError.  As in the reference, any non-zero result means "no
 * output was produced"; imop::demosaic() then falls back to its CPU path
 * (src/algorithms/dispatch.cpp:96-111).
 *
 * Enumerations are the reference's: pattern = imop::BayerPattern
 * (RGGB 0, BGGR 1, GRBG 2, GBRG 3), algorithm = imop::DemosaicAlgorithm
 * (SUPER_FAST 0, HQLI 1, MG 2, L7 3, DFPD 4, AHD 5, AMAZE 6, RCD 7, PRISM 8;
 * every one has a kernel).
 *
 * Raw input layouts (include/imop/pixel_utils.hpp:21-189): bit_depth <= 8 one
 * byte per pixel; bit_depth > 8 unpacked is little-endian u16 per pixel;
 * packed 12-bit is 2 px / 3 B, packed 10-bit 4 px / 5 B, other packed depths
 * an LSB-first bitstream, all over the linear pixel index with no row
 * padding; is_packed = 2 selects the GigE Vision 10-bit packing (2 px / 3 B).
 * Output is interleaved RGB, u8 for bit_depth <= 8, else u16.
 *
 * Black level (black_level/global.cpp:6-24, per_channel.cpp:6-33):
 * black_level_mode 0 = none; 1 = global (offsets[0] for every site);
 * 2 = per channel (offsets = R, Gr, Gb, B).  When enabled the result equals
 * unpack -> process_black_level -> demosaic of the unpacked buffer.
 *
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 */
#ifndef IMOP_HIP_H
#define IMOP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMOP_HIP_ABI_VERSION 1

/* Replaces imop::has_cuda (src/algorithms/cuda_dispatch.cpp:26-33):
 * number of visible HIP devices, 0 when none. */
int imop_hip_device_count(void);

/* Replaces imop::cuda_device_name (src/algorithms/cuda_dispatch.cpp:35-49):
 * name of device 0, "N/A" without a device. */
const char* imop_hip_device_name(void);

/* Replaces imop::cuda_synchronize (src/algorithms/cuda_dispatch.cpp:211-213). */
int imop_hip_synchronize(void);

/* Stream-ordered demosaic of one frame resident in device memory.
 * Replaces the kernel launches behind imop::demosaic_cuda
 * (src/algorithms/cuda_dispatch.cpp:102-135 -> cuda_launch_hqli / _mg /
 * _prism, src/algorithms/cuda_kernels.cu:1690-1866), fused with unpack and
 * black level.  d_in holds in_bytes bytes, d_out out_bytes bytes; both must
 * be at least the sizes pixel::compute_bayer_byte_size /
 * compute_rgb_byte_size give.  black_level points to 4 host floats (may be
 * NULL when black_level_mode is 0).  is_packed: 0 unpacked, 1 the
 * reference's packed layouts (pixel_utils.hpp:58-118), 2 the GigE Vision
 * 10-bit packing (IMOP_HIP_PACKED_GIGE10; bit depth 10 only, else
 * InvalidBitDepth; 3 bytes per pixel pair), which every entry point taking
 * is_packed or a frame descriptor accepts.  Returns after enqueueing.
 *
 * PRISM and RCD keep four FP32 planes (16 bytes per pixel) between their two
 * passes.  Each call takes them as its own stream-ordered allocation on
 * `stream` and releases them on `stream` after its kernels, so calls on
 * different streams or from different host threads never share planes.
 * Eager calls draw on a library-owned memory pool per device that keeps
 * released memory reserved (the process holds the largest concurrent need
 * until it ends); under stream capture the allocation and release become
 * nodes of the graph, which owns its planes.  The other algorithms use no
 * device memory beyond d_in and d_out. */
#define IMOP_HIP_PACKED_GIGE10 2
int imop_hip_demosaic(const void* d_in, uint64_t in_bytes, void* d_out, uint64_t out_bytes, int width,
                      int height, int pattern, int algorithm, int bit_depth, int is_packed,
                      const float* black_level, int black_level_mode, void* stream);

/* imop::demosaic_cuda (src/algorithms/cuda_dispatch.cpp:68-141) with the same
 * host-buffer, synchronous contract, plus packed input and black level.
 * Buffers are caller-owned host memory sized by compute_bayer_byte_size /
 * compute_rgb_byte_size; the result is valid on return. */
int imop_hip_demosaic_host(const uint8_t* bayer, uint8_t* rgb, int width, int height, int pattern,
                           int algorithm, int bit_depth, int is_packed, const float* black_level,
                           int black_level_mode);

/* Black level in place on an unpacked Bayer frame in device memory
 * (u8 for bit_depth <= 8, else u16).  algorithm = black_level::
 * BlackLevelAlgorithm (PER_CHANNEL 0, GLOBAL 1); offsets = R, Gr, Gb, B
 * (host floats).  Replaces the kernels behind
 * black_level::process_black_level_cuda (src/gpu_dispatch.cpp:421-448 ->
 * src/gpu_ops_kernels.cu:592-627) with the CPU module's semantics
 * (src/black_level/global.cpp:6-24, per_channel.cpp:6-33).  Returns a
 * black_level::BlackLevelError value (same numbering, 4 = InvalidChannels
 * unused). */
int imop_hip_black_level(void* d_data, int width, int height, int pattern, int algorithm, int bit_depth,
                         const float* offsets, void* stream);

/* black_level::process_black_level_cuda (src/gpu_dispatch.cpp:421-448):
 * host buffer, synchronous. */
int imop_hip_black_level_host(uint8_t* data, int width, int height, int pattern, int algorithm, int bit_depth,
                              const float* offsets);

/* Launch shape of the register-resident HQLI / MG kernel, chosen by the
 * autotuner (im-operator_amd/imop_amd/autotune.py): rows each wave
 * computes (2 or 4), waves per workgroup stacked down one strip (4),
 * and the workgroup order: bit 0 remaps workgroups so each XCD works on a
 * contiguous run of them (else launch order), bit 1 orders them
 * strip-fastest (adjacent workgroups cover adjacent columns of the same
 * rows; else they walk down a strip).  Out-of-range values (e.g. -1) leave
 * a setting unchanged.  rows 0 is the automatic band height, the initial
 * setting: 2 rows (the fused-epilogue form keeps 4); get reports 0 while it
 * is in effect, so a get / set round trip keeps it.  The initial shape is
 * (0, 4, 3).  Process-wide; has no counterpart in the reference (its CUDA
 * launch shape is fixed at 32x16, src/algorithms/cuda_kernels.cu:1496). */
int imop_hip_set_launch_shape(int rows_per_wave, int waves_per_block, int xcd_remap);
int imop_hip_get_launch_shape(int* rows_per_wave, int* waves_per_block, int* xcd_remap);

/* ---------------------------------------------------------------- batches */

/* One frame of a batch.  Buffers are device memory.  out_elem_size 0 keeps
 * demosaic_cuda's output element (u8 for bit_depth <= 8, else u16); 2 writes
 * u16 for every depth (8-bit values widened, same numbers); 1 is accepted for
 * bit_depth <= 8 only.  black_level / black_level_mode as in
 * imop_hip_demosaic.  pattern / algorithm may differ between frames. */
typedef struct imop_hip_frame_desc {
    const void* in;
    uint64_t in_bytes;
    void* out;
    uint64_t out_bytes;
    int32_t width, height;
    int32_t pattern;
    int32_t algorithm;
    int32_t bit_depth;
    int32_t is_packed;
    int32_t out_elem_size;
    int32_t black_level_mode;
    float black_level[4];
} imop_hip_frame_desc;

/* Stream-ordered demosaic of n frames of any size, layout, bit depth and
 * black level.  HQLI / MG frames of even width with u8 / u16 / 12-bit /
 * 10-bit packed input run in one persistent launch per (pattern, algorithm)
 * (the register-resident kernel walking a work list of every frame's bands);
 * other frames (PRISM, odd widths, other packed depths) get their
 * single-frame launches on the same stream.  Every frame is checked first,
 * as imop_hip_demosaic checks it; on the first invalid frame nothing is
 * enqueued, its index goes to *failed_frame (may be NULL) and its error code
 * is returned.  The work list is copied from host memory of the call, so
 * the call returns InternalError on a capturing stream (a graph replay would
 * read that memory after the call); graphs use imop_hip_batch_plan_create.
 * Replaces imop::demosaic_cuda_batch
 * (src/algorithms/cuda_dispatch.cpp:150-209: one size, one depth, a launch
 * and two copies per frame). */
int imop_hip_demosaic_batch(const imop_hip_frame_desc* frames, int n, void* stream, int* failed_frame);

/* The same ragged batch with frames in caller-owned host memory (pageable;
 * in / out are host pointers sized as in imop_hip_frame_desc, with no
 * alignment required).  Synchronous: outputs are valid on return.  Frames
 * are staged in order, in chunks of about 96 MB, through three pinned host
 * and device slots on three streams, so the copies of neighbouring chunks
 * overlap the current chunk's kernels, and the host-side gathers and
 * scatters (threaded) overlap the GPU.  Every frame is
 * checked first as imop_hip_demosaic_batch checks it; on the first invalid
 * frame nothing is copied or written, its index goes to *failed_frame (may
 * be NULL) and its error code is returned.  Per host thread, the staging
 * grows to the largest chunk and is kept.  Replaces the double-buffered host
 * path of imop::demosaic_cuda_batch (src/algorithms/cuda_dispatch.cpp:150-209,
 * one frame size and depth, one frame in flight per buffer). */
int imop_hip_demosaic_batch_host(const imop_hip_frame_desc* frames, int n, int* failed_frame);

/* The same batch captured once into a HIP graph, for repeated launches over
 * the same buffers (replaces cuda_graph_capture, src/algorithms/
 * cuda_kernels.cu:1520-1688, which captured one frame size).  The plan owns
 * the device work list, a PRISM workspace when needed, and the graph; the
 * launch shape current at creation is used.  Launches on one plan are
 * stream-ordered by the caller; the buffers must stay valid until destroy. */
typedef struct imop_hip_batch_plan imop_hip_batch_plan;
int imop_hip_batch_plan_create(const imop_hip_frame_desc* frames, int n, imop_hip_batch_plan** plan,
                               int* failed_frame);
int imop_hip_batch_plan_launch(imop_hip_batch_plan* plan, void* stream);
int imop_hip_batch_plan_destroy(imop_hip_batch_plan* plan);

/* Launch shape of the batched kernel, chosen by the autotuner.
 * waves_per_block 4: work items are the single-frame kernel's workgroups,
 * 4 waves stacked down one strip sharing decoded rows through LDS, each wave
 * computing rows_per_wave rows (2 or 4); waves_per_block 1: every wave is
 * its own item, a strip by a band of rows_per_wave rows (4 or 8).  A rows
 * value the form does not take runs as its nearest (group form: 8 -> 4;
 * wave form: 2 -> 4).  blocks_per_cu: workgroups per CU of the persistent
 * grid (1..64; 0 = one workgroup per item); order: bit 1 strip-fastest
 * within a frame, bit 2 frame-major placement (one-workgroup-per-item grids:
 * the frames run one after another, each spread over the 8 XCDs as a
 * single-frame launch spreads it; else the whole list is cut into 8
 * contiguous per-XCD runs).  Initial shape (4, 4, 0, 6).  Out-of-range values
 * leave a setting unchanged.  Process-wide; no counterpart in the reference. */
int imop_hip_set_batch_shape(int rows_per_wave, int waves_per_block, int blocks_per_cu, int order);
int imop_hip_get_batch_shape(int* rows_per_wave, int* waves_per_block, int* blocks_per_cu, int* order);

/* ---------------------------------------------- differentiable ISP chain */

/* FP32 chain over n frames of one size, all buffers device memory:
 *   x = raw / max_value, a = D x (bilinear demosaic: hqli.cpp:7-44 with exact
 *   averages, fill_rgb_borders of common.hpp:47-96, corners 0),
 *   u = clamp01(gains .* a)  (white balance, gpu_ops_kernels.cu:119-125),
 *   z = clamp01(ccm u)       (CCM, gpu_ops_kernels.cu:92-96),
 *   out = z^(1/gamma)        (gamma, gpu_ops_kernels.cu:57-64).
 * raw [n][h][w]; out, target, grad_out [n][h][w][3]; gains[3] and ccm[9]
 * (row-major) are device pointers.  The pair follows the
 * torch.autograd.Function contract: forward keeps nothing; backward
 * recomputes the forward from raw and the parameters.  No counterpart in the
 * reference (its chain, gpu_ops_kernels.cu:40-135, has no backward).
 *
 * Forward: writes out (may be NULL) and, with target, the mean squared
 * error to *d_loss (one device float). */
int imop_hip_isp_chain_fwd(const float* d_raw, int n, int width, int height, int pattern, float max_value,
                           const float* d_gains, const float* d_ccm, float gamma, float* d_out,
                           const float* d_target, float* d_loss, void* stream);

/* Backward (vector-Jacobian product): the cotangent is d_grad_out, or, when
 * it is NULL, the gradient of the forward's mean squared error against
 * d_target.  Writes dL/draw to d_grad_raw (may be NULL) and the 12 parameter
 * gradients (gains[3] then ccm[9]) to d_grad_params (may be NULL), summed
 * over frames.  Clamps pass the gradient where 0 <= v <= 1; the gamma
 * derivative is 0 at 0.  Chain calls on one device share a partial-sum
 * buffer: run them in stream order. */
int imop_hip_isp_chain_bwd(const float* d_raw, int n, int width, int height, int pattern, float max_value,
                           const float* d_gains, const float* d_ccm, float gamma, const float* d_grad_out,
                           const float* d_target, float* d_grad_raw, float* d_grad_params, void* stream);

/* imop_hip_isp_chain_bwd without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight training or serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Same arguments, checks, kernels and results as imop_hip_isp_chain_bwd; the
 * per-workgroup partial sums live in a fixed 96 KiB device array
 * (2048 workgroups x 12 floats) that each GPU holds from the library's load
 * until the process ends, whose address is looked up on the first call on a
 * device.  The backward never launches more than 2048 workgroups, so every
 * size the chain accepts fits; there is no further size limit.  Every call
 * on one device shares that device's array: calls that overlap on one device
 * (different streams, or host threads without ordering) race. */
int imop_hip_isp_chain_bwd_noalloc(const float* d_raw, int n, int width, int height, int pattern, float max_value,
                                   const float* d_gains, const float* d_ccm, float gamma, const float* d_grad_out,
                                   const float* d_target, float* d_grad_raw, float* d_grad_params, void* stream);

/* The chain's demosaic operator alone, FP32, one frame: out[h][w][3] = D x
 * and out[h][w] = D^T v (for checks against the transposed operator). */
int imop_hip_bilinear_fwd(const float* d_x, int width, int height, int pattern, float* d_out, void* stream);
int imop_hip_bilinear_vjp(const float* d_v, int width, int height, int pattern, float* d_out, void* stream);

/* ------------------------------------------------------ RGB epilogue */

/* The integer-domain RGB stages that follow the demosaic, applied in this
 * order with the reference CPU semantics (each stage rounds and clamps to
 * [0, max] of the bit depth, as its separate pass does):
 *   flag 1: manual white balance, white_balance::process_manual_wb
 *           (src/white_balance/manual.cpp:8-41), gains r, g, b;
 *   flag 2: linear 3x3 colour matrix, ccm::process_linear_3x3
 *           (src/ccm/linear_3x3.cpp:9-46), row-major, out = M * in;
 *   flag 4: gamma, tone::process_gamma (src/tone/gamma.cpp:10-34), the
 *           exponent clamped to [0.1, 10] like the reference.
 * Outputs equal those three CPU functions run one after the other. */
typedef struct imop_hip_rgb_epilogue {
    int32_t flags;
    float wb_gains[3];
    float ccm[9];
    float gamma;
} imop_hip_rgb_epilogue;

/* imop_hip_demosaic followed by the epilogue, in one launch on the
 * register-resident path (HQLI / MG of even widths) and as one extra
 * in-place pass otherwise.  Replaces the reference's demosaic_cuda
 * (src/algorithms/cuda_dispatch.cpp:68-141) followed by
 * process_white_balance_cuda / process_ccm_cuda / process_tone_cuda
 * (src/gpu_dispatch.cpp:288-360), each a host round trip of the RGB image.
 * epilogue NULL or flags 0: imop_hip_demosaic.  The first call with a given
 * (bit depth, gamma) builds that gamma table on the device with a blocking
 * copy, so it must not be made during stream capture; up to 64 tables are
 * kept per process. */
int imop_hip_demosaic_isp(const void* d_in, uint64_t in_bytes, void* d_out, uint64_t out_bytes, int width,
                          int height, int pattern, int algorithm, int bit_depth, int is_packed,
                          const float* black_level, int black_level_mode, const imop_hip_rgb_epilogue* epilogue,
                          void* stream);

/* The epilogue alone, in place, on an interleaved RGB image of width x
 * height (u8 if bit_depth <= 8, else u16): the reference's
 * process_white_balance_cuda / process_ccm_cuda / process_tone_cuda
 * (src/gpu_dispatch.cpp:288-360) with CPU semantics, device buffer. */
int imop_hip_rgb_epilogue_apply(void* d_rgb, uint64_t rgb_bytes, int width, int height, int bit_depth,
                                const imop_hip_rgb_epilogue* epilogue, void* stream);

/* ------------------------------------------------------ lens shading */

/* lens_shading::LensShadingParams of the polynomial correction: coef[c] =
 * {a2, a4, a6} of ShadingPolynomial for c = R, Gr, Gb, B (r_coef, gr_coef,
 * gb_coef, b_coef); center_x / center_y the normalised optical centre. */
typedef struct imop_hip_lens_shading {
    float coef[4][3];
    float center_x, center_y;
} imop_hip_lens_shading;

/* Polynomial lens shading in place on an unpacked Bayer frame in device
 * memory (u8 for bit_depth <= 8, else u16; bytes >= width * height * 1 or 2):
 * gain = 1 + a2 r^2 + a4 r^4 + a6 r^6 of the radius from the optical centre
 * normalised by the centre's distance to the origin, per CFA channel.
 * Replaces lens_shading::process_lens_shading_cuda (src/gpu_dispatch.cpp:
 * 455-490, POLYNOMIAL) with the CPU semantics of process_polynomial
 * (src/lens_shading/polynomial.cpp:11-52): same bytes.  Returns a
 * lens_shading::LensShadingError value (same numbering as DemosaicError;
 * 4 = InvalidChannels unused). */
int imop_hip_lens_shading_poly(void* d_data, uint64_t bytes, int width, int height, int pattern, int bit_depth,
                               const imop_hip_lens_shading* params, void* stream);

/* Bayer-domain stages between the black level and the demosaic.
 *   flag 1: polynomial lens shading with `lens`,
 *           lens_shading::process_polynomial (src/lens_shading/
 *           polynomial.cpp:11-52).
 * Other flag bits are reserved (InternalError). */
typedef struct imop_hip_bayer_prologue {
    int32_t flags;
    struct imop_hip_lens_shading lens;
} imop_hip_bayer_prologue;

/* The whole single-frame path in one call: unpack -> black level -> the
 * prologue -> demosaic -> the RGB epilogue, equal to the reference's CPU
 * functions run in that order on the unpacked buffer (unpack,
 * process_black_level, process_polynomial, demosaic_cpu, process_manual_wb,
 * process_linear_3x3, process_gamma).  HQLI / MG frames on the
 * register-resident path apply the lens gain to each sample as it is
 * decoded, in the same launch as the demosaic and the epilogue (no unpacked
 * frame is written); every other frame gets one extra pass that writes the
 * prologue's result to a stream-ordered unpacked buffer, demosaiced from
 * there.  prologue or epilogue NULL (or flags 0): that side is absent.
 * Replaces lens_shading::process_lens_shading_cuda (src/gpu_dispatch.cpp:
 * 455-486) + imop::demosaic_cuda + the RGB stages' GPU hooks, each a host
 * round trip in the reference. */
int imop_hip_demosaic_pipeline(const void* d_in, uint64_t in_bytes, void* d_out, uint64_t out_bytes, int width,
                               int height, int pattern, int algorithm, int bit_depth, int is_packed,
                               const float* black_level, int black_level_mode,
                               const imop_hip_bayer_prologue* prologue, const imop_hip_rgb_epilogue* epilogue,
                               void* stream);

/* ---------------------------------------------- adaptive defect correction */

/* Adaptive defect correction in place on an unpacked Bayer frame in device
 * memory (u8 for bit_depth <= 8, else u16; bytes >= width * height * 1 or 2):
 * each interior pixel ([2, w-2) x [2, h-2)) whose distance from the median of
 * its same-colour neighbours at distance 2 exceeds
 * clamp(threshold, 0.05, 1) * max becomes that median, in the reference's
 * raster order (neighbours above and to the left as already corrected).
 * Replaces defect_correct::process_defect_correct_cuda (src/gpu_dispatch.cpp:
 * 529-571, ADAPTIVE) with the CPU semantics of process_adaptive
 * (src/defect_correct/adaptive.cpp:8-61): same bytes.  `pattern` is accepted
 * and unused, as there (neighbours at distance 2 share the pixel's colour).
 * Synchronises `stream` (the number of fixed-point passes depends on the
 * frame) and returns when the frame is final.  Returns a
 * defect_correct::DefectCorrectError value (0 Ok, 1 NullInput,
 * 2 InvalidDimensions, 3 InvalidBitDepth, 5 ImageTooSmall below 5 x 5,
 * 6 InternalError). */
int imop_hip_defect_adaptive(void* d_data, uint64_t bytes, int width, int height, int pattern, int bit_depth,
                             float threshold, void* stream);

/* The same correction, stream-ordered: enqueues the input copy and
 * max_passes fixed-point passes (1..65536) on `stream` and returns without
 * waiting, so it can be captured into a HIP graph.  Each pass after one that
 * changed nothing returns at once.  When the passes end it writes *d_status
 * (one int of device memory): 0 -- the last pass changed nothing, so the frame
 * equals imop_hip_defect_adaptive's result; 1 -- the frame still changed in
 * the last pass (it holds an intermediate iterate; run the synchronous form
 * on the original frame, or allow more passes); 2 -- internal error.  Natural
 * frames settle in 2-4 passes; a 12 MP frame of pure noise at threshold 0.05
 * needs 28.  Scratch (a copy of the frame, control words, two bytes per
 * 128 x 32 tile) is this call's own stream-ordered allocation, graph-owned
 * under capture, so calls on different streams never share it.  Checks and
 * error codes as imop_hip_defect_adaptive, plus 1 for a NULL d_status and 6
 * for max_passes out of range.  No counterpart in the reference (its
 * correction is a synchronous CPU loop). */
int imop_hip_defect_adaptive_async(void* d_data, uint64_t bytes, int width, int height, int pattern,
                                   int bit_depth, float threshold, int max_passes, int* d_status, void* stream);

/* --------------------------------------------------------- raw analyzer */

/* imop::DataInfo (include/imop/types.hpp:76-86) with possible_dimensions as
 * up to 8 (width, height) pairs. */
typedef struct imop_hip_data_info {
    int32_t detected_bit_depth;
    int32_t suggested_width, suggested_height;
    int32_t pixel_count;
    int32_t max_value, min_value;
    int32_t is_likely_16bit, is_packed;
    int32_t n_dimensions;
    int32_t dimensions[8][2];
} imop_hip_data_info;

/* imop::analyze_data (src/analyzer.cpp:151-214) of a raw buffer in device
 * memory: the statistics its heuristics read (maxima, counts, the whole
 * buffer's range) are gathered on the GPU, the decisions are the reference's;
 * same DataInfo.  Synchronises `stream`.  Returns 0, 1 if info is null, 6 on a
 * HIP failure or a buffer of 2^31 bytes or more. */
int imop_hip_analyze_data(const void* d_data, uint64_t byte_size, imop_hip_data_info* info, void* stream);

/* imop::guess_pattern (src/analyzer.cpp:216-321) of a frame in device memory
 * (u8 for bit_depth <= 8, else u16 or packed; is_packed 2 with bit depth 10
 * reads the GigE Vision packing): *pattern = the BayerPattern
 * value it returns.  Synchronises `stream`.  Returns 0, 1 if pattern is null,
 * 3 for a bit depth outside 1..16 or is_packed 2 with a bit depth other than
 * 10, 6 for is_packed outside 0..2, if the buffer is smaller than the frame
 * (depths up to 8 are read one byte per pixel, packed or not) or on a HIP
 * failure.  A null buffer or a side below 4 gives RGGB and 0, as the
 * reference returns RGGB there (analyzer.cpp:218). */
int imop_hip_guess_pattern(const void* d_data, uint64_t byte_size, int width, int height, int bit_depth, int is_packed,
                           int* pattern, void* stream);

/* ------------------------------------------------------------ measurement */

/* The single-rate vector-ALU issue rate this GPU sustains (the rate of FMA
 * and integer instructions), in wave64 VALU instructions per second, from a
 * probe kernel of independent integer chains on every SIMD (best of five
 * timed launches after three untimed).
 * No counterpart in the reference: it calibrates VALU rooflines, since GPUs
 * of one model run at different sustained clocks.  Synchronises `stream`.  Returns 0, 1 if the pointer is null, 6 on
 * a HIP failure. */
int imop_hip_valu_issue_rate(double* wave_instructions_per_second, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* IMOP_HIP_H */
