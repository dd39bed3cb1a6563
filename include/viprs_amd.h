/*
 * viprs_amd: MI355X (gfx950) executor for viprs's resample and separable
 * convolution path -- ShrinkH/ShrinkV box pre-shrink, ReduceH/ReduceV
 * Lanczos/cubic/linear reducers (thumbnail / resize / reduce_h / reduce_v),
 * gauss_blur and convsep, and the thumbnail / resize tails around them
 * (Affine, Zoom, ExtractArea).
 *
 * C ABI: plain pointers, sizes and status codes; no exceptions cross it.
 * Conventions follow the reference's existing C ABI (include/viprs.h:7-20,
 * src/ffi.rs:24-230): opaque handles created by *_new / *_create, released by
 * the matching *_destroy, NULL-safe destroy, negative status on failure.
 *
 * Image memory is packed, band-interleaved, row-major, exactly as the
 * reference's tiles (crates/viprs-core/src/image/region.rs:154-230).
 * Unless a function says otherwise, image pointers are DEVICE pointers and the
 * work is enqueued on the context's stream (asynchronous).
 */
#ifndef VIPRS_AMD_H
#define VIPRS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BandFormatId (crates/viprs-core/src/format/band.rs:19-35), numbered by this
 * library (U8, U16, F32 first). */
typedef enum {
    VGPU_U8 = 0,
    VGPU_U16 = 1,
    VGPU_F32 = 2,
    VGPU_I16 = 3,
    VGPU_U32 = 4,
    VGPU_I32 = 5,
    VGPU_F64 = 6
} vgpu_format;

/* InterpolationKernel (crates/viprs-core/src/kernel.rs:22-49). LBB, VSQBS and
 * Nohalo are rejected by the reduce ops like the reference does
 * (reduce_common.rs:94-124). */
typedef enum {
    VGPU_NEAREST = 0,
    VGPU_BILINEAR = 1,
    VGPU_BICUBIC = 2,
    VGPU_LANCZOS2 = 3,
    VGPU_LANCZOS3 = 4,
    VGPU_CATMULLROM = 5,
    VGPU_LBB = 6,
    VGPU_VSQBS = 7,
    VGPU_NOHALO = 8
} vgpu_kernel;

/* ThumbnailTarget (crates/viprs-ops-spatial/src/resample/thumbnail.rs:125-145). */
typedef enum {
    VGPU_FITBOX = 0,
    VGPU_FORCE_EXACT = 1,
    VGPU_WIDTH = 2,
    VGPU_HEIGHT = 3
} vgpu_thumbnail_target;

/* Status codes.  BuildError::InvalidReduceParameters / SourceHint ->
 * INVALID_ARGUMENT; BuildError::InvalidKernel -> INVALID_KERNEL;
 * BuildError::UnsupportedFormat -> UNSUPPORTED_FORMAT
 * (crates/viprs-core/src/error.rs). */
typedef enum {
    VGPU_OK = 0,
    VGPU_E_INVALID_ARGUMENT = -1,
    VGPU_E_INVALID_KERNEL = -2,
    VGPU_E_UNSUPPORTED_FORMAT = -3,
    VGPU_E_UNSUPPORTED = -4,  /* plan needs a node outside this path (an op after an Affine tail) */
    VGPU_E_BUFFER_SIZE = -5,  /* a buffer length does not match the region */
    VGPU_E_DEVICE = -6,       /* HIP runtime error (sticky on the context) */
    VGPU_E_NO_MEMORY = -7,
    VGPU_E_CANCELLED = -8,
    VGPU_E_ADMISSION = -9,    /* request exceeds the device memory budget */
    VGPU_E_NOT_READY = -10,   /* a wait timed out before the request finished */
    VGPU_E_SCRATCH_LIMIT = -11 /* plan tables + intermediates exceed VGPU_NOALLOC_SCRATCH_BYTES */
} vgpu_status;

/* Region (crates/viprs-core/src/image/region.rs:18-27). */
typedef struct {
    int32_t x, y;
    uint32_t width, height;
} vgpu_region;

typedef struct vgpu_context vgpu_context;
typedef struct vgpu_pipeline vgpu_pipeline;
typedef struct vgpu_op vgpu_op;

const char *vgpu_status_string(int status);

/* ---------------------------------------------------------------- context */

/* Device executor context: one HIP device, one stream, a device buffer pool
 * and the coefficient/plan cache.  No reference counterpart; it is the state
 * the device executor keeps for RayonScheduler
 * (crates/viprs-runtime/src/scheduler/rayon_scheduler.rs:53-75). */
int vgpu_context_create(int device, vgpu_context **out);
/* Destroy every pipeline, op, batch and buffer made on the context first:
 * they hold memory from its pool. */
void vgpu_context_destroy(vgpu_context *ctx);
/* Use a caller-owned hipStream_t (NULL = the context's own stream). */
int vgpu_context_set_stream(vgpu_context *ctx, void *hip_stream);
void *vgpu_context_stream(vgpu_context *ctx);
int vgpu_synchronize(vgpu_context *ctx);
/* Bytes held by the device buffer pool. */
size_t vgpu_context_pool_bytes(const vgpu_context *ctx);
/* Release every pooled buffer not in use. */
int vgpu_context_trim(vgpu_context *ctx);

/* Device buffers from the context's pool (the device buffer pool that
 * replaces the scheduler's per-thread ThreadBufferPool,
 * crates/viprs-runtime/src/pipeline/compiled.rs:475-).  Allocation only
 * calls hipMalloc on a pool miss; free returns the buffer to the pool after
 * the context stream has drained.  The copies are synchronous. */
int vgpu_buffer_alloc(vgpu_context *ctx, size_t bytes, void **out);
int vgpu_buffer_free(vgpu_context *ctx, void *ptr);
int vgpu_memcpy_h2d(vgpu_context *ctx, void *dst_device, const void *src_host, size_t bytes);
int vgpu_memcpy_d2h(vgpu_context *ctx, void *dst_host, const void *src_device, size_t bytes);
/* Seeded synthetic bytes on the device, enqueued on the context stream: byte
 * 8i+k is byte k (little-endian) of splitmix64(seed + (i+1)*0x9E3779B97F4A7C15).
 * No reference counterpart (the reference's benches seed inputs on the host);
 * it lets multi-GB benchmark inputs be made in HBM and regenerated on the
 * host for checking. */
int vgpu_fill_random_u8(vgpu_context *ctx, void *dst_device, size_t bytes, uint64_t seed);
/* Measured streaming peaks of this device, for the roofline reports: mean ms
 * of a nontemporal 16-byte streaming read of `bytes` (the box shrink's load
 * form) and of a device-to-device copy of `bytes`, each over `iterations`
 * back-to-back launches between HIP events (buffers from the context pool).
 * No reference counterpart (measurement helper). */
int vgpu_measure_stream(vgpu_context *ctx, size_t bytes, int iterations, float *read_ms, float *copy_ms);

/* ----------------------------------------------------- pipeline (builder) */

/* PipelineBuilder::from_source(MemorySource) (crates/viprs-runtime/src/
 * pipeline/builder/core.rs:40; sources/memory.rs:149-235): an in-memory source
 * of width x height x bands samples of `format`, read with clamp-to-edge. */
int vgpu_pipeline_new(vgpu_context *ctx, uint32_t width, uint32_t height, uint32_t bands,
                      int format, vgpu_pipeline **out);
void vgpu_pipeline_destroy(vgpu_pipeline *p);

/* PipelineBuilder::reduce_h / reduce_v (builder/resample.rs:16-123,270-339). */
int vgpu_pipeline_reduce_h(vgpu_pipeline *p, double factor, int kernel);
int vgpu_pipeline_reduce_v(vgpu_pipeline *p, double factor, int kernel);
/* PipelineBuilder::reduce (builder/resample.rs:349-382): the fused f64
 * ReduceOp (resample/reduce.rs:23-262), filters never bound to the input
 * length (offset 0), H results rounded to the sample type before the V pass. */
int vgpu_pipeline_reduce(vgpu_pipeline *p, double h_factor, double v_factor, int kernel);
/* PipelineBuilder::extract_area (builder/geometry.rs:40-80): a width x height
 * rectangle at (x, y), inside the current image. */
int vgpu_pipeline_extract_area(vgpu_pipeline *p, uint32_t x, uint32_t y, uint32_t width, uint32_t height);
/* PipelineBuilder::shrink_h_with_ceil / shrink_v_with_ceil
 * (builder/resample.rs:133-209,392-444). */
int vgpu_pipeline_shrink_h(vgpu_pipeline *p, uint32_t factor, int ceil_mode);
int vgpu_pipeline_shrink_v(vgpu_pipeline *p, uint32_t factor, int ceil_mode);
/* PipelineBuilder::shrink(hshrink, vshrink) -> Shrink<F> (builder/resample.rs:
 * 219-268; resample/shrink.rs:37-293): ShrinkH then ShrinkV inside one op. */
int vgpu_pipeline_shrink(vgpu_pipeline *p, uint32_t hshrink, uint32_t vshrink);
/* PipelineBuilder::gauss_blur (builder/colour.rs:130-197): U8 stays U8, other
 * formats produce F32. */
int vgpu_pipeline_gauss_blur(vgpu_pipeline *p, float sigma);
/* ConvSep::with_mask via PipelineBuilder::then (convolution/convsep.rs:218-233):
 * H pass with (scale, offset), V pass with (scale, 0); output F32. */
int vgpu_pipeline_convsep(vgpu_pipeline *p, const double *coeffs, uint32_t n, double scale,
                          double offset);
/* PipelineBuilder::premultiply / unpremultiply (builder/resample.rs:721-794). */
int vgpu_pipeline_premultiply(vgpu_pipeline *p);
int vgpu_pipeline_unpremultiply(vgpu_pipeline *p);
/* PipelineBuilder::thumbnail(Thumbnail::new(target, kernel))
 * (builder/resample.rs:871-979; thumbnail.rs:229-524), in-memory source,
 * every plan including the Affine tail (affine/core.rs).  Ops added after an
 * Affine tail other than extract_area / unpremultiply make the pipeline's
 * run return VGPU_E_UNSUPPORTED. */
int vgpu_pipeline_thumbnail(vgpu_pipeline *p, int target, uint32_t width, uint32_t height,
                            int kernel);
/* The same with Thumbnail::with_crop (thumbnail.rs:185-190) when crop != 0:
 * the image is reduced to cover the target box, then a centred ExtractArea
 * of the box's size is taken (thumbnail.rs:246-285, 421-428). */
int vgpu_pipeline_thumbnail_crop(vgpu_pipeline *p, int target, uint32_t width, uint32_t height,
                                 int kernel, int crop);
/* Thumbnail::into_pipeline_nodes(...).shrink_factor (thumbnail.rs:326-330,
 * 514-524): the shrink-on-load factor (2, 4, 16; 0 = none) a decoder source
 * would be asked for (ImageSource::set_thumbnail_shrink_on_load,
 * builder/resample.rs:878-906).  A caller that decodes at 1/factor builds
 * the pipeline over the decoded image and calls vgpu_pipeline_thumbnail on
 * it, which plans without the hint from those dimensions, as the builder
 * does.  Host-only (no device work). */
int vgpu_thumbnail_shrink_on_load_factor(uint32_t width, uint32_t height, uint32_t bands, int target,
                                         uint32_t target_width, uint32_t target_height, int crop,
                                         uint32_t *factor);
/* The PNG source's box pre-shrink on the device, as the pipeline's first
 * step (the decoded image stays in HBM; the chain reads its shrunk copy):
 * ceil_mode == 0 is software_box_shrink_generic (sources/decoder_source/
 * shrink.rs:47-117), what DecoderSource::set_thumbnail_shrink_on_load does to
 * an eager backing (source_impl.rs:245-275): output max(src / f, 1) per axis,
 * u8 rounds (sum + n/2) / n, other formats from_f64_clamped(sum / n);
 * ceil_mode == 1 (u8) is decode_png_with_box_shrink_u8 (codecs/src/png/
 * decode_full.rs:320-452): output ceil(src / f), the last column and row
 * divided by their partial windows' pixel counts.  Only before any other
 * node (VGPU_E_INVALID_ARGUMENT otherwise). */
int vgpu_pipeline_box_shrink(vgpu_pipeline *p, uint32_t factor, int ceil_mode);
/* PipelineBuilder::thumbnail over a PNG source with an eager backing
 * (builder/resample.rs:871-906): when the thumbnail is the first node and
 * the plan's shrink-on-load hint is set, vgpu_pipeline_box_shrink(hint, 0)
 * replaces the source, and the thumbnail is planned without the hint over
 * the shrunk dimensions (and is "large", for its tiling, by those). */
int vgpu_pipeline_thumbnail_shrink_on_load(vgpu_pipeline *p, int target, uint32_t width, uint32_t height,
                                           int kernel, int crop);
/* PipelineBuilder::resize(Resize::new(hscale, vscale, kernel))
 * (builder/resample.rs:637-711; resize.rs:163-284): shrink and reduce for
 * downscales, Zoom and the Affine fix-up for upscales.  LBB is one Affine
 * with the locally bounded bicubic stencil over the whole image
 * (resize.rs:176-199); VSQBS never zooms and its Affine takes the 4 x 4
 * quadratic B-spline stencil, while a VSQBS downscale that needs a reduce
 * step fails with VGPU_E_INVALID_KERNEL, as ReduceH / ReduceV refuse the
 * kernel (reduce_common.rs:107-113). */
int vgpu_pipeline_resize(vgpu_pipeline *p, double hscale, double vscale, int kernel);

/* Kernel choice for u8 separable pairs, a bit mask: bit 0 puts ReduceH ->
 * ReduceV and GaussBlurH -> GaussBlurV on the int8-limb MFMA kernel (the
 * default, 1), bit 1 also the ShrinkH -> ShrinkV box pair (slower than the
 * plain box kernel on MI355X; kept for comparison); 0 selects the plain
 * integer kernels.  All choices are bit-exact.  No reference counterpart. */
int vgpu_pipeline_set_integer_mfma(vgpu_pipeline *p, int enable);
/* Arithmetic of f32 reduce / Gaussian pairs: 0 = the reference's FP64
 * accumulation in its operation order (default; bit-identical), 1 = plain
 * FP32 (f32 weights, FMA accumulation, f32 intermediate), 2 = FP16 hi/lo
 * split on MFMA with FP32 accumulation (falls back to 1 where its windows do
 * not fit).  Modes 1 and 2 differ from the reference in the last bits.  No
 * reference counterpart. */
int vgpu_pipeline_set_float_mode(vgpu_pipeline *p, int mode);
/* Kernel tuning knobs (results do not change).  No reference counterpart. */
typedef enum {
    VGPU_OPT_SEP_ROWS_PER_TASK = 2, /* output rows per MFMA wavefront task: 0 = auto, else a multiple of 16 */
    VGPU_OPT_SHRINK_RUN = 3,        /* box-shrink output rows per wavefront run: 0 = auto, 1..64 */
    VGPU_OPT_GAUSS_ROWS = 4,        /* u8 3/4-band Gaussian VALU tile rows: 0 = auto, 16 or 32; -1 = the MFMA kernel */
    VGPU_OPT_BATCH_REDUCE_CHAIN = 5, /* batch only, the launch-chain layout: 2 (default) = each reduce launch
                                        right after the shrink launch that wrote its inputs on the same stream,
                                        shrink launches with the most images first; 0 = level by level on each
                                        group's stream; 1 = reduce stages on a stream of their own */
    VGPU_OPT_SEP_FORM = 6,          /* int8 MFMA pairs: 0 = block tiles staged in LDS where they fit (default),
                                       1 = one wavefront per strip task streaming its operands */
    VGPU_OPT_REDUCE_FORM = 7        /* f32 ReduceH -> ReduceV pairs (FP64 order and FP32): 0 = blocks walking
                                       down strips of output columns where the strip fits (default), 1 = one
                                       block per output tile */
} vgpu_option;
int vgpu_pipeline_set_option(vgpu_pipeline *p, int option, int value);
/* CompiledPipeline output geometry (pipeline/compiled.rs:338-). */
int vgpu_pipeline_output_info(const vgpu_pipeline *p, uint32_t *width, uint32_t *height,
                              uint32_t *bands, int *format);
/* Number of GPU launches one run makes after fusion (diagnostic). */
int vgpu_pipeline_stage_count(vgpu_pipeline *p);
/* Name of the kernel stage `index` launches ("sep_mfma_i8", "reduce_f32",
 * "sep_mfma_f16x3", "shrink", "reduce_hv", ...), NUL-terminated into
 * name[0 .. cap).  Diagnostic; no reference counterpart. */
int vgpu_pipeline_stage_kernel(vgpu_pipeline *p, int index, char *name, size_t cap);

/* TileScheduler::run (crates/viprs-ports/src/scheduler.rs:63-98) with the
 * device region executor: runs the whole node chain over the full output on
 * the GPU, bit-identical to the reference's tile-by-tile execution.
 * src / dst are device pointers; enqueued on the context stream. */
int vgpu_pipeline_run(vgpu_pipeline *p, const void *src, size_t src_bytes, void *dst,
                      size_t dst_bytes);
/* CancellationToken (crates/viprs-core/src/cancel.rs:19-72): a flag shared
 * between the caller and a run; cancel() may be called from any thread. */
typedef struct vgpu_cancel_token vgpu_cancel_token;
int vgpu_cancel_token_new(vgpu_cancel_token **out);
void vgpu_cancel_token_destroy(vgpu_cancel_token *t);
void vgpu_cancel_token_cancel(vgpu_cancel_token *t);
int vgpu_cancel_token_is_cancelled(const vgpu_cancel_token *t);
/* TileScheduler::run_cancellable (crates/viprs-ports/src/scheduler.rs:132-
 * 142; RayonScheduler checks between tiles, scheduler/rayon_scheduler/
 * runtime.rs:1237): as vgpu_pipeline_run, polling `token` (may be NULL)
 * before the run and before each fused stage launch; returns
 * VGPU_E_CANCELLED when it is set (stages already enqueued still run, so the
 * destination may be partly written, as the reference's sink is). */
int vgpu_pipeline_run_cancellable(vgpu_pipeline *p, const void *src, size_t src_bytes, void *dst,
                                  size_t dst_bytes, const vgpu_cancel_token *token);
/* RayonScheduler::run_with_profile (crates/viprs-runtime/src/scheduler/
 * rayon_scheduler/runtime.rs:424-; NodeRunProfile): runs the pipeline
 * `iterations` times, timing every fused stage with HIP events on the context
 * stream, and writes the mean milliseconds per stage.  Returns the stage
 * count (>= 1) or a negative status. */
int vgpu_pipeline_run_profiled(vgpu_pipeline *p, const void *src, size_t src_bytes, void *dst,
                               size_t dst_bytes, int iterations, float *stage_ms,
                               int max_stages);
/* Mean duration of fused stage `stage` alone: one full run (fills the
 * intermediates), then `iterations` back-to-back launches of that stage
 * between two HIP events on the context stream.  No reference counterpart. */
int vgpu_pipeline_time_stage(vgpu_pipeline *p, const void *src, size_t src_bytes, void *dst,
                             size_t dst_bytes, int stage, int iterations, float *mean_ms);
/* As vgpu_pipeline_time_stage, the stage reading srcs[i % nsrc] on launch i
 * (all of src_bytes): with sources that together exceed the last-level
 * cache, the timing shows the kernel streaming from HBM rather than
 * re-reading a cached image.  No reference counterpart. */
int vgpu_pipeline_time_stage_sources(vgpu_pipeline *p, const void *const *srcs, int nsrc,
                                     size_t src_bytes, void *dst, size_t dst_bytes, int stage,
                                     int iterations, float *mean_ms);
/* As vgpu_pipeline_benchmark, step i reading srcs[i % nsrc].  No reference
 * counterpart. */
int vgpu_pipeline_benchmark_sources(vgpu_pipeline *p, const void *const *srcs, int nsrc,
                                    size_t src_bytes, void *dst, size_t dst_bytes, int warmup,
                                    int iterations, float *total_ms, float *step_ms);
/* Timing harness on the context stream: `warmup` untimed runs, then
 * `iterations` back-to-back runs bracketed by HIP events (total_ms), then
 * `iterations` more with an event after each (step_ms[i], may be NULL).
 * No reference counterpart (the reference's criterion benches). */
int vgpu_pipeline_benchmark(vgpu_pipeline *p, const void *src, size_t src_bytes, void *dst,
                            size_t dst_bytes, int warmup, int iterations, float *total_ms,
                            float *step_ms);
/* Capacity of the fixed device scratch array used by
 * vgpu_pipeline_run_noalloc (one copy per GPU the library runs on, held until
 * the process ends). */
#define VGPU_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
/* vgpu_pipeline_run without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The plan tables and every intermediate live in the library's fixed scratch
 * array of VGPU_NOALLOC_SCRATCH_BYTES per GPU instead of the context's pool;
 * the caller passes the source and destination device buffers as for
 * vgpu_pipeline_run, and gets bit-identical results.  A pipeline whose tables
 * plus intermediates (each rounded up to 256 bytes) do not fit returns
 * VGPU_E_SCRATCH_LIMIT before enqueuing anything.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads, or the caller's own queued calls on other streams) race
 * and must be serialised by the caller.  Calls on one stream are ordered and
 * safe.  The pipeline must stay unmodified until the call's work has run. */
int vgpu_pipeline_run_noalloc(vgpu_pipeline *p, const void *src, size_t src_bytes, void *dst,
                              size_t dst_bytes);
/* Same with host pointers: H2D copy, run, D2H copy, synchronize. */
int vgpu_pipeline_run_host(vgpu_pipeline *p, const void *src, size_t src_bytes, void *dst,
                           size_t dst_bytes);

/* ------------------------------------------------- op-level (DynOperation) */

/* Bridges (DynOperation, crates/viprs-core/src/op/dynamic_op.rs:16-260). */
/* ReduceHBridge::new(factor, kernel, bands, input_w, hint) (reduceh.rs:187-206) */
int vgpu_op_reduce_h_new(vgpu_context *ctx, double factor, int kernel, uint32_t bands,
                         uint32_t input_w, int format, vgpu_op **out);
/* ReduceVBridge::new(factor, kernel, bands, input_h, hint) (reducev.rs:186-205) */
int vgpu_op_reduce_v_new(vgpu_context *ctx, double factor, int kernel, uint32_t bands,
                         uint32_t input_h, int format, vgpu_op **out);
/* ReduceBridge::new(h_factor, v_factor, kernel, bands) (reduce.rs:273-287):
 * ReduceOp, both passes on one call; the input region's rows must be
 * required_input_region_v(output)'s rows, as the scheduler derives them. */
int vgpu_op_reduce_new(vgpu_context *ctx, double h_factor, double v_factor, int kernel,
                       uint32_t bands, int format, vgpu_op **out);
/* ShrinkHBridge::new_with_ceil_and_source_width (shrinkh.rs:1354-1383);
 * source_width = 0 means None. */
int vgpu_op_shrink_h_new(vgpu_context *ctx, uint32_t factor, int ceil_mode, uint32_t bands,
                         uint32_t source_width, int format, vgpu_op **out);
/* ShrinkVBridge::new_with_ceil (shrinkv.rs:680-699) */
int vgpu_op_shrink_v_new(vgpu_context *ctx, uint32_t factor, int ceil_mode, uint32_t bands,
                         int format, vgpu_op **out);
/* OperationBridge<GaussBlurH<F>> / <GaussBlurV<F>> (gauss_blur.rs:1722-1900). */
int vgpu_op_gauss_blur_h_new(vgpu_context *ctx, float sigma, uint32_t bands, int format,
                             vgpu_op **out);
int vgpu_op_gauss_blur_v_new(vgpu_context *ctx, float sigma, uint32_t bands, int format,
                             vgpu_op **out);
/* OperationBridge<Premultiply<U8>> / <Unpremultiply<U8>>. */
/* ConvSepH::with_mask / ConvSepV::with_mask (convolution/convsep.rs:23-52,
 * 120-149; process_region 53-110, 150-207): acc = fma chain over the mask in
 * FP64, then (acc / scale + offset) as f32; any input format, F32 output. */
int vgpu_op_convsep_h_new(vgpu_context *ctx, const double *coeffs, uint32_t n, double scale, double offset,
                          uint32_t bands, int format, vgpu_op **out);
int vgpu_op_convsep_v_new(vgpu_context *ctx, const double *coeffs, uint32_t n, double scale, double offset,
                          uint32_t bands, int format, vgpu_op **out);
int vgpu_op_premultiply_new(vgpu_context *ctx, uint32_t bands, int format, vgpu_op **out);
int vgpu_op_unpremultiply_new(vgpu_context *ctx, uint32_t bands, int format, vgpu_op **out);
void vgpu_op_destroy(vgpu_op *op);

/* DynOperation::required_input_region (dynamic_op.rs:16-260). */
int vgpu_op_required_input_region(const vgpu_op *op, vgpu_region output, vgpu_region *input);
/* DynOperation::output_width / output_height. */
int vgpu_op_output_size(const vgpu_op *op, uint32_t in_w, uint32_t in_h, uint32_t *out_w,
                        uint32_t *out_h);
int vgpu_op_output_format(const vgpu_op *op, int *format);
/* DynOperation::dyn_process_region(state, input, output, input_region,
 * output_region) (dynamic_op.rs:222-229) for a device scheduler: DEVICE
 * pointers (host pointers are rejected with VGPU_E_INVALID_ARGUMENT),
 * enqueued on the context stream.  Unlike OperationBridge
 * (bridge.rs:208-236), a length that does not match the region returns
 * VGPU_E_BUFFER_SIZE instead of silently writing nothing. */
int vgpu_op_process_region(vgpu_op *op, const void *input, size_t input_len, void *output,
                           size_t output_len, vgpu_region input_region,
                           vgpu_region output_region);
/* The same for HOST buffers, as the unchanged RayonScheduler hands them to
 * dyn_process_region (dynamic_op.rs:222-229; execute.rs:323-325,675-677):
 * the input tile is copied to pooled device memory, processed, and the
 * output copied back before the call returns (synchronous, like the
 * reference's process_region).  Concurrent calls on one op are serialised;
 * distinct ops may be called from different threads. */
int vgpu_op_process_region_host(vgpu_op *op, const void *input, size_t input_len, void *output,
                                size_t output_len, vgpu_region input_region,
                                vgpu_region output_region);

/* ------------------------------------------------------ batched thumbnails */

/* A batch of independent thumbnails of u8 images (1..4 bands) with one
 * target.  Each image gets exactly the plan and output that
 * vgpu_pipeline_thumbnail gives it alone (thumbnail.rs:229-524); the batch
 * runs all of them as a few persistent launches over device task lists (one
 * per plan level and kernel family), captured once into a HIP graph.  No
 * reference counterpart: it replaces `count` separate thumbnail() pipelines
 * run through TileScheduler::run (crates/viprs-ports/src/scheduler.rs:63-98).
 * Images whose plan has stages the batched kernels do not take (an Affine
 * tail, a lone copy) run stage by stage after the batched launches. */
typedef struct vgpu_batch vgpu_batch;

typedef struct {
    uint32_t images;
    uint32_t launches;         /* persistent launches per run (batched levels) */
    uint32_t shrink_tasks;     /* blocks of work in the fused-shrink launches */
    uint32_t reduce_tasks;     /* fused-reduce blocks plus MFMA wavefront tasks */
    uint32_t fallback_images;  /* images run one stage launch at a time */
    int32_t grid_cap;          /* persistent grid size of the shrink launches, 0 = one block per task */
    int32_t use_graph;
    int32_t reduce_first_tile; /* largest reduce tile candidate (0 = 64x16) */
    int32_t reduce_grid_cap;   /* the same for the reduce and MFMA separable launches */
    uint64_t input_bytes;
    uint64_t output_bytes;
    uint64_t intermediate_bytes;
    int32_t reduce_chain;      /* VGPU_OPT_BATCH_REDUCE_CHAIN in effect */
} vgpu_batch_stats;

int vgpu_batch_thumbnail_new(vgpu_context *ctx, uint32_t count, const uint32_t *widths,
                             const uint32_t *heights, const uint32_t *bands, int target,
                             uint32_t width, uint32_t height, int kernel, vgpu_batch **out);
void vgpu_batch_destroy(vgpu_batch *b);
int vgpu_batch_size(const vgpu_batch *b);
int vgpu_batch_output_info(const vgpu_batch *b, uint32_t index, uint32_t *width,
                           uint32_t *height, uint32_t *bands);
int vgpu_batch_info(const vgpu_batch *b, vgpu_batch_stats *stats);
/* Launch shape: persistent grid size of the shrink launches and of the
 * reduce / MFMA separable launches (0 = one block per task), HIP graph on or
 * off, largest reduce tile candidate (re-plans the batch). */
int vgpu_batch_configure(vgpu_batch *b, int grid_cap, int reduce_grid_cap, int use_graph, int reduce_first_tile);
/* As vgpu_pipeline_set_option, for every image; re-plans the batch. */
int vgpu_batch_set_option(vgpu_batch *b, int option, int value);
/* As vgpu_pipeline_set_integer_mfma, for every image; re-plans the batch. */
int vgpu_batch_set_integer_mfma(vgpu_batch *b, int enable);
/* Enqueue the whole batch on the context stream.  srcs[i] / dsts[i] are device
 * pointers (sources 16-byte aligned) with byte lengths src_bytes[i] /
 * dst_bytes[i]; runs of one batch must stay on one stream. */
int vgpu_batch_run(vgpu_batch *b, const void *const *srcs, const size_t *src_bytes,
                   void *const *dsts, const size_t *dst_bytes);
/* Runs the batched launches (no graph, no stage-by-stage images)
 * `iterations` times with HIP events between launches; writes the mean ms,
 * the algorithmic bytes (input buffers + outputs of its jobs) and the kernel
 * family (0 = fused shrink, 1 = fused reduce, 2 = int8-limb MFMA separable
 * pair, 3 = its block-tile form) of each launch; launch_bytes and
 * launch_family may be NULL.  Returns the launch count or a negative status. */
int vgpu_batch_run_profiled(vgpu_batch *b, const void *const *srcs, const size_t *src_bytes,
                            void *const *dsts, const size_t *dst_bytes, int iterations,
                            float *launch_ms, uint64_t *launch_bytes, int *launch_family,
                            int max_launches);
/* Timing harness, as vgpu_pipeline_benchmark. */
int vgpu_batch_benchmark(vgpu_batch *b, const void *const *srcs, const size_t *src_bytes,
                         void *const *dsts, const size_t *dst_bytes, int warmup, int iterations,
                         float *total_ms, float *step_ms);
/* Autotuner: times every (reduce tile, grid sizes, reduce chain on / off)
 * candidate on these buffers and keeps the fastest; best_ms (may be NULL)
 * receives its ms per run. */
int vgpu_batch_tune(vgpu_batch *b, const void *const *srcs, const size_t *src_bytes,
                    void *const *dsts, const size_t *dst_bytes, int iterations, float *best_ms);

/* ------------------------------------------------------------------------
 * Request loop (serving).  The reference runs concurrent pipelines under a
 * cap (ExecutionSemaphore, crates/viprs-runtime/src/limits.rs:290-360;
 * RayonScheduler::with_max_concurrent_pipelines) and rejects requests over
 * its resource limits (ResourceLimits::validate_output, limits.rs:238-275)
 * with a typed error.  Here a server admits requests by device-memory
 * footprint into a preallocated arena of `hbm_budget` bytes and a number of
 * slots, as soon as both free up (continuous batching: requests that arrive
 * while others run are launched next to them on several HIP streams), and
 * reports per-request latency.  A request whose footprint exceeds the whole
 * budget finishes with VGPU_E_ADMISSION. */
typedef struct vgpu_server vgpu_server;

typedef struct {
    uint64_t hbm_budget;      /* arena bytes: input copies, tables, intermediates, outputs */
    uint32_t max_inflight;    /* requests on the device at once */
    uint32_t streams;         /* HIP streams requests are spread over (1..16) */
    uint32_t planner_threads; /* host threads building plans (1..64) */
} vgpu_server_config;

/* One request: thumbnail(target, target_w, target_h, kernel) when
 * thumbnail = 1 (arguments as vgpu_pipeline_thumbnail), then
 * gauss_blur(blur_sigma) when blur_sigma > 0, of an image of
 * width x height x bands samples of `format`.  src is a device pointer
 * (src_on_device = 1), pinned host memory (src_pinned = 1, e.g. from
 * vgpu_host_alloc) or pageable host memory (copied at submit).  The caller
 * keeps device / pinned sources valid until the request finishes.
 * arrival_ns = 0 means "now" (steady clock); latency counts from it. */
typedef struct {
    const void *src;
    uint32_t width, height, bands;
    int format;
    int src_on_device, src_pinned;
    int thumbnail;
    int target;
    uint32_t target_w, target_h;
    int kernel;
    float blur_sigma;
    int64_t arrival_ns;
} vgpu_request;

typedef struct {
    int status;                   /* VGPU_OK or the request's error */
    uint32_t width, height, bands; /* output geometry */
    int format;
    uint64_t out_bytes;
    uint64_t arena_bytes;         /* footprint held from admission to release */
    int64_t plan_ns;              /* arrival -> plan built */
    int64_t queue_ns;             /* arrival -> admission (launch) */
    int64_t run_ns;               /* admission -> completion */
    int64_t latency_ns;           /* arrival -> completion seen by the scheduler */
} vgpu_request_result;

typedef struct {
    uint64_t completed, rejected, inflight, peak_inflight, queued;
    uint64_t arena_bytes, arena_used, arena_peak;
} vgpu_server_counters;

int vgpu_server_create(vgpu_context *ctx, const vgpu_server_config *cfg, vgpu_server **out);
/* Waits for every submitted request, then frees the arena and streams. */
void vgpu_server_destroy(vgpu_server *s);
/* Queue a request; *id names it.  auto_release = 1 frees its memory as soon
 * as it finishes (its output cannot be read then). */
int vgpu_server_submit(vgpu_server *s, const vgpu_request *r, int auto_release, uint64_t *id);
/* Block until request `id` finishes (timeout_us < 0: no limit); returns
 * VGPU_E_NOT_READY on timeout. */
int vgpu_server_wait(vgpu_server *s, uint64_t id, int64_t timeout_us, vgpu_request_result *res);
/* Device pointer and length of a finished request's output (valid until release). */
int vgpu_server_output(vgpu_server *s, uint64_t id, const void **dev_ptr, size_t *bytes);
/* Copy a finished request's output to host memory (synchronous). */
int vgpu_server_fetch(vgpu_server *s, uint64_t id, void *host_dst, size_t bytes);
/* Cancel a request (CancellationToken per request, cancel.rs:6-10): a
 * request not yet on the device is dropped without running; one already
 * launched runs to the end on the device but its result is discarded.  Either
 * way it finishes with status VGPU_E_CANCELLED.  Returns VGPU_OK, or
 * VGPU_E_INVALID_ARGUMENT for an unknown or already finished id. */
int vgpu_server_cancel(vgpu_server *s, uint64_t id);
/* Free a finished request's memory and record. */
int vgpu_server_release(vgpu_server *s, uint64_t id);
/* Counters; returns the server's sticky status. */
int vgpu_server_stats(vgpu_server *s, vgpu_server_counters *st);
/* Open-loop load: request i is submitted at start + arrival_offset_ns[i] and
 * released when done; latency_ns[i] and status[i] are filled in; wall_ns,
 * plan_ns[i] and queue_ns[i] (each may be NULL) receive start -> last
 * completion and the request's arrival -> plan built / -> launch.  No
 * reference counterpart (measurement harness). */
int vgpu_server_replay(vgpu_server *s, const vgpu_request *reqs, uint32_t n,
                       const int64_t *arrival_offset_ns, int64_t *latency_ns, int32_t *status,
                       int64_t *wall_ns, int64_t *plan_ns, int64_t *queue_ns);
/* ---------------------------------------------------------------------
 * Encode after the path and the bytes-in / bytes-out helper
 * (crates/viprs-runtime/src/process.rs:147-260; jpeg/encode.rs through
 * turbojpeg; png via zlib).
 * ------------------------------------------------------------------- */
/* JpegSubsampling (viprs-core codec_options): Auto is 4:2:0 below quality
 * 90, else none (jpeg/common.rs:115-131). */
typedef enum {
    VGPU_JPEG_SUBSAMPLE_AUTO = 0,
    VGPU_JPEG_SUBSAMPLE_OFF = 1,
    VGPU_JPEG_SUBSAMPLE_420 = 2
} vgpu_jpeg_subsampling;
/* Baseline JFIF JPEG of a u8 image in device memory (1 band: grey; 3 bands:
 * YCbCr; 4 bands: YCbCr of the first three, alpha dropped), quality 0..100
 * (0 encodes as 1), the standard tables and no Huffman optimisation -- the
 * bytes libjpeg-turbo's encoder writes for the same settings
 * (JpegCodec::encode_with_options, jpeg/encode.rs:169-258).  The
 * per-block colour conversion, downsampling, DCT and quantisation run on the
 * device, the entropy coding on the host.  *out_len receives the stream's
 * size; VGPU_E_BUFFER_SIZE when out_cap is smaller (out may be NULL to ask). */
int vgpu_jpeg_encode(vgpu_context *ctx, const void *pixels, uint32_t width, uint32_t height, uint32_t bands,
                     int quality, int subsampling, void *out, size_t out_cap, size_t *out_len);
/* PNG decode of host bytes to u8 (png/decode_full.rs:101-108, 470-520:
 * grey, grey + alpha, RGB, RGBA, palette -> RGB; bit depths 1, 2, 4 expanded
 * to 8; tRNS -> an alpha band; Adam7 interlacing undone; 16-bit files return
 * VGPU_E_UNSUPPORTED_FORMAT): the size into *width, *height, *bands, and, when
 * pixels is not NULL, the samples (VGPU_E_BUFFER_SIZE if cap is short). */
int vgpu_png_decode(const void *data, size_t len, uint32_t *width, uint32_t *height, uint32_t *bands, void *pixels,
                    size_t cap);
/* PNG encode of host u8 pixels (1..4 bands) at zlib level 0..9, lossless. */
int vgpu_png_encode(const void *pixels, uint32_t width, uint32_t height, uint32_t bands, int compression, void *out,
                    size_t out_cap, size_t *out_len);
/* EncodeOptions (process.rs:52-111): JPEG (quality, subsampling) or PNG
 * (compression). */
typedef enum { VGPU_ENCODE_JPEG = 0, VGPU_ENCODE_PNG = 1 } vgpu_encode_format;
typedef struct {
    int format;
    int quality;      /* JPEG: 1..100 */
    int subsampling;  /* JPEG: vgpu_jpeg_subsampling */
    int compression;  /* PNG: 0..9 */
} vgpu_encode_options;
/* ProcessResult (process.rs:125-136). */
typedef struct {
    char input_format[8];
    uint32_t input_width, input_height;
    char output_format[8];
    uint64_t bytes_written;
} vgpu_process_result;
/* The `ops` closure of process(): adds nodes to a pipeline over the decoded
 * image (U8, its size and bands) and returns VGPU_OK, or an error that
 * process returns. */
typedef int (*vgpu_process_build_fn)(vgpu_pipeline *p, void *user);
/* process() / process_pipeline() (process.rs:147-260): validate the encode
 * options, decode the input bytes (PNG), run the caller's pipeline on the
 * device, encode (JPEG on the device + host, PNG on the host), copy the
 * stream to out (*out_len = its size; VGPU_E_BUFFER_SIZE when short).  The
 * token is checked between the steps and by the pipeline run
 * (VGPU_E_CANCELLED). */
int vgpu_process(vgpu_context *ctx, const void *input, size_t input_len, vgpu_process_build_fn build, void *user,
                 const vgpu_encode_options *enc, const vgpu_cancel_token *token, void *out, size_t out_cap,
                 size_t *out_len, vgpu_process_result *result);

/* Pinned host memory for request sources (hipHostMalloc). */
int vgpu_host_alloc(size_t bytes, void **out);
void vgpu_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
