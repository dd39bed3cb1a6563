// Device kernel launchers for the resample / separable-convolution path.
//
// Every launcher reads its input through an InView: a packed band-interleaved
// buffer that covers `buf` (a rectangle of Z^2), plus the rectangle `clip` the
// reference's tile buffer would have held.  A read of column x is served from
// clamp(clamp(x, clip), buf): the inner clamp restates the in-tile clamp_axis
// of the reduce kernels (reduce_common.rs:318-322), the outer one the
// clamp-to-edge source read (sources/memory.rs:149-235).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vgpu {

struct InView {
    const void* ptr = nullptr;
    int32_t bx = 0, by = 0, bw = 0, bh = 0;  // buffer rectangle
    int32_t cx0 = 0, cy0 = 0, cx1 = 0, cy1 = 0;  // clip rectangle, inclusive
};

struct OutView {
    void* ptr = nullptr;
    int32_t x = 0, y = 0, w = 0, h = 0;  // output rectangle (dense buffer)
};

// Per-axis plan uploaded to the device: absolute start column/row of tap 0 and
// the phase index (0..64) of every output column/row.
struct DevPlan {
    const int32_t* start = nullptr;
    const uint8_t* phase = nullptr;
    const int16_t* wi16 = nullptr;  // 65 x taps
    const double* wf64 = nullptr;   // 65 x taps
    const float* wf32 = nullptr;    // 65 x taps (float modes)
    int taps = 1;
};

enum Alpha : int { ALPHA_NONE = 0, ALPHA_PREMULT = 1, ALPHA_UNPREMULT = 2 };

// One output column (or row) of an axis-aligned Affine: the first source
// index read (trunc(x) for Nearest and Bilinear, floor(x) - S + 1 for the
// separable kernels of half-width S, floor(x) - 1 for the 4 x 4 stencils),
// the phase (x - trunc(x) for Bilinear, x - floor(x) for the stencils), and
// the separable weights of source indices i0 .. i0 + 2S - 1.
enum AffineKernel : int {
    AFF_NEAREST = 0,
    AFF_BILINEAR = 1,
    AFF_SEP4 = 2,   // Bicubic, CatmullRom, Lanczos2: interp_separable<2>
    AFF_SEP6 = 3,   // Lanczos3: interp_separable<3>
    AFF_VSQBS = 4,  // interp_vsqbs: 4 x 4 stencil, quadratic B-spline (vertex split)
    AFF_LBB = 5,    // interp_lbb: 4 x 4 stencil, locally bounded bicubic
};
struct AffineTap {
    int32_t i0 = 0, pad = 0;
    double f = 0.0;
    double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
};
struct AffineJob {
    InView in;
    OutView out;
    int bands = 3, fmt = 0, kernel = 2;
    int32_t in_w = 0, in_h = 0;  // the Affine's input image: ExtendMode::Copy clamps to it
    int32_t zx = 1, zy = 1;      // a Zoom folded in: `in` holds the image before it
    // u8 with 2 or 4 bands: a Premultiply before the Affine (and its Zoom)
    // and an Unpremultiply after it folded into the sample reads and stores
    int premult = 0, unpremult = 0;
    const AffineTap* xs = nullptr;  // out.w entries
    const AffineTap* ys = nullptr;  // out.h entries
};

// Fused [Premultiply] -> ShrinkH -> ShrinkV -> [Unpremultiply] on u8, one
// image.  plan_shrink_job fills the launch shape (tx, S, m, lsl, mh, mv, gx,
// gy); it returns false when the wavefront kernel cannot take the job.
struct ShrinkJob {
    InView in;
    OutView out;
    int bands = 3, fh = 1, fv = 1, premult = 0, unpremult = 0;
    int tx = 64, S = 1;
    int m = 1, lsl = 8;  // output rows per wavefront run; log2 chunks per source-row slot
    uint32_t mh = 0, mv = 0;
    // ShrinkH / ShrinkV roundings as ((sum + add) * mul) >> shift (every
    // u8 rounding form of the reference fits: pair, exact /5, libvips
    // multiply-shift; mul < 2^24 and the product < 2^32).
    uint32_t rh_add = 0, rh_mul = 1, rh_shift = 0, rv_add = 0, rv_mul = 1, rv_shift = 0;
    int wnd_t = 16;  // lean forms: window dwords rounded up to 4, 8, 12 or 16 (0: longer)
    int lean = 1;    // lean forms allowed (32-bit buffer offsets: source < 4 GiB)
    int ostride = 0; // output row pitch in pixels (0: out.w)
    int gx = 0, gy = 0;
    int src_slot = -1, dst_slot = -1;  // batch: pointer taken from the per-run tables
};

// Fused [Premultiply] -> ReduceH -> ReduceV -> [Unpremultiply] (bands <= 4;
// the alpha steps on u8 with 2 or 4 bands only); a missing axis is
// an identity plan (factor 1, one 4096 tap).  ph plans the intermediate's
// columns (indexed from out.x), pv the output rows; [vclip0, vclip1] is
// ReduceV's input row range (its reference tile rectangle); nr_max bounds the
// staged rows; wide_acc selects int64 fixed-point sums.
struct ReduceJob {
    InView in;
    OutView out;
    DevPlan ph, pv;
    int bands = 3, fmt = 0, tx = 16, ty = 8, nr_max = 1;
    int vclip0 = 0, vclip1 = 0, wide_acc = 0, premult = 0, unpremult = 0;
    int gx = 0, gy = 0;
    int src_slot = -1, dst_slot = -1;
    int op = 0;  // 1: ReduceOp arithmetic (f64 mul_add chains over wf64, from_f64 rounding)
    // H source staging: span_max = the widest span of source columns a tile's
    // H taps read, hr = source rows staged in LDS per chunk, rb = H rows a
    // thread of the f32 path evaluates at once (hr = rb * 256 / tx).
    int span_max = 1, hr = 1, rb = 1;
    // Streaming form of the f32 pair (stream = 1): strips of st_tx output
    // columns whose H taps read at most st_span source columns, runs of
    // st_run output rows, at most st_open output rows whose V windows are
    // open at once, st_gx x st_gy blocks, st_pad >= the most intermediate
    // rows a run reads.
    int stream = 0, st_tx = 0, st_span = 0, st_run = 0, st_open = 0, st_gx = 0, st_gy = 0, st_pad = 0;
};

// Separable u8 pass pair on MFMA (int8 limbs): H then V with a u8
// intermediate, each pass a banded contraction on v_mfma_i32_16x16x64_i8.
// Family 0 (reduce): fused [Premultiply] -> ReduceH -> ReduceV ->
// [Unpremultiply]; the H weights of a 16-sample strip are folded in the
// kernel from ph (absolute first-tap column per output column, indexed from
// out.x; 65 x taps i16 rows), rounding (sum + 2048) >> 12 clamped.  Family 1
// (tabled): GaussBlurH -> GaussBlurV and [Premultiply] -> ShrinkH -> ShrinkV
// -> [Unpremultiply]; the H operands come from host-folded patterns, rounding
// q = mul_hi(x, mul) >> shift or (x * mul) >> shift per pass, the bias already
// in the seeds.  Both families take the V operands from host-folded patterns
// per 16 output rows.  A wavefront task is one 16-sample strip (jx) x
// rows_per_task output rows (ty).
struct SepJob {
    InView in;
    OutView out;
    DevPlan ph;                         // family 0: H plan
    const unsigned char* htab = nullptr;  // family 1: H operand patterns
    const int32_t* hstrip = nullptr;      // family 1: per strip (pattern, first span pixel)
    int htab_bytes = 0;
    // V operand patterns: kv k-steps x limbs A operands (64 lanes x 16
    // bytes; {hi, lo} when limbs = 2), then 16 int32 seeds 128 * sum(w) + bias.
    const unsigned char* vtab = nullptr;
    const int32_t* vtile = nullptr;  // per 16-row tile: (pattern, base16, hi)
    int vtab_bytes = 0;
    int family = 0, limbs = 2, bands = 3;
    uint32_t h_mul = 1, v_mul = 1;  // family 1 rounding
    int h_shift = 0, v_shift = 0, h_hi = 0, v_hi = 0;
    int h_bias = 0;                  // family 0: H seed bias
    int kh = 1, kv = 1, ring = 128, rows_per_task = 16;
    int premult = 0, unpremult = 0;
    int gx = 0, gy = 0;  // column strips, row chunks
    int src_slot = -1, dst_slot = -1;
    // Block-tile form (tile != 0): a block of four wavefronts takes four
    // adjacent strips x rows_per_task output rows, with its input rows (each
    // t_row_stride bytes, at most t_rows of them), the intermediate columns
    // (t_inter_stride bytes per sample) and the V operands of its 16-row
    // tiles staged in t_lds bytes of LDS; H operands come from htab / hstrip
    // for both families.  t_chunk_magic = ceil(2^32 / (t_row_stride / 16)).
    int tile = 0;
    int t_row_stride = 0, t_rows = 0, t_inter_stride = 0, t_lds = 0;
    uint32_t t_chunk_magic = 0;
};

struct BatchTask {
    int job, bx, by;
};

hipError_t launch_shrink_hv_u8(const InView& in, const OutView& out, int bands, int fh, int fv,
                               int premult, hipStream_t s);
// The PNG box pre-shrinks over the output rectangle [x0, x0 + rw) x
// [y0, y0 + rh) of a dense dw-wide output (any format, window cut at the
// source's edges).
hipError_t launch_box_generic(const void* src, int sw, int sh, int bands, int fmt, int f, void* dst, int dw,
                              int x0, int y0, int rw, int rh, hipStream_t s);
hipError_t launch_shrink_generic(const InView& in, const OutView& out, int bands, int fmt,
                                 int factor, int vertical, hipStream_t s);
// op = 1: ReduceOp's passes instead of ReduceH / ReduceV.
hipError_t launch_reduce_h(const InView& in, const OutView& out, int bands, int fmt,
                           const DevPlan& plan, int op, hipStream_t s);
hipError_t launch_reduce_v(const InView& in, const OutView& out, int bands, int fmt,
                           const DevPlan& plan, int op, hipStream_t s);
bool plan_shrink_job(ShrinkJob* job, bool row_split = true, int run = 0);
// The same kernel as the u8 PNG box pre-shrink (fh == fv == f): window sums
// of each row unrounded, one rounding (sum + f*f/2) / (f*f) of the total.
// False when the wavefront kernel cannot take the job or f*f has no exact
// multiply-shift.
bool plan_box_shrink_job(ShrinkJob* job, bool row_split = true);
size_t shrink_job_lds();
hipError_t launch_shrink_job(const ShrinkJob& job, hipStream_t s);
// Kernel variant of a planned job (its band count); a batch launch runs jobs
// of one variant.
int shrink_job_variant(const ShrinkJob& job);
hipError_t launch_shrink_batch(int variant, const ShrinkJob* jobs, const BatchTask* tasks, int ntasks,
                               const void* const* srcs, void* const* dsts, int grid, hipStream_t s);
size_t reduce_job_lds(const ReduceJob& job);
hipError_t launch_reduce_job(const ReduceJob& job, hipStream_t s);
hipError_t launch_reduce_batch(const ReduceJob* jobs, const BatchTask* tasks, int ntasks,
                               const void* const* srcs, void* const* dsts, int grid, size_t lds,
                               hipStream_t s);
// LDS bytes one wavefront of the MFMA separable kernel needs for this job;
// whether the kernel takes the job's shape.
size_t sep_job_wave_lds(const SepJob& job);
bool sep_job_supported(const SepJob& job);
hipError_t launch_sep_job(const SepJob& job, hipStream_t s);
// Kernel variant of a job (bands, limbs, mode, premultiply); a batch launch
// runs jobs of one variant, one (job, jx, ty) task per wavefront.
int sep_job_variant(const SepJob& job);
hipError_t launch_sep_batch(int variant, const SepJob* jobs, const BatchTask* tasks, int ntasks,
                            const void* const* srcs, void* const* dsts, int grid, size_t wave_lds,
                            hipStream_t s);
// The block-tile form over a task list: one (job, bx, by) task per block,
// bx = strip group of four, by = row chunk; lds = the largest t_lds.
hipError_t launch_sep_tile_batch(int variant, const SepJob* jobs, const BatchTask* tasks, int ntasks,
                                 const void* const* srcs, void* const* dsts, int grid, size_t lds,
                                 hipStream_t s);
// Float separable pair in plain FP32: f32 weights, FMA accumulation, f32
// intermediate (ReduceJob tiling; ph / pv carry wf32).
hipError_t launch_red_f32_job(const ReduceJob& job, hipStream_t s);
// Streaming f32 reduce pair (job.stream): mode 0 = the reference's FP64
// order, 2 = plain FP32.  reduce_stream_lds = its dynamic LDS bytes.
hipError_t launch_reduce_stream(const ReduceJob& job, int mode, hipStream_t s);
size_t reduce_stream_lds(const ReduceJob& job, int mode);

// Float separable pair on MFMA (v_mfma_f32_16x16x32_f16) with an FP16 hi/lo
// split of pixels and weights (x = hi + lo, products hi*hi + hi*lo + lo*hi)
// and FP32 accumulation.  Operand patterns are host-folded like SepJob family
// 1: per 16-sample strip (pattern, first span pixel), per 16 output rows
// (pattern, base8, hi); K steps of 32 samples / rows.
struct F16Job {
    InView in;
    OutView out;
    const unsigned char* htab = nullptr;
    const int32_t* hstrip = nullptr;
    int htab_bytes = 0;
    const unsigned char* vtab = nullptr;
    const int32_t* vtile = nullptr;
    int vtab_bytes = 0;
    int bands = 3, kh = 1, kv = 1, ring = 128, rows_per_task = 16;
    int gx = 0, gy = 0;
    int h_exp = 0, v_exp = 0;  // the tables hold weights * 2^h_exp (H), * 2^v_exp (V)
};
size_t f16_job_wave_lds(const F16Job& job);
hipError_t launch_f16_job(const F16Job& job, hipStream_t s);

// GaussBlurH -> GaussBlurV on 3- or 4-band u8 on the VALU (packed 16-bit
// sums): a block walks tiles_per_block 64 x th output tiles down a column.  taps2[t] holds tap t in both
// 16-bit halves, half2 = (scale / 2) in both halves, mul = ceil(2^32 / scale)
// (0 when scale == 1).
constexpr int kGaussU8MaxRadius = 6;
struct GaussJob {
    InView in;
    OutView out;
    int bands = 4, radius = 1, th = 16;
    uint32_t taps2[2 * kGaussU8MaxRadius + 1] = {};
    uint32_t half2 = 0, mul = 0;
    int tiles_per_block = 1;  // consecutive tile rows one block walks down
    int gx = 0, gy = 0;
};
// Baseline JPEG block stage (jpeg.hip): the quantised DCT coefficients of
// every 8 x 8 block of every component, zigzag order, components one after
// another (component c's blocks at [off[c], off[c + 1]), raster order within).
// mode 0: one grey component; 1: YCbCr 4:4:4; 2: YCbCr 4:2:0.
struct JpegJob {
    const unsigned char* px = nullptr;  // u8 pixels, w x h x bands (1 or 3)
    int w = 0, h = 0, bands = 1, mode = 0;
    int bw[3] = {}, bh[3] = {};  // blocks per row / column of each component
    int off[4] = {};             // first block of each component; off[ncomp] = total
    uint16_t div[2][64] = {};    // luma / chroma divisors (8 x table entry), natural order
    int16_t* out = nullptr;
};
hipError_t launch_jpeg_blocks(const JpegJob& j, hipStream_t s);
bool gauss_u8_supported(int bands, int taps, int32_t scale);
hipError_t launch_gauss_u8(const GaussJob& job, hipStream_t s);
// Fused float H -> V convolution pair (conv_pair.hip): a strip of tw output
// columns (tw * bands <= 256) walking down runs of `run` output rows; the
// masks have 2 * rad + 1 taps; kind 0 = Gaussian (FP64, no FMA), 1 = ConvSep
// (FP64 FMA chains, / scale + offset), 2 = Gaussian in plain FP32.
constexpr int kConvPairThreads = 256;  // threads per block = samples per strip row, at most
struct ConvPairJob {
    InView in;
    OutView out;
    int bands = 1, tw = 64, rad = 0, run = 64;
    const double* wh = nullptr;
    const double* wv = nullptr;
    const float* wh32 = nullptr;
    const float* wv32 = nullptr;
    double hscale = 1.0, hoffset = 0.0, vscale = 1.0, voffset = 0.0;
};
bool conv_pair_supported(int bands, int rad);
size_t conv_pair_lds(const ConvPairJob& job);
hipError_t launch_conv_pair(const ConvPairJob& job, int in_fmt, int kind, hipStream_t s);
hipError_t launch_conv(const InView& in, const OutView& out, int bands, int in_fmt, int out_fmt,
                       int kind, int vertical, const int16_t* itaps, int32_t iscale,
                       const double* wts, int ntaps, double cscale, double coffset,
                       hipStream_t s);
hipError_t launch_alpha_u8(const InView& in, const OutView& out, int bands, int mode,
                           hipStream_t s);
hipError_t launch_copy(const InView& in, const OutView& out, int bands, int fmt, hipStream_t s);
hipError_t launch_affine(const AffineJob& job, hipStream_t s);
// Zoom (structural/zoom.rs:364-400): output (x, y) repeats input (x / zx, y / zy).
hipError_t launch_zoom(const InView& in, const OutView& out, int bands, int fmt, int zx, int zy, hipStream_t s);
hipError_t launch_fill_random(void* dst, size_t bytes, uint64_t seed, hipStream_t s);
// Streaming read of bytes (a multiple of 16) by `blocks` blocks; sink holds
// one word per block.
hipError_t launch_stream_read(const void* src, size_t bytes, uint32_t* sink, int blocks, hipStream_t s);

}  // namespace vgpu
