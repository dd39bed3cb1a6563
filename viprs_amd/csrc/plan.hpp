// Host-side planning for the resample / separable-convolution path.
//
// Everything here is FP64 host arithmetic whose rounding decides bit-exactness
// on the device: the 65-phase reduce coefficient tables, the per-axis
// (start, phase) plans, the integer Gaussian masks and the Thumbnail / Resize
// planners.  Device kernels only consume the tables and plans built here.
#pragma once

#include <cstdint>
#include <string>
#include <cmath>
#include <vector>

namespace vgpu {

// Band formats (crates/viprs-core/src/format/band.rs:19-35); the numbers are
// this library's own (the three first ones predate the others).
enum Format : int { U8 = 0, U16 = 1, F32 = 2, I16 = 3, U32 = 4, I32 = 5, F64 = 6 };
// The formats the fused kernels take; the others run on the per-node kernels.
inline bool base_format(int f) { return f == U8 || f == U16 || f == F32; }
enum Kernel : int {
    NEAREST = 0,
    BILINEAR = 1,
    BICUBIC = 2,
    LANCZOS2 = 3,
    LANCZOS3 = 4,
    CATMULLROM = 5,
    LBB = 6,
    VSQBS = 7,
    NOHALO = 8,
};

enum Status : int {
    OK = 0,
    E_INVALID_ARGUMENT = -1,
    E_INVALID_KERNEL = -2,
    E_UNSUPPORTED_FORMAT = -3,
    E_UNSUPPORTED = -4,
    E_BUFFER_SIZE = -5,
    E_DEVICE = -6,
    E_NO_MEMORY = -7,
    E_CANCELLED = -8,
    E_ADMISSION = -9,
};

inline size_t format_size(int f) {
    switch (f) {
    case U8: return 1;
    case U16:
    case I16: return 2;
    case F64: return 8;
    default: return 4;
    }
}

// Region on Z^2 (crates/viprs-core/src/image/region.rs:18-27), widened to 64
// bits so halo arithmetic never wraps on the host.
struct Rect {
    int64_t x = 0, y = 0, w = 0, h = 0;
    bool empty() const { return w <= 0 || h <= 0; }
};

constexpr int kPhases = 64;  // REDUCE_PHASES (reduce_common.rs:5)

// ReduceKernel (reduce_common.rs:87-255) for one axis.
struct ReduceAxis {
    int kernel = LANCZOS3;
    double factor = 1.0;
    int taps = 1;
    int64_t pad = 0;
    double offset = 0.0;
    std::vector<double> wf64;  // (kPhases + 1) x taps, normalised weights
    std::vector<int16_t> wi16; // (kPhases + 1) x taps, trunc(w * 4096)

    double source_position(double index) const;
    void plan(double sp, int64_t* start, int* phase) const;
    // required_input_region_{h,v} along this axis.
    void required(int64_t origin, int64_t len, int64_t* start, int64_t* n) const;
};

int reduce_tap_count(int kernel, double factor);  // -1 when over the 2000-tap limit
double kernel_weight(int kernel, double x);
Status validate_reduce(int kernel, double factor);
Status build_reduce_axis(int kernel, double factor, uint32_t input_len, ReduceAxis* out);
uint32_t reduce_output_len(double factor, uint32_t input_len);

// Integer Gaussian mask (gauss_blur.rs:80-126).
struct GaussMask {
    std::vector<int16_t> taps;
    int32_t scale = 1;
    std::vector<double> weights;  // taps / scale (gauss_kernel.rs:58-65)
    int radius() const { return (int)(taps.size() - 1) / 2; }
};
GaussMask build_gauss_mask(float sigma);

// Pipeline node kinds (the reference's bridge set on this path).
enum NodeKind : int {
    N_SHRINK_H = 1,
    N_SHRINK_V = 2,
    N_REDUCE_H = 3,
    N_REDUCE_V = 4,
    N_GAUSS_H = 5,
    N_GAUSS_V = 6,
    N_CONVSEP_H = 7,
    N_CONVSEP_V = 8,
    N_PREMULT = 9,
    N_UNPREMULT = 10,
    N_COPY = 11,
    // ReduceOp (reduce.rs:23-262) as two nodes, H then V, both axes with
    // offset 0: the backward walk then yields ReduceOp's scratch region
    // (required_input_region_v of the output) as the V node's input region.
    N_REDUCE_OP_H = 12,
    N_REDUCE_OP_V = 13,
    N_EXTRACT = 14,  // ExtractArea (structural/extract_area.rs:23-60)
    N_AFFINE = 15,   // Affine: [a, 0, 0, d], ExtendMode::Copy (affine/core.rs)
    N_ZOOM = 16,     // Zoom (structural/zoom.rs:339-400)
    // The PNG box pre-shrink that replaces the source before any node
    // (decoder_source/shrink.rs:47-117; png/decode_full.rs:320-452): only
    // ever node 0; its output is read as the chain's source.
    N_BOX_SHRINK = 17,
};

struct Node {
    NodeKind kind = N_COPY;
    int in_fmt = U8, out_fmt = U8;
    uint32_t bands = 1;
    uint32_t in_w = 0, in_h = 0, out_w = 0, out_h = 0;
    uint32_t factor = 1;        // shrink
    uint32_t source_width = 0;  // ShrinkH full-row borrow width (0 = none)
    ReduceAxis axis;            // reduce
    std::vector<double> conv;   // gauss float weights / convsep coefficients
    std::vector<int16_t> itaps; // gauss integer taps
    int32_t iscale = 1;
    double cscale = 1.0, coffset = 0.0;
    int32_t ex = 0, ey = 0;     // extract area: offset of the output in the input
    // affine: matrix [a, b, c, d], translation, kernel (NEAREST / BILINEAR /
    // BICUBIC), and whether the scheduler runs it on 128 x 128 tiles
    // (SmallTile) rather than full-width strips
    double am[4] = {1.0, 0.0, 0.0, 1.0}, atx = 0.0, aty = 0.0;
    int akern = 0;
    bool asmall = false;
    uint32_t zx = 1, zy = 1;  // zoom factors
    int box_ceil = 0;         // box pre-shrink: ceil output size, partial edge windows

    int radius() const { return conv.empty() ? 0 : (int)(conv.size() - 1) / 2; }
    Rect required_input(const Rect& out) const;
    // Affine::map_to_input (affine/core.rs:78-84)
    void affine_map(double x, double y, double* mx, double* my) const {
        *mx = std::fma(am[1], y, am[0] * x) + atx;
        *my = std::fma(am[3], y, am[2] * x) + aty;
    }
};

// Per-column (or per-row) source positions of an Affine node over
// [first, first + n) of its output: Affine::process_region (affine/core.rs:
// 1593-1632) starts each scheduler tile at its origin and steps one matrix
// entry per pixel, so a position depends on where its tile starts.  Tiles
// are `tile` wide in the pipeline's output, which sits at `shift` in the
// Affine's output.  Matrix diagonal: columns use a / tx, rows d / ty.
void affine_positions(const Node& nd, bool columns, int64_t first, int64_t n, int64_t tile, int64_t shift,
                      std::vector<double>* pos);

// Thumbnail planner (thumbnail.rs:229-524), in-memory source (no
// shrink-on-load), crop = false.
enum ThumbTarget : int { FITBOX = 0, FORCE_EXACT = 1, WIDTH = 2, HEIGHT = 3 };
struct PlanStep {
    NodeKind kind;
    double param;
    uint32_t ex = 0, ey = 0, ew = 0, eh = 0;  // N_EXTRACT rectangle; N_AFFINE output size (ew x eh)
    double a = 1.0, d = 1.0, tx = 0.0, ty = 0.0;  // N_AFFINE
    int akern = 0;
    uint32_t zx = 1, zy = 1;  // N_ZOOM
};
// crop: Thumbnail::with_crop (thumbnail.rs:185-190), a centred ExtractArea
// of the covering geometry.
Status plan_thumbnail(uint32_t in_w, uint32_t in_h, uint32_t bands, int target, uint32_t tw,
                      uint32_t th, std::vector<PlanStep>* steps, uint32_t* out_w,
                      uint32_t* out_h, bool crop = false, int kernel = 4, uint32_t* shrink_on_load = nullptr);
// ResizeOp::into_pipeline_nodes, downscale subset (resize.rs:163-284).
Status plan_resize(uint32_t in_w, uint32_t in_h, double hscale, double vscale, int kernel,
                   std::vector<PlanStep>* steps, int* eff_kernel);

}  // namespace vgpu
