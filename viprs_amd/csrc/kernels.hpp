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
    int taps = 1;
};

enum Alpha : int { ALPHA_NONE = 0, ALPHA_PREMULT = 1, ALPHA_UNPREMULT = 2 };

hipError_t launch_shrink_hv_u8(const InView& in, const OutView& out, int bands, int fh, int fv,
                               int premult, hipStream_t s);
hipError_t launch_shrink_generic(const InView& in, const OutView& out, int bands, int fmt,
                                 int factor, int vertical, hipStream_t s);
hipError_t launch_reduce_h(const InView& in, const OutView& out, int bands, int fmt,
                           const DevPlan& plan, int post_alpha, hipStream_t s);
hipError_t launch_reduce_v(const InView& in, const OutView& out, int bands, int fmt,
                           const DevPlan& plan, int post_alpha, hipStream_t s);
// Fused ReduceH -> ReduceV (bands <= 4): ph plans the ReduceH output columns
// (indexed from out.x), pv the ReduceV output rows; [vclip0, vclip1] is
// ReduceV's input row range (its reference tile rectangle).  nr_max bounds
// the staged rows.  wide_acc selects int64 fixed-point sums (int32 otherwise).
hipError_t launch_reduce_hv(const InView& in, const OutView& out, int bands, int fmt,
                            const DevPlan& ph, const DevPlan& pv, int tx, int ty, int nr_max,
                            int vclip0, int vclip1, int wide_acc, hipStream_t s);
hipError_t launch_conv(const InView& in, const OutView& out, int bands, int in_fmt, int out_fmt,
                       int kind, int vertical, const int16_t* itaps, int32_t iscale,
                       const double* wts, int ntaps, double cscale, double coffset,
                       hipStream_t s);
hipError_t launch_alpha_u8(const InView& in, const OutView& out, int bands, int mode,
                           hipStream_t s);
hipError_t launch_copy(const InView& in, const OutView& out, int bands, int fmt, hipStream_t s);

}  // namespace vgpu
