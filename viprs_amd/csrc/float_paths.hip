// Fast float paths for f32 images: the same separable H -> V pair as the
// reference-exact kernel, computed in plain FP32 (this file) instead of the
// reference's FP64 accumulation.  Results differ from the reference in the
// last bits; the FP64 error study bounds them.
#include "device.hpp"

namespace vgpu {
namespace {

// Block = TX output columns x TY output rows (ReduceJob tiling).  Pass 1
// evaluates the H outputs of the rows the tile's V taps read into LDS (f32),
// pass 2 applies V.  Both passes accumulate with fmaf in tap order.
template <int B>
__global__ __launch_bounds__(kThreads) void k_red_f32(ReduceJob job) {
    extern __shared__ __align__(16) float hbuf[];
    const InView& in = job.in;
    const OutView& out = job.out;
    const DevPlan& ph = job.ph;
    const DevPlan& pv = job.pv;
    const int TX = job.tx, TY = job.ty;
    const int tx0 = blockIdx.x * TX, ty0 = blockIdx.y * TY;
    const int ncols = min(TX, out.w - tx0), nrows = min(TY, out.h - ty0);
    const int ys0 = pv.start[ty0];
    const int nr = min(job.nr_max, pv.start[ty0 + nrows - 1] + pv.taps - ys0);
    const int x = threadIdx.x % TX, rl = threadIdx.x / TX, nrl = kThreads / TX;
    const int pitch = TX * B;
    const float* src = static_cast<const float*>(in.ptr);
    if (x < ncols) {
        const int xo = tx0 + x;
        const int hs = ph.start[xo];
        const float* w = ph.wf32 + (size_t)ph.phase[xo] * ph.taps;
        for (int r = rl; r < nr; r += nrl) {
            const float* row = src + (size_t)row_of(in, ys0 + r) * in.bw * B;
            float acc[B];
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = 0.f;
            for (int t = 0; t < ph.taps; ++t) {
                const float* px = row + (size_t)col_of(in, hs + t) * B;
                const float wt = w[t];
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] = fmaf(wt, px[b], acc[b]);
            }
#pragma unroll
            for (int b = 0; b < B; ++b) hbuf[r * pitch + x * B + b] = acc[b];
        }
    }
    __syncthreads();
    for (int ry = rl; x < ncols && ry < nrows; ry += nrl) {
        const int yo = ty0 + ry;
        const int vs = pv.start[yo];
        const float* w = pv.wf32 + (size_t)pv.phase[yo] * pv.taps;
        float acc[B];
#pragma unroll
        for (int b = 0; b < B; ++b) acc[b] = 0.f;
        for (int t = 0; t < pv.taps; ++t) {
            const int yy = clampi(clampi(vs + t, job.vclip0, job.vclip1) - ys0, 0, nr - 1);
            const float* hp = hbuf + yy * pitch + x * B;
            const float wt = w[t];
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = fmaf(wt, hp[b], acc[b]);
        }
        float* o = static_cast<float*>(out.ptr) + ((size_t)yo * out.w + tx0 + x) * B;
#pragma unroll
        for (int b = 0; b < B; ++b) o[b] = acc[b];
    }
}

}  // namespace

hipError_t launch_red_f32_job(const ReduceJob& job, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    const dim3 grid((unsigned)job.gx, (unsigned)job.gy);
    const size_t lds = (size_t)job.nr_max * job.tx * job.bands * sizeof(float);
    switch (job.bands) {
    case 1: hipLaunchKernelGGL(k_red_f32<1>, grid, dim3(kThreads), lds, s, job); break;
    case 2: hipLaunchKernelGGL(k_red_f32<2>, grid, dim3(kThreads), lds, s, job); break;
    case 3: hipLaunchKernelGGL(k_red_f32<3>, grid, dim3(kThreads), lds, s, job); break;
    case 4: hipLaunchKernelGGL(k_red_f32<4>, grid, dim3(kThreads), lds, s, job); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace vgpu
