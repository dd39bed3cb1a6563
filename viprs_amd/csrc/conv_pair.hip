// Fused separable float convolution pair: GaussBlurH -> GaussBlurV on the
// float path (u16 / f32 inputs, F32 out) and ConvSepH -> ConvSepV (any
// input, F32 out), in the reference's own FP64 operation order, or in plain
// FP32 for the Gaussian's fast mode.
//
// A block owns a strip of TW output columns (TW * B = up to 256 samples, one
// thread per sample) and walks down a run of output rows, four source rows
// per step.  It stages the rows' spans (TW + 2r columns, clamp-to-edge) in
// LDS as f32; every thread evaluates the H outputs of its sample for the four
// rows (four independent accumulation chains), rounds them to f32 as the
// reference's intermediate tile holds them, and keeps them in its own column
// of an LDS ring of H rows; then it evaluates V, in tap order, for each of
// the four rows whose window the ring holds.  Each thread reads only the ring
// entries it wrote, so only the staged rows need block barriers.  The next
// four rows are loaded into registers while the current ones are computed.
//
// Arithmetic per pass (gauss_blur.rs:410-464; convsep.rs:87-107,184-204;
// common.rs:150-153):
//   KIND 0 (Gaussian): acc = acc + f64(f32(x)) * w, tap by tap, no FMA;
//   KIND 1 (ConvSep):  acc = fma(f64(x), w, acc), then acc / scale + offset;
//   KIND 2 (Gaussian, plain FP32): acc = fmaf(w32, x, acc).
// The result of each pass is `as f32`.
#include <type_traits>

#include "device.hpp"

namespace vgpu {
static_assert(kConvPairThreads == kThreads, "one thread per strip sample");
namespace {

template <typename T>
__device__ __forceinline__ float sample_f32(const void* base, size_t i) {
    return (float)gmem<T>(base)[i];
}

constexpr int kRowsPerStep = 4;  // source rows staged and evaluated per block step

// Tap-ordered accumulation of RB independent outputs at once: output k sums
// x(k, t) * w[t] for t < n in tap order (KIND 0: separate multiply and add
// in FP64; KIND 1: FP64 FMA; KIND 2: FP32 FMA), so each chain keeps the
// reference's rounding sequence while the RB chains overlap their latencies.
template <int KIND, int RB, typename Acc>
__device__ __forceinline__ void chain_step(const float (&xv)[RB], double w64, float w32, Acc (&acc)[RB]) {
    if constexpr (KIND == 2) {
#pragma unroll
        for (int k = 0; k < RB; ++k) acc[k] = fmaf(w32, xv[k], acc[k]);
    } else if constexpr (KIND == 1) {
#pragma unroll
        for (int k = 0; k < RB; ++k) acc[k] = fma((double)xv[k], w64, acc[k]);
    } else {
#pragma unroll
        for (int k = 0; k < RB; ++k) acc[k] = acc[k] + (double)xv[k] * w64;
    }
}

template <int KIND, int RB, typename X, typename Acc>
__device__ __forceinline__ void chains(X x, const double* w64, const float* w32, int t0, int count,
                                       Acc (&acc)[RB]) {
    // Taps t0 .. t0 + count - 1 in blocks of TB: all their LDS reads are
    // issued before the first dependent step, so the chains do not wait on
    // LDS latency tap by tap.
    constexpr int TB = 4;
    const int n = t0 + count;
    int t = t0;
    for (; t + TB <= n; t += TB) {
        float xv[TB][RB];
        double wd[TB];
        float wf[TB];
#pragma unroll
        for (int u = 0; u < TB; ++u) {
#pragma unroll
            for (int k = 0; k < RB; ++k) xv[u][k] = x(k, t + u);
            if constexpr (KIND == 2) wf[u] = w32[t + u];
            else wd[u] = w64[t + u];
        }
#pragma unroll
        for (int u = 0; u < TB; ++u) chain_step<KIND, RB>(xv[u], KIND == 2 ? 0.0 : wd[u], KIND == 2 ? wf[u] : 0.f, acc);
    }
    for (; t < n; ++t) {
        float xv[RB];
#pragma unroll
        for (int k = 0; k < RB; ++k) xv[k] = x(k, t);
        chain_step<KIND, RB>(xv, KIND == 2 ? 0.0 : w64[t], KIND == 2 ? w32[t] : 0.f, acc);
    }
}

template <typename Tin, int KIND, int B>
__global__ __launch_bounds__(kThreads) void k_conv_pair(ConvPairJob job) {
    using Acc = typename std::conditional<KIND == 2, float, double>::type;
    constexpr int RB = kRowsPerStep;
    constexpr int TW = kThreads / B;
    constexpr int tws = TW * B;             // samples per strip row
    extern __shared__ __align__(16) float smem[];
    const InView& in = job.in;
    const OutView& out = job.out;
    const int r = job.rad, n = 2 * job.rad + 1;
    // H ring slots: a step writes RB rows before the V windows of its rows
    // are read, so the ring keeps the n rows of the oldest window plus RB - 1;
    // slots RS .. RS + RB - 2 repeat slots 0 .. RB - 2, so the RB rows of any
    // window start q are the contiguous slots q .. q + RB - 1.
    const int RS = n + RB - 1;
    const int span = (TW + 2 * r) * B;      // staged source samples per row
    const int s = threadIdx.x;              // this thread's sample in the strip
    // Neighbouring strips share their 2r halo columns: XCD-aware order keeps
    // them on one L2 (a contiguous range of strips per XCD).
    const int lb = xcd_block((int)(blockIdx.y * gridDim.x + blockIdx.x), (int)(gridDim.x * gridDim.y));
    const int bx = lb % (int)gridDim.x, by = lb / (int)gridDim.x;
    const int x0 = out.x + bx * TW;  // first output column of the strip
    const int ncols = min(TW, out.x + out.w - x0);
    const bool active = s < ncols * B;
    const int yo0 = out.y + by * job.run;
    const int nrows = min(job.run, out.y + out.h - yo0);
    if (nrows <= 0) return;
    // LDS: the H and V weights (read per tap as broadcasts; from global
    // memory every tap would wait on a scalar load), two groups of RB staged
    // source rows, then the H ring (RS rows x tws samples).
    double* wh = reinterpret_cast<double*>(smem);
    double* wv = wh + n;
    float* wh32 = reinterpret_cast<float*>(wv + n);
    float* wv32 = wh32 + n;
    // (The two staging groups are addressed arithmetically from the shared
    // base: a pointer picked from an array loses the LDS address space and
    // turns every tap read into a flat access.)
    float* const stage0 = wv32 + n;
    float* ring = stage0 + 2 * RB * span;
    for (int t = s; t < n; t += kThreads) {
        wh[t] = job.wh[t];
        wv[t] = job.wv[t];
        wh32[t] = job.wh32[t];
        wv32[t] = job.wv32[t];
    }
    const size_t row_samples = (size_t)in.bw * B;

    // Row j of the run is source row yo0 - r + j; the thread stages samples
    // e = s, s + 256, ... of each row's span (clamp-to-edge columns).
    const int total = nrows + 2 * r;
    const int ngroups = (total + RB - 1) / RB;
    float nxt[RB][4];
    auto load_group = [&](int g) {
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            const int j = min(g * RB + k, total - 1);
            const size_t rowoff = (size_t)row_of(in, yo0 - r + j) * row_samples;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = s + q * kThreads;
                nxt[k][q] = 0.f;
                if (e < span) {
                    const int cc = B == 3 ? (e * 21846) >> 16 : e >> (B >> 1);  // e / B for e < 1024
                    const int col = col_of(in, x0 - r + cc);
                    nxt[k][q] = sample_f32<Tin>(in.ptr, rowoff + (size_t)col * B + (e - cc * B));
                }
            }
        }
    };
    auto park_group = [&](float* dst) {
#pragma unroll
        for (int k = 0; k < RB; ++k)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = s + q * kThreads;
                if (e < span) dst[k * span + e] = nxt[k][q];
            }
    };

    load_group(0);
    for (int g = 0; g < ngroups; ++g) {
        float* cur = stage0 + (g & 1) * (RB * span);
        park_group(cur);
        __syncthreads();
        if (g + 1 < ngroups) load_group(g + 1);
        if (active) {
            const int j0 = g * RB;
            // H outputs of rows j0 .. j0 + RB - 1 of this sample.
            Acc hacc[RB];
#pragma unroll
            for (int k = 0; k < RB; ++k) hacc[k] = 0;
            const float* px = cur + s;  // s = c * B + b
            chains<KIND, RB>([&](int k, int t) { return px[k * span + t * B]; }, wh, wh32, 0, n, hacc);
#pragma unroll
            for (int k = 0; k < RB; ++k) {
                float h;
                if constexpr (KIND == 1) h = (float)(hacc[k] / job.hscale + job.hoffset);
                else h = (float)hacc[k];
                if (j0 + k < total) {
                    const int slot = (j0 + k) % RS;
                    ring[slot * tws + s] = h;
                    if (slot < RB - 1) ring[(slot + RS) * tws + s] = h;
                }
            }
            // V outputs of rows j0 + k >= 2r (output row yo0 + j0 + k - 2r):
            // ring rows j0 + k - 2r + t.  Row j0 - 2r + t sits at slot
            // q0 + t for t < RS - q0, then at q0 + t - RS; row k more is k
            // slots on (the repeated slots cover the wrap).
            if (j0 + RB - 1 >= 2 * r) {
                const int q0 = ((j0 - 2 * r) % RS + RS) % RS;
                const int na = min(n, RS - q0);
                Acc vacc[RB];
#pragma unroll
                for (int k = 0; k < RB; ++k) vacc[k] = 0;
                const float* seg_a = ring + q0 * tws + s;
                const float* seg_b = ring + (q0 - RS) * tws + s;  // indexed by t >= RS - q0
                chains<KIND, RB>([&](int k, int t) { return seg_a[(k + t) * tws]; }, wv, wv32, 0, na, vacc);
                chains<KIND, RB>([&](int k, int t) { return seg_b[(k + t) * tws]; }, wv, wv32, na, n - na, vacc);
#pragma unroll
                for (int k = 0; k < RB; ++k) {
                    const int j = j0 + k;
                    if (j < 2 * r || j >= total) continue;
                    float v;
                    if constexpr (KIND == 1) v = (float)(vacc[k] / job.vscale + job.voffset);
                    else v = (float)vacc[k];
                    static_cast<float*>(out.ptr)[((size_t)(yo0 + j - 2 * r - out.y) * out.w + (x0 - out.x)) * B + s] =
                        v;
                }
            }
        }
        __syncthreads();  // the staged group is overwritten two steps on
    }
}

}  // namespace

size_t conv_pair_lds(const ConvPairJob& j) {
    const size_t span = (size_t)(j.tw + 2 * j.rad) * j.bands;
    const size_t n = 2 * (size_t)j.rad + 1;
    return 2 * n * (sizeof(double) + sizeof(float)) +
           (2 * (size_t)kRowsPerStep * span + (n + 2 * (kRowsPerStep - 1)) * j.tw * j.bands) * sizeof(float);
}

bool conv_pair_supported(int bands, int rad) {
    if (bands < 1 || bands > 4 || rad < 0) return false;
    const int tw = kThreads / bands;
    // Staged span loads take up to four samples per thread; the ring must fit LDS.
    if ((tw + 2 * rad) * bands > 4 * kThreads) return false;
    ConvPairJob j;
    j.bands = bands;
    j.tw = tw;
    j.rad = rad;
    return conv_pair_lds(j) <= 64 * 1024;
}

namespace {
template <typename T, int K>
hipError_t launch_cp(const ConvPairJob& job, dim3 grid, size_t lds, hipStream_t s) {
    switch (job.bands) {
    case 1: hipLaunchKernelGGL((k_conv_pair<T, K, 1>), grid, dim3(kThreads), lds, s, job); break;
    case 2: hipLaunchKernelGGL((k_conv_pair<T, K, 2>), grid, dim3(kThreads), lds, s, job); break;
    case 3: hipLaunchKernelGGL((k_conv_pair<T, K, 3>), grid, dim3(kThreads), lds, s, job); break;
    case 4: hipLaunchKernelGGL((k_conv_pair<T, K, 4>), grid, dim3(kThreads), lds, s, job); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
}  // namespace

hipError_t launch_conv_pair(const ConvPairJob& job, int in_fmt, int kind, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    if (job.tw != kThreads / job.bands) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((job.out.w + job.tw - 1) / job.tw), (unsigned)((job.out.h + job.run - 1) / job.run));
    const size_t lds = conv_pair_lds(job);
    // The Gaussian reads every sample as f32 (ToF32, gauss_blur.rs:248-293),
    // so any non-u8 format stages exactly; ConvSep reads f64, which the f32
    // staging holds exactly for u8, u16 and f32 only.
    if (kind == 1) {
        if (in_fmt == 0) return launch_cp<uint8_t, 1>(job, grid, lds, s);
        if (in_fmt == 1) return launch_cp<uint16_t, 1>(job, grid, lds, s);
        if (in_fmt == 2) return launch_cp<float, 1>(job, grid, lds, s);
        return hipErrorInvalidValue;
    }
    if (in_fmt == 0 || (kind != 0 && kind != 2)) return hipErrorInvalidValue;
#define VGPU_CPK(T) return kind == 0 ? launch_cp<T, 0>(job, grid, lds, s) : launch_cp<T, 2>(job, grid, lds, s)
    VGPU_FOR_FORMAT(in_fmt, VGPU_CPK)
#undef VGPU_CPK
    return hipErrorInvalidValue;
}

}  // namespace vgpu
