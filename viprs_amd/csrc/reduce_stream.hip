// Streaming form of the fused f32 ReduceH -> ReduceV pair, in the
// reference's FP64 operation order (bit-identical to the tiled kernel) or in
// plain FP32.
//
// A block owns a strip of TX output columns (TX * B <= 256 samples, one
// thread per sample) and walks down a run of output rows.  The intermediate
// rows the run's V taps read (rows a .. z of the H output, clamped to the
// intermediate image as ReduceV clamps them) are evaluated in order, G = 4
// source rows per step: the step's rows are staged in LDS (their tap spans,
// clamp-to-edge columns) and every thread evaluates its sample's H output of
// the G rows (G independent chains), rounded to f32 as the reference's
// intermediate holds it.  V is pushed: each H row, as it is produced, is
// added into the register accumulators of every output row whose window
// holds it (at most P open at once, P a power of two the planner checks).
// Rows come in ascending order, so every output receives its taps in tap
// order -- the reference's own sequence of roundings -- with the taps that
// ReduceV's clamp maps to the first or last intermediate row applied, in
// order, when that row arrives.  No H row is kept: the intermediate never
// leaves registers.  The next step's rows come in by LDS-direct loads
// (one sample a lane, gathered from its clamp-to-edge address) into the
// second of two staging buffers while the current step is evaluated; one
// barrier a step publishes them and retires the buffer just read.
//
// Against the tiled kernel (one output tile of 32 x 16 per block) this reads
// each source row once per strip and run instead of once per 16-row tile,
// and evaluates each intermediate row's H once per strip.
//
// Arithmetic (MODE 0, reduceh.rs / reducev.rs f32 paths, as reduce_h_f32_rows
// and reduce_v_f32_rows in kernels.hip):
//   H, B == 1: per group of 4 taps g = 0 + x0 w0 + x1 w1 + x2 w2 + x3 w3
//              (separate multiply and add), acc = acc + g; remaining taps
//              acc = fma(x, w, acc);
//   H, B > 1:  acc = fma(x, w, acc) over the taps;
//   V:         samples before vec_len (the row's length rounded down to a
//              multiple of 4) acc = acc + x w, the rest acc = fma(x, w, acc);
// every chain in FP64 from f64(f32) samples, each pass's result `as f32`.
// MODE 2 (plain FP32): acc = fmaf(w32, x, acc) in both passes.
#include <type_traits>

#include "device.hpp"

namespace vgpu {
namespace {

constexpr int kStreamRows = 4;  // intermediate rows per block step
constexpr int kStreamQ = 4;     // staged samples per row at most kStreamQ * 256 (span * B <= 1024)
constexpr int kPhaseRows = 65;  // rows of a reduce weight table (phases 0 .. 64)
constexpr int kHRegTaps = 24;   // H taps held in registers (more: the 65-phase table in LDS)

template <int MODE>
using StreamAcc = typename std::conditional<MODE == 2, float, double>::type;

// LDS bytes before the two staging buffers (weight tables, V starts),
// rounded to 16.
__host__ __device__ inline size_t stream_stage_offset(const ReduceJob& job, int mode) {
    const size_t wsz = mode == 2 ? sizeof(float) : sizeof(double);
    const size_t htab = job.ph.taps <= kHRegTaps ? 0 : (size_t)kPhaseRows * job.ph.taps * wsz;
    return (htab + ((size_t)job.st_run * job.pv.taps + job.st_pad) * wsz +
            (size_t)job.st_run * 3 * sizeof(int) + 15) & ~(size_t)15;
}

template <int B, int MODE, int P, int HREG>
__global__ __launch_bounds__(kThreads) void k_reduce_stream(ReduceJob job) {
    using W = StreamAcc<MODE>;
    constexpr int G = kStreamRows;
    extern __shared__ __align__(16) unsigned char smem[];
    const InView& in = job.in;
    const OutView& out = job.out;
    const DevPlan& ph = job.ph;
    const DevPlan& pv = job.pv;
    const int TX = job.st_tx, run = job.st_run;
    const int hT = ph.taps, vT = pv.taps;
    const int tx0 = blockIdx.x * TX;
    const int ncols = min(TX, out.w - tx0);
    const int yo0 = blockIdx.y * run;
    const int nrows = min(run, out.h - yo0);
    if (ncols <= 0 || nrows <= 0) return;
    const int srow = (job.st_span * B + 63) & ~63;  // staged row stride (samples)
    // LDS: H weight table (65 phases), the run's V weight rows (then
    // st_pad zeros), the run's V starts and clamped first / last rows, the
    // two staging buffers.
    const int vtab = nrows * vT;
    W* wl_h = reinterpret_cast<W*>(smem);  // HREG: no table (the thread's H weights live in registers)
    W* wl_v = wl_h + (HREG ? 0 : kPhaseRows * hT);
    int* vstart = reinterpret_cast<int*>(wl_v + run * vT + job.st_pad);
    int* vlo = vstart + run;
    int* vhi = vlo + run;
    float* stage = reinterpret_cast<float*>(smem + stream_stage_offset(job, MODE));
    const W* tab_h;
    const W* tab_v;
    if constexpr (MODE == 2) {
        tab_h = ph.wf32;
        tab_v = pv.wf32;
    } else {
        tab_h = ph.wf64;
        tab_v = pv.wf64;
    }
    const int vclip0 = job.vclip0, vclip1 = job.vclip1;
    if constexpr (!HREG)
        for (int i = threadIdx.x; i < kPhaseRows * hT; i += kThreads) wl_h[i] = tab_h[i];
    for (int i = threadIdx.x; i < vtab + job.st_pad; i += kThreads) {
        W w = 0;
        if (i < vtab) {
            const int r = i / vT;
            w = tab_v[(size_t)pv.phase[yo0 + r] * vT + (i - r * vT)];
        }
        wl_v[i] = w;
    }
    for (int i = threadIdx.x; i < nrows; i += kThreads) {
        const int v0 = pv.start[yo0 + i];
        vstart[i] = v0;
        vlo[i] = clampi(v0, vclip0, vclip1);
        vhi[i] = clampi(v0 + vT - 1, vclip0, vclip1);
    }

    const int a = clampi(pv.start[yo0], vclip0, vclip1);
    const int z = clampi(pv.start[yo0 + nrows - 1] + vT - 1, vclip0, vclip1);
    const int total = z - a + 1;  // intermediate rows of the run
    const int c0 = ph.start[tx0];
    const int spn = (ph.start[tx0 + ncols - 1] + hT - c0) * B;  // staged samples of this strip's rows
    const int s = threadIdx.x;
    const bool active = s < ncols * B;
    const int xs = min(s / B, ncols - 1);
    const int hoff = (ph.start[tx0 + xs] - c0) * B + (s - (s / B) * B);
    const int hph = ph.phase[tx0 + xs];
    const W* wh = wl_h + hph * hT;
    W whr[HREG > 0 ? HREG : 1];  // HREG: this thread's H weights (taps < hT)
    if constexpr (HREG > 0) {
#pragma unroll
        for (int t = 0; t < HREG; ++t) whr[t] = t < hT ? tab_h[(size_t)hph * hT + t] : W(0);
    }
    const size_t row_samples = (size_t)in.bw * B;
    const int flat = tx0 * B + s;  // the thread's sample index within an output row
    const int vec_len = out.w * B / 4 * 4;
    // ReduceV's multiply-add / FMA split, decided per wavefront where every
    // lane agrees (all but the wavefront holding vec_len).
    const int wfirst = __builtin_amdgcn_readfirstlane(tx0 * B + (s & ~63));
    const bool all_mul = wfirst + 63 < vec_len, all_fma = wfirst >= vec_len;
    const bool mul_add = flat < vec_len;
    float* const orow = static_cast<float*>(out.ptr) + (size_t)yo0 * out.w * B + flat;
    const size_t ostride = (size_t)out.w * B;

    // Staging by LDS-direct loads: each lane gathers one sample (its
    // clamp-to-edge address) straight into LDS, so nothing in flight holds a
    // register.  Two buffers: group g + 1 is loaded while group g is
    // evaluated; rows are srow64 samples apart (a multiple of 64, so a
    // wavefront's 64 landing slots never reach the next row).
    const int lane = s & 63, wave = s >> 6;
    const bool inside = c0 >= max(in.cx0, in.bx) && c0 + spn / B - 1 <= min(in.cx1, in.bx + in.bw - 1);
    auto dma_group = [&](int g, float* dst) {
        for (int k = wave; k < G; k += kThreads / 64) {
            const int j = min(g * G + k, total - 1);
            const float* src = static_cast<const float*>(in.ptr) + (size_t)row_of(in, a + j) * row_samples;
            if (inside) {
                // The span lies inside the source: one contiguous run.
                const float* run0 = src + (size_t)(c0 - in.bx) * B;
                for (int d0 = 0; d0 < spn; d0 += 64)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void*)(run0 + min(d0 + lane, spn - 1)),
                        (__attribute__((address_space(3))) void*)(dst + k * srow + d0), 4, 0, 0);
            } else {
                for (int d0 = 0; d0 < spn; d0 += 64) {
                    const int e = min(d0 + lane, spn - 1);
                    const int cc = B == 3 ? (e * 21846) >> 16 : e >> (B >> 1);  // e / B for e < 1024
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void*)(src + (size_t)col_of(in, c0 + cc) * B + (e - cc * B)),
                        (__attribute__((address_space(3))) void*)(dst + k * srow + d0), 4, 0, 0);
                }
            }
        }
    };

    // V accumulators: slot j holds output row i with i % P == j while the
    // row's window is open; wp[j] = the wl_v index of its next tap.  Every
    // interior intermediate row is pushed into every slot (one tap each,
    // wp advancing): a slot whose window is closed gathers values nobody
    // reads (it is reset when it opens again; the zero padding after the
    // weight rows keeps its pointer inside LDS), so no slot needs a mask.
    W vacc[P];
    int wp[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        vacc[j] = 0;
        wp[j] = 0;
    }
    int i_first = 0, i_next = 0;  // first output not yet stored / not yet opened
    constexpr int kNever = 0x7fffffff;
    int lo_next = 0, hi_first = 0;  // vlo[i_next] / vhi[i_first] (kNever past the run)
    auto vstep = [&](W acc, float x, W w) -> W {
        if constexpr (MODE == 2) {
            return fmaf(w, x, acc);
        } else {
            const double xd = (double)x;
            if (all_mul) return acc + xd * w;
            if (all_fma) return fma(xd, w, acc);
            return mul_add ? acc + xd * w : fma(xd, w, acc);
        }
    };
    // Push intermediate row r (this thread's sample x) into the outputs.
    auto push_row = [&](int r, float x) {
        while (lo_next <= r) {
            const int jn = i_next & (P - 1);
#pragma unroll
            for (int j = 0; j < P; ++j)
                if (j == jn) {
                    vacc[j] = 0;
                    wp[j] = i_next * vT;
                }
            ++i_next;
            lo_next = i_next < nrows ? __builtin_amdgcn_readfirstlane(vlo[i_next]) : kNever;
        }
        if (r != vclip0 && r != vclip1) {
            W wv[P];
#pragma unroll
            for (int j = 0; j < P; ++j) wv[j] = wl_v[wp[j]];
            if constexpr (MODE == 2) {
#pragma unroll
                for (int j = 0; j < P; ++j) vacc[j] = fmaf(wv[j], x, vacc[j]);
            } else {
                const double xd = (double)x;
                if (all_mul) {
#pragma unroll
                    for (int j = 0; j < P; ++j) vacc[j] = vacc[j] + xd * wv[j];
                } else if (all_fma) {
#pragma unroll
                    for (int j = 0; j < P; ++j) vacc[j] = fma(xd, wv[j], vacc[j]);
                } else {
#pragma unroll
                    for (int j = 0; j < P; ++j) vacc[j] = mul_add ? vacc[j] + xd * wv[j] : fma(xd, wv[j], vacc[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < P; ++j) ++wp[j];
        } else {
            // First / last intermediate row: every tap ReduceV's clamp maps
            // here, in tap order, for each open output.
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int i = i_first + ((j - i_first) & (P - 1));
                if (i < i_next) {
                    const int v0 = __builtin_amdgcn_readfirstlane(vstart[i]);
                    const int t0 = max(r == vclip0 ? 0 : r - v0, 0);
                    const int t1 = min(r == vclip1 ? vT - 1 : r - v0, vT - 1);
                    for (int t = t0; t <= t1; ++t) vacc[j] = vstep(vacc[j], x, wl_v[i * vT + t]);
                    wp[j] = i * vT + t1 + 1;
                }
            }
        }
        while (hi_first <= r) {
            const int jf = i_first & (P - 1);
#pragma unroll
            for (int j = 0; j < P; ++j)
                if (j == jf && active) orow[(size_t)i_first * ostride] = (float)vacc[j];
            ++i_first;
            hi_first = i_first < nrows ? __builtin_amdgcn_readfirstlane(vhi[i_first]) : kNever;
        }
    };

    const int ngroups = (total + G - 1) / G;
    dma_group(0, stage);
    __syncthreads();  // weight tables, starts
    lo_next = __builtin_amdgcn_readfirstlane(vlo[0]);
    hi_first = __builtin_amdgcn_readfirstlane(vhi[0]);
    for (int g = 0; g < ngroups; ++g) {
        const float* cur = stage + (g & 1) * (G * srow);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wavefront's share of group g has landed
        __syncthreads();  // every share landed; group g - 1's buffer is read
        if (g + 1 < ngroups) dma_group(g + 1, stage + ((g + 1) & 1) * (G * srow));
        // H of rows g * G .. g * G + G - 1, tap blocks of 8 with all their
        // LDS reads issued before the dependent steps.
        W acc[G];
#pragma unroll
        for (int k = 0; k < G; ++k) acc[k] = 0;
        const float* px = cur + (active ? hoff : 0);
        if constexpr (HREG > 0) {
            // Groups of four taps (all their LDS reads first), the weights
            // from registers; a last partial group tap by tap.
#pragma unroll
            for (int t4 = 0; t4 < HREG; t4 += 4) {
                if (t4 + 4 <= hT) {
                    float xv[4][G];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int k = 0; k < G; ++k) xv[u][k] = px[k * srow + (t4 + u) * B];
                    if constexpr (MODE == 2) {
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int k = 0; k < G; ++k) acc[k] = fmaf(whr[t4 + u], xv[u][k], acc[k]);
                    } else if constexpr (B == 1) {
#pragma unroll
                        for (int k = 0; k < G; ++k) {
                            double gsum = 0.0;
#pragma unroll
                            for (int u = 0; u < 4; ++u) gsum = gsum + (double)xv[u][k] * whr[t4 + u];
                            acc[k] = acc[k] + gsum;
                        }
                    } else {
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int k = 0; k < G; ++k) acc[k] = fma((double)xv[u][k], whr[t4 + u], acc[k]);
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (t4 + u < hT) {
#pragma unroll
                            for (int k = 0; k < G; ++k) {
                                const float xx = px[k * srow + (t4 + u) * B];
                                if constexpr (MODE == 2) acc[k] = fmaf(whr[t4 + u], xx, acc[k]);
                                else acc[k] = fma((double)xx, whr[t4 + u], acc[k]);
                            }
                        }
                }
            }
        }
        int t = HREG > 0 ? hT : 0;
        for (; t + 8 <= hT; t += 8) {
            float xv[8][G];
            W wv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                wv[u] = wh[t + u];
#pragma unroll
                for (int k = 0; k < G; ++k) xv[u][k] = px[k * srow + (t + u) * B];
            }
            if constexpr (MODE == 2) {
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int k = 0; k < G; ++k) acc[k] = fmaf(wv[u], xv[u][k], acc[k]);
            } else if constexpr (B == 1) {
#pragma unroll
                for (int h4 = 0; h4 < 8; h4 += 4)
#pragma unroll
                    for (int k = 0; k < G; ++k) {
                        double gsum = 0.0;
#pragma unroll
                        for (int u = h4; u < h4 + 4; ++u) gsum = gsum + (double)xv[u][k] * wv[u];
                        acc[k] = acc[k] + gsum;
                    }
            } else {
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int k = 0; k < G; ++k) acc[k] = fma((double)xv[u][k], wv[u], acc[k]);
            }
        }
        if constexpr (MODE == 0 && B == 1 && HREG == 0) {
            // A last group of four (AVX2 grouping) before the FMA tail.
            if (t + 4 <= hT) {
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    double gsum = 0.0;
#pragma unroll
                    for (int u = 0; u < 4; ++u) gsum = gsum + (double)px[k * srow + (t + u) * B] * wh[t + u];
                    acc[k] = acc[k] + gsum;
                }
                t += 4;
            }
        }
        for (; t < hT; ++t) {
            const W wt = wh[t];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                if constexpr (MODE == 2) acc[k] = fmaf(wt, px[k * srow + t * B], acc[k]);
                else acc[k] = fma((double)px[k * srow + t * B], wt, acc[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (g * G + k < total) push_row(a + g * G + k, (float)acc[k]);
    }
}

}  // namespace

size_t reduce_stream_lds(const ReduceJob& job, int mode) {
    return stream_stage_offset(job, mode) + (size_t)2 * kStreamRows * ((job.st_span * job.bands + 63) & ~63) * sizeof(float);
}

hipError_t launch_reduce_stream(const ReduceJob& job, int mode, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    if (!job.stream || job.bands < 1 || job.bands > 4 || job.st_tx * job.bands > kThreads ||
        job.st_span * job.bands > kStreamQ * kThreads || job.st_open < 1 || job.st_open > 16)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)job.st_gx, (unsigned)job.st_gy);
    const size_t lds = reduce_stream_lds(job, mode);
    const bool p8 = job.st_open <= 8;
    const bool hreg = job.ph.taps <= kHRegTaps;
#define VGPU_RSQ(BB, MM)                                                                                       \
    if (p8 && hreg) hipLaunchKernelGGL((k_reduce_stream<BB, MM, 8, kHRegTaps>), grid, dim3(kThreads), lds, s, job); \
    else if (p8) hipLaunchKernelGGL((k_reduce_stream<BB, MM, 8, 0>), grid, dim3(kThreads), lds, s, job);            \
    else if (hreg) hipLaunchKernelGGL((k_reduce_stream<BB, MM, 16, kHRegTaps>), grid, dim3(kThreads), lds, s, job); \
    else hipLaunchKernelGGL((k_reduce_stream<BB, MM, 16, 0>), grid, dim3(kThreads), lds, s, job);
#define VGPU_RS(BB)          \
    if (mode == 2) {         \
        VGPU_RSQ(BB, 2)      \
    } else {                 \
        VGPU_RSQ(BB, 0)      \
    }
    switch (job.bands) {
    case 1: VGPU_RS(1) break;
    case 2: VGPU_RS(2) break;
    case 3: VGPU_RS(3) break;
    default: VGPU_RS(4) break;
    }
#undef VGPU_RS
#undef VGPU_RSQ
    return hipGetLastError();
}

}  // namespace vgpu
