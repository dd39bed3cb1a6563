// Separable u8 passes on MFMA with int8 limbs (gfx950 v_mfma_i32_16x16x64_i8).
//
// A u8 horizontal pass out[x][b] = sum_t w[phase(x)][t] * in[col(start(x)+t)][b]
// is a banded matrix product: over the contiguous band-interleaved bytes of
// the source span of 16 output samples (K), with a weight matrix W[k][n] that
// is zero except where byte k is the (band-matching) pixel a tap of sample n
// reads.  Clamped reads at image edges fold into W by adding the weights of
// taps that land on the same pixel.  The vertical pass is the same product
// over intermediate rows.  The reference's integer arithmetic is reproduced
// exactly:
//   - pixels p in 0..255 enter as p' = p ^ 0x80 = p - 128 (signed int8);
//   - each i16 weight w splits as w = 128 a + c with c in [-64, 63], so a and
//     c are int8 and sum w p = 128 sum a p' + sum c p' + 128 sum w;
//   - the int32 sums are exact, then the reference rounding is applied:
//     (sum + 2048) >> 12 clamped to 0..255 for the reducers
//     (reduce_common.rs fixed-point path), (sum + scale/2) / scale for the
//     Gaussian (gauss_blur.rs:405-518).
//
// Wavefront task: 16 output samples (one MFMA N tile) x rows_per_task output
// rows.  The H weights of the strip are folded once per task; the task then
// walks its output rows 16 at a time: it computes the intermediate rows the V
// window needs (16-row MFMA M tiles, kept as u8 in a per-wavefront LDS ring,
// column-major so a V operand is one 16-byte LDS read), folds the V weights
// of the 16 output rows, and applies the V product.  Wavefronts never wait on
// each other.
#include "device.hpp"

#include <algorithm>

namespace vgpu {
namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int kSepWaves = kThreads / 64;

__device__ __forceinline__ i32x4 mfma_i8(i32x4 a, i32x4 b, i32x4 c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Split 16 folded i32 weights into the two int8 limb operands.
__device__ __forceinline__ void split_limbs(const int (&v)[16], i32x4& hi, i32x4& lo) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t h = 0, l = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int w = v[4 * j + e];
            const int c = ((w + 64) & 127) - 64;
            const int a = (w - c) >> 7;
            l |= (uint32_t)(c & 255) << (8 * e);
            h |= (uint32_t)(a & 255) << (8 * e);
        }
        hi[j] = (int)h;
        lo[j] = (int)l;
    }
}

// 16 bytes of the input buffer starting at byte `a` (any alignment); bytes
// past the end of the buffer read as 0.
__device__ __forceinline__ void load16(const unsigned char* base, size_t a, size_t total, uint32_t (&o)[4]) {
    const size_t a4 = a & ~(size_t)3;
    const int al = (int)(a & 3);
    if (a4 + 20 <= total) {
        const u32x4 v = gmem<u32x4>(base + a4)[0];
        const uint32_t x = gmem<uint32_t>(base + a4 + 16)[0];
        o[0] = __builtin_amdgcn_alignbyte(v[1], v[0], al);
        o[1] = __builtin_amdgcn_alignbyte(v[2], v[1], al);
        o[2] = __builtin_amdgcn_alignbyte(v[3], v[2], al);
        o[3] = __builtin_amdgcn_alignbyte(x, v[3], al);
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t w = 0;
            for (int e = 0; e < 4; ++e) {
                const size_t off = a + 4 * j + e;
                if (off < total) w |= (uint32_t)gmem<unsigned char>(base)[off] << (8 * e);
            }
            o[j] = w;
        }
    }
}

template <int B>
__device__ __forceinline__ uint32_t premult_dword(uint32_t v) {
    if constexpr (B == 4) {
        const uint32_t a = v >> 24;
        return premult_u8(v & 255u, a) | (premult_u8((v >> 8) & 255u, a) << 8) |
               (premult_u8((v >> 16) & 255u, a) << 16) | (a << 24);
    } else {
        const uint32_t a0 = (v >> 8) & 255u, a1 = v >> 24;
        return premult_u8(v & 255u, a0) | (a0 << 8) | (premult_u8((v >> 16) & 255u, a1) << 16) | (a1 << 24);
    }
}

__device__ __forceinline__ int sep_round(int acc, int mode, int scale) {
    int v = mode == 0 ? (acc + 2048) >> 12 : (acc + (scale >> 1)) / scale;
    return v < 0 ? 0 : (v > 255 ? 255 : v);
}

// Fold the taps of 16 rows j (lane j + 16 tq adds taps tq, tq + 4, ...) into
// the wavefront's scratch [16][ks] and corr[16]; `pos` maps (j, t) to the K
// index (or -1 for no contribution).
template <typename Pos>
__device__ __forceinline__ void fold_weights(int* scratch, int* corr, int ks, int taps, const int16_t* wrow,
                                             bool active, Pos pos) {
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, tq = lane >> 4;
    // Zero scratch (16 * ks ints, ks a multiple of 64) and corr.
    for (int i = lane * 4; i < 16 * ks; i += 256) *reinterpret_cast<i32x4*>(scratch + i) = i32x4{0, 0, 0, 0};
    if (lane < 16) corr[lane] = 0;
    wave_lds_sync();
    if (active) {
        for (int t = tq; t < taps; t += 4) {
            const int k = pos(t);
            const int w = wrow[t];
            if (k >= 0 && k < ks) atomicAdd(scratch + j * ks + k, w);
            atomicAdd(corr + j, w);
        }
    }
    wave_lds_sync();
}

// Read lane (j, q)'s 16 folded weights of k-step s and split into limbs.
__device__ __forceinline__ void read_limbs(const int* scratch, int ks, int s, i32x4& hi, i32x4& lo) {
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, q = lane >> 4;
    const int* p = scratch + j * ks + 64 * s + 16 * q;
    int v[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const i32x4 x = *reinterpret_cast<const i32x4*>(p + 4 * i);
        v[4 * i] = x[0];
        v[4 * i + 1] = x[1];
        v[4 * i + 2] = x[2];
        v[4 * i + 3] = x[3];
    }
    split_limbs(v, hi, lo);
}

constexpr int kMaxSteps = 2;

template <int B>
__device__ void sep_task(const SepJob& job, int jx, int ty, unsigned char* lds, size_t wave_lds) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ks = 64 * max(job.kh, job.kv);
    int* scratch = reinterpret_cast<int*>(lds + (size_t)wave * wave_lds);
    int* corr = scratch + 16 * ks;
    unsigned char* ring = reinterpret_cast<unsigned char*>(corr + 16);
    const InView& in = job.in;
    const OutView& out = job.out;
    const DevPlan& ph = job.ph;
    const DevPlan& pv = job.pv;
    const int KH = job.kh, KV = job.kv, RING = job.ring;
    const int nsamp = out.w * B;
    const int s0 = jx * 16;
    const int n = lane & 15, q = lane >> 4;
    const unsigned char* base = static_cast<const unsigned char*>(in.ptr);
    const size_t row_bytes = (size_t)in.bw * B;
    const size_t total = row_bytes * (size_t)in.bh;

    // ---- H weights of the strip: B operand, rows j = samples, K = span bytes.
    const int x_first = s0 / B;
    const int span0 = col_of(in, ph.start[x_first]);  // buffer-relative first span pixel
    {
        const int smp = s0 + n;
        const bool active = smp < nsamp;
        const int x = active ? smp / B : 0;
        const int b = smp - x * B;
        const int st = ph.start[x];
        fold_weights(scratch, corr, ks, ph.taps, ph.wi16 + (size_t)ph.phase[x] * ph.taps, active,
                     [&](int t) { return (col_of(in, st + t) - span0) * B + b; });
    }
    i32x4 whi[kMaxSteps], wlo[kMaxSteps];
#pragma unroll
    for (int s = 0; s < kMaxSteps; ++s)
        if (s < KH) read_limbs(scratch, ks, s, whi[s], wlo[s]);
    const int corr_h = 128 * corr[n];
    wave_lds_sync();

    const int y_begin = ty * job.rows_per_task;
    const int y_end = min(out.h, y_begin + job.rows_per_task);
    const int mode = job.mode, scale = job.scale;
    int have = INT32_MIN;  // intermediate rows below `have` (a multiple of 16) are in the ring
    for (int y0 = y_begin; y0 < y_end; y0 += 16) {
        const int nrow = min(16, y_end - y0);
        const int lo = clampi(pv.start[y0], job.vclip0, job.vclip1);
        const int hi = clampi(pv.start[y0 + nrow - 1] + pv.taps - 1, job.vclip0, job.vclip1);
        const int base16 = lo & ~15;
        // ---- intermediate rows [max(have, base16), hi], 16 at a time.
        for (int r0 = have > base16 ? have : base16; r0 <= hi; r0 += 16) {
            const size_t rb = (size_t)row_of(in, r0 + n) * row_bytes + (size_t)span0 * B;
            i32x4 dh = {0, 0, 0, 0}, dl = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < kMaxSteps; ++s) {
                if (s < KH) {
                    uint32_t a[4];
                    load16(base, rb + 64 * s + 16 * q, total, a);
                    i32x4 av;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        uint32_t v = a[j];
                        if constexpr (B == 2 || B == 4) {
                            if (job.premult) v = premult_dword<B>(v);
                        }
                        av[j] = (int)(v ^ 0x80808080u);
                    }
                    dh = mfma_i8(av, whi[s], dh);
                    dl = mfma_i8(av, wlo[s], dl);
                }
            }
            // D[m = row][n = sample]: this lane holds rows r0 + 4q .. + 3 of sample n.
            uint32_t packed = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                packed |= (uint32_t)sep_round(128 * dh[e] + dl[e] + corr_h, mode, scale) << (8 * e);
            *reinterpret_cast<uint32_t*>(ring + n * RING + ((r0 + 4 * q) & (RING - 1))) = packed;
            have = r0 + 16;
        }
        // ---- V weights of the 16 output rows: A operand, rows j = output rows,
        // K = intermediate rows from base16.
        {
            const int m = n;
            const bool active = m < nrow;
            const int y = active ? y0 + m : y0;
            const int vs = pv.start[y];
            const int v0 = job.vclip0, v1 = job.vclip1;
            fold_weights(scratch, corr, ks, pv.taps, pv.wi16 + (size_t)pv.phase[y] * pv.taps, active,
                         [&](int t) { return clampi(vs + t, v0, v1) - base16; });
        }
        i32x4 vhi[kMaxSteps], vlo[kMaxSteps];
#pragma unroll
        for (int s = 0; s < kMaxSteps; ++s)
            if (s < KV) read_limbs(scratch, ks, s, vhi[s], vlo[s]);
        const i32x4 cv = *reinterpret_cast<const i32x4*>(corr + 4 * q);
        wave_lds_sync();
        i32x4 dh = {0, 0, 0, 0}, dl = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < kMaxSteps; ++s) {
            if (s < KV) {
                u32x4 bv = *reinterpret_cast<const u32x4*>(ring + n * RING + ((base16 + 64 * s + 16 * q) & (RING - 1)));
                i32x4 bi;
#pragma unroll
                for (int j = 0; j < 4; ++j) bi[j] = (int)(bv[j] ^ 0x80808080u);
                dh = mfma_i8(vhi[s], bi, dh);
                dl = mfma_i8(vlo[s], bi, dl);
            }
        }
        // D[m = output row][n = sample]: rows y0 + 4q + e of sample s0 + n.
        int v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = sep_round(128 * dh[e] + dl[e] + 128 * cv[e], mode, scale);
        if constexpr (B == 2 || B == 4) {
            if (job.unpremult) {
                // The pixel's alpha sample is in lane (lane | (B - 1)) of this group.
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int a = __shfl(v[e], lane | (B - 1), 64);
                    if ((n & (B - 1)) != B - 1) v[e] = (int)unpremult_u8((uint32_t)v[e], (uint32_t)a);
                }
            }
        }
        const int smp = s0 + n;
        if (smp < nsamp) {
            unsigned char* o = static_cast<unsigned char*>(out.ptr) + smp;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int yy = y0 + 4 * q + e;
                if (yy < y_end) o[(size_t)yy * nsamp] = (unsigned char)v[e];
            }
        }
    }
}

template <int B>
__global__ __launch_bounds__(kThreads) void k_sep_job(SepJob job, int wave_lds) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int t = blockIdx.x * kSepWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (t >= job.gx * job.gy) return;
    sep_task<B>(job, t % job.gx, t / job.gx, lds, (size_t)wave_lds);
}

template <int B>
__global__ __launch_bounds__(kThreads) void k_sep_batch(const SepJob* __restrict__ jobs,
                                                        const BatchTask* __restrict__ tasks, int ntasks,
                                                        const void* const* __restrict__ srcs,
                                                        void* const* __restrict__ dsts, int wave_lds) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int t = blockIdx.x * kSepWaves + wave; t < ntasks; t += gridDim.x * kSepWaves) {
        const BatchTask task = tasks[t];
        SepJob job = jobs[task.job];
        if (job.src_slot >= 0) job.in.ptr = srcs[job.src_slot];
        if (job.dst_slot >= 0) job.out.ptr = dsts[job.dst_slot];
        sep_task<B>(job, task.bx, task.by, lds, (size_t)wave_lds);
    }
}

}  // namespace

size_t sep_job_wave_lds(const SepJob& job) {
    const size_t ks = 64 * (size_t)std::max(job.kh, job.kv);
    return (16 * ks + 16) * sizeof(int) + 16 * (size_t)job.ring;
}

hipError_t launch_sep_job(const SepJob& job, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    const int wl = (int)sep_job_wave_lds(job);
    const int ntasks = job.gx * job.gy;
    const dim3 grid((unsigned)((ntasks + kSepWaves - 1) / kSepWaves));
    const size_t lds = (size_t)wl * kSepWaves;
    switch (job.bands) {
    case 1: hipLaunchKernelGGL(k_sep_job<1>, grid, dim3(kThreads), lds, s, job, wl); break;
    case 2: hipLaunchKernelGGL(k_sep_job<2>, grid, dim3(kThreads), lds, s, job, wl); break;
    case 3: hipLaunchKernelGGL(k_sep_job<3>, grid, dim3(kThreads), lds, s, job, wl); break;
    case 4: hipLaunchKernelGGL(k_sep_job<4>, grid, dim3(kThreads), lds, s, job, wl); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sep_batch(int bands, const SepJob* jobs, const BatchTask* tasks, int ntasks,
                            const void* const* srcs, void* const* dsts, int grid, size_t wave_lds,
                            hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    const size_t lds = wave_lds * kSepWaves;
    const int wl = (int)wave_lds;
    switch (bands) {
    case 1: hipLaunchKernelGGL(k_sep_batch<1>, dim3((unsigned)grid), dim3(kThreads), lds, s, jobs, tasks, ntasks, srcs, dsts, wl); break;
    case 2: hipLaunchKernelGGL(k_sep_batch<2>, dim3((unsigned)grid), dim3(kThreads), lds, s, jobs, tasks, ntasks, srcs, dsts, wl); break;
    case 3: hipLaunchKernelGGL(k_sep_batch<3>, dim3((unsigned)grid), dim3(kThreads), lds, s, jobs, tasks, ntasks, srcs, dsts, wl); break;
    case 4: hipLaunchKernelGGL(k_sep_batch<4>, dim3((unsigned)grid), dim3(kThreads), lds, s, jobs, tasks, ntasks, srcs, dsts, wl); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace vgpu
