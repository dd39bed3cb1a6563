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
    // H over staged chunks of job.hr source rows (the tile's source span in
    // LDS, clamped reads), then V from the H rows.
    float* stage = hbuf + reduce_hbuf_bytes(job, sizeof(float)) / sizeof(float);
    const int c0 = ph.start[tx0];
    const int spx = ph.start[tx0 + ncols - 1] + ph.taps - c0;
    const int srow = spx * B;
    const int xo = tx0 + min(x, ncols - 1);
    const int hoff = (ph.start[xo] - c0) * B;
    const float* w = ph.wf32 + (size_t)ph.phase[xo] * ph.taps;
    const int hstep = max(job.hr, 1);
    for (int rc = 0; rc < nr; rc += hstep) {
        const int hr = min(hstep, nr - rc);
        stage_h_rows<float, B>(in, ys0 + rc, hr, c0, spx, stage);
        __syncthreads();
        for (int r = rl; x < ncols && r < hr; r += nrl) {
            const float* row = stage + r * srow + hoff;
            float acc[B];
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = 0.f;
            for (int t = 0; t < ph.taps; ++t) {
                const float wt = w[t];
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] = fmaf(wt, row[t * B + b], acc[b]);
            }
#pragma unroll
            for (int b = 0; b < B; ++b) hbuf[(rc + r) * pitch + x * B + b] = acc[b];
        }
        __syncthreads();
    }
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

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));

constexpr int kF16MaxH = 4;  // H K steps (128 samples)
constexpr int kF16MaxV = 2;  // V K steps (64 rows)

__device__ __forceinline__ f32x4 mfma_f16(u32x4v a, u32x4v b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                  0, 0, 0);
}

// Split 8 floats into FP16 hi and lo operands: hi = f16(x), lo = f16(x - hi),
// both rounded to nearest; x - hi is exact in FP32.  The caller scales x by a
// power of two first so that |x| < 2^14 (no overflow) and the lo limbs of the
// values that matter stay normal.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split8(const float (&x)[8], u32x4v& hi, u32x4v& lo) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const f32x2 v = {x[2 * j], x[2 * j + 1]};
        const f16x2 h = __builtin_convertvector(v, f16x2);
        const f32x2 r = v - __builtin_convertvector(h, f32x2);
        const f16x2 l = __builtin_convertvector(r, f16x2);
        hi[j] = __builtin_bit_cast(uint32_t, h);
        lo[j] = __builtin_bit_cast(uint32_t, l);
    }
}

// Exponent s with max|x| * 2^s < 2^14 over the whole wavefront (wave-uniform),
// from m = this lane's max |x|.  Non-negative floats order like their bit
// patterns: max within each 16-lane row by DPP (xor 1, xor 2, half-row mirror,
// row mirror), then across the four rows with readlane.
__device__ __forceinline__ int wave_scale_exp(float m) {
    int v = __float_as_int(m);
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false));  // row_mirror
    const int r = max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
                      max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
    // frexp exponent of the max: biased exponent - 126 (zero and subnormal
    // maxima give a large scale, harmless for such values).
    return 14 - (((r >> 23) & 0xff) - 126);
}

// 8 consecutive floats of the input buffer from element `e` (bytes past the
// end read as 0).
__device__ __forceinline__ void load8f(const float* base, uint32_t e, uint32_t total, float (&x)[8]) {
    if (e + 8 <= total) {
        // Loaded as float vectors: __builtin_bit_cast of a vector element
        // (a[i]) reads element 0 whatever i is with this compiler.
        const f32x4 a = gmem<f32x4>(base + e)[0];
        const f32x4 b = gmem<f32x4>(base + e + 4)[0];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            x[i] = a[i];
            x[4 + i] = b[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = e + i < total ? gmem<float>(base)[e + i] : 0.f;
    }
}

// One wavefront: a 16-sample strip `jx` over rows_per_task output rows.
// H: A = input pixels [16 rows][K = span samples], B = folded weights
// [K][16 samples]; D rows go to a per-wave LDS ring of f32 intermediate rows.
// V: A = folded V weights [16 output rows][K = ring rows from base8], B = the
// ring.  Each product x * w is x_hi w_hi + x_hi w_lo + x_lo w_hi; the hi * hi
// MFMA of each K step starts from zero and is added to the sum on the VALU
// (round to nearest), and the two correction products share their own
// accumulator, so no large partial sum goes through the MFMA accumulator.
// MAXH: H K steps the registers hold (KH <= MAXH); PF: the next M tile's
// samples are loaded while the current one is worked (up to three K steps:
// with four the extra registers cost more occupancy than the overlap gains).
template <int B, int MAXH, bool PF>
__device__ void f16_task(const F16Job& job, int jx, int ty, unsigned char* lds, size_t wave_lds) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int KH = job.kh, KV = job.kv, RING = job.ring;
    const int RS = RING + 8;  // ring row stride in floats (padded against bank conflicts)
    float* ring = reinterpret_cast<float*>(lds + (size_t)wave * wave_lds);
    const InView& in = job.in;
    const OutView& out = job.out;
    const int nsamp = out.w * B;
    const int s0 = jx * 16;
    const int n = lane & 15, q = lane >> 4;
    const float* base = static_cast<const float*>(in.ptr);
    const uint32_t row_elems = (uint32_t)in.bw * B;
    const uint32_t total = row_elems * (uint32_t)in.bh;

    // H operands of the strip (B operand: rows j = samples, K = span samples).
    const int pat = __builtin_amdgcn_readfirstlane(job.hstrip[2 * jx]);
    const int span0 = __builtin_amdgcn_readfirstlane(job.hstrip[2 * jx + 1]);
    const unsigned char* ht = job.htab + (size_t)pat * job.htab_bytes;
    u32x4v whi[MAXH], wlo[MAXH];
#pragma unroll
    for (int s = 0; s < MAXH; ++s) {
        if (s < KH) {
            const u32x4 a = gmem<u32x4>(ht + ((size_t)(2 * s) * 64 + lane) * 16)[0];
            const u32x4 b = gmem<u32x4>(ht + ((size_t)(2 * s + 1) * 64 + lane) * 16)[0];
            whi[s] = u32x4v{a[0], a[1], a[2], a[3]};
            wlo[s] = u32x4v{b[0], b[1], b[2], b[3]};
        }
    }
    const uint32_t lane_off = (uint32_t)span0 * B + 8 * q;

    const int y_begin = ty * job.rows_per_task;
    const int y_end = min(out.h, y_begin + job.rows_per_task);
    int have = INT32_MIN;
    // The next M tile's samples are loaded while the current one is split
    // and multiplied (xn, for row r_next).
    float xn[MAXH][8];
    int r_next = INT32_MIN;
    auto load_m = [&](int r) {
        const uint32_t rb = (uint32_t)row_of(in, r + n) * row_elems + lane_off;
#pragma unroll
        for (int s = 0; s < MAXH; ++s)
            if (s < KH) load8f(base, rb + 32 * s, total, xn[s]);
        r_next = r;
    };
    for (int y0 = y_begin, ti = y_begin >> 4; y0 < y_end; y0 += 16, ++ti) {
        const int vpat = job.vtile[3 * ti];
        const int base8 = job.vtile[3 * ti + 1], hi = job.vtile[3 * ti + 2];
        // V operands of this output tile, loaded ahead of its H work.
        const unsigned char* vt = job.vtab + (size_t)vpat * job.vtab_bytes;
        u32x4 vop[kF16MaxV][2];
#pragma unroll
        for (int s = 0; s < kF16MaxV; ++s) {
            if (s < KV) {
                vop[s][0] = gmem<u32x4>(vt + ((size_t)(2 * s) * 64 + lane) * 16)[0];
                vop[s][1] = gmem<u32x4>(vt + ((size_t)(2 * s + 1) * 64 + lane) * 16)[0];
            }
        }
        // ---- intermediate rows from max(have, base8 rounded to 16) through hi.
        int r0 = max(have, base8 & ~15);
        if (PF && r0 <= hi && r_next != r0) load_m(r0);
        for (; r0 <= hi; r0 += 16) {
            float x[MAXH][8];
            float mx = 0.f;
            if constexpr (PF) {
#pragma unroll
                for (int s = 0; s < MAXH; ++s) {
                    if (s < KH) {
#pragma unroll
                        for (int e = 0; e < 8; ++e) x[s][e] = xn[s][e];
                    }
                }
                // Prefetch the following M tile (this output tile's next
                // one, or the first one the next output tile computes).
                if (r0 + 16 <= hi || y0 + 16 < y_end) load_m(r0 + 16);
            } else {
                const uint32_t rb = (uint32_t)row_of(in, r0 + n) * row_elems + lane_off;
#pragma unroll
                for (int s = 0; s < MAXH; ++s)
                    if (s < KH) load8f(base, rb + 32 * s, total, x[s]);
            }
#pragma unroll
            for (int s = 0; s < MAXH; ++s) {
                if (s < KH) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) mx = fmaxf(mx, fabsf(x[s][e]));
                }
            }
            const int sx = wave_scale_exp(mx);
            f32x4 d = {0.f, 0.f, 0.f, 0.f}, dc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < MAXH; ++s) {
                if (s < KH) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) x[s][e] = __builtin_amdgcn_ldexpf(x[s][e], sx);
                    u32x4v ah, al;
                    split8(x[s], ah, al);
                    d += mfma_f16(ah, whi[s], f32x4{0.f, 0.f, 0.f, 0.f});
                    dc = mfma_f16(ah, wlo[s], dc);
                    dc = mfma_f16(al, whi[s], dc);
                }
            }
            d += dc;
            const int ux = -(sx + job.h_exp);
#pragma unroll
            for (int e = 0; e < 4; ++e) d[e] = __builtin_amdgcn_ldexpf(d[e], ux);
            // D[m = row][n = sample]: rows r0 + 4q .. + 3 of sample n.
            *reinterpret_cast<f32x4*>(ring + n * RS + ((r0 + 4 * q) & (RING - 1))) = d;
        }
        have = max(have, r0);
        // Ring rows of the V window from `have` on are not computed yet and
        // carry zero weight: clear them so that stale LDS (possibly not a
        // finite value) reaches neither the sum nor the scale.
        for (int r = have; r < base8 + 32 * KV; r += 16)
            *reinterpret_cast<f32x4*>(ring + n * RS + ((r + 4 * q) & (RING - 1))) = f32x4{0.f, 0.f, 0.f, 0.f};
        // ---- V: A = folded V weights (rows j = output rows, K = rows from base8).
        float yv[kF16MaxV][8];
        float my = 0.f;
#pragma unroll
        for (int s = 0; s < kF16MaxV; ++s) {
            if (s < KV) {
                const float* rp = ring + n * RS + ((base8 + 32 * s + 8 * q) & (RING - 1));
                const f32x4 r0v = *reinterpret_cast<const f32x4*>(rp);
                const f32x4 r1v = *reinterpret_cast<const f32x4*>(rp + 4);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    yv[s][e] = r0v[e];
                    yv[s][4 + e] = r1v[e];
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) my = fmaxf(my, fabsf(yv[s][e]));
            }
        }
        const int sy = wave_scale_exp(my);
        f32x4 d = {0.f, 0.f, 0.f, 0.f}, dc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < kF16MaxV; ++s) {
            if (s < KV) {
                const u32x4 a = vop[s][0], b = vop[s][1];
#pragma unroll
                for (int e = 0; e < 8; ++e) yv[s][e] = __builtin_amdgcn_ldexpf(yv[s][e], sy);
                u32x4v bh, bl;
                split8(yv[s], bh, bl);
                const u32x4v vh = u32x4v{a[0], a[1], a[2], a[3]}, vl = u32x4v{b[0], b[1], b[2], b[3]};
                d += mfma_f16(vh, bh, f32x4{0.f, 0.f, 0.f, 0.f});
                dc = mfma_f16(vl, bh, dc);
                dc = mfma_f16(vh, bl, dc);
            }
        }
        d += dc;
        const int uy = -(sy + job.v_exp);
#pragma unroll
        for (int e = 0; e < 4; ++e) d[e] = __builtin_amdgcn_ldexpf(d[e], uy);
        // D[m = output row][n = sample]: rows y0 + 4q + e of sample s0 + n.
        const int smp = s0 + n;
        if (smp < nsamp) {
            float* o = static_cast<float*>(out.ptr) + (size_t)(y0 + 4 * q) * nsamp + smp;
            const int rows_left = y_end - (y0 + 4 * q);
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (e < rows_left) o[(size_t)e * nsamp] = d[e];
        }
    }
}

template <int B, int MAXH, bool PF>
__global__ __launch_bounds__(kThreads) void k_f16_job(F16Job job, int wave_lds) {
    extern __shared__ __align__(16) unsigned char lds[];
    // Neighbouring strips overlap by their tap halo: keep them on one XCD.
    const int t = xcd_block(blockIdx.x, gridDim.x) * (kThreads / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (t >= job.gx * job.gy) return;
    f16_task<B, MAXH, PF>(job, t % job.gx, t / job.gx, lds, (size_t)wave_lds);
}

}  // namespace

size_t f16_job_wave_lds(const F16Job& job) { return 16 * ((size_t)job.ring + 8) * sizeof(float); }

hipError_t launch_f16_job(const F16Job& job, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    if (job.kh < 1 || job.kh > kF16MaxH || job.kv < 1 || job.kv > kF16MaxV || job.ring < 32 * job.kv + 16)
        return hipErrorInvalidValue;
    const int wl = (int)f16_job_wave_lds(job);
    const int ntasks = job.gx * job.gy;
    const dim3 grid((unsigned)((ntasks + 3) / 4));
    const size_t lds = (size_t)wl * 4;
    const bool pf = job.kh <= 3;  // H spans up to 3 K steps: prefetch the next M tile
#define VGPU_F16(BB)                                                                                   \
    if (pf) hipLaunchKernelGGL((k_f16_job<BB, 3, true>), grid, dim3(kThreads), lds, s, job, wl);       \
    else hipLaunchKernelGGL((k_f16_job<BB, kF16MaxH, false>), grid, dim3(kThreads), lds, s, job, wl);
    switch (job.bands) {
    case 1: VGPU_F16(1) break;
    case 2: VGPU_F16(2) break;
    case 3: VGPU_F16(3) break;
    case 4: VGPU_F16(4) break;
    default: return hipErrorInvalidValue;
    }
#undef VGPU_F16
    return hipGetLastError();
}

hipError_t launch_red_f32_job(const ReduceJob& job, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    if (job.stream) return launch_reduce_stream(job, 2, s);
    const dim3 grid((unsigned)job.gx, (unsigned)job.gy);
    const size_t lds = reduce_hbuf_bytes(job, sizeof(float)) + (size_t)job.hr * job.span_max * job.bands * sizeof(float);
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
