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

// Read lane (j, q)'s 16 folded weights of k-step s as one int8 operand
// (every weight fits int8).
__device__ __forceinline__ i32x4 read_single(const int* scratch, int ks, int s) {
    const int lane = threadIdx.x & 63;
    const int j = lane & 15, q = lane >> 4;
    const int* p = scratch + j * ks + 64 * s + 16 * q;
    i32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const i32x4 x = *reinterpret_cast<const i32x4*>(p + 4 * i);
        r[i] = (int)((uint32_t)(x[0] & 255) | ((uint32_t)(x[1] & 255) << 8) | ((uint32_t)(x[2] & 255) << 16) |
                     ((uint32_t)x[3] << 24));
    }
    return r;
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

// Plan tables are read-only for the kernel's life: constant-address loads
// go through the scalar cache.
template <typename T>
__device__ __forceinline__ T cload(const T* p, int i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Register budget of the H operands: up to 6 K steps with one limb, 3 with two.
template <int LIMBS>
constexpr int max_steps() {
    return LIMBS == 1 ? 6 : 3;
}

// One A-operand step of an M tile: lane (m, q) holds bytes 16q .. 16q + 15
// of K step s of the span of intermediate row r0 + m (pixels as p ^ 0x80).
template <int B, bool PM>
__device__ __forceinline__ i32x4 load_a(const unsigned char* base, uint32_t off, uint32_t total) {
    uint32_t w[4];
    if constexpr (B == 4) {
        // 4-band spans start on pixel (4-byte) boundaries.
        if (off + 16 <= total) {
            const u32x4 v = gmem<u32x4>(base + off)[0];
            w[0] = v[0];
            w[1] = v[1];
            w[2] = v[2];
            w[3] = v[3];
        } else {
            load16(base, off, total, w);
        }
    } else {
        load16(base, off, total, w);
    }
    i32x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t v = w[j];
        if constexpr ((B == 2 || B == 4) && PM) v = premult_dword<B>(v);
        a[j] = (int)(v ^ 0x80808080u);
    }
    return a;
}

// Reference rounding of an exact, biased sum.  Family 0: (sum + 2048) >> 12
// clamped to 0..255.  Family 1: a quotient by multiply-shift, exact for the
// operand ranges the host admits (Gaussian: mul = ceil(2^32 / scale), scale <
// 4096; ShrinkH f = 5: mul_hi by ceil(2^34 / 5) >> 2; other shrinks: the
// reference's own 32-bit ((sum + f/2) * floor(2^24 / f)) >> 24, or the pair
// rounding (sum + 1) >> 1); these quotients lie in 0..255.
template <int FAM>
__device__ __forceinline__ uint32_t sep_round(int acc, uint32_t mul, int shift, int hi) {
    if constexpr (FAM == 0) {
        return (uint32_t)min(max(acc >> 12, 0), 255);
    } else {
        const uint32_t x = (uint32_t)acc;
        return (hi ? __umulhi(x, mul) : x * mul) >> shift;
    }
}

template <int B, int LIMBS, int FAM, bool PM>
__device__ void sep_task(const SepJob& job, int jx, int ty, unsigned char* lds, size_t wave_lds) {
    constexpr int KMAX = max_steps<LIMBS>();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int KH = job.kh, KV = job.kv, RING = job.ring;
    const int RS = RING + 16;  // ring row stride: padded against LDS bank conflicts
    const int ks = 64 * KH;
    int* scratch = reinterpret_cast<int*>(lds + (size_t)wave * wave_lds);
    int* corr = scratch + 16 * ks;
    unsigned char* ring = FAM == 0 ? reinterpret_cast<unsigned char*>(corr + 16)
                                   : reinterpret_cast<unsigned char*>(scratch);
    const InView& in = job.in;
    const OutView& out = job.out;
    const int nsamp = out.w * B;
    const int s0 = jx * 16;
    const int n = lane & 15, q = lane >> 4;
    const unsigned char* base = static_cast<const unsigned char*>(in.ptr);
    const uint32_t row_bytes = (uint32_t)in.bw * B;
    const uint32_t total = row_bytes * (uint32_t)in.bh;  // < 2^31 (checked on the host)

    // ---- H operands of the strip (B operand: rows j = samples, K = span bytes).
    i32x4 whi[KMAX], wlo[KMAX];
    int span0, hseed;
    if constexpr (FAM == 0) {
        const DevPlan& ph = job.ph;
        span0 = col_of(in, cload(ph.start, s0 / B));  // buffer-relative first span pixel
        const int smp = s0 + n;
        const bool active = smp < nsamp;
        const int x = active ? smp / B : 0;
        const int b = smp - x * B;
        const int st = ph.start[x];
        const int sp0 = span0;
        fold_weights(scratch, corr, ks, ph.taps, ph.wi16 + (size_t)ph.phase[x] * ph.taps, active,
                     [&](int t) { return (col_of(in, st + t) - sp0) * B + b; });
#pragma unroll
        for (int s = 0; s < KMAX; ++s) {
            if (s < KH) {
                if constexpr (LIMBS == 2) read_limbs(scratch, ks, s, whi[s], wlo[s]);
                else wlo[s] = read_single(scratch, ks, s);
            }
        }
        hseed = 128 * corr[n] + job.h_bias;
        wave_lds_sync();
    } else {
        const int pat = cload(job.hstrip, 2 * jx);
        span0 = cload(job.hstrip, 2 * jx + 1);
        const unsigned char* ht = job.htab + (size_t)pat * job.htab_bytes;
#pragma unroll
        for (int s = 0; s < KMAX; ++s) {
            if (s < KH) {
                if constexpr (LIMBS == 2) {
                    whi[s] = gmem<i32x4>(ht + ((size_t)(2 * s) * 64 + lane) * 16)[0];
                    wlo[s] = gmem<i32x4>(ht + ((size_t)(2 * s + 1) * 64 + lane) * 16)[0];
                } else {
                    wlo[s] = gmem<i32x4>(ht + ((size_t)s * 64 + lane) * 16)[0];
                }
            }
        }
        hseed = gmem<int>(ht + (size_t)KH * LIMBS * 64 * 16)[n];
    }

    const int y_begin = ty * job.rows_per_task;
    const int y_end = min(out.h, y_begin + job.rows_per_task);
    const uint32_t span_off = (uint32_t)span0 * B;
    const uint32_t lane_off = span_off + 16 * q;
    auto row_addr = [&](int r) -> uint32_t { return (uint32_t)row_of(in, r + n) * row_bytes + lane_off; };

    i32x4 acur, anxt;
    int have = INT32_MIN;  // intermediate rows below `have` (a multiple of 16) are in the ring
    int pre = INT32_MIN;   // M tile whose first-step A operand is in acur
    int ti = y_begin >> 4;
    int base16 = cload(job.vtile, 3 * ti + 1), hi = cload(job.vtile, 3 * ti + 2);
    for (int y0 = y_begin; y0 < y_end; y0 += 16, ++ti) {
        // ---- V operands of the 16 output rows (A operand, K = intermediate
        // rows from base16), host-folded; loaded ahead of the H work.
        const unsigned char* vt = job.vtab + (size_t)cload(job.vtile, 3 * ti) * job.vtab_bytes;
        const i32x4 vseed = gmem<i32x4>(vt + (size_t)KV * LIMBS * 64 * 16 + 16 * q)[0];
        i32x4 vop0h = {0, 0, 0, 0}, vop0l;
        if constexpr (LIMBS == 2) {
            vop0h = gmem<i32x4>(vt + (size_t)lane * 16)[0];
            vop0l = gmem<i32x4>(vt + ((size_t)64 + lane) * 16)[0];
        } else {
            vop0l = gmem<i32x4>(vt + (size_t)lane * 16)[0];
        }
        // ---- intermediate rows [max(have, base16), hi], 16 at a time; each
        // next operand's load is in flight while the current one is used.
        int r0 = have > base16 ? have : base16;
        uint32_t rb = row_addr(r0);
        if (r0 <= hi && pre != r0) acur = load_a<B, PM>(base, rb, total);
        for (; r0 <= hi; r0 += 16) {
            i32x4 dh = {0, 0, 0, 0}, dl = {hseed, hseed, hseed, hseed};
            const uint32_t rb_next = r0 + 16 <= hi ? row_addr(r0 + 16) : 0;
#pragma unroll
            for (int s = 0; s < KMAX; ++s) {
                if (s < KH) {
                    if (s + 1 < KH) anxt = load_a<B, PM>(base, rb + 64 * (s + 1), total);
                    else if (r0 + 16 <= hi) anxt = load_a<B, PM>(base, rb_next, total);
                    if constexpr (LIMBS == 2) dh = mfma_i8(acur, whi[s], dh);
                    dl = mfma_i8(acur, wlo[s], dl);
                    acur = anxt;
                }
            }
            rb = rb_next;
            // D[m = row][n = sample]: rows r0 + 4q .. + 3 of sample n, stored
            // in the ring already offset (p ^ 0x80) for the V operand.
            uint32_t packed = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int acc = LIMBS == 2 ? 128 * dh[e] + dl[e] : dl[e];
                packed |= sep_round<FAM>(acc, job.h_mul, job.h_shift, job.h_hi) << (8 * e);
            }
            *reinterpret_cast<uint32_t*>(ring + n * RS + ((r0 + 4 * q) & (RING - 1))) = packed ^ 0x80808080u;
        }
        have = r0 > have ? r0 : have;
        const int cur_base16 = base16;
        // Prefetch the first new M tile of the next row tile.
        if (y0 + 16 < y_end) {
            base16 = cload(job.vtile, 3 * ti + 4);
            hi = cload(job.vtile, 3 * ti + 5);
            const int r1 = have > base16 ? have : base16;
            if (r1 <= hi) {
                acur = load_a<B, PM>(base, row_addr(r1), total);
                pre = r1;
            }
        }
        i32x4 dh = {0, 0, 0, 0}, dl = vseed;
        for (int s = 0; s < KV; ++s) {
            const i32x4 bv =
                *reinterpret_cast<const i32x4*>(ring + n * RS + ((cur_base16 + 64 * s + 16 * q) & (RING - 1)));
            if constexpr (LIMBS == 2) {
                dh = mfma_i8(s == 0 ? vop0h : gmem<i32x4>(vt + ((size_t)(2 * s) * 64 + lane) * 16)[0], bv, dh);
                dl = mfma_i8(s == 0 ? vop0l : gmem<i32x4>(vt + ((size_t)(2 * s + 1) * 64 + lane) * 16)[0], bv, dl);
            } else {
                dl = mfma_i8(s == 0 ? vop0l : gmem<i32x4>(vt + ((size_t)s * 64 + lane) * 16)[0], bv, dl);
            }
        }
        // D[m = output row][n = sample]: rows y0 + 4q + e of sample s0 + n.
        uint32_t v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            v[e] = sep_round<FAM>(LIMBS == 2 ? 128 * dh[e] + dl[e] : dl[e], job.v_mul, job.v_shift, job.v_hi);
        if constexpr (B == 2 || B == 4) {
            if (job.unpremult) {
                // The pixel's alpha sample is in lane (lane | (B - 1)) of this group.
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t al = (uint32_t)__shfl((int)v[e], lane | (B - 1), 64);
                    if ((n & (B - 1)) != B - 1) v[e] = unpremult_u8(v[e], al);
                }
            }
        }
        const int smp = s0 + n;
        if (smp < nsamp) {
            unsigned char* o = static_cast<unsigned char*>(out.ptr) + (size_t)(y0 + 4 * q) * nsamp + smp;
            const int rows_left = y_end - (y0 + 4 * q);
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (e < rows_left) o[(size_t)e * nsamp] = (unsigned char)v[e];
        }
    }
}

// One kernel per (bands, limbs, family, premultiply) so each carries only its
// own registers.
template <int B, int LIMBS, int FAM, bool PM>
__global__ __launch_bounds__(kThreads) void k_sep_job(SepJob job, int wave_lds) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int t = blockIdx.x * kSepWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (t >= job.gx * job.gy) return;
    sep_task<B, LIMBS, FAM, PM>(job, t % job.gx, t / job.gx, lds, (size_t)wave_lds);
}

template <int B, int LIMBS, int FAM, bool PM>
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
        sep_task<B, LIMBS, FAM, PM>(job, task.bx, task.by, lds, (size_t)wave_lds);
    }
}

}  // namespace

size_t sep_job_wave_lds(const SepJob& job) {
    const size_t ring = 16 * ((size_t)job.ring + 16);
    if (job.family != 0) return ring;
    return (16 * 64 * (size_t)job.kh + 16) * sizeof(int) + ring;
}

int sep_job_variant(const SepJob& job) {
    const int pm = (job.bands == 2 || job.bands == 4) && job.premult ? 1 : 0;
    return (((job.bands - 1) * 2 + (job.limbs - 1)) * 2 + job.family) * 2 + pm;
}

bool sep_job_supported(const SepJob& job) {
    if (job.bands < 1 || job.bands > 4 || job.limbs < 1 || job.limbs > 2 || job.family < 0 || job.family > 1)
        return false;
    const int kmax = job.limbs == 1 ? 6 : 3;
    if (job.kh < 1 || job.kh > kmax || job.kv < 1 || job.kv > 7) return false;
    if (job.family == 0 && job.kh > 2) return false;
    return job.ring >= 64 * job.kv + 16 && (job.ring & (job.ring - 1)) == 0 && job.rows_per_task % 16 == 0;
}

// Every (bands, limbs, family, premultiply) instantiation, indexed by
// sep_job_variant.
#define VGPU_SEP_B(X, B)                                                                               \
    X(B, 1, 0, false) X(B, 1, 0, true) X(B, 1, 1, false) X(B, 1, 1, true) X(B, 2, 0, false) X(B, 2, 0, true) \
        X(B, 2, 1, false) X(B, 2, 1, true)
#define VGPU_SEP_ALL(X) VGPU_SEP_B(X, 1) VGPU_SEP_B(X, 2) VGPU_SEP_B(X, 3) VGPU_SEP_B(X, 4)

hipError_t launch_sep_job(const SepJob& job, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    if (!sep_job_supported(job)) return hipErrorInvalidValue;
    const int wl = (int)sep_job_wave_lds(job);
    const int ntasks = job.gx * job.gy;
    const dim3 grid((unsigned)((ntasks + kSepWaves - 1) / kSepWaves));
    const size_t lds = (size_t)wl * kSepWaves;
    const int var = sep_job_variant(job);
    int idx = 0;
#define X(B, L, F, P)                                                                                  \
    if (idx++ == var) hipLaunchKernelGGL((k_sep_job<B, L, F, P>), grid, dim3(kThreads), lds, s, job, wl);
    VGPU_SEP_ALL(X)
#undef X
    return hipGetLastError();
}

hipError_t launch_sep_batch(int variant, const SepJob* jobs, const BatchTask* tasks, int ntasks,
                            const void* const* srcs, void* const* dsts, int grid, size_t wave_lds,
                            hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    if (variant < 0 || variant >= 64) return hipErrorInvalidValue;
    const size_t lds = wave_lds * kSepWaves;
    const int wl = (int)wave_lds;
    int idx = 0;
#define X(B, L, F, P)                                                                                   \
    if (idx++ == variant)                                                                               \
        hipLaunchKernelGGL((k_sep_batch<B, L, F, P>), dim3((unsigned)grid), dim3(kThreads), lds, s, jobs, tasks, \
                           ntasks, srcs, dsts, wl);
    VGPU_SEP_ALL(X)
#undef X
    return hipGetLastError();
}
#undef VGPU_SEP_ALL
#undef VGPU_SEP_B

}  // namespace vgpu
