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
// H A operands stream through a per-wavefront LDS ring of slots by
// LDS-direct loads (global_load_lds_dwordx4: lane L's 16 bytes land at
// slot + 16 L; global_load_lds_dword: lane L's dword at slot + 1024 + 4 L),
// so several M-tile steps are in flight without holding registers.  A slot
// carries the 16 bytes from the lane's dword-aligned start (4 bands: rows
// start on dwords) plus, for other band counts, the next dword, realigned by
// alignbyte.
template <int B>
constexpr int a_dwords() {
    return B == 4 ? 4 : 5;
}
template <int FAM>
constexpr int sep_slots() {
    return FAM == 0 ? 3 : 4;  // deeper rings measured slower (LDS occupancy)
}
// A slot holds an H operand (1 KiB, plus 256 bytes when realigned) or a V
// tile's first K-step operand (1 KiB per limb).
template <int B, int LIMBS>
constexpr size_t a_slot_bytes() {
    return (size_t)(a_dwords<B>() > 4 * LIMBS ? a_dwords<B>() : 4 * LIMBS) * 256;
}
// Bytes of the per-wavefront region before the ring: family 0's
// weight-folding scratch, which doubles as the slot ring once the H weights
// are in registers, or the slot ring alone.
template <int B, int LIMBS, int FAM>
__host__ __device__ constexpr size_t sep_slot_region(int kh) {
    const size_t slots = sep_slots<FAM>() * a_slot_bytes<B, LIMBS>();
    const size_t fold = FAM == 0 ? (16 * 64 * (size_t)kh + 16) * sizeof(int) : 0;
    return fold > slots ? fold : slots;
}

// LDS-direct loads: every lane's 16 (4) bytes from src land at dst + 16 lane
// (dst + 4 lane); dst is wavefront-uniform.
__device__ __forceinline__ void lds_load16(const unsigned char* src, unsigned char* dst) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}
__device__ __forceinline__ void lds_load4(const unsigned char* src, unsigned char* dst) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)dst, 4, 0, 0);
}

// Wait until at most N vector-memory operations are outstanding (they retire
// in issue order); the memory clobber keeps LDS reads after it.
template <int N>
__device__ __forceinline__ void wait_vmem() {
    static_assert(N >= 0 && N < 16, "vmcnt immediate");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

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
    unsigned char* ring = lds + (size_t)wave * wave_lds + sep_slot_region<B, LIMBS, FAM>(KH);
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

    // ---- Operand stream, in the order the loop consumes it: for each 16-row
    // output tile, its new H (M tile, K step) operands, then its first V
    // operand; NS slots ahead, through LDS-direct loads.  Waiting for a slot
    // needs an immediate vmcnt: every element is at least one load, so while
    // NS - 1 newer elements are in flight, NS - 1 outstanding loads leave the
    // slot's own landed.
    constexpr int ND = a_dwords<B>(), NS = sep_slots<FAM>();
    constexpr size_t SLOT = a_slot_bytes<B, LIMBS>();
    unsigned char* aring = lds + (size_t)wave * wave_lds;
    const int last_ti = (y_end - 1) >> 4;
    int pf_ti = y_begin >> 4;
    int pf_r0 = cload(job.vtile, 3 * pf_ti + 1), pf_hi = cload(job.vtile, 3 * pf_ti + 2), pf_s = 0;
    bool pf_v = false;  // the cursor is on tile pf_ti's V operand
    bool pf_done = y_begin >= y_end;
    int issued = 0, consumed = 0;

    auto pf_issue = [&]() {
        if (pf_done) return;
        unsigned char* slot = aring + (size_t)(issued % NS) * SLOT;
        ++issued;
        if (pf_v) {
            const unsigned char* vt = job.vtab + (size_t)cload(job.vtile, 3 * pf_ti) * job.vtab_bytes;
#pragma unroll
            for (int l = 0; l < LIMBS; ++l) lds_load16(vt + (size_t)l * 1024 + lane * 16, slot + 1024 * l);
            pf_v = false;
            if (++pf_ti > last_ti) {
                pf_done = true;
                return;
            }
            pf_r0 = max(pf_r0, cload(job.vtile, 3 * pf_ti + 1));
            pf_hi = cload(job.vtile, 3 * pf_ti + 2);
            pf_v = pf_r0 > pf_hi;
            return;
        }
        const uint32_t a4 = (row_addr(pf_r0) + 64 * pf_s) & ~3u;
        const bool whole = a4 + 4 * ND <= total;
        const unsigned char* src = base + (whole ? a4 : 0u);
        lds_load16(src, slot);
        if constexpr (ND == 5) lds_load4(src + 16, slot + 1024);
        if (!whole) {
            // The buffer's last bytes (a lane whose dwords run past its end
            // loaded from its start above): read bytewise, zero past the end,
            // and written over the slot once those loads (and so the slot's,
            // issued before them) have landed.
            uint32_t d[ND];
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                uint32_t v = 0;
                for (int e = 0; e < 4; ++e) {
                    const uint32_t off = a4 + 4 * j + e;
                    if (off < total) v |= (uint32_t)gmem<unsigned char>(base)[off] << (8 * e);
                }
                d[j] = v;
            }
            wait_vmem<0>();
#pragma unroll
            for (int j = 0; j < 4; ++j) *reinterpret_cast<uint32_t*>(slot + 16 * lane + 4 * j) = d[j];
            if constexpr (ND == 5) *reinterpret_cast<uint32_t*>(slot + 1024 + 4 * lane) = d[4];
        }
        if (++pf_s == KH) {
            pf_s = 0;
            pf_r0 += 16;
            pf_v = pf_r0 > pf_hi;
        }
    };
    // Wait for the next element's slot; returns it.  next() then refills it
    // with the element NS ahead once its reads are done.
    auto slot_in = [&]() -> const unsigned char* {
        if (issued - consumed >= NS) wait_vmem<NS - 1>();
        else wait_vmem<0>();
        return aring + (size_t)(consumed % NS) * SLOT;
    };
    auto next = [&]() {
        ++consumed;
        // (The slot's reads are in registers before the load that overwrites it.)
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        pf_issue();
    };
    auto take_h = [&](int r0, int s) -> i32x4 {
        const unsigned char* slot = slot_in();
        const u32x4 d4 = *reinterpret_cast<const u32x4*>(slot + 16 * lane);
        uint32_t d[ND];
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = d4[j];
        if constexpr (ND == 5) d[4] = *reinterpret_cast<const uint32_t*>(slot + 1024 + 4 * lane);
        uint32_t w[4];
        if constexpr (B == 4) {
#pragma unroll
            for (int k = 0; k < 4; ++k) w[k] = d[k];
        } else {
            const int al = (int)((row_addr(r0) + 64 * s) & 3);
#pragma unroll
            for (int k = 0; k < 4; ++k) w[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], al);
        }
        i32x4 av;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t v = w[k];
            if constexpr ((B == 2 || B == 4) && PM) v = premult_dword<B>(v);
            av[k] = (int)(v ^ 0x80808080u);
        }
        next();
        return av;
    };
    // (When a step's first element is consumed NS elements after the
    // previous ones were issued, slot_in's wait count stays valid.)
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // family 0's fold reads are done
    for (int i = 0; i < NS; ++i) pf_issue();

    int have = INT32_MIN;  // intermediate rows below `have` (a multiple of 16) are in the ring
    int ti = y_begin >> 4;
    int base16 = cload(job.vtile, 3 * ti + 1), hi = cload(job.vtile, 3 * ti + 2);
    for (int y0 = y_begin; y0 < y_end; y0 += 16, ++ti) {
        const unsigned char* vt = job.vtab + (size_t)cload(job.vtile, 3 * ti) * job.vtab_bytes;
        // ---- intermediate rows [max(have, base16), hi], 16 at a time.
        int r0 = have > base16 ? have : base16;
        for (; r0 <= hi; r0 += 16) {
            i32x4 dh = {0, 0, 0, 0}, dl = {hseed, hseed, hseed, hseed};
#pragma unroll
            for (int s = 0; s < KMAX; ++s) {
                if (s < KH) {
                    const i32x4 acur = take_h(r0, s);
                    if constexpr (LIMBS == 2) dh = mfma_i8(acur, whi[s], dh);
                    dl = mfma_i8(acur, wlo[s], dl);
                }
            }
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
        if (y0 + 16 < y_end) {
            base16 = cload(job.vtile, 3 * ti + 4);
            hi = cload(job.vtile, 3 * ti + 5);
        }
        // ---- V operands of the 16 output rows (A operand, K = intermediate
        // rows from base16), host-folded: the first K step from the stream,
        // the seeds (128 * sum of the row's weights + bias, per output row
        // 4q + e) through the scalar cache.
        i32x4 vop0h = {0, 0, 0, 0}, vop0l;
        {
            const unsigned char* slot = slot_in();
            if constexpr (LIMBS == 2) {
                vop0h = *reinterpret_cast<const i32x4*>(slot + 16 * lane);
                vop0l = *reinterpret_cast<const i32x4*>(slot + 1024 + 16 * lane);
            } else {
                vop0l = *reinterpret_cast<const i32x4*>(slot + 16 * lane);
            }
            next();
        }
        const int* seeds = reinterpret_cast<const int*>(vt + (size_t)KV * LIMBS * 64 * 16);
        i32x4 vseed;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int s0v = cload(seeds, e), s1v = cload(seeds, 4 + e), s2v = cload(seeds, 8 + e),
                      s3v = cload(seeds, 12 + e);
            vseed[e] = q == 0 ? s0v : (q == 1 ? s1v : (q == 2 ? s2v : s3v));
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

// ---- Block-tile form ----------------------------------------------------
// A block of four wavefronts owns four adjacent 16-sample strips x R output
// rows (R = rows_per_task).  Everything the block reads -- the input rows its
// V windows need (intermediate rows r_lo .. r_hi, each row of input
// row_of(r)) across the four strips' H spans, and the V operand patterns of
// its 16-row tiles -- goes into LDS first by LDS-direct loads, all in flight
// at once, so the block waits one memory latency.  Each wavefront then runs
// its strip from LDS: the H M tiles into its own intermediate columns (16
// samples x the block's rows, column-major, already offset for the V
// operand), then the V tiles; both two tiles per step so that one tile's LDS
// reads and MFMAs overlap the other's.  No wavefront reads another's
// columns, so the only block barrier is the one after the loads.
//
// Staged row i (intermediate row r_lo + i) holds the bytes of input row
// row_of(r_lo + i) from the dword below the tile's first span byte, so the
// row's byte j sits at offset mis(i) + j, mis(i) = (row address + tile byte
// 0) & 3.  Chunks past the end of the buffer are read bytewise, zero beyond
// it.
template <int B, int LIMBS, int FAM, bool PM>
__device__ void sep_tile_task(const SepJob& job, int bx, int by, unsigned char* lds) {
    constexpr int KMAX = max_steps<LIMBS>();
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const InView& in = job.in;
    const OutView& out = job.out;
    const int KH = job.kh, KV = job.kv;
    const int nsamp = out.w * B;
    const int jx0 = 4 * bx;
    const int nst = min(4, job.gx - jx0);
    const int R = job.rows_per_task;
    const int y_begin = by * R, y_end = min(out.h, y_begin + R);
    const int t0 = y_begin >> 4, t1 = (y_end - 1) >> 4;
    const int r_lo = cload(job.vtile, 3 * t0 + 1);
    const int r_hi = cload(job.vtile, 3 * t1 + 2);
    const int nrows = r_hi - r_lo + 1;
    const int c_first = cload(job.hstrip, 2 * jx0 + 1);
    const unsigned char* base = static_cast<const unsigned char*>(in.ptr);
    const uint32_t row_bytes = (uint32_t)in.bw * B;
    const uint32_t total = row_bytes * (uint32_t)in.bh;  // < 2^31 (checked on the host)
    const uint32_t byte0 = (uint32_t)c_first * B;
    const int RSB = job.t_row_stride, RS = job.t_inter_stride;
    const int VSZ = KV * LIMBS * 1024;
    unsigned char* tile = lds;
    unsigned char* inter = lds + (size_t)job.t_rows * RSB;
    unsigned char* vstage = inter + (size_t)64 * RS;

    // ---- Stage the input rows: chunk e = (row e / nchunk_row, 16-byte
    // column e % nchunk_row); a wavefront's 64 chunks land contiguously.
    const int nchunk_row = RSB >> 4;
    const int nchunks = nrows * nchunk_row;
    auto chunk_addr = [&](int e) -> uint32_t {
        const int i = (int)__umulhi((uint32_t)e, job.t_chunk_magic);
        const int c = e - i * nchunk_row;
        const uint32_t ra = (uint32_t)row_of(in, r_lo + min(i, nrows - 1)) * row_bytes + byte0;
        return (ra & ~3u) + 16u * (uint32_t)c;
    };
    for (int e0 = wave * 64; e0 < nchunks; e0 += 256) {
        const int e = e0 + lane;
        const uint32_t a = chunk_addr(e);
        if (e < nchunks) lds_load16(base + (a + 16 <= total ? a : 0u), tile + (size_t)e0 * 16);
    }
    // Whether some chunk of the block runs past the buffer end (rows are
    // clamped, so the last staged row has the highest address).
    const bool edge = ((((uint32_t)row_of(in, r_hi) * row_bytes + byte0) & ~3u) + (uint32_t)RSB) > total;
    // ---- Stage the V operands of the block's 16-row tiles.
    for (int t = t0; t <= t1; ++t) {
        const unsigned char* vt = job.vtab + (size_t)cload(job.vtile, 3 * t) * job.vtab_bytes;
        for (int k = wave; k < KV * LIMBS; k += 4)
            lds_load16(vt + (size_t)k * 1024 + lane * 16, vstage + (size_t)(t - t0) * VSZ + (size_t)k * 1024);
    }
    // ---- This wavefront's H operands (host-folded pattern of its strip).
    const bool have = wave < nst;
    const int jx = jx0 + (have ? wave : 0);
    i32x4 whi[KMAX], wlo[KMAX];
    const int pat = cload(job.hstrip, 2 * jx);
    const int span0 = cload(job.hstrip, 2 * jx + 1);
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
    const int n = lane & 15, q = lane >> 4;
    const int hseed = gmem<int>(ht + (size_t)KH * LIMBS * 64 * 16)[n];
    wait_vmem<0>();
    if (edge) {
        // Chunks that cross the buffer end were loaded from its start: write
        // the buffer's last bytes over them, zero past its end.
        for (int e = wave * 64 + lane; e < nchunks; e += 256) {
            const uint32_t a = chunk_addr(e);
            if (a + 16 <= total) continue;
            uint32_t d[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint32_t v = 0;
                for (int b = 0; b < 4; ++b) {
                    const uint32_t off = a + 4 * j + b;
                    if (off < total) v |= (uint32_t)gmem<unsigned char>(base)[off] << (8 * b);
                }
                d[j] = v;
            }
            *reinterpret_cast<u32x4*>(tile + (size_t)e * 16) = u32x4{d[0], d[1], d[2], d[3]};
        }
    }
    __syncthreads();
    if (!have) return;

    // ---- H: intermediate rows r_lo .. r_hi, 32 at a time (two M tiles),
    // into this wavefront's columns (sample w * 16 + n at
    // inter + (w * 16 + n) * RS).
    unsigned char* col = inter + (size_t)(wave * 16 + n) * RS;
    const uint32_t soff = (uint32_t)(span0 - c_first) * B + 16u * q;
    auto a_operand = [&](int r0, int s) -> i32x4 {
        const int i = min(r0 - r_lo + n, nrows - 1);  // lane (m = n, q): staged row of M row m
        const uint32_t mis = ((uint32_t)row_of(in, r_lo + i) * row_bytes + byte0) & 3u;
        const uint32_t bo = soff + 64u * s + mis;
        const uint32_t* p = reinterpret_cast<const uint32_t*>(tile + (size_t)i * RSB + (bo & ~3u));
        uint32_t w[4];
        if constexpr (B == 4) {
#pragma unroll
            for (int k = 0; k < 4; ++k) w[k] = p[k];
        } else {
            uint32_t d[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) d[k] = p[k];
            const int al = (int)(bo & 3u);
#pragma unroll
            for (int k = 0; k < 4; ++k) w[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], al);
        }
        i32x4 av;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t v = w[k];
            if constexpr ((B == 2 || B == 4) && PM) v = premult_dword<B>(v);
            av[k] = (int)(v ^ 0x80808080u);
        }
        return av;
    };
    auto h_pack = [&](const i32x4& dh, const i32x4& dl) -> uint32_t {
        uint32_t packed = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int acc = LIMBS == 2 ? 128 * dh[e] + dl[e] : dl[e];
            packed |= sep_round<FAM>(acc, job.h_mul, job.h_shift, job.h_hi) << (8 * e);
        }
        return packed ^ 0x80808080u;
    };
    for (int r0 = r_lo; r0 <= r_hi; r0 += 32) {
        const bool two = r0 + 16 <= r_hi;
        const int r1 = two ? r0 + 16 : r0;
        i32x4 dh0 = {0, 0, 0, 0}, dl0 = {hseed, hseed, hseed, hseed};
        i32x4 dh1 = {0, 0, 0, 0}, dl1 = {hseed, hseed, hseed, hseed};
#pragma unroll
        for (int s = 0; s < KMAX; ++s) {
            if (s < KH) {
                const i32x4 a0 = a_operand(r0, s);
                const i32x4 a1 = a_operand(r1, s);
                if constexpr (LIMBS == 2) {
                    dh0 = mfma_i8(a0, whi[s], dh0);
                    dh1 = mfma_i8(a1, whi[s], dh1);
                }
                dl0 = mfma_i8(a0, wlo[s], dl0);
                dl1 = mfma_i8(a1, wlo[s], dl1);
            }
        }
        *reinterpret_cast<uint32_t*>(col + (r0 - r_lo + 4 * q)) = h_pack(dh0, dl0);
        if (two) *reinterpret_cast<uint32_t*>(col + (r1 - r_lo + 4 * q)) = h_pack(dh1, dl1);
    }

    // ---- V: the 16-row output tiles from the staged operands, two per step.
    const int s0 = jx * 16;
    const int smp = s0 + n;
    auto v_tile = [&](int t, i32x4& dh, i32x4& dl) {
        const int base16 = cload(job.vtile, 3 * t + 1);
        const unsigned char* vt = job.vtab + (size_t)cload(job.vtile, 3 * t) * job.vtab_bytes;
        const int* seeds = reinterpret_cast<const int*>(vt + (size_t)KV * LIMBS * 64 * 16);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int s0v = cload(seeds, e), s1v = cload(seeds, 4 + e), s2v = cload(seeds, 8 + e),
                      s3v = cload(seeds, 12 + e);
            dl[e] = q == 0 ? s0v : (q == 1 ? s1v : (q == 2 ? s2v : s3v));
            dh[e] = 0;
        }
        const unsigned char* vs = vstage + (size_t)(t - t0) * VSZ;
        for (int s = 0; s < KV; ++s) {
            const i32x4 bv = *reinterpret_cast<const i32x4*>(col + (base16 - r_lo + 64 * s + 16 * q));
            if constexpr (LIMBS == 2) {
                dh = mfma_i8(*reinterpret_cast<const i32x4*>(vs + (size_t)(2 * s) * 1024 + lane * 16), bv, dh);
                dl = mfma_i8(*reinterpret_cast<const i32x4*>(vs + (size_t)(2 * s + 1) * 1024 + lane * 16), bv, dl);
            } else {
                dl = mfma_i8(*reinterpret_cast<const i32x4*>(vs + (size_t)s * 1024 + lane * 16), bv, dl);
            }
        }
    };
    auto v_store = [&](int t, const i32x4& dh, const i32x4& dl) {
        const int y0 = 16 * t;
        uint32_t v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            v[e] = sep_round<FAM>(LIMBS == 2 ? 128 * dh[e] + dl[e] : dl[e], job.v_mul, job.v_shift, job.v_hi);
        if constexpr (B == 2 || B == 4) {
            if (job.unpremult) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t al = (uint32_t)__shfl((int)v[e], lane | (B - 1), 64);
                    if ((n & (B - 1)) != B - 1) v[e] = unpremult_u8(v[e], al);
                }
            }
        }
        if (smp < nsamp) {
            unsigned char* o = static_cast<unsigned char*>(out.ptr) + (size_t)(y0 + 4 * q) * nsamp + smp;
            const int rows_left = y_end - (y0 + 4 * q);
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (e < rows_left) o[(size_t)e * nsamp] = (unsigned char)v[e];
        }
    };
    for (int t = t0; t <= t1; t += 2) {
        i32x4 dh0, dl0, dh1, dl1;
        v_tile(t, dh0, dl0);
        if (t + 1 <= t1) v_tile(t + 1, dh1, dl1);
        v_store(t, dh0, dl0);
        if (t + 1 <= t1) v_store(t + 1, dh1, dl1);
    }
}

template <int B, int LIMBS, int FAM, bool PM>
__global__ __launch_bounds__(kThreads) void k_sep_tile(SepJob job) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int gxb = (job.gx + 3) >> 2;
    sep_tile_task<B, LIMBS, FAM, PM>(job, (int)blockIdx.x % gxb, (int)blockIdx.x / gxb, lds);
}

template <int B, int LIMBS, int FAM, bool PM>
__global__ __launch_bounds__(kThreads) void k_sep_tile_batch(const SepJob* __restrict__ jobs,
                                                             const BatchTask* __restrict__ tasks, int ntasks,
                                                             const void* const* __restrict__ srcs,
                                                             void* const* __restrict__ dsts) {
    extern __shared__ __align__(16) unsigned char lds[];
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const BatchTask task = tasks[t];
        SepJob job = jobs[task.job];
        if (job.src_slot >= 0) job.in.ptr = srcs[job.src_slot];
        if (job.dst_slot >= 0) job.out.ptr = dsts[job.dst_slot];
        sep_tile_task<B, LIMBS, FAM, PM>(job, task.bx, task.by, lds);
        __syncthreads();  // the next task's staging overwrites this one's LDS
    }
}

}  // namespace

size_t sep_job_wave_lds(const SepJob& job) {
    const size_t ring = 16 * ((size_t)job.ring + 16);
    const int nd = job.bands == 4 ? 4 : 5, vd = 4 * job.limbs;
    const size_t slots = (size_t)(nd > vd ? nd : vd) * 256 * (job.family == 0 ? 3 : 4);
    const size_t fold = job.family == 0 ? (16 * 64 * (size_t)job.kh + 16) * sizeof(int) : 0;
    return std::max(fold, slots) + ring;
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
    if (job.tile)
        return job.rows_per_task % 16 == 0 && job.rows_per_task > 0 && job.t_row_stride % 16 == 0 &&
               job.t_row_stride > 0 && job.t_inter_stride % 16 == 0 && job.t_lds > 0 && job.t_lds <= 64 * 1024 &&
               job.htab && job.hstrip;
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
    if (job.tile) {
        const int blocks = ((job.gx + 3) / 4) * job.gy;
        const int var = sep_job_variant(job);
        int idx = 0;
#define X(B, L, F, P)                                                                                   \
    if (idx++ == var)                                                                                   \
        hipLaunchKernelGGL((k_sep_tile<B, L, F, P>), dim3((unsigned)blocks), dim3(kThreads), (size_t)job.t_lds, s, job);
        VGPU_SEP_ALL(X)
#undef X
        return hipGetLastError();
    }
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
hipError_t launch_sep_tile_batch(int variant, const SepJob* jobs, const BatchTask* tasks, int ntasks,
                                 const void* const* srcs, void* const* dsts, int grid, size_t lds,
                                 hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    if (variant < 0 || variant >= 64) return hipErrorInvalidValue;
    int idx = 0;
#define X(B, L, F, P)                                                                                   \
    if (idx++ == variant)                                                                               \
        hipLaunchKernelGGL((k_sep_tile_batch<B, L, F, P>), dim3((unsigned)grid), dim3(kThreads), lds, s, jobs, tasks, \
                           ntasks, srcs, dsts);
    VGPU_SEP_ALL(X)
#undef X
    return hipGetLastError();
}
#undef VGPU_SEP_ALL
#undef VGPU_SEP_B

}  // namespace vgpu
