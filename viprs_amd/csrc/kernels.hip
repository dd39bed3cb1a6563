// gfx950 kernels for the resample / separable-convolution path.
//
// Layout everywhere: packed, band-interleaved, row-major samples
// ((y * w + x) * bands + b), as in the reference's tiles
// (crates/viprs-core/src/image/region.rs:154-230).
//
// Integer formats reproduce the reference's fixed-point sums exactly (the
// sums are exact integers, so summation order is free).  The f32 kernels here
// are the reference-exact ones: they carry the reference's FP64 accumulation
// in the reference's operation order (FMA where it uses mul_add, separate
// multiply/add where it does not), so they are bit-identical to it.  This
// file must be compiled with -ffp-contract=off.
#include "device.hpp"

#include <algorithm>

namespace vgpu {
namespace {

// ShrinkH u8 rounding on x86 (shrinkh.rs:381-397 dispatch): f = 2 rounds a
// pair, f = 5 with 1/3/4 bands divides exactly, every other factor uses the
// libvips multiply-shift ((sum + f/2) * floor(2^24 / f)) >> 24.  The product
// stays below 2^32 for u8 sums, so 32-bit arithmetic is exact.
__device__ __forceinline__ uint32_t shrink_round_h(uint32_t sum, uint32_t f, uint32_t bands,
                                                   uint32_t mult) {
    if (f == 2) return (sum + 1) >> 1;
    if (f == 5 && bands != 2 && bands <= 4) return (sum + 2) / 5;
    return ((sum + (f >> 1)) * mult) >> 24;
}
// ShrinkV u8 (shrinkv.rs:206-339): pair rounding at f = 2, else multiply-shift.
__device__ __forceinline__ uint32_t shrink_round_v(uint32_t sum, uint32_t f, uint32_t mult) {
    if (f == 2) return (sum + 1) >> 1;
    return ((sum + (f >> 1)) * mult) >> 24;
}

// ---------------------------------------------------------------------------
// Fused ShrinkH -> ShrinkV for u8 (the thumbnail's box pre-shrink).
//
// Block = TX output columns of one output row.  The block walks the fv source
// rows of that output row RL = 256 / TX rows at a time: it stages the
// contiguous byte span the TX box windows cover into LDS with 16-byte loads,
// then every (column, row-lane) thread sums its fh-pixel window per band,
// applies the ShrinkH rounding, and accumulates the rounded value.  The
// row-lane partial sums are combined through LDS and the ShrinkV rounding is
// applied once per output sample.  Premultiply (4-band / 2-band thumbnails)
// is applied to each source pixel as it is read.
// ---------------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(kThreads) void k_shrink_hv_u8(InView in, OutView out, int fh, int fv,
                                                           int tx, int RL, int premult,
                                                           uint32_t mh, uint32_t mv,
                                                           int aligned) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lanes = kThreads / tx;  // row lanes that hold partial sums
    const int t = threadIdx.x;
    const int c = t % tx, r = t / tx;
    const int col0 = blockIdx.x * tx;
    const int ncols = min(tx, out.w - col0);
    const int oy = out.y + (int)blockIdx.y;
    const int cx0 = (out.x + col0) * fh;  // absolute first source column
    const int span = ncols * fh * B;
    const int pitch = ((tx * fh * B + 32) + 15) & ~15;
    const unsigned char* base = static_cast<const unsigned char*>(in.ptr);
    const size_t row_bytes = (size_t)in.bw * B;
    const size_t total = row_bytes * (size_t)in.bh;
    const bool inside = aligned && cx0 >= in.bx && cx0 + ncols * fh <= in.bx + in.bw;

    uint32_t acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) acc[b] = 0;

    for (int k0 = 0; k0 < fv; k0 += RL) {
        const int nrows = min(RL, fv - k0);
        if (inside) {
            // Interior: one contiguous byte run per row, read as 16-byte chunks
            // starting at the 16-byte boundary at or before the run.
            int nchunks_max = (span + 15 + 15) / 16;
            for (int it = t; it < nrows * nchunks_max; it += kThreads) {
                const int rr = it / nchunks_max, i = it - rr * nchunks_max;
                const int row = row_of(in, oy * fv + k0 + rr);
                const size_t a = row * row_bytes + (size_t)(cx0 - in.bx) * B;
                const size_t a16 = a & ~(size_t)15;
                const int nchunks = (int)((a - a16 + span + 15) / 16);
                if (i >= nchunks) continue;
                const size_t off = a16 + (size_t)i * 16;
                uint4 v;
                if (off + 16 <= total) {
                    v = *reinterpret_cast<const uint4*>(base + off);
                } else {
                    unsigned char tmp[16];
                    for (int q = 0; q < 16; ++q) tmp[q] = (off + q < total) ? base[off + q] : 0;
                    v = *reinterpret_cast<const uint4*>(tmp);
                }
                *reinterpret_cast<uint4*>(lds + rr * pitch + i * 16) = v;
            }
        } else {
            // Border: byte gather with clamp-to-edge columns.
            for (int it = t; it < nrows * span; it += kThreads) {
                const int rr = it / span, j = it - rr * span;
                const int row = row_of(in, oy * fv + k0 + rr);
                const int col = col_of(in, cx0 + j / B);
                lds[rr * pitch + j] = base[(size_t)row * row_bytes + (size_t)col * B + (j % B)];
            }
        }
        __syncthreads();
        if (r < nrows && c < ncols) {
            int shift = 0;
            if (inside) {
                const int row = row_of(in, oy * fv + k0 + r);
                const size_t a = row * row_bytes + (size_t)(cx0 - in.bx) * B;
                shift = (int)(a & 15);
            }
            const unsigned char* w = lds + r * pitch + shift + c * fh * B;
            uint32_t s[B];
#pragma unroll
            for (int b = 0; b < B; ++b) s[b] = 0;
            if (premult && (B == 2 || B == 4)) {
                for (int j = 0; j < fh; ++j) {
                    const uint32_t a = w[j * B + B - 1];
#pragma unroll
                    for (int b = 0; b < B - 1; ++b) s[b] += premult_u8(w[j * B + b], a);
                    s[B - 1] += a;
                }
            } else {
                for (int j = 0; j < fh; ++j) {
#pragma unroll
                    for (int b = 0; b < B; ++b) s[b] += w[j * B + b];
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] += shrink_round_h(s[b], (uint32_t)fh, B, mh);
        }
        __syncthreads();
    }
    uint32_t* part = reinterpret_cast<uint32_t*>(lds);
    if (c < ncols) {
#pragma unroll
        for (int b = 0; b < B; ++b) part[(r * tx + c) * B + b] = acc[b];
    }
    __syncthreads();
    if (r == 0 && c < ncols) {
        unsigned char* o = static_cast<unsigned char*>(out.ptr) + ((size_t)blockIdx.y * out.w + col0 + c) * B;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            uint32_t sum = 0;
            for (int q = 0; q < lanes; ++q) sum += part[(q * tx + c) * B + b];
            o[b] = (unsigned char)shrink_round_v(sum, (uint32_t)fv, mv);
        }
    }
}

// ---------------------------------------------------------------------------
// Fused ShrinkH -> ShrinkV for u8, wavefront-run form.
//
// Each wavefront owns TX (<= 64) output columns of a run of m consecutive
// output rows (or, with S > 1, every S-th source row of one output row, the
// S partial sums meeting in LDS at the end).  The source rows of the run are
// contiguous, so the wavefront streams them in fetch groups: a group is RP
// source rows, each given a 2^lsl-chunk slot of 16 bytes (RP * 2^lsl = 256
// chunks = 4 KiB), loaded branch-free as four 16-byte loads per lane one
// group ahead in registers, then parked in the wavefront's own LDS buffer.
// Small box windows therefore still keep about 4 KiB per wavefront in flight.
// Each lane sums its fh-pixel box window with dword LDS reads realigned by
// alignbyte and per-band byte masks folded in by udot4; windows that reach
// past the clamped span (the Z^2 halo columns of a thumbnail) take a
// per-pixel clamped path.  Rounding is the reference's: ShrinkH rounding per
// source row, then ShrinkV rounding of the sum.  Wavefronts synchronise only
// with themselves, except for the S-way partial-sum combine.
// ---------------------------------------------------------------------------
constexpr int kSpanBytes = 4096;  // LDS bytes per wavefront group buffer

template <int B>
__device__ __forceinline__ uint32_t band_mask(int b, int kk) {
    if constexpr (B == 1) return 0x01010101u;
    if constexpr (B == 2) return 0x00010001u << (8 * b);
    if constexpr (B == 4) return 1u << (8 * b);
    // B == 3: bytes of aligned dword k carry bands (4k + j) % 3; kk = k % 3.
    const int m = (b - kk + 3) % 3;
    return m == 0 ? 0x01000001u : (m == 1 ? 0x00000100u : 0x00010000u);
}

// Sum of one box window of wbytes bytes starting at LDS byte sbyte, per band;
// rbuf must be 4-byte aligned.  Dword d of the window is realigned from two
// LDS dwords by alignbyte, masked (bytes past the window end -> 0; the mask
// depends only on d and the window size, so it is scalar), and folded into
// the band sums by udot4 with per-band byte masks.  The loop is unrolled by U
// dwords (a multiple of 3 for 3 bands, so the band pattern restarts) with two
// accumulator chains, and may read up to U dwords past the window; those
// bytes are masked.
template <int B>
__device__ __forceinline__ void window_sums(const unsigned char* rbuf, int sbyte, int wbytes,
                                            uint32_t (&sum)[B]) {
    const uint32_t* l32 = reinterpret_cast<const uint32_t*>(rbuf) + (sbyte >> 2);
    const int al = sbyte & 3;
    const int nd = (wbytes + 3) >> 2;
    const int rem_last = wbytes - 4 * (nd - 1);  // 1..4 valid bytes in the last dword
    const uint32_t last_mask = rem_last >= 4 ? 0xffffffffu : ((1u << (8 * rem_last)) - 1u);
    constexpr int U = B == 3 ? 6 : 4;
    // Even and odd dwords accumulate separately: two independent udot4 chains
    // per band.
    uint32_t odd[B];
#pragma unroll
    for (int b = 0; b < B; ++b) odd[b] = 0;
    uint32_t cur = l32[0];
    int d0 = 0;
    // Blocks of U dwords that lie wholly before the last dword: no masks.
    for (; d0 + U <= nd - 1; d0 += U) {
        uint32_t w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = l32[d0 + u + 1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t v = __builtin_amdgcn_alignbyte(w[u], cur, al);
            cur = w[u];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (u & 1) odd[b] = __builtin_amdgcn_udot4(v, band_mask<B>(b, u % 3), odd[b], false);
                else sum[b] = __builtin_amdgcn_udot4(v, band_mask<B>(b, u % 3), sum[b], false);
            }
        }
    }
    // Tail block: the last 1..U dwords, masked.  Unrolled, so that all its
    // LDS reads are in flight together (a dword-at-a-time loop exposes the LDS
    // latency on every step and is slower even for small windows).
    {
        uint32_t w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = l32[d0 + u + 1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int d = d0 + u;
            const uint32_t m = d < nd - 1 ? 0xffffffffu : (d == nd - 1 ? last_mask : 0u);
            const uint32_t v = __builtin_amdgcn_alignbyte(w[u], cur, al) & m;
            cur = w[u];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (u & 1) odd[b] = __builtin_amdgcn_udot4(v, band_mask<B>(b, u % 3), odd[b], false);
                else sum[b] = __builtin_amdgcn_udot4(v, band_mask<B>(b, u % 3), sum[b], false);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) sum[b] += odd[b];
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Premultiplied box sums of a 4-band window of fh pixels starting at the
// 4-byte-aligned LDS byte sbyte.  Pixels go in pairs: v_perm gathers each
// colour channel of the two pixels into one packed 16-bit pair [c0, c1] and
// their alphas into [a0, a1]; premultiply = round(c * a / 255) =
// (t + (t >> 8)) >> 8 with t = c * a + 128, exact for all 8-bit c, a and equal
// to the (2ca + 255) / 510 of premult_u8; the >> 8 of both halves is their
// high bytes, which udot4 adds into the channel's 32-bit sum.  Alpha passes
// through premultiply unchanged, so its sum is the plain byte sum.  An odd
// last pixel pairs with a zero pixel, which adds 0 to every sum.
__device__ __forceinline__ void premult_window_sums4(const unsigned char* rbuf, int sbyte, int fh,
                                                     uint32_t (&sum)[4]) {
    const uint32_t* l32 = reinterpret_cast<const uint32_t*>(rbuf + sbyte);
    const u16x2 half = {128, 128};
    const u16x2 eight = {8, 8};
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    auto pair = [&](uint32_t v0, uint32_t v1) {
        // perm(hi, lo, sel): selector bytes 0-3 pick lo's bytes, 4-7 hi's, 0x0c gives 0.
        const u16x2 aa = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(v1, v0, 0x0c070c03u));
        const u16x2 r = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(v1, v0, 0x0c040c00u));
        const u16x2 g = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(v1, v0, 0x0c050c01u));
        const u16x2 b = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(v1, v0, 0x0c060c02u));
        const u16x2 tr = r * aa + half, tg = g * aa + half, tb = b * aa + half;
        const uint32_t ur = __builtin_bit_cast(uint32_t, (u16x2)(tr + (tr >> eight)));
        const uint32_t ug = __builtin_bit_cast(uint32_t, (u16x2)(tg + (tg >> eight)));
        const uint32_t ub = __builtin_bit_cast(uint32_t, (u16x2)(tb + (tb >> eight)));
        s0 = __builtin_amdgcn_udot4(ur, 0x01000100u, s0, false);
        s1 = __builtin_amdgcn_udot4(ug, 0x01000100u, s1, false);
        s2 = __builtin_amdgcn_udot4(ub, 0x01000100u, s2, false);
        s3 = __builtin_amdgcn_udot4(__builtin_bit_cast(uint32_t, aa), 0x00010001u, s3, false);
    };
    int j = 0;
    for (; j + 4 <= fh; j += 4) {
        const uint32_t v0 = l32[j], v1 = l32[j + 1], v2 = l32[j + 2], v3 = l32[j + 3];
        pair(v0, v1);
        pair(v2, v3);
    }
    for (; j + 2 <= fh; j += 2) pair(l32[j], l32[j + 1]);
    if (j < fh) pair(l32[j], 0u);
    sum[0] = s0;
    sum[1] = s1;
    sum[2] = s2;
    sum[3] = s3;
}

// window_sums for a window of exactly ND dwords (ND = ceil(wbytes / 4),
// known per job): fully unrolled, the ND + 1 dwords it spans read at once,
// only the last dword masked (last_mask keeps its valid bytes).
template <int B, int ND>
__device__ __forceinline__ void window_sums_n(const unsigned char* rbuf, int sbyte, uint32_t last_mask,
                                              uint32_t (&sum)[B]) {
    const uint32_t* l32 = reinterpret_cast<const uint32_t*>(rbuf) + (sbyte >> 2);
    const int al = sbyte & 3;
    uint32_t w[ND + 1];
#pragma unroll
    for (int d = 0; d <= ND; ++d) w[d] = l32[d];
    uint32_t odd[B];
#pragma unroll
    for (int b = 0; b < B; ++b) odd[b] = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        uint32_t v = __builtin_amdgcn_alignbyte(w[d + 1], w[d], al);
        if (d == ND - 1) v &= last_mask;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (d & 1) odd[b] = __builtin_amdgcn_udot4(v, band_mask<B>(b, d % 3), odd[b], false);
            else sum[b] = __builtin_amdgcn_udot4(v, band_mask<B>(b, d % 3), sum[b], false);
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) sum[b] += odd[b];
}

// Window sums with the window length (in dwords) dispatched once per call
// to an unrolled form; longer windows take the looped form.
template <int B>
__device__ __forceinline__ void window_sums_any(const unsigned char* rbuf, int sbyte, int wbytes, int nd,
                                                uint32_t last_mask, uint32_t (&sum)[B]) {
    switch (nd) {
#define VGPU_WN(N) \
    case N: window_sums_n<B, N>(rbuf, sbyte, last_mask, sum); return;
        VGPU_WN(1) VGPU_WN(2) VGPU_WN(3) VGPU_WN(4) VGPU_WN(5) VGPU_WN(6) VGPU_WN(7) VGPU_WN(8)
        VGPU_WN(9) VGPU_WN(10) VGPU_WN(11) VGPU_WN(12) VGPU_WN(13) VGPU_WN(14) VGPU_WN(15) VGPU_WN(16)
#undef VGPU_WN
    default: window_sums<B>(rbuf, sbyte, wbytes, sum); return;
    }
}

// A rounding ((sum + add) * mul) >> shift with the planner's per-job
// constants (sum + add and mul below 2^24, product below 2^32).
__device__ __forceinline__ uint32_t round_mul(uint32_t sum, uint32_t add, uint32_t mul, uint32_t shift) {
    return (uint32_t)__umul24(sum + add, mul) >> shift;
}

template <int B>
__device__ __forceinline__ void shrink_rows_block(const ShrinkJob& job, int bx, int by,
                                                  unsigned char* lds) {
    const InView& in = job.in;
    const OutView& out = job.out;
    const int fh = job.fh, fv = job.fv, tx = job.tx, S = job.S, m = job.m, lsl = job.lsl;
    const int RP = 256 >> lsl;  // source rows per fetch group
    const bool premult = (B == 2 || B == 4) && job.premult;
    const uint32_t rh_add = job.rh_add, rh_mul = job.rh_mul, rh_shift = job.rh_shift;
    const uint32_t rv_add = job.rv_add, rv_mul = job.rv_mul, rv_shift = job.rv_shift;
    const int ostride = job.ostride ? job.ostride : out.w;
    // The wavefront index is uniform; saying so keeps all per-wavefront
    // geometry and row addressing in scalar registers.
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int part = wave % S;
    const int oyi0 = (by * ((kThreads / 64) / S) + wave / S) * m;  // first output row of the run
    const int nout = max(0, min(m, out.h - oyi0));
    const int nsrc = S > 1 ? ((nout > 0 && part < fv) ? (fv - 1 - part) / S + 1 : 0) : nout * fv;
    const int col0 = bx * tx;
    const int ncols = min(tx, out.w - col0);
    const int ybase = (out.y + oyi0) * fv + part;  // source row of k = 0 (row k: ybase + k * S)
    const int cx0 = (out.x + col0) * fh;
    const int cx1 = cx0 + ncols * fh;
    const int lo = clampi(cx0, in.bx, in.bx + in.bw - 1);
    const int hi = clampi(cx1 - 1, in.bx, in.bx + in.bw - 1) + 1;
    const int nbytes = (hi - lo) * B;
    const unsigned char* base = static_cast<const unsigned char*>(in.ptr);
    const size_t row_bytes = (size_t)in.bw * B;
    const size_t total = row_bytes * (size_t)in.bh;
    const size_t col_off = (size_t)(lo - in.bx) * B;
    unsigned char* rbuf = lds + wave * (kSpanBytes + 32);
    const int wbytes = fh * B;
    const int slot_mask = (1 << lsl) - 1;

    // Fetch group g into p0..p3; vm holds 2 bits per chunk: 1 = full 16-byte
    // chunk, 2 = the buffer's last, partial chunk (copied bytewise at park
    // time), 0 = unused.  Unused or partial chunks reload the buffer's first
    // 16 bytes, so every load is unconditional.  One group is in flight in
    // registers while the parked one is summed (a second register set costs
    // more occupancy than it buys).
#define VGPU_CHUNK(j, g, pj, ROW, VM)                                                               \
    do {                                                                                            \
        const int k_ = (g) * RP + (ROW);                                                            \
        const int c_ = (lane + 64 * (j)) & slot_mask;                                               \
        const size_t a_ = (size_t)row_of(in, ybase + k_ * S) * row_bytes + col_off;                 \
        const size_t a16_ = a_ & ~(size_t)15;                                                       \
        const int nch_ = ((int)(a_ - a16_) + nbytes + 15) >> 4;                                     \
        const size_t off_ = a16_ + (size_t)c_ * 16;                                                 \
        const bool use_ = k_ < nsrc && c_ < nch_;                                                   \
        const bool full_ = use_ && off_ + 16 <= total;                                              \
        VM |= (uint32_t)(full_ ? 1 : (use_ ? 2 : 0)) << (2 * (j));                                  \
        pj = ld_stream16(base + (full_ ? off_ : 0));                                                \
    } while (0)
// The row of register j's chunks is wave-uniform when a slot spans at least
// one register (lsl >= 6): that branch keeps the address math scalar.
#define VGPU_FETCH(g, P0, P1, P2, P3, VM)                                               \
    do {                                                                                \
        VM = 0;                                                                         \
        if (lsl >= 6) {                                                                 \
            VGPU_CHUNK(0, g, P0, 0, VM);                                                \
            VGPU_CHUNK(1, g, P1, 1 >> (lsl - 6), VM);                                   \
            VGPU_CHUNK(2, g, P2, 2 >> (lsl - 6), VM);                                   \
            VGPU_CHUNK(3, g, P3, 3 >> (lsl - 6), VM);                                   \
        } else {                                                                        \
            VGPU_CHUNK(0, g, P0, (lane >> lsl), VM);                                    \
            VGPU_CHUNK(1, g, P1, (lane >> lsl) + (64 >> lsl), VM);                      \
            VGPU_CHUNK(2, g, P2, (lane >> lsl) + 2 * (64 >> lsl), VM);                  \
            VGPU_CHUNK(3, g, P3, (lane >> lsl) + 3 * (64 >> lsl), VM);                  \
        }                                                                               \
    } while (0)
#define VGPU_PARK1(j, g, pj, VM)                                                                    \
    do {                                                                                            \
        const uint32_t st_ = (VM >> (2 * (j))) & 3u;                                                \
        const int idx_ = lane + 64 * (j);                                                           \
        if (st_ == 1u) {                                                                            \
            *reinterpret_cast<u32x4*>(rbuf + idx_ * 16) = pj;                                       \
        } else if (st_ == 2u) {                                                                     \
            const int k_ = (g) * RP + (idx_ >> lsl);                                                \
            const size_t a_ = (size_t)row_of(in, ybase + k_ * S) * row_bytes + col_off;             \
            const size_t off_ = (a_ & ~(size_t)15) + (size_t)(idx_ & slot_mask) * 16;               \
            for (int q_ = 0; q_ < 16 && off_ + q_ < total; ++q_) rbuf[idx_ * 16 + q_] = gmem<unsigned char>(base)[off_ + q_]; \
        }                                                                                           \
    } while (0)
#define VGPU_PARK(g, P0, P1, P2, P3, VM)  \
    do {                                  \
        VGPU_PARK1(0, g, P0, VM);         \
        VGPU_PARK1(1, g, P1, VM);         \
        VGPU_PARK1(2, g, P2, VM);         \
        VGPU_PARK1(3, g, P3, VM);         \
    } while (0)

    uint32_t acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) acc[b] = 0;
    // Box sums of the parked group g's source rows.
    auto compute = [&](int g) {
        const int nr = min(RP, nsrc - g * RP);
        for (int r = 0; r < nr; ++r) {
            const int k = g * RP + r;
            if (lane < ncols) {
                const size_t a = (size_t)row_of(in, ybase + k * S) * row_bytes + col_off;
                // The slot is 16-byte aligned; the row's bytes start `shift` in.
                const unsigned char* slot = rbuf + (r << lsl) * 16;
                const int shift = (int)(a & 15);
                const int X0 = cx0 + lane * fh;
                uint32_t sum[B];
#pragma unroll
                for (int b = 0; b < B; ++b) sum[b] = 0;
                if (!premult && X0 >= lo && X0 + fh <= hi) {
                    window_sums<B>(slot, shift + (X0 - lo) * B, wbytes, sum);
                } else {
                    for (int j = 0; j < fh; ++j) {
                        const unsigned char* px = slot + shift + (clampi(X0 + j, lo, hi - 1) - lo) * B;
                        if (premult) {
                            const uint32_t al = px[B - 1];
#pragma unroll
                            for (int b = 0; b < B - 1; ++b) sum[b] += premult_u8(px[b], al);
                            sum[B - 1] += al;
                        } else {
#pragma unroll
                            for (int b = 0; b < B; ++b) sum[b] += px[b];
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] += round_mul(sum[b], rh_add, rh_mul, rh_shift);
            }
            if (S == 1 && (k + 1) % fv == 0) {
                // Output row oyi0 + k / fv is complete.
                if (lane < ncols) {
                    unsigned char* o = static_cast<unsigned char*>(out.ptr) +
                                       ((size_t)(oyi0 + k / fv) * ostride + col0 + lane) * B;
                    uint32_t v[B];
#pragma unroll
                    for (int b = 0; b < B; ++b) v[b] = round_mul(acc[b], rv_add, rv_mul, rv_shift);
                    if ((B == 2 || B == 4) && job.unpremult) {
#pragma unroll
                        for (int b = 0; b < B - 1; ++b) v[b] = unpremult_u8(v[b], v[B - 1]);
                    }
#pragma unroll
                    for (int b = 0; b < B; ++b) o[b] = (unsigned char)v[b];
                }
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] = 0;
            }
        }
    };
    const int ng = (nsrc + RP - 1) / RP;
    u32x4 p0 = {0u, 0u, 0u, 0u}, p1 = p0, p2 = p0, p3 = p0;
    uint32_t vma = 0;
    if (ng > 0) VGPU_FETCH(0, p0, p1, p2, p3, vma);
    for (int g = 0; g < ng; ++g) {
        VGPU_PARK(g, p0, p1, p2, p3, vma);
        wave_lds_sync();
        if (g + 1 < ng) VGPU_FETCH(g + 1, p0, p1, p2, p3, vma);
        compute(g);
        wave_lds_sync();
    }
#undef VGPU_PARK
#undef VGPU_PARK1
#undef VGPU_FETCH
#undef VGPU_CHUNK
    if (S > 1) {
        // Partial sums of the S wavefronts of each output row.  Every
        // wavefront of the block reaches both barriers (S is per job).
        __syncthreads();
        uint32_t* part_sums = reinterpret_cast<uint32_t*>(lds);
#pragma unroll
        for (int b = 0; b < B; ++b) part_sums[(wave * 64 + lane) * B + b] = acc[b];
        __syncthreads();
        if (part == 0 && nout > 0 && lane < ncols) {
            for (int q = 1; q < S; ++q) {
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] += part_sums[((wave + q) * 64 + lane) * B + b];
            }
            unsigned char* o = static_cast<unsigned char*>(out.ptr) + ((size_t)oyi0 * ostride + col0 + lane) * B;
            uint32_t v[B];
#pragma unroll
            for (int b = 0; b < B; ++b) v[b] = round_mul(acc[b], rv_add, rv_mul, rv_shift);
            if ((B == 2 || B == 4) && job.unpremult) {
#pragma unroll
                for (int b = 0; b < B - 1; ++b) v[b] = unpremult_u8(v[b], v[B - 1]);
            }
#pragma unroll
            for (int b = 0; b < B; ++b) o[b] = (unsigned char)v[b];
        }
    }
}

// Window sums of a lean-form window of nd dwords (ND - 3 <= nd <= ND, ND a
// multiple of 4 fixed per kernel path): the first ND - 4 dwords are whole,
// the last four are masked by tail[0..3] (per wavefront, computed once), so
// no per-row dispatch on the window length is needed.  Reads ND + 1 dwords
// from the 4-byte-aligned position below sbyte (the LDS group buffer is
// padded for the read past the last window).
template <int B, int ND>
__device__ __forceinline__ void window_sums_masked(const unsigned char* rbuf, int sbyte, const uint32_t (&tail)[4],
                                                   uint32_t (&sum)[B]) {
    const uint32_t* l32 = reinterpret_cast<const uint32_t*>(rbuf) + (sbyte >> 2);
    const int al = sbyte & 3;
    uint32_t w[ND + 1];
#pragma unroll
    for (int d = 0; d <= ND; ++d) w[d] = l32[d];
    uint32_t odd[B];
#pragma unroll
    for (int b = 0; b < B; ++b) odd[b] = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        uint32_t v = __builtin_amdgcn_alignbyte(w[d + 1], w[d], al);
        if (d >= ND - 4) v &= tail[d - (ND - 4)];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (d & 1) odd[b] = __builtin_amdgcn_udot4(v, band_mask<B>(b, d % 3), odd[b], false);
            else sum[b] = __builtin_amdgcn_udot4(v, band_mask<B>(b, d % 3), sum[b], false);
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) sum[b] += odd[b];
}

// __umul24 returns the low 32 bits of the 48-bit product as an int: shift it
// as unsigned (products reach 2^32 - 1).

// Lean form of the same kernel for spans of at least 128 chunks (lsl >= 7):
// a fetch group is then one or two source rows and each 64-lane register
// holds part of one of them (register j: row j >> (lsl - 6)), so a row's
// address, chunk limit and 16-byte shift are wave-uniform.  The per-row
// scalar work is kept small, since the scalar unit is shared by the CU's
// four SIMDs: 32-bit buffer offsets (the planner sends images of 4 GiB and
// more to the general form), the window length fixed per kernel path (ND),
// both roundings as one multiply-shift with per-job constants.  Loads take
// the row's 16-byte-aligned start plus a per-lane 32-bit offset; lanes past
// the row's whole chunks re-read its last one.  Parking is four
// unconditional 16-byte LDS stores; a chunk cut by the end of the buffer
// (scalar flag) is then copied bytewise.  Output rows complete on a counter.
template <int B, bool TWO, int ND>
__device__ __forceinline__ void shrink_rows_lean_nd(const ShrinkJob& job, int bx, int by, unsigned char* lds) {
    const InView& in = job.in;
    const OutView& out = job.out;
    const int fh = job.fh, fv = job.fv, tx = job.tx, S = job.S, m = job.m, lsl = job.lsl;
    constexpr bool two = TWO;  // lsl == 7: two source rows per fetch group (registers 0-1, 2-3)
    constexpr int RP = two ? 2 : 1;
    const bool premult = (B == 2 || B == 4) && job.premult;
    const uint32_t rh_add = job.rh_add, rh_mul = job.rh_mul, rh_shift = job.rh_shift;
    const uint32_t rv_add = job.rv_add, rv_mul = job.rv_mul, rv_shift = job.rv_shift;
    const int ostride = job.ostride ? job.ostride : out.w;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int part = wave % S;
    const int oyi0 = (by * ((kThreads / 64) / S) + wave / S) * m;
    const int nout = max(0, min(m, out.h - oyi0));
    const int nsrc = S > 1 ? ((nout > 0 && part < fv) ? (fv - 1 - part) / S + 1 : 0) : nout * fv;
    const int col0 = bx * tx;
    const int ncols = min(tx, out.w - col0);
    const int ybase = (out.y + oyi0) * fv + part;
    const int cx0 = (out.x + col0) * fh;
    const int cx1 = cx0 + ncols * fh;
    const int lo = clampi(cx0, in.bx, in.bx + in.bw - 1);
    const int hi = clampi(cx1 - 1, in.bx, in.bx + in.bw - 1) + 1;
    const uint32_t nbytes = (uint32_t)(hi - lo) * B;
    const unsigned char* base = static_cast<const unsigned char*>(in.ptr);
    const uint32_t row_bytes = (uint32_t)in.bw * B;
    const uint32_t total = row_bytes * (uint32_t)in.bh;  // < 2^32 (planner)
    const uint32_t col_off = (uint32_t)(lo - in.bx) * B;
    unsigned char* rbuf = lds + wave * (kSpanBytes + 32);
    const int wbytes = fh * B;
    const int X0 = cx0 + lane * fh;
    // Windows that run past the buffer's right end read its last pixel
    // repeated (the clamped read): park() writes those copies after the
    // span, so such windows take the fast paths too.  Windows wholly left of
    // the buffer (the Z^2 halo columns left of a source) are fh copies of its
    // first pixel.  Only a window cut by the left end takes the per-pixel
    // loop.
    const int over = max(0, cx1 - hi);  // pixels past the end (wave-uniform)
    const bool fast = !premult && lane < ncols && X0 >= lo;
    // 4-band premultiplied windows inside the span: packed 16-bit path.
    const bool fast_pm4 = B == 4 && premult && lane < ncols && X0 >= lo;
    const bool left_out = lane < ncols && X0 + fh <= lo;
    const int lane_byte = (X0 - lo) * B;
    // Masks of the window's last four dwords ND - 4 .. ND - 1 (ND = 0: the
    // looped window sums of windows longer than 16 dwords).
    uint32_t tail[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int valid = wbytes - 4 * (ND - 4 + t);
        tail[t] = valid >= 4 ? 0xffffffffu : (valid <= 0 ? 0u : ((1u << (8 * valid)) - 1u));
    }

    // Row k of the wavefront's run: a16 = its span's 16-byte-aligned buffer
    // offset, lb = where its chunks are loaded from (a16, or 0 -- the
    // buffer's first chunk -- for a row past the run or with no whole chunk),
    // lim = chunks to load (>= 1), cut = the chunk cut by the buffer end or
    // -1, shift = the span's offset past a16.
    auto row_fetch = [&](int k, uint32_t& a16, uint32_t& lb, int& lim, int& cut, int& shift) {
        a16 = 0;
        lb = 0;
        lim = 1;
        cut = -1;
        shift = 0;
        if (k < nsrc) {
            const int y = row_of(in, ybase + k * S);
            const uint32_t a = (uint32_t)y * row_bytes + col_off;
            a16 = a & ~15u;
            shift = (int)(a & 15u);
            const int nch = (int)((a & 15u) + nbytes + 15u) >> 4;
            int w = nch;
            if (a16 + (uint32_t)nch * 16u > total) {
                w = (int)((total - a16) >> 4);
                cut = w;
            }
            if (w > 0) {
                lb = a16;
                lim = w;
            }
        }
    };

    // A fetch group in flight: its four 16-byte registers and, per row, the
    // span offset, cut chunk and shift.
    struct Group {
        u32x4 p0, p1, p2, p3;
        uint32_t a0, a1;
        int c0, c1, s0, s1;
    };
    auto fetch = [&](int g) {
        Group G;
        const int k0 = g * RP;
        uint32_t b0 = 0, b1 = 0;
        int l0 = 1, l1 = 1;
        G.a1 = 0;
        G.c1 = -1;
        G.s1 = 0;
        row_fetch(k0, G.a0, b0, l0, G.c0, G.s0);
        if (two) row_fetch(k0 + 1, G.a1, b1, l1, G.c1, G.s1);
        if (two) {
            G.p0 = ld_stream16(base + b0 + (uint32_t)min(lane, l0 - 1) * 16u);
            G.p1 = ld_stream16(base + b0 + (uint32_t)min(lane + 64, l0 - 1) * 16u);
            G.p2 = ld_stream16(base + b1 + (uint32_t)min(lane, l1 - 1) * 16u);
            G.p3 = ld_stream16(base + b1 + (uint32_t)min(lane + 64, l1 - 1) * 16u);
        } else {
            G.p0 = ld_stream16(base + b0 + (uint32_t)min(lane, l0 - 1) * 16u);
            G.p1 = ld_stream16(base + b0 + (uint32_t)min(lane + 64, l0 - 1) * 16u);
            G.p2 = ld_stream16(base + b0 + (uint32_t)min(lane + 128, l0 - 1) * 16u);
            G.p3 = ld_stream16(base + b0 + (uint32_t)min(lane + 192, l0 - 1) * 16u);
        }
        return G;
    };
    // The cut chunk c of a row whose slot starts at LDS chunk slot0: its
    // bytes before the end of the buffer.
    auto fix_cut = [&](uint32_t a16, int c, int slot0) {
        if (lane != (c & 63)) return;  // the lane that stored that slot
        const uint32_t off = a16 + (uint32_t)c * 16u;
        unsigned char* d = rbuf + (slot0 + c) * 16;
        for (int q = 0; q < 16; ++q) d[q] = off + q < total ? gmem<unsigned char>(base)[off + q] : 0;
    };
    // The over * B bytes after a parked row's span: its last pixel again.
    auto replicate = [&](int shift, int slot0) {
        unsigned char* row = rbuf + slot0 * 16 + shift;
        const int e0 = (hi - lo) * B;
        for (int k = lane; k < over * B; k += 64) row[e0 + k] = row[e0 - B + k % B];
    };
    auto park = [&](const Group& G) {
        u32x4* dst = reinterpret_cast<u32x4*>(rbuf) + lane;
        dst[0] = G.p0;
        dst[64] = G.p1;
        dst[128] = G.p2;
        dst[192] = G.p3;
        if (G.c0 >= 0 || (two && G.c1 >= 0)) {
            // Lanes whose chunk index is the cut one rewrite its slot (the
            // same lane stored it, so LDS order holds).
            if (G.c0 >= 0) fix_cut(G.a0, G.c0, 0);
            if (two && G.c1 >= 0) fix_cut(G.a1, G.c1, 128);
        }
        if (over > 0) {
            wave_lds_sync();  // the span's last pixel is in LDS
            replicate(G.s0, 0);
            if (two) replicate(G.s1, 128);
        }
    };

    uint32_t acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) acc[b] = 0;
    int krow = 0, orow = 0;  // source rows into the current output row; output rows done
    auto store = [&](int oy) {
        unsigned char* o = static_cast<unsigned char*>(out.ptr) + ((size_t)oy * ostride + col0 + lane) * B;
        uint32_t v[B];
#pragma unroll
        for (int b = 0; b < B; ++b) v[b] = round_mul(acc[b], rv_add, rv_mul, rv_shift);
        if ((B == 2 || B == 4) && job.unpremult) {
#pragma unroll
            for (int b = 0; b < B - 1; ++b) v[b] = unpremult_u8(v[b], v[B - 1]);
        }
#pragma unroll
        for (int b = 0; b < B; ++b) o[b] = (unsigned char)v[b];
    };
    auto row_sum = [&](int r, int shift) {
        const unsigned char* slot = rbuf + (r << lsl) * 16;
        uint32_t sum[B];
#pragma unroll
        for (int b = 0; b < B; ++b) sum[b] = 0;
        if (fast) {
            if constexpr (ND == 0) window_sums<B>(slot, shift + lane_byte, wbytes, sum);
            else window_sums_masked<B, ND>(slot, shift + lane_byte, tail, sum);
        } else if (fast_pm4) {
            if constexpr (B == 4) premult_window_sums4(slot, shift + lane_byte, fh, sum);
        } else if (left_out) {
            const unsigned char* px = slot + shift;
            if (premult) {
                const uint32_t al = px[B - 1];
#pragma unroll
                for (int b = 0; b < B - 1; ++b) sum[b] = premult_u8(px[b], al) * (uint32_t)fh;
                sum[B - 1] = al * (uint32_t)fh;
            } else {
#pragma unroll
                for (int b = 0; b < B; ++b) sum[b] = px[b] * (uint32_t)fh;
            }
        } else if (lane < ncols) {
            for (int j = 0; j < fh; ++j) {
                const unsigned char* px = slot + shift + (clampi(X0 + j, lo, hi - 1) - lo) * B;
                if (premult) {
                    const uint32_t al = px[B - 1];
#pragma unroll
                    for (int b = 0; b < B - 1; ++b) sum[b] += premult_u8(px[b], al);
                    sum[B - 1] += al;
                } else {
#pragma unroll
                    for (int b = 0; b < B; ++b) sum[b] += px[b];
                }
            }
        }
#pragma unroll
        for (int b = 0; b < B; ++b) acc[b] += round_mul(sum[b], rh_add, rh_mul, rh_shift);
        if (S == 1 && ++krow == fv) {
            if (lane < ncols) store(oyi0 + orow);
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = 0;
            krow = 0;
            ++orow;
        }
    };
    const int ng = (nsrc + RP - 1) / RP;
    Group GA;
    if (ng > 0) GA = fetch(0);
    for (int g = 0; g < ng; ++g) {
        park(GA);
        wave_lds_sync();
        const int a0 = GA.s0, a1 = GA.s1;
        if (g + 1 < ng) GA = fetch(g + 1);
        row_sum(0, a0);
        if (two && g * RP + 1 < nsrc) row_sum(1, a1);
        wave_lds_sync();
    }
    if (S > 1) {
        // Partial sums of the S wavefronts of each output row; every
        // wavefront of the block reaches both barriers (S is per job).
        __syncthreads();
        uint32_t* part_sums = reinterpret_cast<uint32_t*>(lds);
#pragma unroll
        for (int b = 0; b < B; ++b) part_sums[(wave * 64 + lane) * B + b] = acc[b];
        __syncthreads();
        if (part == 0 && nout > 0 && lane < ncols) {
            for (int q = 1; q < S; ++q) {
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] += part_sums[((wave + q) * 64 + lane) * B + b];
            }
            store(oyi0);
        }
    }
}

// The lean form with its window length fixed once per wavefront: two-row
// groups (lsl = 7, 64 columns) have windows of 4..8 dwords, one-row groups
// (lsl = 8) of 8..16 dwords with 64 columns and more with fewer columns
// (wnd_t = 0: looped sums); 4-band premultiplied windows take their own path.
template <int B, bool TWO>
__device__ __forceinline__ void shrink_rows_lean(const ShrinkJob& job, int bx, int by, unsigned char* lds) {
    if constexpr (TWO) {
        if (job.wnd_t == 4) shrink_rows_lean_nd<B, true, 4>(job, bx, by, lds);
        else if (job.wnd_t == 8) shrink_rows_lean_nd<B, true, 8>(job, bx, by, lds);
        else shrink_rows_lean_nd<B, true, 0>(job, bx, by, lds);
    } else if constexpr (B == 4) {
        if (job.premult || job.wnd_t == 0) shrink_rows_lean_nd<B, false, 0>(job, bx, by, lds);
        else if (job.wnd_t == 16) shrink_rows_lean_nd<B, false, 16>(job, bx, by, lds);
        else if (job.wnd_t == 12) shrink_rows_lean_nd<B, false, 12>(job, bx, by, lds);
        else shrink_rows_lean_nd<B, false, 8>(job, bx, by, lds);
    } else {
        if (job.wnd_t == 16) shrink_rows_lean_nd<B, false, 16>(job, bx, by, lds);
        else if (job.wnd_t == 12) shrink_rows_lean_nd<B, false, 12>(job, bx, by, lds);
        else if (job.wnd_t == 8) shrink_rows_lean_nd<B, false, 8>(job, bx, by, lds);
        else shrink_rows_lean_nd<B, false, 0>(job, bx, by, lds);
    }
}

// One kernel per band count and form (FORM 0: general; 1: lean, one source
// row per fetch group (lsl = 8); 2: lean, two rows (lsl = 7)), so each keeps
// only its own registers.
template <int B, int FORM>
__device__ __forceinline__ void shrink_block(const ShrinkJob& job, int bx, int by, unsigned char* lds) {
    if constexpr (FORM == 1) shrink_rows_lean<B, false>(job, bx, by, lds);
    else if constexpr (FORM == 2) shrink_rows_lean<B, true>(job, bx, by, lds);
    else shrink_rows_block<B>(job, bx, by, lds);
}

template <int B, int FORM>
__global__ __launch_bounds__(kThreads) void k_shrink_job(ShrinkJob job) {
    extern __shared__ __align__(16) unsigned char lds[];
    shrink_block<B, FORM>(job, blockIdx.x, blockIdx.y, lds);
}

// Persistent form: the grid walks a task list of (job, bx, by) over many
// images (all of one variant); jobs reading the caller's source or writing
// the caller's output take those pointers from per-run slot tables.
template <int B, int FORM>
__global__ __launch_bounds__(kThreads) void k_shrink_batch(const ShrinkJob* __restrict__ jobs,
                                                           const BatchTask* __restrict__ tasks, int ntasks,
                                                           const void* const* __restrict__ srcs,
                                                           void* const* __restrict__ dsts) {
    extern __shared__ __align__(16) unsigned char lds[];
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const BatchTask task = tasks[t];
        ShrinkJob job = jobs[task.job];
        if (job.src_slot >= 0) job.in.ptr = srcs[job.src_slot];
        if (job.dst_slot >= 0) job.out.ptr = dsts[job.dst_slot];
        shrink_block<B, FORM>(job, task.bx, task.by, lds);
        // With one wavefront per output row (S == 1, every batch job) each
        // wavefront parks only in its own LDS span, so the wavefronts of a
        // block go on to the next task without waiting for each other; split
        // rows share LDS and wait.
        if (job.S > 1) __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Generic single-axis kernels: one thread per output sample.
// ---------------------------------------------------------------------------

// ShrinkH/ShrinkV for every format other than u8 (and u8 without the fused
// kernel): f64 sum * (1/f) then ShrinkSample::from_f64_clamped
// (shrinkh.rs:44-128,271-300; shrinkv.rs:168-187); u8 uses the integer
// roundings above.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_shrink_generic(InView in, OutView out, int B, int f,
                                                             int vertical, uint32_t mult) {
    const size_t n = (size_t)out.w * out.h * B;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const int b = (int)(i % B);
        const size_t px = i / B;
        const int x = out.x + (int)(px % out.w), y = out.y + (int)(px / out.w);
        T* o = static_cast<T*>(out.ptr);
        if constexpr (sizeof(T) == 1) {
            uint32_t sum = 0;
            for (int j = 0; j < f; ++j) {
                const int row = row_of(in, vertical ? y * f + j : y);
                const int col = col_of(in, vertical ? x : x * f + j);
                sum += load_at<uint8_t>(in, row, col, B, b);
            }
            o[i] = (T)(vertical ? shrink_round_v(sum, f, mult) : shrink_round_h(sum, f, B, mult));
        } else {
            double sum = 0.0;
            for (int j = 0; j < f; ++j) {
                const int row = row_of(in, vertical ? y * f + j : y);
                const int col = col_of(in, vertical ? x : x * f + j);
                sum += (double)load_at<T>(in, row, col, B, b);
            }
            o[i] = shrink_from_f64<T>(sum * (1.0 / (double)f));
        }
    }
}

// The PNG box pre-shrinks, one thread per output pixel of the rectangle
// [x0, x0 + rw) x [y0, y0 + rh) of a dense dw-wide output: the window
// [d*f, min(d*f + f, src)) per axis, divided by its own pixel count n
// (software_box_shrink_generic, decoder_source/shrink.rs:47-117, and its
// partial-window form in decode_png_with_box_shrink_u8, png/decode_full.rs:
// 320-452).  u8: (sum + n/2) / n in integers; every other format: the f64
// sum in row order, divided, ShrinkSample::from_f64_clamped.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_box_generic(const T* __restrict__ src, int sw, int sh, int B, int f,
                                                          T* __restrict__ dst, int dw, int x0, int y0, int rw,
                                                          int rh) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= (int64_t)rw * rh) return;
    const int dx = x0 + (int)(i % rw), dy = y0 + (int)(i / rw);
    const int64_t sx0 = (int64_t)dx * f, sy0 = (int64_t)dy * f;
    const int64_t sx1 = min(sx0 + f, (int64_t)sw), sy1 = min(sy0 + f, (int64_t)sh);
    const uint64_t n = max<int64_t>((sy1 - sy0) * (sx1 - sx0), 1);
    T* o = dst + ((size_t)dy * dw + dx) * B;
    for (int b = 0; b < B; ++b) {
        if constexpr (sizeof(T) == 1) {
            uint64_t sum = 0;
            for (int64_t y = sy0; y < sy1; ++y)
                for (int64_t x = sx0; x < sx1; ++x) sum += src[((size_t)y * sw + x) * B + b];
            o[b] = (T)((sum + n / 2) / n);
        } else {
            double sum = 0.0;
            for (int64_t y = sy0; y < sy1; ++y)
                for (int64_t x = sx0; x < sx1; ++x) sum += (double)src[((size_t)y * sw + x) * B + b];
            o[b] = shrink_from_f64<T>(sum / (double)n);
        }
    }
}

// One ReduceH output sample at (row, output column xo, band b).
template <typename T>
__device__ __forceinline__ T reduce_h_sample(const InView& in, int row, int B, int b, int start,
                                             int ph, const DevPlan& p) {
    const int T_ = p.taps;
    if constexpr (std::is_same<T, double>::value) {
        // reduce_h_scalar (reduce_simd.rs:47-90): the FMA chain in tap order.
        const double* w = p.wf64 + (size_t)ph * T_;
        double acc = 0.0;
        for (int t = 0; t < T_; ++t) acc = fma(load_at<double>(in, row, col_of(in, start + t), B, b), w[t], acc);
        return acc;
    } else if constexpr (std::is_same<T, float>::value) {
        const double* w = p.wf64 + (size_t)ph * T_;
        double acc = 0.0;
        int t = 0;
        if (B == 1) {
            for (; t + 4 <= T_; t += 4) {
                double s = 0.0;
                for (int l = 0; l < 4; ++l) {
                    const double prod = (double)load_at<float>(in, row, col_of(in, start + t + l), 1, 0) * w[t + l];
                    s = s + prod;
                }
                acc = acc + s;
            }
        }
        for (; t < T_; ++t) acc = fma((double)load_at<float>(in, row, col_of(in, start + t), B, b), w[t], acc);
        return (T)(float)acc;
    } else {
        const int16_t* w = p.wi16 + (size_t)ph * T_;
        int64_t acc = 0;
        for (int t = 0; t < T_; ++t)
            acc += (int64_t)w[t] * (int64_t)load_at<T>(in, row, col_of(in, start + t), B, b);
        return fixed_from_acc<T, int64_t>(acc);
    }
}

// reduce_h_sample's arithmetic over an accessor get(t) -> the sample under
// tap t (already clamped), for the staged fused kernels.
template <typename T, typename Get>
__device__ __forceinline__ T reduce_h_from(Get get, int B, int ph, const DevPlan& p) {
    const int T_ = p.taps;
    if constexpr (std::is_same<T, float>::value) {
        const double* w = p.wf64 + (size_t)ph * T_;
        double acc = 0.0;
        int t = 0;
        if (B == 1) {
            for (; t + 4 <= T_; t += 4) {
                double s = 0.0;
                for (int l = 0; l < 4; ++l) {
                    const double prod = (double)get(t + l) * w[t + l];
                    s = s + prod;
                }
                acc = acc + s;
            }
        }
        for (; t < T_; ++t) acc = fma((double)get(t), w[t], acc);
        return (T)(float)acc;
    } else {
        static_assert(std::is_same<T, double>::value, "integer H passes sum inline");
        const double* w = p.wf64 + (size_t)ph * T_;
        double acc = 0.0;
        for (int t = 0; t < T_; ++t) acc = fma(get(t), w[t], acc);
        return acc;
    }
}

// One ReduceV output sample from a column accessor get(row) -> T.
template <typename T, typename Get>
__device__ __forceinline__ T reduce_v_sample(Get get, int start, int ph, const DevPlan& p, bool vec) {
    const int T_ = p.taps;
    if constexpr (std::is_same<T, double>::value) {
        // reduce_v_scalar (reduce_simd.rs:93-144): the FMA chain in tap order.
        const double* w = p.wf64 + (size_t)ph * T_;
        double acc = 0.0;
        for (int t = 0; t < T_; ++t) acc = fma(get(start + t), w[t], acc);
        return acc;
    } else if constexpr (std::is_same<T, float>::value) {
        const double* w = p.wf64 + (size_t)ph * T_;
        double acc = 0.0;
        if (vec) {
            for (int t = 0; t < T_; ++t) {
                const double prod = (double)get(start + t) * w[t];
                acc = acc + prod;
            }
        } else {
            for (int t = 0; t < T_; ++t) acc = fma((double)get(start + t), w[t], acc);
        }
        return (T)(float)acc;
    } else {
        const int16_t* w = p.wi16 + (size_t)ph * T_;
        int64_t acc = 0;
        for (int t = 0; t < T_; ++t) acc += (int64_t)w[t] * (int64_t)get(start + t);
        return fixed_from_acc<T, int64_t>(acc);
    }
}


// One ReduceOp pass sample (reduce.rs:104-166): acc = fma(x_t, w_t, acc) in
// tap order over the f64 phase row, from acc = 0.
template <typename T, typename Get>
__device__ __forceinline__ T reduce_op_sample(Get get, int start, int ph, const DevPlan& p) {
    const double* w = p.wf64 + (size_t)ph * p.taps;
    double acc = 0.0;
    for (int t = 0; t < p.taps; ++t) acc = fma((double)get(start + t), w[t], acc);
    return op_from_f64<T>(acc);
}

// ReduceH (reduceh.rs:101-163; reduce_simd.rs:47-90,2463-2693): fixed-point
// i16 taps for integer formats -> from_fixed_i64; f64 the scalar FMA chain;
// f32 carries FP64:
// bands == 1 follows the AVX2 4-lane grouping (products summed per group, then
// an FMA tail), other band counts the scalar FMA chain.
// OP = true: ReduceOp's horizontal pass instead (reduce.rs:104-133).
template <typename T, bool OP>
__global__ __launch_bounds__(kThreads) void k_reduce_h(InView in, OutView out, int B, DevPlan p) {
    const size_t n = (size_t)out.w * out.h * B;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const int b = (int)(i % B);
        const size_t px = i / B;
        const int xo = (int)(px % out.w);
        const int row = row_of(in, out.y + (int)(px / out.w));
        if constexpr (OP) {
            auto get = [&](int x) -> T { return load_at<T>(in, row, col_of(in, x), B, b); };
            static_cast<T*>(out.ptr)[i] = reduce_op_sample<T>(get, p.start[xo], p.phase[xo], p);
        } else {
            static_cast<T*>(out.ptr)[i] = reduce_h_sample<T>(in, row, B, b, p.start[xo], p.phase[xo], p);
        }
    }
}

// ReduceV (reducev.rs:100-162; reduce_simd.rs:2700-2888): direct per-row plan
// (x86 AVX2 recomputes the source position per row); f32 rows use separate
// multiply/add for the first (row_len / 4) * 4 samples and an FMA chain for
// the tail, as the AVX2 kernel does.
// OP = true: ReduceOp's vertical pass instead (reduce.rs:136-166).
template <typename T, bool OP>
__global__ __launch_bounds__(kThreads) void k_reduce_v(InView in, OutView out, int B, DevPlan p) {
    const int row_len = out.w * B;
    const int vec_len = row_len / 4 * 4;
    const size_t n = (size_t)row_len * out.h;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const int s = (int)(i % row_len);
        const int yo = (int)(i / row_len);
        const int b = s % B;
        const int col = col_of(in, out.x + s / B);
        auto get = [&](int y) -> T { return load_at<T>(in, row_of(in, y), col, B, b); };
        if constexpr (OP)
            static_cast<T*>(out.ptr)[i] = reduce_op_sample<T>(get, p.start[yo], p.phase[yo], p);
        else
            static_cast<T*>(out.ptr)[i] = reduce_v_sample<T>(get, p.start[yo], p.phase[yo], p, s < vec_len);
    }
}

// Reference-order f32 ReduceH of RB staged rows at once (rows[k] at the
// thread's first tap): per row and band the f64 chain of reduce_h_from --
// for one band the AVX2 4-lane grouping (each group's products summed, then
// added), else the FMA chain -- with the RB x B chains interleaved.
template <int B, int RB>
__device__ __forceinline__ void reduce_h_f32_rows(const float* const (&rows)[RB], int ph, const DevPlan& p,
                                                  double (&acc)[RB][B]) {
    const int T_ = p.taps;
    const double* w = p.wf64 + (size_t)ph * T_;
#pragma unroll
    for (int k = 0; k < RB; ++k)
#pragma unroll
        for (int b = 0; b < B; ++b) acc[k][b] = 0.0;
    int t = 0;
    if constexpr (B == 1) {
        for (; t + 4 <= T_; t += 4) {
            const double w0 = w[t], w1 = w[t + 1], w2 = w[t + 2], w3 = w[t + 3];
#pragma unroll
            for (int k = 0; k < RB; ++k) {
                double g = 0.0;
                g = g + (double)rows[k][t] * w0;
                g = g + (double)rows[k][t + 1] * w1;
                g = g + (double)rows[k][t + 2] * w2;
                g = g + (double)rows[k][t + 3] * w3;
                acc[k][0] = acc[k][0] + g;
            }
        }
    }
    for (; t < T_; ++t) {
        const double wt = w[t];
#pragma unroll
        for (int k = 0; k < RB; ++k)
#pragma unroll
            for (int b = 0; b < B; ++b) acc[k][b] = fma((double)rows[k][t * B + b], wt, acc[k][b]);
    }
}

// Reference-order f32 ReduceV of RB output rows at once from the H rows in
// LDS (col = the thread's first sample, rows `pitch` floats apart): per row
// its own start and phase; samples before vec_len (a multiple of 4) take
// separate multiply and add, the rest the FMA chain (reduce_v_sample).
template <int B, int RB>
__device__ __forceinline__ void reduce_v_f32_rows(const float* col, int pitch, const int (&vs)[RB],
                                                  const int (&vph)[RB], const DevPlan& p, int vclip0, int vclip1,
                                                  int ys0, int nr, int s0, int vec_len, double (&acc)[RB][B]) {
    const int T_ = p.taps;
    const double* w[RB];
#pragma unroll
    for (int k = 0; k < RB; ++k) {
        w[k] = p.wf64 + (size_t)vph[k] * T_;
#pragma unroll
        for (int b = 0; b < B; ++b) acc[k][b] = 0.0;
    }
    for (int t = 0; t < T_; ++t) {
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            const int yy = clampi(clampi(vs[k] + t, vclip0, vclip1) - ys0, 0, nr - 1);
            const float* hp = col + yy * pitch;
            const double wt = w[k][t];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const double xv = (double)hp[b];
                if (s0 + b < vec_len) acc[k][b] = acc[k][b] + xv * wt;
                else acc[k][b] = fma(xv, wt, acc[k][b]);
            }
        }
    }
}

// Fixed-point accumulator: int32 when the host has proven
// max_phase(sum |w|) * max_sample < 2^31, else int64.
template <typename T, typename Acc>
__device__ __forceinline__ T fixed_round(Acc acc) {
    static_assert(sizeof(T) <= 2, "the fused reduce takes u8 and u16 fixed-point sums");
    return fixed_from_acc<T, Acc>(acc);
}

// Fused ReduceH -> ReduceV for B <= 4 bands.  Block = TX output columns x TY
// output rows; thread = (column x, row lane).  Pass 1 evaluates the ReduceH
// outputs the tile's ReduceV taps need (rows start_v[first] .. start_v[last] +
// taps_v - 1, on Z^2 exactly as the unfused ReduceH) into LDS in the
// intermediate's sample type (so the reference's rounding between passes is
// kept); pass 2 applies ReduceV from LDS.
template <typename T, int B, typename Acc>
__device__ __forceinline__ void reduce_hv_block(const ReduceJob& job, int bx, int by,
                                                unsigned char* smem) {
    const InView& in = job.in;
    const OutView& out = job.out;
    const DevPlan& ph = job.ph;
    const DevPlan& pv = job.pv;
    const int TX = job.tx, TY = job.ty, nr_max = job.nr_max;
    const int vclip0 = job.vclip0, vclip1 = job.vclip1;
    T* hbuf = reinterpret_cast<T*>(smem);
    const int tx0 = bx * TX, ty0 = by * TY;
    const int ncols = min(TX, out.w - tx0), nrows = min(TY, out.h - ty0);
    const int ys0 = pv.start[ty0];
    const int nr = min(nr_max, pv.start[ty0 + nrows - 1] + pv.taps - ys0);
    const int x = threadIdx.x % TX, rl = threadIdx.x / TX, nrl = kThreads / TX;
    const int pitch = TX * B;
    // Pass 1 (f32) in chunks of job.hr source rows: the chunk's span
    // (absolute columns c0 .. c0 + spx - 1, clamped reads) staged in LDS,
    // then every thread's H samples of those rows from it.
    T* stage = reinterpret_cast<T*>(smem + reduce_hbuf_bytes(job, sizeof(T)));
    const int c0 = ph.start[tx0];
    const int spx = ph.start[tx0 + ncols - 1] + ph.taps - c0;
    const int srow = spx * B;
    const int xo = tx0 + min(x, ncols - 1);
    const int hoff = (ph.start[xo] - c0) * B;
    const int hph = ph.phase[xo];
    const int hstep = max(job.hr, 1);  // rows per staged chunk (>= 1 on every staged path)
    for (int rc = 0; sizeof(T) == 4 && rc < nr; rc += hstep) {
        const int hr = min(hstep, nr - rc);
        stage_h_rows<T, B>(in, ys0 + rc, hr, c0, spx, stage);
        __syncthreads();
        if constexpr (sizeof(T) == 4) {
            // job.rb rows per thread at a time (a chunk holds rb * nrl
            // rows): rb x B independent FP64 chains interleaved tap by tap,
            // each in the reference's order.
            auto h_rows = [&](auto rbc) {
                constexpr int RB = decltype(rbc)::value;
                for (int r = rl; x < ncols && r < hr; r += RB * nrl) {
                    const T* rows[RB];
#pragma unroll
                    for (int k = 0; k < RB; ++k) rows[k] = stage + min(r + k * nrl, hr - 1) * srow + hoff;
                    double acc[RB][B];
                    reduce_h_f32_rows<B, RB>(rows, hph, ph, acc);
#pragma unroll
                    for (int k = 0; k < RB; ++k) {
                        if (r + k * nrl >= hr) break;
                        T* dst = hbuf + (rc + r + k * nrl) * pitch + x * B;
#pragma unroll
                        for (int b = 0; b < B; ++b) dst[b] = (T)(float)acc[k][b];
                    }
                }
            };
            if (job.rb >= 4) h_rows(std::integral_constant<int, 4>());
            else if (job.rb == 2) h_rows(std::integral_constant<int, 2>());
            else h_rows(std::integral_constant<int, 1>());
        }
        __syncthreads();
    }
    if constexpr (sizeof(T) != 4) {
        // Integer samples read the clamped source directly (job.hr == 0):
        // the fixed-point sums are cheap, and a staging round trip would
        // lengthen the small, latency-bound tiles these jobs run.
        const int16_t* w = ph.wi16 + (size_t)hph * ph.taps;
        const int hs = ph.start[xo];
        for (int r = rl; x < ncols && r < nr; r += nrl) {
            const T* src = static_cast<const T*>(in.ptr) + (size_t)row_of(in, ys0 + r) * in.bw * B;
            T* dst = hbuf + r * pitch + x * B;
            Acc acc[B];
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = 0;
            if (sizeof(T) == 1 && (B == 2 || B == 4) && job.premult) {
                // Premultiply on load (pixel-local, so it commutes with the
                // clamped source read).
                for (int t = 0; t < ph.taps; ++t) {
                    const T* px = src + (size_t)col_of(in, hs + t) * B;
                    const Acc wt = w[t];
                    const uint32_t a = px[B - 1];
#pragma unroll
                    for (int b = 0; b < B - 1; ++b) acc[b] += wt * (Acc)premult_u8(px[b], a);
                    acc[B - 1] += wt * (Acc)a;
                }
            } else {
                for (int t = 0; t < ph.taps; ++t) {
                    const T* px = src + (size_t)col_of(in, hs + t) * B;
                    const Acc wt = w[t];
#pragma unroll
                    for (int b = 0; b < B; ++b) acc[b] += wt * (Acc)px[b];
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) dst[b] = fixed_round<T, Acc>(acc[b]);
        }
        __syncthreads();
    }
    const int row_len = out.w * B;
    const int vec_len = row_len / 4 * 4;
    if constexpr (sizeof(T) == 4) {
        // Up to four output rows per thread at a time (RB x B FP64 chains).
        auto v_rows = [&](auto rbc) {
            constexpr int RB = decltype(rbc)::value;
            for (int ry = rl; x < ncols && ry < nrows; ry += RB * nrl) {
                int vs[RB], vph[RB];
#pragma unroll
                for (int k = 0; k < RB; ++k) {
                    const int yo = ty0 + min(ry + k * nrl, nrows - 1);
                    vs[k] = pv.start[yo];
                    vph[k] = pv.phase[yo];
                }
                double acc[RB][B];
                reduce_v_f32_rows<B, RB>(hbuf + x * B, pitch, vs, vph, pv, vclip0, vclip1, ys0, nr,
                                         (tx0 + x) * B, vec_len, acc);
#pragma unroll
                for (int k = 0; k < RB; ++k) {
                    if (ry + k * nrl >= nrows) break;
                    T* o = static_cast<T*>(out.ptr) + (size_t)(ty0 + ry + k * nrl) * row_len + (size_t)(tx0 + x) * B;
#pragma unroll
                    for (int b = 0; b < B; ++b) o[b] = (T)(float)acc[k][b];
                }
            }
        };
        const int rows_per_thread = (TY + nrl - 1) / nrl;
        if (rows_per_thread >= 4) v_rows(std::integral_constant<int, 4>());
        else if (rows_per_thread >= 2) v_rows(std::integral_constant<int, 2>());
        else v_rows(std::integral_constant<int, 1>());
    }
    if constexpr (sizeof(T) != 4) {
    for (int ry = rl; x < ncols && ry < nrows; ry += nrl) {
        const int yo = ty0 + ry;
        const int vs = pv.start[yo];
        const int vph = pv.phase[yo];
        T* o = static_cast<T*>(out.ptr) + (size_t)yo * row_len + (size_t)(tx0 + x) * B;
        {
            const int16_t* w = pv.wi16 + (size_t)vph * pv.taps;
            Acc acc[B];
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = 0;
            for (int t = 0; t < pv.taps; ++t) {
                const int yy = clampi(clampi(vs + t, vclip0, vclip1) - ys0, 0, nr - 1);
                const T* hp = hbuf + yy * pitch + x * B;
                const Acc wt = w[t];
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] += wt * (Acc)hp[b];
            }
            T v[B];
#pragma unroll
            for (int b = 0; b < B; ++b) v[b] = fixed_round<T, Acc>(acc[b]);
            if constexpr (sizeof(T) == 1 && (B == 2 || B == 4)) {
                if (job.unpremult) {
#pragma unroll
                    for (int b = 0; b < B - 1; ++b) v[b] = (T)unpremult_u8(v[b], v[B - 1]);
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) o[b] = v[b];
        }
    }
    }
}

// ReduceOp (reduce.rs:224-250) on the fused tile: the H pass into LDS over
// the rows the tile's V taps need, rounded by from_f64 as the reference's
// scratch is, then the V pass from LDS; both f64 mul_add chains.
template <typename T, int B>
__device__ __forceinline__ void reduce_op_block(const ReduceJob& job, int bx, int by, unsigned char* smem) {
    const InView& in = job.in;
    const OutView& out = job.out;
    const DevPlan& ph = job.ph;
    const DevPlan& pv = job.pv;
    const int TX = job.tx, TY = job.ty;
    T* hbuf = reinterpret_cast<T*>(smem);
    const int tx0 = bx * TX, ty0 = by * TY;
    const int ncols = min(TX, out.w - tx0), nrows = min(TY, out.h - ty0);
    const int ys0 = pv.start[ty0];
    const int nr = min(job.nr_max, pv.start[ty0 + nrows - 1] + pv.taps - ys0);
    const int x = threadIdx.x % TX, rl = threadIdx.x / TX, nrl = kThreads / TX;
    const int pitch = TX * B;
    // H over staged chunks of job.hr source rows, as in reduce_hv_block.
    T* stage = reinterpret_cast<T*>(smem + reduce_hbuf_bytes(job, sizeof(T)));
    const int c0 = ph.start[tx0];
    const int spx = ph.start[tx0 + ncols - 1] + ph.taps - c0;
    const int srow = spx * B;
    const int xo = tx0 + min(x, ncols - 1);
    const int hoff = (ph.start[xo] - c0) * B;
    const double* w = ph.wf64 + (size_t)ph.phase[xo] * ph.taps;
    const int hstep = max(job.hr, 1);
    for (int rc = 0; rc < nr; rc += hstep) {
        const int hr = min(hstep, nr - rc);
        stage_h_rows<T, B>(in, ys0 + rc, hr, c0, spx, stage);
        __syncthreads();
        for (int r = rl; x < ncols && r < hr; r += nrl) {
            const T* row = stage + r * srow + hoff;
            double acc[B];
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = 0.0;
            for (int t = 0; t < ph.taps; ++t) {
                const T* px = row + t * B;
                const double wt = w[t];
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] = fma((double)px[b], wt, acc[b]);
            }
            T* dst = hbuf + (rc + r) * pitch + x * B;
#pragma unroll
            for (int b = 0; b < B; ++b) dst[b] = op_from_f64<T>(acc[b]);
        }
        __syncthreads();
    }
    const int vclip0 = job.vclip0, vclip1 = job.vclip1;
    for (int ry = rl; x < ncols && ry < nrows; ry += nrl) {
        const int yo = ty0 + ry;
        const int vs = pv.start[yo];
        const double* w = pv.wf64 + (size_t)pv.phase[yo] * pv.taps;
        double acc[B];
#pragma unroll
        for (int b = 0; b < B; ++b) acc[b] = 0.0;
        for (int t = 0; t < pv.taps; ++t) {
            const int yy = clampi(clampi(vs + t, vclip0, vclip1) - ys0, 0, nr - 1);
            const T* hp = hbuf + yy * pitch + x * B;
            const double wt = w[t];
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = fma((double)hp[b], wt, acc[b]);
        }
        T* o = static_cast<T*>(out.ptr) + (size_t)yo * out.w * B + (size_t)(tx0 + x) * B;
#pragma unroll
        for (int b = 0; b < B; ++b) o[b] = op_from_f64<T>(acc[b]);
    }
}

__device__ __forceinline__ void reduce_job_block(const ReduceJob& job, int bx, int by,
                                                 unsigned char* smem) {
    if (job.op) {
#define VGPU_OB(TT)                                            \
    switch (job.bands) {                                       \
    case 1: reduce_op_block<TT, 1>(job, bx, by, smem); break;  \
    case 2: reduce_op_block<TT, 2>(job, bx, by, smem); break;  \
    case 3: reduce_op_block<TT, 3>(job, bx, by, smem); break;  \
    default: reduce_op_block<TT, 4>(job, bx, by, smem); break; \
    }
        if (job.fmt == 0) { VGPU_OB(uint8_t) } else if (job.fmt == 1) { VGPU_OB(uint16_t) } else { VGPU_OB(float) }
#undef VGPU_OB
        return;
    }
#define VGPU_RB(TT, AA)                                            \
    switch (job.bands) {                                           \
    case 1: reduce_hv_block<TT, 1, AA>(job, bx, by, smem); break;  \
    case 2: reduce_hv_block<TT, 2, AA>(job, bx, by, smem); break;  \
    case 3: reduce_hv_block<TT, 3, AA>(job, bx, by, smem); break;  \
    default: reduce_hv_block<TT, 4, AA>(job, bx, by, smem); break; \
    }
    if (job.fmt == 0) {
        if (job.wide_acc) { VGPU_RB(uint8_t, int64_t) } else { VGPU_RB(uint8_t, int32_t) }
    } else if (job.fmt == 1) {
        if (job.wide_acc) { VGPU_RB(uint16_t, int64_t) } else { VGPU_RB(uint16_t, int32_t) }
    } else {
        VGPU_RB(float, int64_t)
    }
#undef VGPU_RB
}

__global__ __launch_bounds__(kThreads) void k_reduce_job(ReduceJob job) {
    extern __shared__ __align__(16) unsigned char smem[];
    reduce_job_block(job, blockIdx.x, blockIdx.y, smem);
}

__global__ __launch_bounds__(kThreads) void k_reduce_batch(const ReduceJob* __restrict__ jobs,
                                                           const BatchTask* __restrict__ tasks, int ntasks,
                                                           const void* const* __restrict__ srcs,
                                                           void* const* __restrict__ dsts) {
    extern __shared__ __align__(16) unsigned char smem[];
    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const BatchTask task = tasks[t];
        ReduceJob job = jobs[task.job];
        if (job.src_slot >= 0) job.in.ptr = srcs[job.src_slot];
        if (job.dst_slot >= 0) job.out.ptr = dsts[job.dst_slot];
        reduce_job_block(job, task.bx, task.by, smem);
        __syncthreads();
    }
}

// Separable convolution pass.  kind 0 = Gaussian: u8 -> u8 integer taps with
// (sum + scale/2) / scale (gauss_blur.rs:405-518); other formats -> f32 with
// acc += f64(f32(x)) * w, no FMA (gauss_blur.rs:410-464).  kind 1 = ConvSep:
// FMA chain, acc / scale + offset -> f32 (convsep.rs:87-107,184-205).
template <typename Tin, typename Tout>
__global__ __launch_bounds__(kThreads) void k_conv(InView in, OutView out, int B, int kind,
                                                   int vertical, const int16_t* itaps,
                                                   int32_t iscale, const double* wts, int ntaps,
                                                   double cscale, double coffset) {
    const size_t n = (size_t)out.w * out.h * B;
    const int rad = (ntaps - 1) / 2;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const int b = (int)(i % B);
        const size_t px = i / B;
        const int x = out.x + (int)(px % out.w), y = out.y + (int)(px / out.w);
        Tout* o = static_cast<Tout*>(out.ptr);
        if (kind == 0 && std::is_same<Tin, uint8_t>::value && std::is_same<Tout, uint8_t>::value) {
            int32_t sum = 0;
            for (int t = 0; t < ntaps; ++t) {
                const int row = row_of(in, vertical ? y - rad + t : y);
                const int col = col_of(in, vertical ? x : x - rad + t);
                sum += (int32_t)itaps[t] * (int32_t)load_at<uint8_t>(in, row, col, B, b);
            }
            const int32_t v = (sum + iscale / 2) / iscale;
            o[i] = (Tout)(v < 0 ? 0 : (v > 255 ? 255 : v));
        } else if (kind == 0) {
            double acc = 0.0;
            for (int t = 0; t < ntaps; ++t) {
                const int row = row_of(in, vertical ? y - rad + t : y);
                const int col = col_of(in, vertical ? x : x - rad + t);
                const double prod = (double)(float)load_at<Tin>(in, row, col, B, b) * wts[t];
                acc = acc + prod;
            }
            o[i] = (Tout)(float)acc;
        } else {
            double acc = 0.0;
            for (int t = 0; t < ntaps; ++t) {
                const int row = row_of(in, vertical ? y - rad + t : y);
                const int col = col_of(in, vertical ? x : x - rad + t);
                acc = fma((double)load_at<Tin>(in, row, col, B, b), wts[t], acc);
            }
            o[i] = (Tout)(float)(acc / cscale + coffset);
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_alpha_u8(InView in, OutView out, int B, int mode) {
    const size_t n = (size_t)out.w * out.h;
    for (size_t px = (size_t)blockIdx.x * kThreads + threadIdx.x; px < n; px += (size_t)gridDim.x * kThreads) {
        const int x = out.x + (int)(px % out.w), y = out.y + (int)(px / out.w);
        const int row = row_of(in, y), col = col_of(in, x);
        const unsigned char* s = static_cast<const unsigned char*>(in.ptr) + ((size_t)row * in.bw + col) * B;
        unsigned char* d = static_cast<unsigned char*>(out.ptr) + px * B;
        const uint32_t a = s[B - 1];
        for (int b = 0; b + 1 < B; ++b)
            d[b] = (unsigned char)(mode == ALPHA_PREMULT ? premult_u8(s[b], a) : unpremult_u8(s[b], a));
        d[B - 1] = (unsigned char)a;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_copy(InView in, OutView out, int B) {
    const size_t n = (size_t)out.w * out.h * B;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const int b = (int)(i % B);
        const size_t px = i / B;
        const int x = out.x + (int)(px % out.w), y = out.y + (int)(px / out.w);
        static_cast<T*>(out.ptr)[i] = load_at<T>(in, row_of(in, y), col_of(in, x), B, b);
    }
}

// Seeded synthetic bytes: 8-byte word i is splitmix64 of seed + (i + 1) *
// golden gamma, little-endian (a counter-based generator, so any sub-range
// can be regenerated on the host).
__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kThreads) void k_fill_random(unsigned char* dst, size_t bytes, uint64_t seed) {
    const size_t words = bytes / 8;
    const bool aligned = (reinterpret_cast<uintptr_t>(dst) & 7) == 0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < words; i += (size_t)gridDim.x * kThreads) {
        const uint64_t v = splitmix64_at(seed, i);
        if (aligned) {
            *reinterpret_cast<uint64_t*>(dst + i * 8) = v;
        } else {
            for (int k = 0; k < 8; ++k) dst[i * 8 + k] = (unsigned char)(v >> (8 * k));
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (bytes & 7)) {
        const uint64_t v = splitmix64_at(seed, words);
        for (size_t k = 0; k < (bytes & 7); ++k) dst[words * 8 + k] = (unsigned char)(v >> (8 * k));
    }
}

// Streaming read of n16 16-byte words (nontemporal, as the box shrink
// reads its source); each block leaves the xor of what it read in sink[block]
// so the loads stay live.
__global__ __launch_bounds__(kThreads) void k_stream_read(const unsigned char* __restrict__ src, size_t n16,
                                                         uint32_t* __restrict__ sink) {
    u32x4 acc = {0u, 0u, 0u, 0u};
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n16; i += (size_t)gridDim.x * kThreads)
        acc ^= ld_stream16(src + i * 16);
    uint32_t v = acc[0] ^ acc[1] ^ acc[2] ^ acc[3];
    for (int o = 32; o > 0; o >>= 1) v ^= (uint32_t)__shfl_xor((int)v, o, 64);
    __shared__ uint32_t part[kThreads / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) sink[blockIdx.x] = part[0] ^ part[1] ^ part[2] ^ part[3];
}

unsigned grid_for(size_t n) {
    size_t g = (n + kThreads - 1) / kThreads;
    if (g < 1) g = 1;
    if (g > 1u << 20) g = 1u << 20;
    return (unsigned)g;
}

}  // namespace

bool plan_shrink_job(ShrinkJob* job, bool row_split, int run) {
    const int fh = job->fh, fv = job->fv, bands = job->bands;
    if (bands < 1 || bands > 4 || fh < 1 || fv < 1) return false;
    const size_t src_bytes = (size_t)job->in.bw * job->in.bh * bands;
    const int aligned = ((uintptr_t)job->in.ptr % 16) == 0 || job->src_slot >= 0;
    const int wb = fh * bands;
    int tx = 64;
    while (tx > 1 && tx * wb + 15 > kSpanBytes) tx /= 2;
    if (!aligned || src_bytes < 16 || tx * wb + 15 > kSpanBytes) return false;
    // Fetch-group geometry: a source row needs at most (15 + span + 15) / 16
    // chunks; its slot is the next power of two, RP = 256 / slot rows a group.
    const int cpr = (15 + tx * wb + 15) / 16;
    int lsl = 0;
    while ((1 << lsl) < cpr) ++lsl;
    const int rp = 256 >> lsl;
    const int gx = (job->out.w + tx - 1) / tx;
    // Run length: enough output rows for two fetch groups per wavefront.
    int m = std::max(1, std::min(32, (2 * rp + fv - 1) / fv));
    auto waves = [&](int mm) { return (int64_t)gx * ((job->out.h + mm - 1) / mm); };
    int S = 1;
    if (row_split) {
        // One image alone: keep about 16 wavefronts per CU busy before
        // lengthening runs, and split the source rows of an output row over
        // S wavefronts while the row tasks leave room on the chip.  A source
        // whose row pitch is not a multiple of 16 bytes gets up to 16
        // wavefronts per SIMD instead of 4: its waves run the same
        // instructions and fetch the same bytes, yet each row's loads take
        // longer (measured: 5845x5551x4 at shrink 14 went from 39 to 27 us,
        // the 16-byte-aligned 5844 wide image stays at 24 us either way), so
        // it needs more rows in flight.
        while (m > 1 && waves(m) < 256 * 16) m = (m + 1) / 2;
        const int64_t wave_cap = ((uint64_t)job->in.bw * bands) % 16 ? 256 * 4 * 16 : 256 * 4 * 4;
        if (m == 1)
            while (S < 4 && 2 * S <= fv && waves(1) * 2 * S <= wave_cap) S *= 2;
    }
    if (run > 0) {  // caller's run length: no row split
        m = std::min(run, 64);
        S = 1;
    }
    job->tx = tx;
    job->S = S;
    job->m = S > 1 ? 1 : m;
    job->lsl = lsl;
    job->mh = (uint32_t)((1ull << 32) / (256ull * (uint64_t)fh));
    job->mv = (uint32_t)((1ull << 32) / (256ull * (uint64_t)fv));
    // ((sum + add) * mul) >> shift forms of the roundings (shrinkh.rs:381-471,
    // shrinkv.rs:206-339): identity at f = 1, the pair rounding at f = 2, the
    // exact x86 division by 5 for 1, 3 or 4 bands (ceil(2^24 / 5) is exact
    // for every sum + 2 <= 1277), else the libvips multiply-shift.
    auto rounding = [&](int f, bool horizontal, uint32_t* add, uint32_t* mul, uint32_t* shift) {
        if (f == 1) {
            *add = 0, *mul = 1, *shift = 0;
        } else if (f == 2) {
            *add = 1, *mul = 1, *shift = 1;
        } else if (horizontal && f == 5 && bands != 2) {
            *add = 2, *mul = (uint32_t)(((1u << 24) + 4) / 5), *shift = 24;
        } else {
            *add = (uint32_t)(f >> 1), *mul = (uint32_t)((1u << 24) / (uint32_t)f), *shift = 24;
        }
    };
    rounding(fh, true, &job->rh_add, &job->rh_mul, &job->rh_shift);
    rounding(fv, false, &job->rv_add, &job->rv_mul, &job->rv_shift);
    // Lean forms: window length in dwords rounded up to a multiple of 4
    // (0 beyond 16: looped sums); 32-bit buffer offsets need < 4 GiB.
    const int wnd = (wb + 3) >> 2;
    job->wnd_t = wnd <= 16 ? (wnd + 3) / 4 * 4 : 0;
    job->lean = src_bytes + 8192 < (1ull << 32) ? 1 : 0;
    const int rows_per_block = ((kThreads / 64) / S) * job->m;
    job->gx = gx;
    job->gy = (job->out.h + rows_per_block - 1) / rows_per_block;
    return true;
}

bool plan_box_shrink_job(ShrinkJob* job, bool row_split) {
    if (job->fh != job->fv || job->premult || job->unpremult) return false;
    if (!plan_shrink_job(job, row_split, 0)) return false;
    // Rows of raw window sums (identity ShrinkH rounding), one rounding of the
    // f x f total: (sum + d/2) / d as a multiply-shift, exact for every
    // sum <= 255 d (a power-of-two d shifts; others are checked here).
    const uint32_t d = (uint32_t)job->fh * (uint32_t)job->fv, add = d / 2;
    const uint64_t xmax = 255ull * d + add;
    if (xmax >= (1u << 24)) return false;
    uint32_t mul = 0, shift = 0;
    if ((d & (d - 1)) == 0) {
        mul = 1;
        while ((1u << shift) < d) ++shift;
    } else {
        for (uint32_t sh = 24; sh < 40 && !mul; ++sh) {
            const uint64_t m = ((1ull << sh) + d - 1) / d;
            if (m >= (1u << 24) || xmax * m >= (1ull << 32)) break;
            bool ok = true;
            for (uint64_t x = 0; x <= xmax && ok; ++x) ok = ((x * m) >> sh) == x / d;
            if (ok) mul = (uint32_t)m, shift = sh;
        }
        if (!mul) return false;
    }
    job->rh_add = 0, job->rh_mul = 1, job->rh_shift = 0;
    job->rv_add = add, job->rv_mul = mul, job->rv_shift = shift;
    return true;
}

int shrink_job_variant(const ShrinkJob& job) {
    if (!job.lean) return job.bands - 1;
    return job.bands - 1 + (job.lsl >= 8 ? 4 : (job.lsl == 7 ? 8 : 0));
}

size_t shrink_job_lds() { return (size_t)(kThreads / 64) * (kSpanBytes + 32); }

#define VGPU_SHRINK_VARIANTS(X)                                                                    \
    X(0, 1, 0) X(1, 2, 0) X(2, 3, 0) X(3, 4, 0) X(4, 1, 1) X(5, 2, 1) X(6, 3, 1) X(7, 4, 1) X(8, 1, 2) X(9, 2, 2) \
    X(10, 3, 2) X(11, 4, 2)

hipError_t launch_shrink_job(const ShrinkJob& job, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    const dim3 grid((unsigned)job.gx, (unsigned)job.gy);
    switch (shrink_job_variant(job)) {
#define X(V, B, L) \
    case V: hipLaunchKernelGGL((k_shrink_job<B, L>), grid, dim3(kThreads), shrink_job_lds(), s, job); break;
        VGPU_SHRINK_VARIANTS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_shrink_batch(int variant, const ShrinkJob* jobs, const BatchTask* tasks, int ntasks,
                               const void* const* srcs, void* const* dsts, int grid, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    switch (variant) {
#define X(V, B, L)                                                                                    \
    case V:                                                                                           \
        hipLaunchKernelGGL((k_shrink_batch<B, L>), dim3((unsigned)grid), dim3(kThreads), shrink_job_lds(), s, \
                           jobs, tasks, ntasks, srcs, dsts);                                          \
        break;
        VGPU_SHRINK_VARIANTS(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
#undef VGPU_SHRINK_VARIANTS

hipError_t launch_shrink_hv_u8(const InView& in, const OutView& out, int bands, int fh, int fv,
                               int premult, hipStream_t s) {
    if (out.w <= 0 || out.h <= 0) return hipSuccess;
    const uint32_t mh = (uint32_t)((1ull << 32) / (256ull * (uint64_t)fh));
    const uint32_t mv = (uint32_t)((1ull << 32) / (256ull * (uint64_t)fv));
    const int aligned = ((uintptr_t)in.ptr % 16) == 0;
    int tx = 64;
    // Block-per-row kernel for very wide box windows (fh * bands > 4080
    // bytes) or a source that is not 16-byte aligned, with byte-exact loads.
    const size_t budget = 64 * 1024;
    auto pitch_of = [&](int cols) { return ((size_t)cols * fh * bands + 32 + 15) & ~(size_t)15; };
    while (tx > 1 && pitch_of(tx) > budget) tx /= 2;
    const size_t pitch = pitch_of(tx);
    if (pitch > budget) return hipErrorInvalidValue;
    int rl = kThreads / tx;
    if ((size_t)rl * pitch > budget) rl = (int)(budget / pitch);
    if (rl < 1) rl = 1;
    size_t shmem = (size_t)rl * pitch;
    const size_t part = (size_t)kThreads * bands * 4;
    if (shmem < part) shmem = part;
    dim3 grid((unsigned)((out.w + tx - 1) / tx), (unsigned)out.h);
    switch (bands) {
    case 1: hipLaunchKernelGGL(k_shrink_hv_u8<1>, grid, dim3(kThreads), shmem, s, in, out, fh, fv, tx, rl, premult, mh, mv, aligned); break;
    case 2: hipLaunchKernelGGL(k_shrink_hv_u8<2>, grid, dim3(kThreads), shmem, s, in, out, fh, fv, tx, rl, premult, mh, mv, aligned); break;
    case 3: hipLaunchKernelGGL(k_shrink_hv_u8<3>, grid, dim3(kThreads), shmem, s, in, out, fh, fv, tx, rl, premult, mh, mv, aligned); break;
    case 4: hipLaunchKernelGGL(k_shrink_hv_u8<4>, grid, dim3(kThreads), shmem, s, in, out, fh, fv, tx, rl, premult, mh, mv, aligned); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_shrink_generic(const InView& in, const OutView& out, int bands, int fmt,
                                 int factor, int vertical, hipStream_t s) {
    const size_t n = (size_t)out.w * out.h * bands;
    if (n == 0) return hipSuccess;
    const uint32_t mult = (uint32_t)((1ull << 32) / (256ull * (uint64_t)factor));
    const unsigned g = grid_for(n);
#define VGPU_SG(T) \
    hipLaunchKernelGGL(k_shrink_generic<T>, dim3(g), dim3(kThreads), 0, s, in, out, bands, factor, vertical, mult)
    VGPU_FOR_FORMAT(fmt, VGPU_SG)
#undef VGPU_SG
    return hipGetLastError();
}

hipError_t launch_box_generic(const void* src, int sw, int sh, int bands, int fmt, int f, void* dst, int dw,
                              int x0, int y0, int rw, int rh, hipStream_t s) {
    const int64_t n = (int64_t)rw * rh;
    if (n <= 0) return hipSuccess;
    const unsigned g = (unsigned)((n + kThreads - 1) / kThreads);
#define VGPU_BOX(T)                                                                                          \
    hipLaunchKernelGGL(k_box_generic<T>, dim3(g), dim3(kThreads), 0, s, static_cast<const T*>(src), sw, sh, \
                       bands, f, static_cast<T*>(dst), dw, x0, y0, rw, rh)
    VGPU_FOR_FORMAT(fmt, VGPU_BOX)
#undef VGPU_BOX
    return hipGetLastError();
}

// op = 1: ReduceOp's passes (f64 mul_add chains, from_f64 rounding).
hipError_t launch_reduce_h(const InView& in, const OutView& out, int bands, int fmt,
                           const DevPlan& plan, int op, hipStream_t s) {
    const size_t n = (size_t)out.w * out.h * bands;
    if (n == 0) return hipSuccess;
    const unsigned g = grid_for(n);
#define VGPU_RP(TT)                                                                                          \
    if (op) hipLaunchKernelGGL((k_reduce_h<TT, true>), dim3(g), dim3(kThreads), 0, s, in, out, bands, plan); \
    else hipLaunchKernelGGL((k_reduce_h<TT, false>), dim3(g), dim3(kThreads), 0, s, in, out, bands, plan);
    VGPU_FOR_FORMAT(fmt, VGPU_RP)
#undef VGPU_RP
    return hipGetLastError();
}

hipError_t launch_reduce_v(const InView& in, const OutView& out, int bands, int fmt,
                           const DevPlan& plan, int op, hipStream_t s) {
    const size_t n = (size_t)out.w * out.h * bands;
    if (n == 0) return hipSuccess;
    const unsigned g = grid_for(n);
#define VGPU_RP(TT)                                                                                          \
    if (op) hipLaunchKernelGGL((k_reduce_v<TT, true>), dim3(g), dim3(kThreads), 0, s, in, out, bands, plan); \
    else hipLaunchKernelGGL((k_reduce_v<TT, false>), dim3(g), dim3(kThreads), 0, s, in, out, bands, plan);
    VGPU_FOR_FORMAT(fmt, VGPU_RP)
#undef VGPU_RP
    return hipGetLastError();
}

size_t reduce_job_lds(const ReduceJob& job) {
    const size_t esz = job.fmt == 0 ? 1 : (job.fmt == 1 ? 2 : 4);
    return reduce_hbuf_bytes(job, esz) + (size_t)job.hr * job.span_max * job.bands * esz;
}

hipError_t launch_reduce_job(const ReduceJob& job, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    if (job.bands < 1 || job.bands > 4) return hipErrorInvalidValue;
    if (job.stream && job.fmt == 2 && !job.op) return launch_reduce_stream(job, 0, s);
    hipLaunchKernelGGL(k_reduce_job, dim3((unsigned)job.gx, (unsigned)job.gy), dim3(kThreads),
                       reduce_job_lds(job), s, job);
    return hipGetLastError();
}

hipError_t launch_reduce_batch(const ReduceJob* jobs, const BatchTask* tasks, int ntasks,
                               const void* const* srcs, void* const* dsts, int grid, size_t lds,
                               hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_batch, dim3((unsigned)grid), dim3(kThreads), lds, s, jobs, tasks,
                       ntasks, srcs, dsts);
    return hipGetLastError();
}

hipError_t launch_conv(const InView& in, const OutView& out, int bands, int in_fmt, int out_fmt,
                       int kind, int vertical, const int16_t* itaps, int32_t iscale,
                       const double* wts, int ntaps, double cscale, double coffset,
                       hipStream_t s) {
    const size_t n = (size_t)out.w * out.h * bands;
    if (n == 0) return hipSuccess;
    const unsigned g = grid_for(n);
#define VGPU_CONV(TI, TO) hipLaunchKernelGGL((k_conv<TI, TO>), dim3(g), dim3(kThreads), 0, s, in, out, bands, kind, vertical, itaps, iscale, wts, ntaps, cscale, coffset)
    if (out_fmt == 0) {
        if (in_fmt != 0) return hipErrorInvalidValue;
        VGPU_CONV(uint8_t, uint8_t);
    } else if (out_fmt == 2) {
#define VGPU_CONV_F32(TI) VGPU_CONV(TI, float)
        VGPU_FOR_FORMAT(in_fmt, VGPU_CONV_F32)
#undef VGPU_CONV_F32
    } else {
        return hipErrorInvalidValue;
    }
#undef VGPU_CONV
    return hipGetLastError();
}

hipError_t launch_alpha_u8(const InView& in, const OutView& out, int bands, int mode, hipStream_t s) {
    const size_t n = (size_t)out.w * out.h;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_alpha_u8, dim3(grid_for(n)), dim3(kThreads), 0, s, in, out, bands, mode);
    return hipGetLastError();
}

hipError_t launch_copy(const InView& in, const OutView& out, int bands, int fmt, hipStream_t s) {
    const size_t n = (size_t)out.w * out.h * bands;
    if (n == 0) return hipSuccess;
    const unsigned g = grid_for(n);
    // By sample size: 1 (U8), 2 (U16, I16), 8 (F64) or 4 bytes.
    switch (fmt == 0 ? 1 : (fmt == 1 || fmt == 3) ? 2 : fmt == 6 ? 8 : 4) {
    case 1: hipLaunchKernelGGL(k_copy<uint8_t>, dim3(g), dim3(kThreads), 0, s, in, out, bands); break;
    case 2: hipLaunchKernelGGL(k_copy<uint16_t>, dim3(g), dim3(kThreads), 0, s, in, out, bands); break;
    case 4: hipLaunchKernelGGL(k_copy<uint32_t>, dim3(g), dim3(kThreads), 0, s, in, out, bands); break;
    case 8: hipLaunchKernelGGL(k_copy<uint64_t>, dim3(g), dim3(kThreads), 0, s, in, out, bands); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fill_random(void* dst, size_t bytes, uint64_t seed, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_random, dim3(grid_for(bytes / 8 + 1)), dim3(kThreads), 0, s,
                       static_cast<unsigned char*>(dst), bytes, seed);
    return hipGetLastError();
}

hipError_t launch_stream_read(const void* src, size_t bytes, uint32_t* sink, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_stream_read, dim3((unsigned)blocks), dim3(kThreads), 0, s,
                       static_cast<const unsigned char*>(src), bytes / 16, sink);
    return hipGetLastError();
}

}  // namespace vgpu
