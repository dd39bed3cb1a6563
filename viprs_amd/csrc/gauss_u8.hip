// GaussBlurH -> GaussBlurV for 3- and 4-band u8 images on the VALU, one
// tile per block (gauss_blur.rs:405-518 integer path: sum of c_t * p, then
// (sum + scale / 2) / scale, H rounded to u8 before V).
//
// A pixel travels as two packed 16-bit pairs, (b0, b2) and (b1, b3), so one
// v_pk_mad_u16 applies a tap to two bands.  The sums stay below 2^16 when
// 255 * scale + scale / 2 < 2^16 (scale <= 256), which the host checks.  The
// division by scale is an exact multiply-high: with mul = ceil(2^32 / scale),
// floor(x * mul / 2^32) = floor(x / scale) for every x < 256 * scale.
//
// Block = 64 output columns x TH output rows.  The input tile (TH + 2R rows x
// 64 + 2R columns, read through the clamped view, i.e. the reference's
// clamp-to-edge source and its Z^2 intermediate rows) goes to LDS as packed
// pixels; the H pass (lane = column) writes the TH + 2R intermediate rows the
// V taps need, rounded to bytes, to LDS; the V pass gives each lane one
// column and each wavefront TH / 4 consecutive rows, with the 2R + TH / 4
// intermediate rows it reads held in registers as pairs.
#include "device.hpp"

namespace vgpu {
namespace {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_pair(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// (x + scale/2) / scale for the two 16-bit halves of a pair (the half-scale
// bias is already in x); mul == 0 means scale == 1.
__device__ __forceinline__ uint32_t div_pair(uint32_t x, uint32_t mul) {
    const uint32_t lo = x & 0xffffu, hi = x >> 16;
    if (mul == 0) return lo | (hi << 16);
    return __umulhi(lo, mul) | (__umulhi(hi, mul) << 16);
}

template <int B, int R, int TH>
__global__ __launch_bounds__(kThreads) void k_gauss_u8(GaussJob job) {
    constexpr int TW = 64;
    constexpr int IW = TW + 2 * R;    // input tile columns (<= 128: two per lane)
    constexpr int NR = TH + 2 * R;    // input / intermediate tile rows
    constexpr int VR = TH / 4;        // output rows per wavefront
    constexpr int KR = (NR + 3) / 4;  // input rows per wavefront
    // Pixels stay packed (one dword each) in LDS: the input tile row by row,
    // the H results as rounded bytes; a lane reads its column's dwords, which
    // are consecutive across the wavefront (no bank conflicts), and splits
    // them into the (b0, b2), (b1, b3) pairs in registers.
    // A lane of the H pass makes 4 adjacent columns from NRD 16-byte reads
    // (4 + 2R pixels); tile rows are IWP >= IW dwords, a multiple of 4.
    constexpr int NRD = (2 * R + 7) / 4;
    constexpr int IWP = 60 + 4 * NRD;
    static_assert(IWP >= IW && IWP % 4 == 0, "tile pitch");
    __shared__ __align__(16) uint32_t tile[NR * IWP];
    __shared__ __align__(16) uint32_t mid[NR * TW];
    const InView& in = job.in;
    const OutView& out = job.out;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // XCD-aware order: workgroups go to the eight XCDs round robin, so
    // workgroup p takes logical block (p % 8) * chunk + p / 8 -- each XCD a
    // contiguous run of tile columns, whose shared halo lines then hit in
    // that XCD's L2.  Logical blocks past gx * gy exit (the whole block).
    const int nblk = job.gx * job.gy, chunk = (nblk + 7) / 8;
    const int lb = ((int)blockIdx.x % 8) * chunk + (int)blockIdx.x / 8;
    if (lb >= nblk) return;
    const int bx = lb % job.gx, by = lb / job.gx;
    const int ox0 = bx * TW;
    // This block's tiles: rows [t0, t1) of the column of tiles.
    const int t0 = by * job.tiles_per_block;
    const int t1 = min(t0 + job.tiles_per_block, (out.h + TH - 1) / TH);
    const unsigned char* src = static_cast<const unsigned char*>(in.ptr);
    const size_t pitch = (size_t)in.bw * B;

    // Input tile of tile row tr: rows out.y + tr * TH - R + rr (row
    // addresses wave-uniform), columns out.x + ox0 - R + cc for cc = lane and
    // lane + 64 (byte offsets per lane, computed once).
    const int cc1 = lane + 64;
    const uint32_t cb0 = (uint32_t)col_of(in, out.x + ox0 - R + lane) * B;
    const uint32_t cb1 = (uint32_t)col_of(in, out.x + ox0 - R + min(cc1, IW - 1)) * B;
    // A 3-band pixel is read as the two aligned dwords holding its bytes,
    // then realigned and masked.  For the buffer's last pixel the second
    // dword may run up to 3 bytes past the buffer's end; an aligned dword
    // never crosses a page, so those bytes are mapped, and they are dropped.
    auto load_px = [&](const unsigned char* rowp, uint32_t cb) -> uint32_t {
        if constexpr (B == 4) {
            return gmem<uint32_t>(rowp + cb)[0];
        } else {
            const uintptr_t a = (uintptr_t)(rowp + cb);
            const uint32_t lo = gmem<uint32_t>((const void*)(a & ~(uintptr_t)3))[0];
            const uint32_t hi = gmem<uint32_t>((const void*)((a + 2) & ~(uintptr_t)3))[0];
            return __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)(a & 3)) & 0x00ffffffu;
        }
    };
    // A wavefront's loads of one tile, all issued before any is used; the
    // next tile's are in flight while the current one is computed.
    uint32_t v0[KR], v1[KR];
    auto fetch = [&](int tr) {
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int rr = wave + 4 * k;
            if (rr < NR) {
                const unsigned char* rowp = src + (size_t)row_of(in, out.y + tr * TH - R + rr) * pitch;
                v0[k] = load_px(rowp, cb0);
                v1[k] = load_px(rowp, cb1);
            }
        }
    };
    const u16x2 half = as_pair(job.half2);
    unsigned char* dst = static_cast<unsigned char*>(out.ptr);
    const int ox = ox0 + lane;
    const int r0 = wave * VR;
    if (t0 < t1) fetch(t0);
    for (int tr = t0; tr < t1; ++tr) {
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int rr = wave + 4 * k;
            if (rr < NR) {
                tile[rr * IWP + lane] = v0[k];
                if (cc1 < IW) tile[rr * IWP + cc1] = v1[k];
            }
        }
        __syncthreads();  // the tile is in LDS; the previous V pass is done with mid
        if (tr + 1 < t1) fetch(tr + 1);

        // H pass: a wavefront takes 4 rows at a time, 16 lanes a row; lane q
        // makes intermediate columns 4q .. 4q + 3 of its row from input
        // columns 4q .. 4q + 3 + 2R.
        const int q = lane & 15;
        for (int rb = wave * 4; rb < NR; rb += 16) {
            const int rr = rb + (lane >> 4);
            if (rr < NR) {
                uint32_t pa[4 * NRD], pb[4 * NRD];
                const uint4* rp = reinterpret_cast<const uint4*>(tile + rr * IWP + 4 * q);
#pragma unroll
                for (int k = 0; k < NRD; ++k) {
                    const uint4 v = rp[k];
                    const uint32_t px[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        pa[4 * k + e] = px[e] & 0x00ff00ffu;
                        pb[4 * k + e] = (px[e] >> 8) & 0x00ff00ffu;
                    }
                }
                u16x2 a[4], b[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) a[j] = b[j] = half;
#pragma unroll
                for (int t = 0; t <= 2 * R; ++t) {
                    const u16x2 w = as_pair(job.taps2[t]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        a[j] = as_pair(pa[j + t]) * w + a[j];
                        b[j] = as_pair(pb[j + t]) * w + b[j];
                    }
                }
                uint32_t o[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    o[j] = div_pair(as_u32(a[j]), job.mul) | (div_pair(as_u32(b[j]), job.mul) << 8);
                *reinterpret_cast<uint4*>(mid + rr * TW + 4 * q) = make_uint4(o[0], o[1], o[2], o[3]);
            }
        }
        __syncthreads();  // mid is complete; the input tile may be overwritten

        // V pass: output rows tr * TH + wave * VR .. + VR - 1, column ox.
        uint32_t wa[VR + 2 * R], wb[VR + 2 * R];
#pragma unroll
        for (int j = 0; j < VR + 2 * R; ++j) {
            const uint32_t px = mid[(r0 + j) * TW + lane];
            wa[j] = px & 0x00ff00ffu;
            wb[j] = (px >> 8) & 0x00ff00ffu;
        }
        if (ox < out.w) {
#pragma unroll
            for (int o = 0; o < VR; ++o) {
                const int oy = tr * TH + r0 + o;
                if (oy >= out.h) break;
                u16x2 a = half, b = half;
#pragma unroll
                for (int t = 0; t <= 2 * R; ++t) {
                    const u16x2 w = as_pair(job.taps2[t]);
                    a = as_pair(wa[o + t]) * w + a;
                    b = as_pair(wb[o + t]) * w + b;
                }
                const uint32_t qrb = div_pair(as_u32(a), job.mul), qga = div_pair(as_u32(b), job.mul);
                const uint32_t v = qrb | (qga << 8);  // bytes b0 b1 b2 b3
                unsigned char* o8 = dst + ((size_t)oy * out.w + ox) * B;
                if constexpr (B == 4) {
                    *reinterpret_cast<uint32_t*>(o8) = v;
                } else {
                    o8[0] = (unsigned char)v;
                    o8[1] = (unsigned char)(v >> 8);
                    o8[2] = (unsigned char)(v >> 16);
                }
            }
        }
    }
}

}  // namespace

bool gauss_u8_supported(int bands, int taps, int32_t scale) {
    const int r = (taps - 1) / 2;
    return (bands == 3 || bands == 4) && taps == 2 * r + 1 && r >= 1 && r <= kGaussU8MaxRadius && scale >= 1 &&
           scale <= 256;
}

hipError_t launch_gauss_u8(const GaussJob& job, hipStream_t s) {
    if (job.out.w <= 0 || job.out.h <= 0) return hipSuccess;
    const dim3 grid((unsigned)(8 * ((job.gx * job.gy + 7) / 8)));
#define VGPU_G(BB, RR, TT) \
    hipLaunchKernelGGL((k_gauss_u8<BB, RR, TT>), grid, dim3(kThreads), 0, s, job)
#define VGPU_GR(BB, TT)                 \
    switch (job.radius) {               \
    case 1: VGPU_G(BB, 1, TT); break;   \
    case 2: VGPU_G(BB, 2, TT); break;   \
    case 3: VGPU_G(BB, 3, TT); break;   \
    case 4: VGPU_G(BB, 4, TT); break;   \
    case 5: VGPU_G(BB, 5, TT); break;   \
    case 6: VGPU_G(BB, 6, TT); break;   \
    default: return hipErrorInvalidValue; \
    }
    if (job.bands == 4) {
        if (job.th == 32) { VGPU_GR(4, 32) } else { VGPU_GR(4, 16) }
    } else if (job.bands == 3) {
        if (job.th == 32) { VGPU_GR(3, 32) } else { VGPU_GR(3, 16) }
    } else {
        return hipErrorInvalidValue;
    }
#undef VGPU_GR
#undef VGPU_G
    return hipGetLastError();
}

}  // namespace vgpu
