// Device helpers shared by the kernel files: launch width, clamped image
// reads, global-address-space loads and the u8 alpha arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <limits>
#include <type_traits>

#include <cstdint>

#include "kernels.hpp"

namespace vgpu {

constexpr int kThreads = 256;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Image buffers always live in device global memory; saying so lets the
// compiler use global (not flat) loads for pointers that come from job tables.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const void* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// 16 bytes of a stream read once (nontemporal: the load is not kept in the
// caches for reuse, which raises the sustained HBM read rate).
__device__ __forceinline__ u32x4 ld_stream16(const void* p) {
    return __builtin_nontemporal_load(gmem<u32x4>(p));
}

// Buffer-relative column / row of an absolute coordinate: clip, then clamp to
// the buffer.
__device__ __forceinline__ int col_of(const InView& v, int x) {
    return clampi(clampi(x, v.cx0, v.cx1), v.bx, v.bx + v.bw - 1) - v.bx;
}
__device__ __forceinline__ int row_of(const InView& v, int y) {
    return clampi(clampi(y, v.cy0, v.cy1), v.by, v.by + v.bh - 1) - v.by;
}

// Range of an integer sample type (band.rs: U8, U16, I16, U32, I32).
template <typename T>
struct SampleRange {
    static constexpr double lo = (double)std::numeric_limits<T>::min();
    static constexpr double hi = (double)std::numeric_limits<T>::max();
    static constexpr int64_t ilo = (int64_t)std::numeric_limits<T>::min();
    static constexpr int64_t ihi = (int64_t)std::numeric_limits<T>::max();
};
template <typename T>
constexpr bool is_float_sample = std::is_same<T, float>::value || std::is_same<T, double>::value;

// Launch-time dispatch over the band formats (U8, U16, F32, I16, U32, I32,
// F64 = 0..6): MACRO(T) with T the sample type; an unknown format returns
// hipErrorInvalidValue from the enclosing launcher.
#define VGPU_FOR_FORMAT(fmt, MACRO)                  \
    switch (fmt) {                                   \
    case 0: MACRO(uint8_t); break;                   \
    case 1: MACRO(uint16_t); break;                  \
    case 2: MACRO(float); break;                     \
    case 3: MACRO(int16_t); break;                   \
    case 4: MACRO(uint32_t); break;                  \
    case 5: MACRO(int32_t); break;                   \
    case 6: MACRO(double); break;                    \
    default: return hipErrorInvalidValue;            \
    }

// F::Sample::from_f64 (viprs-core shared_ops/sample_conv.rs:53-99): integer
// samples clamp to the type's range, then round half away from zero; f32 is a
// cast, f64 the value itself.
template <typename T>
__device__ __forceinline__ T op_from_f64(double v) {
    if constexpr (std::is_same<T, float>::value) {
        return (float)v;
    } else if constexpr (std::is_same<T, double>::value) {
        return v;
    } else {
        return (T)round(fmin(fmax(v, SampleRange<T>::lo), SampleRange<T>::hi));
    }
}

// ShrinkSample::from_f64_clamped (shrinkh.rs:44-128): unsigned samples round
// half away from zero, then clamp; signed ones clamp v + 0.5, then truncate
// toward zero; f32 casts, f64 is kept.
template <typename T>
__device__ __forceinline__ T shrink_from_f64(double v) {
    if constexpr (std::is_same<T, float>::value) {
        return (float)v;
    } else if constexpr (std::is_same<T, double>::value) {
        return v;
    } else if constexpr (std::is_signed<T>::value) {
        return (T)trunc(fmin(fmax(v + 0.5, SampleRange<T>::lo), SampleRange<T>::hi));
    } else {
        return (T)fmin(fmax(round(v), SampleRange<T>::lo), SampleRange<T>::hi);
    }
}

// ReduceSample::from_fixed_i64 (resample/sample_conv.rs:18-62) of a 12-bit
// fixed-point sum: unsigned (acc + 2048) >> 12; signed adds 2048 to a
// non-negative sum and subtracts it from a negative one before the
// arithmetic shift; then the type's range.
template <typename T, typename Acc>
__device__ __forceinline__ T fixed_from_acc(Acc acc) {
    Acc v;
    if constexpr (std::is_signed<T>::value) v = acc >= 0 ? (acc + 2048) >> 12 : (acc - 2048) >> 12;
    else v = (acc + 2048) >> 12;
    const Acc lo = (Acc)SampleRange<T>::ilo, hi = (Acc)SampleRange<T>::ihi;
    return (T)(v < lo ? lo : (v > hi ? hi : v));
}

template <typename T>
__device__ __forceinline__ T load_at(const InView& v, int row, int col, int bands, int b) {
    return static_cast<const T*>(v.ptr)[((size_t)row * v.bw + col) * bands + b];
}

// Premultiply<U8> (premultiply.rs:103-115): round(c * (a / 255)) in f64.  For
// integer c, a the exact quotient c*a/255 is never within 1/510 of a half, so
// the f64 expression equals the integer rounding (2ca + 255) / 510.
__device__ __forceinline__ uint32_t premult_u8(uint32_t c, uint32_t a) { return (2u * c * a + 255u) / 510u; }
// Unpremultiply<U8> (unpremultiply.rs:96-109): a == 0 -> 0, else
// clamp(round(c * 255 / a)).  The f64 quotient of two small integers is either
// exact (ties included) or at least 1/(2a) away from a half, so it equals the
// integer rounding (510c + a) / (2a).
__device__ __forceinline__ uint32_t unpremult_u8(uint32_t c, uint32_t a) {
    if (a == 0) return 0;
    const uint32_t v = (510u * c + a) / (2u * a);
    return v > 255u ? 255u : v;
}

// Stage source rows ys .. ys + hr - 1 of the columns c0 .. c0 + spx - 1
// (each through the clamped read) into LDS, spx * B samples per row: the H
// taps of a fused tile then read LDS instead of global memory.  A span that
// no clamp touches is a contiguous run of each source row: 4-byte samples
// then come in by LDS-direct dword loads (one wavefront instruction per 64
// samples, no registers); other spans sample by sample.  Ends with the
// loads landed (the caller's barrier then publishes them).
template <typename T, int B>
__device__ __forceinline__ void stage_h_rows(const InView& in, int ys, int hr, int c0, int spx, T* stage) {
    const int srow = spx * B;
    const int lo = max(in.cx0, in.bx), hi = min(in.cx1, in.bx + in.bw - 1);
    if (sizeof(T) == 4 && c0 >= lo && c0 + spx - 1 <= hi) {
        const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        for (int k = wave; k < hr; k += kThreads / 64) {
            const T* src = static_cast<const T*>(in.ptr) + ((size_t)row_of(in, ys + k) * in.bw + (c0 - in.bx)) * B;
            for (int d0 = 0; d0 < srow; d0 += 64) {
                if (d0 + lane < srow)
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + d0 + lane),
                                                     (__attribute__((address_space(3))) void*)(stage + k * srow + d0),
                                                     4, 0, 0);
            }
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        return;
    }
    for (int k = 0; k < hr; ++k) {
        const T* src = static_cast<const T*>(in.ptr) + (size_t)row_of(in, ys + k) * in.bw * B;
        for (int w = threadIdx.x; w < srow; w += kThreads) {
            const int cc = w / B, b = w - cc * B;
            stage[k * srow + w] = src[(size_t)col_of(in, c0 + cc) * B + b];
        }
    }
}

// Bytes the fused reduce tile keeps before its staged source rows (the H
// results, 16-byte aligned).
__host__ __device__ inline size_t reduce_hbuf_bytes(const ReduceJob& job, size_t esz) {
    return ((size_t)job.nr_max * job.tx * job.bands * esz + 15) & ~(size_t)15;
}

// XCD-aware block order: the dispatcher places block b on XCD b % 8, each
// XCD with its own L2.  Renumbering so that every XCD runs one contiguous
// range of logical blocks keeps blocks that share source lines (neighbouring
// strips) on one L2.  A bijection on 0 .. nb - 1.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    constexpr int kXcds = 8;
    const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// LDS hand-off between the lanes of one wavefront: LDS operations of a
// wavefront execute in issue order, so a compiler barrier suffices.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace vgpu
