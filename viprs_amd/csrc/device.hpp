// Device helpers shared by the kernel files: launch width, clamped image
// reads, global-address-space loads and the u8 alpha arithmetic.
#pragma once

#include <hip/hip_runtime.h>

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

// Buffer-relative column / row of an absolute coordinate: clip, then clamp to
// the buffer.
__device__ __forceinline__ int col_of(const InView& v, int x) {
    return clampi(clampi(x, v.cx0, v.cx1), v.bx, v.bx + v.bw - 1) - v.bx;
}
__device__ __forceinline__ int row_of(const InView& v, int y) {
    return clampi(clampi(y, v.cy0, v.cy1), v.by, v.by + v.bh - 1) - v.by;
}

// F::Sample::from_f64 (sample_conv.rs:58-93): integer samples clamp to the
// type's range, then round half away from zero; f32 is a cast.
template <typename T>
__device__ __forceinline__ T op_from_f64(double v) {
    if constexpr (sizeof(T) == 4) {
        return (T)(float)v;
    } else {
        const double mx = sizeof(T) == 1 ? 255.0 : 65535.0;
        return (T)round(fmin(fmax(v, 0.0), mx));
    }
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

// LDS hand-off between the lanes of one wavefront: LDS operations of a
// wavefront execute in issue order, so a compiler barrier suffices.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace vgpu
