// The Affine tails of thumbnail (thumbnail.rs:292-316,396-419) and resize
// (resize.rs:258-280): an axis-aligned matrix [a, 0, 0, d] with
// ExtendMode::Copy over its whole input image, in the reference's generic
// per-pixel path (affine/core.rs:1593-1632 -> sample_pixel_at, 1022-1055).  The source
// positions and, for Bicubic, the per-axis weights are tabulated on the host
// (AffineTap), so a thread does the f64 sampling of one output pixel in the
// reference's operation order:
//  * Nearest: the sample at (trunc x, trunc y);
//  * Bilinear: bilinear_mix (affine/core.rs:197-201) of the 2 x 2 samples
//    from (trunc x, trunc y);
//  * Bicubic, CatmullRom, Lanczos2 (4 taps) and Lanczos3 (6 taps):
//    interp_separable (affine/core.rs:648-697): rows outer, columns inner,
//    taps with a zero weight skipped, acc = fma(p, wx * wy, acc), then
//    acc / sum of weights;
//  * four bands: sample_pixel_at_with_alpha (affine/core.rs:988-1020), every
//    sample of a colour band multiplied by (its alpha / alpha_max) first, the
//    result divided by the interpolated alpha / alpha_max.
// Sources clamp to the input image (resolve_source_coord,
// affine/interpolation.rs:93-125); from_f64 stores (sample_conv.rs:58-93).
#include "device.hpp"

namespace vgpu {
namespace {

constexpr int kAffineRows = 16;

// NB = band count (1 .. 4), or 0 for any count (read from the job).  The
// weight product wx * wy and a tap's alpha / alpha_max are computed once per
// tap and shared by the bands: the same values the reference recomputes per
// band, so every band's chain of operations is unchanged.
template <typename T, int K, bool ALPHA, int NB>
__global__ __launch_bounds__(kThreads) void k_affine(AffineJob j) {
    const InView& in = j.in;
    const OutView& out = j.out;
    constexpr int CAP = NB ? NB : 1;
    const int B = NB ? NB : j.bands;
    // Affine::alpha_max (affine/core.rs:101-109): the type's maximum for
    // integer samples, 1 for f32 and f64.
    const double amax = is_float_sample<T> ? 1.0 : (double)std::numeric_limits<T>::max();
    const T* src = static_cast<const T*>(in.ptr);
    // Block = 64 columns x kAffineRows rows: a lane keeps its column's tap
    // record for all its rows; a row's record is the same for the wavefront.
    const int xo = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
    if (xo >= out.w) return;
    const int wv = (int)(threadIdx.x >> 6);
    const AffineTap tx = j.xs[xo];
    for (int yb = (int)blockIdx.y * kAffineRows; yb < out.h; yb += (int)gridDim.y * kAffineRows)
    for (int yo = yb + wv; yo < min(yb + kAffineRows, out.h); yo += kThreads / 64) {
        const size_t px = (size_t)yo * out.w + xo;
        const AffineTap ty = j.ys[yo];
        // Copy extend: clamp to the input image, then (through a folded Zoom,
        // which repeats pixel (x / zx, y / zy)) to the buffered rectangle.
        auto col = [&](int ix) { return col_of(in, clampi(ix, 0, j.in_w - 1) / j.zx); };
        auto row = [&](int iy) { return (size_t)row_of(in, clampi(iy, 0, j.in_h - 1) / j.zy) * (size_t)in.bw; };
        T* o = static_cast<T*>(out.ptr) + px * B;
        double v[CAP];
        if constexpr (NB == 0) {
            // any band count: one band at a time, as sample_at does
            for (int b = 0; b < B; ++b) {
                double r;
                if constexpr (K == AFF_NEAREST) {
                    r = (double)src[(row(ty.i0) + col(tx.i0)) * B + b];
                } else if constexpr (K == AFF_BILINEAR) {
                    const size_t r0 = row(ty.i0), r1 = row(ty.i0 + 1);
                    const int c0 = col(tx.i0), c1 = col(tx.i0 + 1);
                    const double fx = tx.f, fy = ty.f;
                    const double top = fma((double)src[(r0 + c1) * B + b], fx, (double)src[(r0 + c0) * B + b] * (1.0 - fx));
                    const double bot = fma((double)src[(r1 + c1) * B + b], fx, (double)src[(r1 + c0) * B + b] * (1.0 - fx));
                    r = top * (1.0 - fy) + bot * fy;
                } else {
                    constexpr int NT = K == AFF_SEP6 ? 6 : 4;
                    double acc = 0.0, wsum = 0.0;
                    for (int ky = 0; ky < NT; ++ky) {
                        const double wy = ty.w[ky];
                        if (wy == 0.0) continue;
                        const size_t rr = row(ty.i0 + ky);
                        for (int kx = 0; kx < NT; ++kx) {
                            const double wx = tx.w[kx];
                            if (wx == 0.0) continue;
                            const double w = wx * wy;
                            wsum += w;
                            acc = fma((double)src[(rr + col(tx.i0 + kx)) * B + b], w, acc);
                        }
                    }
                    r = wsum > 0.0 ? acc / wsum : 0.0;
                }
                o[b] = op_from_f64<T>(r);
            }
            continue;
        } else {
            // The samples of one pixel: a folded Premultiply node's output
            // (premultiply.rs:103-115), then the Affine's own premultiplied
            // sampling (sample_band_premultiplied) on four bands.
            auto load = [&](size_t r, int c, double* p) {
                const T* q = src + (r + c) * NB;
                uint32_t u[NB];
#pragma unroll
                for (int b = 0; b < NB; ++b) u[b] = (uint32_t)q[b];
                if constexpr (sizeof(T) == 1 && (NB == 2 || NB == 4)) {
                    if (j.premult) {
#pragma unroll
                        for (int b = 0; b + 1 < NB; ++b) u[b] = premult_u8(u[b], u[NB - 1]);
                    }
                }
                if constexpr (!std::is_same<T, uint8_t>::value) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) p[b] = (double)q[b];
                } else {
#pragma unroll
                    for (int b = 0; b < NB; ++b) p[b] = (double)u[b];
                }
                if constexpr (ALPHA) {
                    const double f = p[3] / amax;
#pragma unroll
                    for (int b = 0; b < 3; ++b) p[b] = p[b] * f;
                }
            };
            if constexpr (K == AFF_NEAREST) {
                load(row(ty.i0), col(tx.i0), v);
            } else if constexpr (K == AFF_BILINEAR) {
                const size_t r0 = row(ty.i0), r1 = row(ty.i0 + 1);
                const int c0 = col(tx.i0), c1 = col(tx.i0 + 1);
                double p00[NB], p01[NB], p10[NB], p11[NB];
                load(r0, c0, p00);
                load(r0, c1, p01);
                load(r1, c0, p10);
                load(r1, c1, p11);
                const double fx = tx.f, fy = ty.f;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const double top = fma(p01[b], fx, p00[b] * (1.0 - fx));
                    const double bot = fma(p11[b], fx, p10[b] * (1.0 - fx));
                    v[b] = top * (1.0 - fy) + bot * fy;
                }
            } else {
                constexpr int NT = K == AFF_SEP6 ? 6 : 4;
                int cols[NT];
#pragma unroll
                for (int kx = 0; kx < NT; ++kx) cols[kx] = col(tx.i0 + kx);
                double acc[NB], wsum = 0.0;
#pragma unroll
                for (int b = 0; b < NB; ++b) acc[b] = 0.0;
                for (int ky = 0; ky < NT; ++ky) {
                    const double wy = ty.w[ky];
                    const size_t rr = row(ty.i0 + ky);
                    // the row's samples first (clamped, always in the
                    // buffer), then the taps whose weights are non-zero
                    double p[NT][NB];
#pragma unroll
                    for (int kx = 0; kx < NT; ++kx) load(rr, cols[kx], p[kx]);
#pragma unroll
                    for (int kx = 0; kx < NT; ++kx) {
                        const double wx = tx.w[kx];
                        if (wy == 0.0 || wx == 0.0) continue;
                        const double w = wx * wy;
                        wsum += w;
#pragma unroll
                        for (int b = 0; b < NB; ++b) acc[b] = fma(p[kx][b], w, acc[b]);
                    }
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) v[b] = wsum > 0.0 ? acc[b] / wsum : 0.0;
            }
            T r[NB];
            if constexpr (ALPHA) {
                const double alpha = fmin(fmax(v[3], 0.0), amax);
                const double na = alpha / amax;
#pragma unroll
                for (int b = 0; b < 3; ++b) r[b] = op_from_f64<T>(na > 0.0 ? v[b] / na : 0.0);
                r[3] = op_from_f64<T>(alpha);
            } else {
#pragma unroll
                for (int b = 0; b < NB; ++b) r[b] = op_from_f64<T>(v[b]);
            }
            if constexpr (sizeof(T) == 1 && (NB == 2 || NB == 4)) {
                if (j.unpremult) {  // unpremultiply.rs:96-109
#pragma unroll
                    for (int b = 0; b + 1 < NB; ++b) r[b] = (T)unpremult_u8(r[b], r[NB - 1]);
                }
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) o[b] = r[b];
        }
    }
}

template <typename T, int K>
void launch_affine_t(const AffineJob& j, dim3 g, hipStream_t s) {
    switch (j.bands) {
    case 1: hipLaunchKernelGGL((k_affine<T, K, false, 1>), g, dim3(kThreads), 0, s, j); break;
    case 2: hipLaunchKernelGGL((k_affine<T, K, false, 2>), g, dim3(kThreads), 0, s, j); break;
    case 3: hipLaunchKernelGGL((k_affine<T, K, false, 3>), g, dim3(kThreads), 0, s, j); break;
    case 4: hipLaunchKernelGGL((k_affine<T, K, true, 4>), g, dim3(kThreads), 0, s, j); break;
    default: hipLaunchKernelGGL((k_affine<T, K, false, 0>), g, dim3(kThreads), 0, s, j); break;
    }
}

template <typename T>
hipError_t launch_affine_k(const AffineJob& j, dim3 g, hipStream_t s) {
    switch (j.kernel) {
    case AFF_NEAREST: launch_affine_t<T, AFF_NEAREST>(j, g, s); break;
    case AFF_BILINEAR: launch_affine_t<T, AFF_BILINEAR>(j, g, s); break;
    case AFF_SEP4: launch_affine_t<T, AFF_SEP4>(j, g, s); break;
    case AFF_SEP6: launch_affine_t<T, AFF_SEP6>(j, g, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_zoom(InView in, OutView out, int B, int zx, int zy) {
    const size_t n = (size_t)out.w * out.h * B;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const int b = (int)(i % B);
        const size_t px = i / B;
        const int x = out.x + (int)(px % out.w), y = out.y + (int)(px / out.w);
        static_cast<T*>(out.ptr)[i] = load_at<T>(in, row_of(in, y / zy), col_of(in, x / zx), B, b);
    }
}

}  // namespace

hipError_t launch_zoom(const InView& in, const OutView& out, int bands, int fmt, int zx, int zy, hipStream_t s) {
    if (out.w <= 0 || out.h <= 0) return hipSuccess;
    if (zx < 1 || zy < 1 || out.x < 0 || out.y < 0) return hipErrorInvalidValue;
    const size_t n = (size_t)out.w * out.h * bands;
    const unsigned g = (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 65536);
#define VGPU_ZOOM(T) hipLaunchKernelGGL(k_zoom<T>, dim3(g), dim3(kThreads), 0, s, in, out, bands, zx, zy)
    VGPU_FOR_FORMAT(fmt, VGPU_ZOOM)
#undef VGPU_ZOOM
    return hipGetLastError();
}

hipError_t launch_affine(const AffineJob& j, hipStream_t s) {
    if (j.out.w <= 0 || j.out.h <= 0) return hipSuccess;
    if (j.in_w <= 0 || j.in_h <= 0 || j.zx < 1 || j.zy < 1 || !j.xs || !j.ys) return hipErrorInvalidValue;
    const dim3 g((unsigned)((j.out.w + 63) / 64),
                 (unsigned)std::min<int64_t>(((int64_t)j.out.h + kAffineRows - 1) / kAffineRows, 65535));
#define VGPU_AFF(T) return launch_affine_k<T>(j, g, s)
    VGPU_FOR_FORMAT(j.fmt, VGPU_AFF)
#undef VGPU_AFF
}

}  // namespace vgpu
