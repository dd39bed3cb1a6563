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
//  * VSQBS and LBB: the 4 x 4 stencil from (floor x - 1, floor y - 1) and
//    the phases x - floor x, y - floor y (interp_vsqbs, interp_lbb,
//    affine/core.rs:588-610,699-725);
//  * four bands: sample_pixel_at_with_alpha (affine/core.rs:988-1020), every
//    sample of a colour band multiplied by (its alpha / alpha_max) first, the
//    result divided by the interpolated alpha / alpha_max.
// Sources clamp to the input image (resolve_source_coord,
// affine/interpolation.rs:93-125); from_f64 stores (sample_conv.rs:58-93).
#include "device.hpp"

namespace vgpu {
namespace {

constexpr int kAffineRows = 16;

// VSQBS over a 4 x 4 neighbourhood nb[row][col] whose [1][1] sample sits at
// (floor x, floor y) (interpolate_vsqbs_2d, kernel.rs:863-923): per axis the
// phase picks the anchor (column / row 1 below one half, else 2 with the
// stencil mirrored) and the distance x0 to it; from the vertex-split
// quadratic B-spline weights of 2|x0| - 1/2 eight taps around the anchor are
// weighted and summed in pairs, times 1/4.
__device__ __forceinline__ double vsqbs_2d(double xp, double yp, const double (&nb)[4][4]) {
    const bool xlo = xp < 0.5, ylo = yp < 0.5;
    const int ax = xlo ? 1 : 2, sx = xlo ? 1 : -1;
    const int ay = ylo ? 1 : 2, sy = ylo ? 1 : -1;
    const double x0 = xlo ? xp : xp - 1.0, y0 = ylo ? yp : yp - 1.0;
    const double x = (double)(2 * sx) * x0 - 0.5;
    const double y = (double)(2 * sy) * y0 - 0.5;
    const double cent = 0.75 - x * x;
    const double mid = 0.75 - y * y;
    const double left = fma(-0.5, x + cent, 0.5);
    const double top = fma(-0.5, y + mid, 0.5);
    const double left_cent = left + cent, top_mid = top + mid;
    const double cent_right = 1.0 - left, mid_bottom = 1.0 - top;
    const double right = 1.0 - left_cent, bottom = 1.0 - top_mid;
    // four times the weights of: (top, mid) (mid, left) (mid, mid)
    // (mid, right) (bottom, mid) (bottom, right) (top, right) (bottom, left)
    const double w_tm = left_cent * top;
    const double w_ml = left * top_mid;
    const double w_mm = left_cent + top_mid;
    const double w_mr = cent_right * top_mid + right;
    const double w_bm = mid_bottom * left_cent + bottom;
    const double w_br = mid_bottom * right + cent_right * bottom;
    const double w_tr = top - w_tm;
    const double w_bl = left - w_ml;
    const int rt = ay - sy, rm = ay, rb = ay + sy;
    const int cl = ax - sx, cm = ax, cr = ax + sx;
    return (((w_tm * nb[rt][cm] + w_ml * nb[rm][cl]) + (w_mm * nb[rm][cm] + w_mr * nb[rm][cr])) +
            ((w_bm * nb[rb][cm] + w_br * nb[rb][cr]) + (w_tr * nb[rt][cr] + w_bl * nb[rb][cl]))) *
           0.25;
}

// LBB, the locally bounded bicubic Hermite interpolator over s[row][col]
// with [1][1] at (floor x, floor y) and phases (rx, ry) (lbbicubic,
// affine/interpolation.rs:170-441).  Per corner c of the central 2 x 2 cell
// (00, 10, 01, 11: x then y): the local min / max over the corner, its
// neighbours and two arm samples; central-difference slopes limited to six
// times the distance to those bounds; the cross derivative clamped by four
// limits; then the Hermite basis of (rx, ry) over value, slopes and cross
// derivative, each of the three sums in corner order.
__device__ __forceinline__ double lbb_min(double a, double b) { return a <= b ? a : b; }
__device__ __forceinline__ double lbb_max(double a, double b) { return a >= b ? a : b; }
__device__ __forceinline__ double lbb_sign(double a) { return a >= 0.0 ? 1.0 : -1.0; }

__device__ __forceinline__ double lbb_2d(const double (&s)[4][4], double rx, double ry) {
    // the central pair of each middle row, ordered
    const bool r1 = s[1][1] <= s[1][2], r2 = s[2][1] <= s[2][2];
    const double lo1 = r1 ? s[1][1] : s[1][2], hi1 = r1 ? s[1][2] : s[1][1];
    const double lo2 = r2 ? s[2][1] : s[2][2], hi2 = r2 ? s[2][2] : s[2][1];
    double z[4], u[4], v[4], dxi[4], dyi[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int cx = c & 1, cy = c >> 1;
        // the row pair of the corner's row, the middle-row sample below or
        // above it, and the corner's two arm samples (a <= b picks a)
        const double rlo = cy ? lo2 : lo1, rhi = cy ? hi2 : hi1;
        const double other = s[2 - cy][1 + cx];
        const double a = cy ? s[2][cx ? 3 : 0] : s[0][1 + cx];
        const double b = cy ? s[3][1 + cx] : s[1][cx ? 3 : 0];
        const double alo = a <= b ? a : b, ahi = a <= b ? b : a;
        z[c] = s[1 + cy][1 + cx];
        u[c] = z[c] - lbb_min(lbb_min(rlo, other), alo);
        v[c] = lbb_max(lbb_max(rhi, other), ahi) - z[c];
        dxi[c] = s[1 + cy][2 + cx] - s[1 + cy][cx];
        dyi[c] = s[2 + cy][1 + cx] - s[cy][1 + cx];
    }
    double dx[4], dy[4], q[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int cx = c & 1, cy = c >> 1;
        const double qi = cy ? (s[3][cx + 2] - s[3][cx]) - dxi[cx] : (s[0][cx] - s[0][cx + 2]) + dxi[2 + cx];
        const double lim = 6.0 * lbb_min(u[c], v[c]);
        const double sgx = lbb_sign(dxi[c]), sgy = lbb_sign(dyi[c]);
        dx[c] = sgx * dxi[c] <= lim ? dxi[c] : sgx * lim;
        dy[c] = sgy * dyi[c] <= lim ? dyi[c] : sgy * lim;
        const double ts = fabs(6.0 * (dx[c] + dy[c]));
        const double td = fabs(6.0 * (dx[c] - dy[c]));
        const double u36 = 36.0 * u[c], v36 = 36.0 * v[c];
        double qq = lbb_max(qi, ts - u36);
        qq = lbb_min(qq, v36 - ts);
        qq = lbb_max(qq, td - v36);
        q[c] = lbb_min(qq, u36 - td);
    }
    // Hermite basis: per axis the value weight 2(1/2 + x) or 2(3/2 - x), the
    // slope weight x or x - 1, and the squared distance to the far side.
    const double xp = rx, xm = xp - 1.0, yp = ry, ym = yp - 1.0;
    const double opx = 0.5 + xp, omx = 1.5 - xp, opy = 0.5 + yp, omy = 1.5 - yp;
    const double xpsq = xp * xp, ypsq = yp * yp, xmsq = xm * xm, ymsq = ym * ym;
    const double tpx = opx + opx, tpy = opy + opy, tmx = omx + omx, tmy = omy + omy;
    double cv[4], cdx[4], cdy[4], cxy[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int cx = c & 1, cy = c >> 1;
        const double tX = cx ? tmx : tpx, sX = cx ? xm : xp, qX = cx ? xpsq : xmsq;
        const double tY = cy ? tmy : tpy, sY = cy ? ym : yp, qY = cy ? ypsq : ymsq;
        const double sq = qX * qY;
        cv[c] = (tX * tY) * sq;
        cdx[c] = (sX * tY) * sq;
        cdy[c] = (tX * sY) * sq;
        cxy[c] = (sX * sY) * sq;
    }
    const double n1 = cv[0] * z[0] + cv[1] * z[1] + cv[2] * z[2] + cv[3] * z[3];
    const double n2 = cdx[0] * dx[0] + cdx[1] * dx[1] + cdx[2] * dx[2] + cdx[3] * dx[3] + cdy[0] * dy[0] +
                      cdy[1] * dy[1] + cdy[2] * dy[2] + cdy[3] * dy[3];
    const double n3 = cxy[0] * q[0] + cxy[1] * q[1] + cxy[2] * q[2] + cxy[3] * q[3];
    return fma(0.25, n3, fma(0.5, n2, n1));
}

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
                } else if constexpr (K == AFF_VSQBS || K == AFF_LBB) {
                    double nb[4][4];
                    for (int ky = 0; ky < 4; ++ky) {
                        const size_t rr = row(ty.i0 + ky);
                        for (int kx = 0; kx < 4; ++kx) nb[ky][kx] = (double)src[(rr + col(tx.i0 + kx)) * B + b];
                    }
                    r = K == AFF_VSQBS ? vsqbs_2d(tx.f, ty.f, nb) : lbb_2d(nb, tx.f, ty.f);
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
            } else if constexpr (K == AFF_VSQBS || K == AFF_LBB) {
                // the 4 x 4 stencil of every band (premultiplied samples
                // for four bands), then each band's interpolation
                double p[4][4][NB];
#pragma unroll
                for (int ky = 0; ky < 4; ++ky) {
                    const size_t rr = row(ty.i0 + ky);
#pragma unroll
                    for (int kx = 0; kx < 4; ++kx) load(rr, col(tx.i0 + kx), p[ky][kx]);
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    double nb[4][4];
#pragma unroll
                    for (int ky = 0; ky < 4; ++ky)
#pragma unroll
                        for (int kx = 0; kx < 4; ++kx) nb[ky][kx] = p[ky][kx][b];
                    v[b] = K == AFF_VSQBS ? vsqbs_2d(tx.f, ty.f, nb) : lbb_2d(nb, tx.f, ty.f);
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
    case AFF_VSQBS: launch_affine_t<T, AFF_VSQBS>(j, g, s); break;
    case AFF_LBB: launch_affine_t<T, AFF_LBB>(j, g, s); break;
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
