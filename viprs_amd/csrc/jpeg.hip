// Baseline JPEG encoding after the path (the step process() runs after the
// pipeline, crates/viprs-runtime/src/process.rs:147-260, through the
// turbojpeg encoder, crates/viprs-codecs/src/jpeg/encode.rs): the per-block
// work on the device, the entropy coding on the host (codec.cpp).
//
// One thread per 8 x 8 block of one component.  The block's samples come
// from the u8 image with the edge replicated (the libjpeg-turbo encoder's
// right-edge expansion and bottom-row padding); three-band images go through
// the JFIF RGB -> YCbCr conversion in 16-bit fixed point; for 4:2:0 a chroma
// sample is the sum of its 2 x 2 converted samples plus a bias alternating
// 1, 2 along the row, shifted down by 2 (below the image's last row pair the
// downsampled row is repeated, as libjpeg pads the downsampled planes).  Then the integer ("islow") forward
// DCT of Loeffler, Ligtenberg and Moschytz with 13-bit constants and 2 extra
// bits between the passes, and quantisation by 8 x the table entry with the
// half-divisor rounding, each sign separately.  Coefficients are stored in
// zigzag order.  These are the published steps of ITU-T T.81 / JFIF as the
// Independent JPEG Group's library implements them.
#include "device.hpp"

namespace vgpu {
namespace {

__constant__ unsigned char kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                          12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                          35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                          58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// 16-bit fixed point of the JFIF conversion weights: round(w * 65536).
constexpr int32_t kFixY_R = 19595, kFixY_G = 38470, kFixY_B = 7471;
constexpr int32_t kFixCb_R = 11059, kFixCb_G = 21709, kFixHalf = 32768;
constexpr int32_t kFixCr_G = 27439, kFixCr_B = 5329;
constexpr int32_t kCbCrOffset = 128 << 16;

__device__ __forceinline__ int32_t conv_y(int32_t r, int32_t g, int32_t b) {
    return (kFixY_R * r + kFixY_G * g + kFixY_B * b + (1 << 15)) >> 16;
}
__device__ __forceinline__ int32_t conv_cb(int32_t r, int32_t g, int32_t b) {
    return (-kFixCb_R * r - kFixCb_G * g + kFixHalf * b + kCbCrOffset + (1 << 15) - 1) >> 16;
}
__device__ __forceinline__ int32_t conv_cr(int32_t r, int32_t g, int32_t b) {
    return (kFixHalf * r - kFixCr_G * g - kFixCr_B * b + kCbCrOffset + (1 << 15) - 1) >> 16;
}

// Forward DCT constants: round(c * 2^13).
constexpr int kConstBits = 13, kPass1Bits = 2;
constexpr int32_t kF0_298 = 2446, kF0_390 = 3196, kF0_541 = 4433, kF0_765 = 6270, kF0_899 = 7373,
                  kF1_175 = 9633, kF1_501 = 12299, kF1_847 = 15137, kF1_961 = 16069, kF2_053 = 16819,
                  kF2_562 = 20995, kF3_072 = 25172;

__device__ __forceinline__ int32_t descale(int32_t x, int n) { return (x + (1 << (n - 1))) >> n; }

// One 1-D pass over v[0..7] (stride s in d): pass 0 keeps kPass1Bits more,
// pass 1 removes them.
template <int PASS>
__device__ __forceinline__ void fdct_1d(int32_t* d, int s) {
    const int32_t t0 = d[0] + d[7 * s], t7 = d[0] - d[7 * s];
    const int32_t t1 = d[s] + d[6 * s], t6 = d[s] - d[6 * s];
    const int32_t t2 = d[2 * s] + d[5 * s], t5 = d[2 * s] - d[5 * s];
    const int32_t t3 = d[3 * s] + d[4 * s], t4 = d[3 * s] - d[4 * s];
    const int32_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
    constexpr int sh = PASS == 0 ? kConstBits - kPass1Bits : kConstBits + kPass1Bits;
    if (PASS == 0) {
        d[0] = (t10 + t11) << kPass1Bits;
        d[4 * s] = (t10 - t11) << kPass1Bits;
    } else {
        d[0] = descale(t10 + t11, kPass1Bits);
        d[4 * s] = descale(t10 - t11, kPass1Bits);
    }
    const int32_t z1e = (t12 + t13) * kF0_541;
    d[2 * s] = descale(z1e + t13 * kF0_765, sh);
    d[6 * s] = descale(z1e - t12 * kF1_847, sh);
    // odd part
    int32_t z1 = t4 + t7, z2 = t5 + t6, z3 = t4 + t6, z4 = t5 + t7;
    const int32_t z5 = (z3 + z4) * kF1_175;
    const int32_t a4 = t4 * kF0_298, a5 = t5 * kF2_053, a6 = t6 * kF3_072, a7 = t7 * kF1_501;
    z1 *= -kF0_899;
    z2 *= -kF2_562;
    z3 = z3 * -kF1_961 + z5;
    z4 = z4 * -kF0_390 + z5;
    d[7 * s] = descale(a4 + z1 + z3, sh);
    d[5 * s] = descale(a5 + z2 + z4, sh);
    d[3 * s] = descale(a6 + z2 + z3, sh);
    d[s] = descale(a7 + z1 + z4, sh);
}

__global__ __launch_bounds__(kThreads) void k_jpeg_blocks(JpegJob j) {
    const int t = (int)(blockIdx.x * kThreads + threadIdx.x);
    const int ncomp = j.mode == 0 ? 1 : 3;
    if (t >= j.off[ncomp]) return;
    const int c = t < j.off[1] || ncomp == 1 ? 0 : (t < j.off[2] ? 1 : 2);
    const int k = t - j.off[c];
    const int bx = k % j.bw[c], by = k / j.bw[c];
    const unsigned char* px = j.px;
    const int W = j.w, H = j.h, B = j.bands;
    int32_t d[64];
    auto pix = [&](int y, int x, int32_t* r, int32_t* g, int32_t* b) {
        const unsigned char* p = px + ((size_t)min(y, H - 1) * W + min(x, W - 1)) * B;
        *r = p[0];
        *g = p[1];
        *b = p[2];
    };
    for (int r = 0; r < 8; ++r) {
        for (int q = 0; q < 8; ++q) {
            int32_t v;
            if (j.mode == 0) {
                v = px[(size_t)min(8 * by + r, H - 1) * W + min(8 * bx + q, W - 1)];
            } else if (c == 0 || j.mode == 1) {
                int32_t R, G, Bv;
                pix(8 * by + r, 8 * bx + q, &R, &G, &Bv);
                v = c == 0 ? conv_y(R, G, Bv) : (c == 1 ? conv_cb(R, G, Bv) : conv_cr(R, G, Bv));
            } else {
                // Columns are replicated before the 2 x 2 sums, rows after:
                // chroma rows below the image's last row pair repeat that
                // pair's row.
                const int cy = min(8 * by + r, (H + 1) / 2 - 1);
                int32_t sum = 0;
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        int32_t R, G, Bv;
                        pix(2 * cy + dy, 2 * (8 * bx + q) + dx, &R, &G, &Bv);
                        sum += c == 1 ? conv_cb(R, G, Bv) : conv_cr(R, G, Bv);
                    }
                v = (sum + 1 + (q & 1)) >> 2;  // bias 1, 2, 1, 2 ... along the row
            }
            d[r * 8 + q] = v - 128;
        }
    }
    for (int r = 0; r < 8; ++r) fdct_1d<0>(d + 8 * r, 1);
    for (int q = 0; q < 8; ++q) fdct_1d<1>(d + q, 8);
    const uint16_t* div = j.div[c == 0 ? 0 : 1];
    int16_t* o = j.out + (size_t)t * 64;
    for (int z = 0; z < 64; ++z) {
        const int n = kZigzag[z];
        const int32_t x = d[n], dv = div[n];
        o[z] = (int16_t)(x < 0 ? -((-x + (dv >> 1)) / dv) : (x + (dv >> 1)) / dv);
    }
}

}  // namespace

hipError_t launch_jpeg_blocks(const JpegJob& j, hipStream_t s) {
    const int total = j.off[j.mode == 0 ? 1 : 3];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jpeg_blocks, dim3((unsigned)((total + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, j);
    return hipGetLastError();
}

}  // namespace vgpu
