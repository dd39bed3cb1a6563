// Encode after the path and the bytes-in / bytes-out helper
// (crates/viprs-runtime/src/process.rs:147-260, 289-360):
//  * JPEG: the device computes every block's quantised coefficients
//    (jpeg.hip); here the host writes the JFIF stream -- the tables of
//    ITU-T T.81 Annex K (quality-scaled quantisation, the standard Huffman
//    tables), markers, the sequential Huffman pass with the DC predictors and
//    the libjpeg-turbo encoder's dummy blocks at the right and bottom edges
//    of 4:2:0 images (DC of the neighbour the encoder copies, AC zero);
//  * PNG: decode (grey, grey + alpha, RGB, RGBA, palette; bit depths 1, 2,
//    4 and 8 expanded to u8, tRNS to alpha, Adam7 interlacing undone) and
//    encode with zlib, lossless;
//  * vgpu_process: decode, the caller's pipeline on the device, encode.
#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.hpp"

using namespace vgpu;

namespace {

const unsigned char kZZ[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                               12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                               35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                               58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
// T.81 Annex K.1, natural order.
const unsigned char kLumaQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                                  14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                                  18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                                  49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const unsigned char kChromaQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                                    24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                                    99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                                    99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
// T.81 Annex K.3: code counts by length 1..16, then the symbols.
const unsigned char kDcLumaBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const unsigned char kDcChromaBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const unsigned char kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const unsigned char kAcLumaBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const unsigned char kAcLumaVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const unsigned char kAcChromaBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const unsigned char kAcChromaVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

struct HuffTable {
    const unsigned char* bits;
    const unsigned char* vals;
    int nvals;
    uint16_t code[256] = {};
    uint8_t len[256] = {};
    HuffTable(const unsigned char* b, const unsigned char* v, int n) : bits(b), vals(v), nvals(n) {
        // canonical codes, shortest first (T.81 C.2)
        uint32_t c = 0;
        int k = 0;
        for (int l = 1; l <= 16; ++l) {
            for (int i = 0; i < bits[l - 1]; ++i, ++k) {
                code[vals[k]] = (uint16_t)c++;
                len[vals[k]] = (uint8_t)l;
            }
            c <<= 1;
        }
    }
};

// JPEG quality scaling: q < 50 -> 5000 / q %, else 200 - 2q %; entries
// rounded, kept in 1..255 (baseline).
void quality_table(const unsigned char* base, int q, unsigned char* out) {
    const long s = q < 50 ? 5000 / q : 200 - 2L * q;
    for (int i = 0; i < 64; ++i) out[i] = (unsigned char)std::min(255L, std::max(1L, (base[i] * s + 50) / 100));
}

class BitWriter {
  public:
    explicit BitWriter(std::vector<unsigned char>& out) : out_(out) {}
    void put(uint32_t bits, int n) {
        acc_ = (acc_ << n) | (bits & ((1u << n) - 1u));
        n_ += n;
        while (n_ >= 8) {
            const unsigned char b = (unsigned char)(acc_ >> (n_ - 8));
            out_.push_back(b);
            if (b == 0xff) out_.push_back(0);  // byte stuffing
            n_ -= 8;
        }
        acc_ &= (1ull << n_) - 1ull;
    }
    void flush() {
        if (n_) put((1u << (8 - n_)) - 1u, 8 - n_);  // pad with 1 bits
    }

  private:
    std::vector<unsigned char>& out_;
    uint64_t acc_ = 0;
    int n_ = 0;
};

int magnitude_bits(int v) {
    unsigned a = (unsigned)(v < 0 ? -v : v);
    int n = 0;
    while (a) {
        ++n;
        a >>= 1;
    }
    return n;
}

// One block in zigzag order: DC difference, then run/size AC codes with ZRL
// and EOB (T.81 F.1.2).
void encode_block(BitWriter& bw, const int16_t* zz, int* pred, const HuffTable& dc, const HuffTable& ac) {
    const int diff = zz[0] - *pred;
    *pred = zz[0];
    int n = magnitude_bits(diff);
    bw.put(dc.code[n], dc.len[n]);
    if (n) bw.put((uint32_t)(diff < 0 ? diff - 1 : diff), n);
    int run = 0;
    for (int k = 1; k < 64; ++k) {
        const int v = zz[k];
        if (v == 0) {
            ++run;
            continue;
        }
        while (run > 15) {
            bw.put(ac.code[0xf0], ac.len[0xf0]);
            run -= 16;
        }
        n = magnitude_bits(v);
        const int sym = (run << 4) | n;
        bw.put(ac.code[sym], ac.len[sym]);
        bw.put((uint32_t)(v < 0 ? v - 1 : v), n);
        run = 0;
    }
    if (run) bw.put(ac.code[0], ac.len[0]);
}

void put16(std::vector<unsigned char>& o, int v) {
    o.push_back((unsigned char)(v >> 8));
    o.push_back((unsigned char)v);
}

void put_dht(std::vector<unsigned char>& o, int cls, int id, const HuffTable& t) {
    o.push_back(0xff);
    o.push_back(0xc4);
    put16(o, 3 + 16 + t.nvals);
    o.push_back((unsigned char)(cls << 4 | id));
    o.insert(o.end(), t.bits, t.bits + 16);
    o.insert(o.end(), t.vals, t.vals + t.nvals);
}

// The JFIF stream around the entropy-coded coefficients (one component: grey;
// else Y, Cb, Cr with Y sampled 2 x 2 for 4:2:0).
int jpeg_stream(const std::vector<int16_t>& coef, const JpegJob& j, const unsigned char* ql,
                const unsigned char* qc, std::vector<unsigned char>* out) {
    static const HuffTable dcl(kDcLumaBits, kDcVals, 12), acl(kAcLumaBits, kAcLumaVals, 162);
    static const HuffTable dcc(kDcChromaBits, kDcVals, 12), acc(kAcChromaBits, kAcChromaVals, 162);
    const bool gray = j.mode == 0;
    std::vector<unsigned char>& o = *out;
    o.clear();
    o.reserve(coef.size() / 2 + 1024);
    const unsigned char soi_app0[] = {0xff, 0xd8, 0xff, 0xe0, 0x00, 0x10, 'J', 'F', 'I', 'F', 0x00,
                                      0x01, 0x01, 0x00, 0x00, 0x01, 0x00, 0x01, 0x00, 0x00};
    o.insert(o.end(), soi_app0, soi_app0 + sizeof(soi_app0));
    for (int t = 0; t < (gray ? 1 : 2); ++t) {
        const unsigned char* q = t == 0 ? ql : qc;
        o.push_back(0xff);
        o.push_back(0xdb);
        put16(o, 67);
        o.push_back((unsigned char)t);
        for (int i = 0; i < 64; ++i) o.push_back(q[kZZ[i]]);
    }
    o.push_back(0xff);
    o.push_back(0xc0);
    put16(o, gray ? 11 : 17);
    o.push_back(8);
    put16(o, j.h);
    put16(o, j.w);
    if (gray) {
        const unsigned char c[] = {1, 1, 0x11, 0};
        o.insert(o.end(), c, c + 4);
    } else {
        const unsigned char c[] = {3, 1, (unsigned char)(j.mode == 2 ? 0x22 : 0x11), 0, 2, 0x11, 1, 3, 0x11, 1};
        o.insert(o.end(), c, c + 10);
    }
    put_dht(o, 0, 0, dcl);
    put_dht(o, 1, 0, acl);
    if (!gray) {
        put_dht(o, 0, 1, dcc);
        put_dht(o, 1, 1, acc);
    }
    o.push_back(0xff);
    o.push_back(0xda);
    if (gray) {
        put16(o, 8);
        const unsigned char c[] = {1, 1, 0x00, 0, 63, 0};
        o.insert(o.end(), c, c + 6);
    } else {
        put16(o, 12);
        const unsigned char c[] = {3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0};
        o.insert(o.end(), c, c + 10);
    }
    BitWriter bw(o);
    int pred[3] = {0, 0, 0};
    auto blk = [&](int c, int by, int bx) { return coef.data() + ((size_t)j.off[c] + (size_t)by * j.bw[c] + bx) * 64; };
    if (j.mode != 2) {
        for (int by = 0; by < j.bh[0]; ++by)
            for (int bx = 0; bx < j.bw[0]; ++bx) {
                encode_block(bw, blk(0, by, bx), &pred[0], dcl, acl);
                if (!gray) {
                    encode_block(bw, blk(1, by, bx), &pred[1], dcc, acc);
                    encode_block(bw, blk(2, by, bx), &pred[2], dcc, acc);
                }
            }
    } else {
        const int mcux = j.bw[1], mcuy = j.bh[1];
        int16_t dummy[64];
        for (int my = 0; my < mcuy; ++my)
            for (int mx = 0; mx < mcux; ++mx) {
                // The MCU's four luma blocks; one outside the image is a
                // dummy: AC zero, DC of the block on its left (right edge) or
                // of the MCU's top-right block (bottom edge).
                int16_t dc[2][2] = {};
                for (int by = 0; by < 2; ++by)
                    for (int bx = 0; bx < 2; ++bx) {
                        const int gy = 2 * my + by, gx = 2 * mx + bx;
                        const int16_t* b;
                        if (gy < j.bh[0] && gx < j.bw[0]) {
                            b = blk(0, gy, gx);
                        } else {
                            std::memset(dummy, 0, sizeof(dummy));
                            dummy[0] = gy < j.bh[0] ? dc[by][bx - 1] : dc[0][1];
                            b = dummy;
                        }
                        dc[by][bx] = b[0];
                        encode_block(bw, b, &pred[0], dcl, acl);
                    }
                encode_block(bw, blk(1, my, mx), &pred[1], dcc, acc);
                encode_block(bw, blk(2, my, mx), &pred[2], dcc, acc);
            }
    }
    bw.flush();
    o.push_back(0xff);
    o.push_back(0xd9);
    return VGPU_OK;
}

int copy_out(const std::vector<unsigned char>& v, void* out, size_t cap, size_t* len) {
    if (len) *len = v.size();
    if (!out || cap < v.size()) return VGPU_E_BUFFER_SIZE;
    std::memcpy(out, v.data(), v.size());
    return VGPU_OK;
}

// Device pixels -> JPEG bytes.
int jpeg_encode_device(vgpu_context* ctx, const void* pixels, uint32_t w, uint32_t h, uint32_t bands, int quality,
                       int subsampling, std::vector<unsigned char>* out) {
    // turbojpeg_quality (jpeg/common.rs:135-145): 0..100, 0 meaning 1.
    if (quality < 0 || quality > 100) return VGPU_E_INVALID_ARGUMENT;
    if (w == 0 || h == 0 || w > 65535 || h > 65535) return VGPU_E_INVALID_ARGUMENT;
    // prepare_encode_input (jpeg/encode.rs:122-158): grey, RGB, or RGBA with
    // the alpha dropped (the kernel reads the first three of `bands`).
    if (bands != 1 && bands != 3 && bands != 4) return VGPU_E_UNSUPPORTED_FORMAT;
    if (subsampling < VGPU_JPEG_SUBSAMPLE_AUTO || subsampling > VGPU_JPEG_SUBSAMPLE_420) return VGPU_E_INVALID_ARGUMENT;
    JpegJob j;
    j.px = static_cast<const unsigned char*>(pixels);
    j.w = (int)w;
    j.h = (int)h;
    j.bands = (int)bands;
    // JpegSubsampling::Auto: 4:2:0 below quality 90, else none (jpeg/common.rs:115-131)
    const bool s420 = subsampling == VGPU_JPEG_SUBSAMPLE_420 || (subsampling == VGPU_JPEG_SUBSAMPLE_AUTO && quality < 90);
    j.mode = bands == 1 ? 0 : (s420 ? 2 : 1);
    const int ncomp = j.mode == 0 ? 1 : 3;
    j.bw[0] = ((int)w + 7) / 8;
    j.bh[0] = ((int)h + 7) / 8;
    for (int c = 1; c < ncomp; ++c) {
        j.bw[c] = j.mode == 2 ? ((int)w + 15) / 16 : j.bw[0];
        j.bh[c] = j.mode == 2 ? ((int)h + 15) / 16 : j.bh[0];
    }
    j.off[0] = 0;
    for (int c = 0; c < ncomp; ++c) j.off[c + 1] = j.off[c] + j.bw[c] * j.bh[c];
    unsigned char ql[64], qc[64];
    quality_table(kLumaQ, std::max(quality, 1), ql);
    quality_table(kChromaQ, std::max(quality, 1), qc);
    for (int i = 0; i < 64; ++i) {
        j.div[0][i] = (uint16_t)(ql[i] << 3);
        j.div[1][i] = (uint16_t)(qc[i] << 3);
    }
    const size_t nblk = (size_t)j.off[ncomp];
    (void)hipSetDevice(ctx->device);
    void* d_coef = ctx->pool.acquire(nblk * 64 * sizeof(int16_t));
    if (!d_coef) return VGPU_E_NO_MEMORY;
    j.out = static_cast<int16_t*>(d_coef);
    std::vector<int16_t> coef(nblk * 64);
    hipError_t e = launch_jpeg_blocks(j, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(coef.data(), d_coef, coef.size() * sizeof(int16_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->sticky = VGPU_E_DEVICE;  // the buffer may still be in use: keep it
    ctx->pool.release(d_coef);
    return jpeg_stream(coef, j, ql, qc, out);
}

// ------------------------------------------------------------------ PNG

uint32_t be32(const unsigned char* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put_chunk(std::vector<unsigned char>& o, const char* type, const unsigned char* data, size_t n) {
    const unsigned char len[4] = {(unsigned char)(n >> 24), (unsigned char)(n >> 16), (unsigned char)(n >> 8),
                                  (unsigned char)n};
    o.insert(o.end(), len, len + 4);
    const size_t at = o.size();
    o.insert(o.end(), type, type + 4);
    if (n) o.insert(o.end(), data, data + n);
    const uLong crc = crc32(0L, o.data() + at, (uInt)(n + 4));
    const unsigned char c[4] = {(unsigned char)(crc >> 24), (unsigned char)(crc >> 16), (unsigned char)(crc >> 8),
                                (unsigned char)crc};
    o.insert(o.end(), c, c + 4);
}

const unsigned char kPngSig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};

// Undo the five PNG row filters of one (sub-)image of `rows` rows of `stride`
// bytes (bpp = bytes per complete pixel, at least 1) in place: `raw` holds
// each row's filter byte followed by its bytes; `out` receives the rows.
bool png_unfilter(const unsigned char* raw, size_t rows, size_t stride, size_t bpp, unsigned char* out) {
    for (size_t y = 0; y < rows; ++y) {
        const unsigned char f = raw[y * (stride + 1)];
        const unsigned char* src = raw + y * (stride + 1) + 1;
        unsigned char* cur = out + y * stride;
        const unsigned char* up = y ? cur - stride : nullptr;
        if (f > 4) return false;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v = 0;
            if (f == 1) {
                v = a;
            } else if (f == 2) {
                v = b;
            } else if (f == 3) {
                v = (a + b) >> 1;
            } else if (f == 4) {
                const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                v = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
            }
            cur[i] = (unsigned char)(src[i] + v);
        }
    }
    return true;
}

// The PNG decode the reference runs for a U8 request: the png crate with
// Transformations::EXPAND (png/decode_full.rs:101-108, 470-520): bit depths
// 1, 2, 4 expanded to 8 (grey scaled by 255 / (2^depth - 1)), palette to RGB,
// a tRNS chunk to an alpha band (palette alpha per entry; grey and RGB alpha
// 0 where the sample equals the key, else 255), Adam7 interlacing undone;
// 16-bit images are refused for U8 ("format/bit-depth mismatch").  Chunk CRCs
// are checked.
int png_decode(const unsigned char* d, size_t n, uint32_t* w, uint32_t* h, uint32_t* bands,
               std::vector<unsigned char>* pixels) {
    if (n < 8 + 25 || std::memcmp(d, kPngSig, 8) != 0) return VGPU_E_UNSUPPORTED_FORMAT;
    size_t p = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    bool have_trns = false, seen_end = false;
    std::vector<unsigned char> idat, plte, trns;
    while (p + 12 <= n) {
        const uint32_t len = be32(d + p);
        if (len > n - p - 12) return VGPU_E_INVALID_ARGUMENT;
        const unsigned char* type = d + p + 4;
        const unsigned char* body = d + p + 8;
        if (crc32(crc32(0L, Z_NULL, 0), type, len + 4) != be32(body + len)) return VGPU_E_INVALID_ARGUMENT;
        if (p == 8 && (std::memcmp(type, "IHDR", 4) != 0 || len != 13)) return VGPU_E_INVALID_ARGUMENT;
        if (!std::memcmp(type, "IHDR", 4)) {
            W = be32(body);
            H = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            if (body[10] != 0 || body[11] != 0 || body[12] > 1) return VGPU_E_INVALID_ARGUMENT;
            interlace = body[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
            have_trns = true;
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            seen_end = true;
            break;
        }
        p += 12 + (size_t)len;
    }
    if (!seen_end) return VGPU_E_INVALID_ARGUMENT;
    const int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    const bool depth_ok = depth == 8 || depth == 16 || ((ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4));
    if (W == 0 || H == 0 || ch == 0 || !depth_ok || (ctype == 3 && depth == 16)) return VGPU_E_INVALID_ARGUMENT;
    if (depth == 16) return VGPU_E_UNSUPPORTED_FORMAT;  // U8 decode of a 16-bit file
    if ((uint64_t)W * H > ((uint64_t)1 << 31)) return VGPU_E_INVALID_ARGUMENT;
    if (ctype == 3 && (plte.empty() || plte.size() % 3 != 0)) return VGPU_E_INVALID_ARGUMENT;
    if (have_trns && (ctype == 4 || ctype == 6)) have_trns = false;  // not allowed there: ignored
    if (have_trns && ((ctype == 0 && trns.size() < 2) || (ctype == 2 && trns.size() < 6)))
        return VGPU_E_INVALID_ARGUMENT;
    // Adam7 passes: origin and step in x and y; one pass when not interlaced.
    static const int kPass[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                    {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int npass = interlace ? 7 : 1;
    const size_t bits_px = (size_t)ch * depth, bpp = std::max<size_t>(1, bits_px / 8);
    // Pass k's sub-image size (0 x 0 when the image is too small for it).
    auto pass_size = [&](int k, size_t* pw, size_t* ph) {
        const uint32_t x0 = interlace ? kPass[k][0] : 0, y0 = interlace ? kPass[k][1] : 0;
        const uint32_t dx = interlace ? kPass[k][2] : 1, dy = interlace ? kPass[k][3] : 1;
        *pw = x0 < W ? (W - x0 + dx - 1) / dx : 0;
        *ph = y0 < H ? (H - y0 + dy - 1) / dy : 0;
        if (!*pw || !*ph) *pw = *ph = 0;
    };
    size_t total = 0;
    for (int k = 0; k < npass; ++k) {
        size_t pw, ph;
        pass_size(k, &pw, &ph);
        total += ph * ((pw * bits_px + 7) / 8 + 1);
    }
    std::vector<unsigned char> raw(total);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) return VGPU_E_NO_MEMORY;
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const bool full = zs.avail_out == 0;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_OK && zr != Z_BUF_ERROR) || !full) return VGPU_E_INVALID_ARGUMENT;
    // Samples at the file's depth, one byte per channel, in image order.
    std::vector<uint8_t> samp((size_t)W * H * ch);
    size_t at = 0;
    std::vector<unsigned char> rows;
    for (int k = 0; k < npass; ++k) {
        const uint32_t x0 = interlace ? kPass[k][0] : 0, y0 = interlace ? kPass[k][1] : 0;
        const uint32_t dx = interlace ? kPass[k][2] : 1, dy = interlace ? kPass[k][3] : 1;
        size_t pw, ph;
        pass_size(k, &pw, &ph);
        if (!pw) continue;
        const size_t stride = (pw * bits_px + 7) / 8;
        rows.resize(stride * ph);
        if (!png_unfilter(raw.data() + at, ph, stride, bpp, rows.data())) return VGPU_E_INVALID_ARGUMENT;
        at += ph * (stride + 1);
        for (size_t y = 0; y < ph; ++y) {
            const unsigned char* r = rows.data() + y * stride;
            for (size_t x = 0; x < pw; ++x)
                for (int c = 0; c < ch; ++c) {
                    const size_t s = x * ch + c;
                    const uint8_t v = depth == 8 ? r[s]
                                                 : (uint8_t)((r[s * depth / 8] >> (8 - depth - (s * depth) % 8)) &
                                                             ((1 << depth) - 1));
                    samp[(((size_t)(y0 + y * dy)) * W + x0 + x * dx) * ch + c] = v;
                }
        }
    }
    const uint32_t ob = ctype == 3 ? (have_trns ? 4 : 3) : (uint32_t)ch + (have_trns ? 1 : 0);
    std::vector<unsigned char> img((size_t)W * H * ob);
    const int scale = ctype == 0 ? 255 / ((1 << depth) - 1) : 1;
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const uint8_t* s = &samp[i * ch];
        unsigned char* o = &img[i * ob];
        if (ctype == 3) {
            const size_t k = (size_t)s[0] * 3;
            if (k + 3 > plte.size()) return VGPU_E_INVALID_ARGUMENT;
            std::memcpy(o, &plte[k], 3);
            if (have_trns) o[3] = s[0] < trns.size() ? trns[s[0]] : 255;
            continue;
        }
        for (int c = 0; c < ch; ++c) o[c] = (unsigned char)(s[c] * scale);
        if (have_trns) {
            bool key = true;  // tRNS keys are 16-bit big-endian per channel
            for (int c = 0; c < ch; ++c) key = key && ((trns[2 * c] << 8 | trns[2 * c + 1]) == s[c]);
            o[ch] = key ? 0 : 255;
        }
    }
    *w = W;
    *h = H;
    *bands = ob;
    pixels->swap(img);
    return VGPU_OK;
}

// Every row with filter 0 (none), one zlib stream at `level`.
int png_encode(const unsigned char* px, uint32_t w, uint32_t h, uint32_t bands, int level,
               std::vector<unsigned char>* out) {
    if (level < 0 || level > 9) return VGPU_E_INVALID_ARGUMENT;  // process.rs:294-296
    if (w == 0 || h == 0 || bands < 1 || bands > 4) return VGPU_E_INVALID_ARGUMENT;
    const size_t stride = (size_t)w * bands;
    std::vector<unsigned char> raw((stride + 1) * h);
    for (uint32_t y = 0; y < h; ++y) {
        raw[y * (stride + 1)] = 0;
        std::memcpy(&raw[y * (stride + 1) + 1], px + y * stride, stride);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), level) != Z_OK) return VGPU_E_NO_MEMORY;
    std::vector<unsigned char>& o = *out;
    o.assign(kPngSig, kPngSig + 8);
    const unsigned char ctype = bands == 1 ? 0 : bands == 2 ? 4 : bands == 3 ? 2 : 6;
    const unsigned char ihdr[13] = {(unsigned char)(w >> 24), (unsigned char)(w >> 16), (unsigned char)(w >> 8),
                                    (unsigned char)w,         (unsigned char)(h >> 24), (unsigned char)(h >> 16),
                                    (unsigned char)(h >> 8),  (unsigned char)h,         8, ctype, 0, 0, 0};
    put_chunk(o, "IHDR", ihdr, 13);
    put_chunk(o, "IDAT", z.data(), zlen);
    put_chunk(o, "IEND", nullptr, 0);
    return VGPU_OK;
}

}  // namespace

extern "C" {

int vgpu_jpeg_encode(vgpu_context* ctx, const void* pixels, uint32_t width, uint32_t height, uint32_t bands,
                     int quality, int subsampling, void* out, size_t out_cap, size_t* out_len) {
    if (!ctx || !pixels || !out_len) return VGPU_E_INVALID_ARGUMENT;
    if (ctx->sticky != VGPU_OK) return ctx->sticky;
    (void)hipSetDevice(ctx->device);
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, pixels) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        return VGPU_E_INVALID_ARGUMENT;  // pixels must be device memory
    }
    std::vector<unsigned char> v;
    const int rc = jpeg_encode_device(ctx, pixels, width, height, bands, quality, subsampling, &v);
    if (rc != VGPU_OK) return rc;
    return copy_out(v, out, out_cap, out_len);
}

int vgpu_png_decode(const void* data, size_t len, uint32_t* width, uint32_t* height, uint32_t* bands, void* pixels,
                    size_t cap) {
    if (!data || !width || !height || !bands) return VGPU_E_INVALID_ARGUMENT;
    std::vector<unsigned char> px;
    const int rc = png_decode(static_cast<const unsigned char*>(data), len, width, height, bands, &px);
    if (rc != VGPU_OK) return rc;
    if (!pixels) return VGPU_OK;  // size query
    if (cap < px.size()) return VGPU_E_BUFFER_SIZE;
    std::memcpy(pixels, px.data(), px.size());
    return VGPU_OK;
}

int vgpu_png_encode(const void* pixels, uint32_t width, uint32_t height, uint32_t bands, int compression, void* out,
                    size_t out_cap, size_t* out_len) {
    if (!pixels || !out_len) return VGPU_E_INVALID_ARGUMENT;
    std::vector<unsigned char> v;
    const int rc = png_encode(static_cast<const unsigned char*>(pixels), width, height, bands, compression, &v);
    if (rc != VGPU_OK) return rc;
    return copy_out(v, out, out_cap, out_len);
}

int vgpu_process(vgpu_context* ctx, const void* input, size_t input_len, vgpu_process_build_fn build, void* user,
                 const vgpu_encode_options* enc, const vgpu_cancel_token* token, void* out, size_t out_cap,
                 size_t* out_len, vgpu_process_result* result) {
    if (!ctx || !input || !enc || !out_len) return VGPU_E_INVALID_ARGUMENT;
    // validate_encode_options (process.rs:289-313)
    if (enc->format == VGPU_ENCODE_JPEG) {
        if (enc->quality < 1 || enc->quality > 100) return VGPU_E_INVALID_ARGUMENT;
    } else if (enc->format == VGPU_ENCODE_PNG) {
        if (enc->compression < 0 || enc->compression > 9) return VGPU_E_INVALID_ARGUMENT;
    } else {
        return VGPU_E_UNSUPPORTED;
    }
    if (vgpu_cancel_token_is_cancelled(token)) return VGPU_E_CANCELLED;
    if (ctx->sticky != VGPU_OK) return ctx->sticky;
    uint32_t w = 0, h = 0, b = 0;
    std::vector<unsigned char> px;
    int rc = png_decode(static_cast<const unsigned char*>(input), input_len, &w, &h, &b, &px);
    if (rc != VGPU_OK) return rc;
    if (result) {
        std::memset(result, 0, sizeof(*result));
        std::strcpy(result->input_format, "png");
        result->input_width = w;
        result->input_height = h;
        std::strcpy(result->output_format, enc->format == VGPU_ENCODE_JPEG ? "jpeg" : "png");
    }
    if (vgpu_cancel_token_is_cancelled(token)) return VGPU_E_CANCELLED;
    (void)hipSetDevice(ctx->device);
    vgpu_pipeline* p = nullptr;
    rc = vgpu_pipeline_new(ctx, w, h, b, VGPU_U8, &p);
    if (rc != VGPU_OK) return rc;
    struct Guard {
        vgpu_context* c;
        vgpu_pipeline* p;
        void* a = nullptr;
        void* o = nullptr;
        ~Guard() {
            (void)hipStreamSynchronize(c->stream);  // no copy or kernel still reads a, o or the host pixels
            c->pool.release(a);
            c->pool.release(o);
            vgpu_pipeline_destroy(p);
        }
    } g{ctx, p};
    if (build && (rc = build(p, user)) != VGPU_OK) return rc;
    if (vgpu_cancel_token_is_cancelled(token)) return VGPU_E_CANCELLED;
    uint32_t ow = 0, oh = 0, ob = 0;
    int of = 0;
    vgpu_pipeline_output_info(p, &ow, &oh, &ob, &of);
    if (of != VGPU_U8) return VGPU_E_UNSUPPORTED_FORMAT;  // process() works on Image<U8>
    const size_t out_bytes = (size_t)ow * oh * ob;
    g.a = ctx->pool.acquire(std::max<size_t>(px.size(), 1));
    g.o = ctx->pool.acquire(std::max<size_t>(out_bytes, 1));
    if (!g.a || !g.o) return VGPU_E_NO_MEMORY;
    rc = from_hip(hipMemcpyAsync(g.a, px.data(), px.size(), hipMemcpyHostToDevice, ctx->stream));
    if (rc == VGPU_OK) rc = vgpu_pipeline_run_cancellable(p, g.a, px.size(), g.o, out_bytes, token);
    if (rc != VGPU_OK) return rc;
    std::vector<unsigned char> v;
    if (enc->format == VGPU_ENCODE_JPEG) {
        rc = jpeg_encode_device(ctx, g.o, ow, oh, ob, enc->quality, enc->subsampling, &v);
    } else {
        std::vector<unsigned char> host(out_bytes);
        rc = from_hip(hipMemcpyAsync(host.data(), g.o, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
        if (rc == VGPU_OK) rc = from_hip(hipStreamSynchronize(ctx->stream));
        if (rc == VGPU_OK) rc = png_encode(host.data(), ow, oh, ob, enc->compression, &v);
    }
    if (rc != VGPU_OK) return rc;
    if (vgpu_cancel_token_is_cancelled(token)) return VGPU_E_CANCELLED;
    if (result) result->bytes_written = v.size();
    return copy_out(v, out, out_cap, out_len);
}

}  // extern "C"
