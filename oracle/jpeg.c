/*
 * CPU oracle for the encode step after the path: baseline JFIF JPEG.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of vgpu_jpeg_encode and of
 * vgpu_process's JPEG output); the product library never links it.
 *
 * The reference encodes through turbojpeg 1.4.0 (Cargo.lock) over the
 * system libjpeg-turbo 3.x (tj3Compress8 with TJPARAM_QUALITY and
 * TJPARAM_SUBSAMP only: jpeg/encode.rs:169-258), so its bytes are
 * libjpeg-turbo's for the default parameters: islow DCT, no Huffman
 * optimisation, no restart markers, JFIF 1.01 APP0.  This file restates
 * that published algorithm (ITU-T T.81 Annex K tables and quality scaling,
 * the JFIF YCbCr conversion in 16-bit fixed point, the 2 x 2 chroma
 * downsampling with alternating 1, 2 bias, the Loeffler-Ligtenberg-Moschytz
 * integer DCT with 13-bit constants and 2 pass-1 bits, rounding quantisation
 * and Huffman coding with dummy blocks at the image's right and bottom
 * edges) a whole plane at a time.
 *
 * Pins: the streams equal the bytes Pillow writes through libjpeg-turbo for
 * the same quality and subsampling (grey, 4:4:4, 4:2:0, odd sizes, quality
 * 1..100), and decode back to the input within the quality's error.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const uint8_t ZZ[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                               41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                               30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

/* T.81 Table K.1 / K.2, natural order. */
static const uint8_t QLUM[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                                 14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                                 18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                                 49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
static const uint8_t QCHR[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                                 24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                                 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                                 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};

/* T.81 Tables K.3 - K.6: code-length counts for lengths 1..16, then values. */
static const uint8_t DCL_BITS[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
static const uint8_t DCC_BITS[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
static const uint8_t DC_VALS[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
static const uint8_t ACL_BITS[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
static const uint8_t ACL_VALS[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
static const uint8_t ACC_BITS[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
static const uint8_t ACC_VALS[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

typedef struct {
    uint16_t code[256];
    uint8_t len[256];
} huff;

/* Canonical codes (T.81 Annex C): consecutive values, doubling per length. */
static void huff_build(huff *h, const uint8_t *bits, const uint8_t *vals) {
    memset(h, 0, sizeof(*h));
    unsigned code = 0, k = 0;
    for (int l = 1; l <= 16; ++l) {
        for (int i = 0; i < bits[l - 1]; ++i, ++k, ++code) {
            h->code[vals[k]] = (uint16_t)code;
            h->len[vals[k]] = (uint8_t)l;
        }
        code <<= 1;
    }
}

typedef struct {
    uint8_t *b;
    size_t n, cap;
    uint64_t acc;
    int nacc;
    int oom;
} stream;

static void put_byte(stream *s, uint8_t v) {
    if (s->n == s->cap) {
        size_t c = s->cap ? 2 * s->cap : 4096;
        uint8_t *nb = (uint8_t *)realloc(s->b, c);
        if (!nb) {
            s->oom = 1;
            return;
        }
        s->b = nb;
        s->cap = c;
    }
    s->b[s->n++] = v;
}
static void put_be16(stream *s, unsigned v) {
    put_byte(s, (uint8_t)(v >> 8));
    put_byte(s, (uint8_t)v);
}
/* Entropy-coded bits, most significant first; a 0xFF byte is followed by a
 * stuffed 0x00 (T.81 F.1.2.3). */
static void put_bits(stream *s, unsigned v, int n) {
    s->acc = (s->acc << n) | (v & ((1u << n) - 1));
    s->nacc += n;
    while (s->nacc >= 8) {
        const uint8_t byte = (uint8_t)(s->acc >> (s->nacc - 8));
        put_byte(s, byte);
        if (byte == 0xff) put_byte(s, 0);
        s->nacc -= 8;
    }
}

static int nbits(int v) {
    unsigned a = (unsigned)(v < 0 ? -v : v);
    int n = 0;
    while (a) {
        ++n;
        a >>= 1;
    }
    return n;
}

/* One block in zigzag order (T.81 F.1.2.1, F.1.2.2). */
static void code_block(stream *s, const int *zz, int *pred, const huff *dc, const huff *ac) {
    const int diff = zz[0] - *pred;
    *pred = zz[0];
    int n = nbits(diff);
    put_bits(s, dc->code[n], dc->len[n]);
    if (n) put_bits(s, (unsigned)(diff < 0 ? diff - 1 : diff), n);
    int run = 0;
    for (int k = 1; k < 64; ++k) {
        if (zz[k] == 0) {
            ++run;
            continue;
        }
        for (; run >= 16; run -= 16) put_bits(s, ac->code[0xf0], ac->len[0xf0]);
        n = nbits(zz[k]);
        const int sym = run << 4 | n;
        put_bits(s, ac->code[sym], ac->len[sym]);
        put_bits(s, (unsigned)(zz[k] < 0 ? zz[k] - 1 : zz[k]), n);
        run = 0;
    }
    if (run) put_bits(s, ac->code[0], ac->len[0]);
}

/* jpeg_quality_scaling + jpeg_add_quant_table with force_baseline. */
static void scaled_table(const uint8_t *base, int q, int *out) {
    const long pct = q < 50 ? 5000 / q : 200 - 2 * q;
    for (int i = 0; i < 64; ++i) {
        long t = (base[i] * pct + 50) / 100;
        out[i] = (int)(t < 1 ? 1 : t > 255 ? 255 : t);
    }
}

#define DESCALE(x, n) (((x) + (1L << ((n)-1))) >> (n))

/* The islow forward DCT (rows, then columns) of one 8 x 8 block of
 * level-shifted samples; output scaled by 8 as libjpeg's quantiser expects. */
static void fdct_islow(long *d) {
    enum { CB = 13, P1 = 2 };
    static const long C0298 = 2446, C0390 = 3196, C0541 = 4433, C0765 = 6270, C0899 = 7373, C1175 = 9633,
                      C1501 = 12299, C1847 = 15137, C1961 = 16069, C2053 = 16819, C2562 = 20995, C3072 = 25172;
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < 8; ++i) {
            long *p = pass == 0 ? d + 8 * i : d + i;
            const int st = pass == 0 ? 1 : 8;
            long x[8];
            for (int k = 0; k < 8; ++k) x[k] = p[k * st];
            const long s07 = x[0] + x[7], d07 = x[0] - x[7], s16 = x[1] + x[6], d16 = x[1] - x[6];
            const long s25 = x[2] + x[5], d25 = x[2] - x[5], s34 = x[3] + x[4], d34 = x[3] - x[4];
            const long e0 = s07 + s34, e3 = s07 - s34, e1 = s16 + s25, e2 = s16 - s25;
            const int sh = pass == 0 ? CB - P1 : CB + P1;
            long y[8];
            y[0] = pass == 0 ? (e0 + e1) * (1L << P1) : DESCALE(e0 + e1, P1);
            y[4] = pass == 0 ? (e0 - e1) * (1L << P1) : DESCALE(e0 - e1, P1);
            const long r = (e2 + e3) * C0541;
            y[2] = DESCALE(r + e3 * C0765, sh);
            y[6] = DESCALE(r - e2 * C1847, sh);
            const long za = d34 + d07, zb = d25 + d16, zc = d34 + d16, zd = d25 + d07;
            const long z5 = (zc + zd) * C1175;
            const long ma = -za * C0899, mb = -zb * C2562, mc = -zc * C1961 + z5, md = -zd * C0390 + z5;
            y[7] = DESCALE(d34 * C0298 + ma + mc, sh);
            y[5] = DESCALE(d25 * C2053 + mb + md, sh);
            y[3] = DESCALE(d16 * C3072 + mb + mc, sh);
            y[1] = DESCALE(d07 * C1501 + ma + md, sh);
            for (int k = 0; k < 8; ++k) p[k * st] = y[k];
        }
    }
}

/* Level shift, DCT, quantise (divisor 8q, rounding half away from zero on
 * the magnitude), zigzag. */
static void block_coeffs(const int *plane, int pw, int by, int bx, const int *qt, int *zz) {
    long d[64];
    for (int r = 0; r < 8; ++r)
        for (int c = 0; c < 8; ++c) d[8 * r + c] = plane[(size_t)(8 * by + r) * pw + 8 * bx + c] - 128;
    fdct_islow(d);
    for (int k = 0; k < 64; ++k) {
        const long v = d[ZZ[k]], q = 8L * qt[ZZ[k]];
        zz[k] = (int)(v < 0 ? -((-v + q / 2) / q) : (v + q / 2) / q);
    }
}

int orc_jpeg_encode(const uint8_t *px, uint32_t w, uint32_t h, uint32_t bands, int quality, int subsampling,
                    uint8_t *out, size_t cap, size_t *len) {
    if (!px || !len || w == 0 || h == 0 || w > 65535 || h > 65535) return ORC_EINVAL;
    if (quality < 0 || quality > 100 || subsampling < 0 || subsampling > 2) return ORC_EINVAL;
    if (bands != 1 && bands != 3 && bands != 4) return ORC_EFORMAT;
    const int q = quality < 1 ? 1 : quality;
    const int ncomp = bands == 1 ? 1 : 3;
    const int s420 = ncomp == 3 && (subsampling == 2 || (subsampling == 0 && quality < 90));
    /* Planes padded (edge replicated) to whole MCUs: 8 (or 16 for 4:2:0). */
    const int mcu = s420 ? 16 : 8;
    const int pw = (int)((w + mcu - 1) / mcu * mcu), ph = (int)((h + mcu - 1) / mcu * mcu);
    int *pl[3] = {0, 0, 0};
    for (int c = 0; c < ncomp; ++c) pl[c] = (int *)malloc(sizeof(int) * (size_t)pw * ph);
    for (int y = 0; y < ph; ++y)
        for (int x = 0; x < pw; ++x) {
            const uint8_t *p = px + ((size_t)(y < (int)h ? y : (int)h - 1) * w + (x < (int)w ? x : (int)w - 1)) * bands;
            const size_t o = (size_t)y * pw + x;
            if (ncomp == 1) {
                pl[0][o] = p[0];
                continue;
            }
            const long R = p[0], G = p[1], B = p[2];
            /* JFIF: weights round(w * 2^16); chroma offset 128, rounding 1/2 - 2^-16 */
            pl[0][o] = (int)((19595 * R + 38470 * G + 7471 * B + 32768) >> 16);
            pl[1][o] = (int)((-11059 * R - 21709 * G + 32768 * B + (128L << 16) + 32767) >> 16);
            pl[2][o] = (int)((32768 * R - 27439 * G - 5329 * B + (128L << 16) + 32767) >> 16);
        }
    int cw[3] = {pw, pw, pw};
    if (s420)
        for (int c = 1; c < 3; ++c) {
            const int dw = pw / 2, dh = ph / 2;
            int *dn = (int *)malloc(sizeof(int) * (size_t)dw * dh);
            /* Rows below the image's last pair repeat the last downsampled
             * row (jcprepct's bottom padding of the downsampled output);
             * columns are padded before downsampling. */
            const int last = (int)((h + 1) / 2) - 1;
            for (int y = 0; y < dh; ++y)
                for (int x = 0; x < dw; ++x) {
                    const int *a = pl[c] + (size_t)(2 * (y < last ? y : last)) * pw + 2 * x;
                    dn[(size_t)y * dw + x] = (a[0] + a[1] + a[pw] + a[pw + 1] + 1 + (x & 1)) >> 2;
                }
            free(pl[c]);
            pl[c] = dn;
            cw[c] = dw;
        }
    int ql[64], qc[64];
    scaled_table(QLUM, q, ql);
    scaled_table(QCHR, q, qc);
    huff dcl, acl, dcc, acc;
    huff_build(&dcl, DCL_BITS, DC_VALS);
    huff_build(&acl, ACL_BITS, ACL_VALS);
    huff_build(&dcc, DCC_BITS, DC_VALS);
    huff_build(&acc, ACC_BITS, ACC_VALS);

    stream s = {0};
    static const uint8_t app0[] = {'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0};
    put_be16(&s, 0xffd8);
    put_be16(&s, 0xffe0);
    put_be16(&s, 16);
    for (size_t i = 0; i < sizeof(app0); ++i) put_byte(&s, app0[i]);
    for (int t = 0; t < (ncomp == 1 ? 1 : 2); ++t) {
        put_be16(&s, 0xffdb);
        put_be16(&s, 67);
        put_byte(&s, (uint8_t)t);
        for (int k = 0; k < 64; ++k) put_byte(&s, (uint8_t)(t ? qc : ql)[ZZ[k]]);
    }
    put_be16(&s, 0xffc0);
    put_be16(&s, 8 + 3 * ncomp);
    put_byte(&s, 8);
    put_be16(&s, h);
    put_be16(&s, w);
    put_byte(&s, (uint8_t)ncomp);
    for (int c = 0; c < ncomp; ++c) {
        put_byte(&s, (uint8_t)(c + 1));
        put_byte(&s, c == 0 && s420 ? 0x22 : 0x11);
        put_byte(&s, c == 0 ? 0 : 1);
    }
    const struct {
        int cls, id;
        const uint8_t *bits, *vals;
        int nv;
    } dht[4] = {{0, 0, DCL_BITS, DC_VALS, 12}, {1, 0, ACL_BITS, ACL_VALS, 162},
                {0, 1, DCC_BITS, DC_VALS, 12}, {1, 1, ACC_BITS, ACC_VALS, 162}};
    for (int t = 0; t < (ncomp == 1 ? 2 : 4); ++t) {
        put_be16(&s, 0xffc4);
        put_be16(&s, 3 + 16 + dht[t].nv);
        put_byte(&s, (uint8_t)(dht[t].cls << 4 | dht[t].id));
        for (int i = 0; i < 16; ++i) put_byte(&s, dht[t].bits[i]);
        for (int i = 0; i < dht[t].nv; ++i) put_byte(&s, dht[t].vals[i]);
    }
    put_be16(&s, 0xffda);
    put_be16(&s, 6 + 2 * ncomp);
    put_byte(&s, (uint8_t)ncomp);
    for (int c = 0; c < ncomp; ++c) {
        put_byte(&s, (uint8_t)(c + 1));
        put_byte(&s, c == 0 ? 0x00 : 0x11);
    }
    put_byte(&s, 0);
    put_byte(&s, 63);
    put_byte(&s, 0);

    /* Blocks holding image samples; the rest of a 4:2:0 MCU is dummy blocks. */
    const int lbw = (int)((w + 7) / 8), lbh = (int)((h + 7) / 8);
    int pred[3] = {0, 0, 0}, zz[64];
    const int mcux = pw / mcu, mcuy = ph / mcu;
    for (int my = 0; my < mcuy; ++my)
        for (int mx = 0; mx < mcux; ++mx) {
            if (!s420) {
                for (int c = 0; c < ncomp; ++c) {
                    block_coeffs(pl[c], cw[c], my, mx, c ? qc : ql, zz);
                    code_block(&s, zz, &pred[c], c ? &dcc : &dcl, c ? &acc : &acl);
                }
                continue;
            }
            int dc[4];
            for (int k = 0; k < 4; ++k) {
                const int by = 2 * my + k / 2, bx = 2 * mx + k % 2;
                if (by < lbh && bx < lbw) {
                    block_coeffs(pl[0], cw[0], by, bx, ql, zz);
                } else {
                    /* libjpeg's compress_data dummy blocks: AC zero, DC of the
                     * previous block of the row, or (a whole dummy row) of the
                     * last block of the row above. */
                    memset(zz, 0, sizeof(zz));
                    zz[0] = by < lbh ? dc[k - 1] : dc[1];
                }
                dc[k] = zz[0];
                code_block(&s, zz, &pred[0], &dcl, &acl);
            }
            for (int c = 1; c < 3; ++c) {
                block_coeffs(pl[c], cw[c], my, mx, qc, zz);
                code_block(&s, zz, &pred[c], &dcc, &acc);
            }
        }
    if (s.nacc & 7) put_bits(&s, 0x7f, 8 - (s.nacc & 7)); /* pad with 1 bits */
    put_be16(&s, 0xffd9);
    for (int c = 0; c < ncomp; ++c) free(pl[c]);
    int rc = ORC_OK;
    if (s.oom) {
        rc = ORC_EINVAL;
    } else {
        *len = s.n;
        if (out) {
            if (cap < s.n) rc = ORC_EINVAL;
            else memcpy(out, s.b, s.n);
        }
    }
    free(s.b);
    return rc;
}
