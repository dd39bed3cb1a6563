/*
 * CPU oracle: scalar C restatement of viprs's resample + separable-convolution
 * path, pinned to the reference's x86-64 + AVX2 dispatch semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for bench.py, as the timed CPU baseline).  The product library never
 * links or calls it.
 *
 * Reference files restated (paths relative to the viprs repository root):
 *   crates/viprs-core/src/kernel.rs:50-221                     kernel weights
 *   crates/viprs-ops-spatial/src/resample/reduce_common.rs     tables, plans, regions
 *   crates/viprs-core/src/resample.rs:62-141                   ReduceConfig
 *   crates/viprs-ops-spatial/src/resample/reduceh.rs:101-163   ReduceH::process_region
 *   crates/viprs-ops-spatial/src/resample/reducev.rs:100-162   ReduceV::process_region
 *   crates/viprs-ops-spatial/src/resample/reduce_simd.rs       scalar + AVX2 kernels
 *   crates/viprs-ops-spatial/src/resample/sample_conv.rs       fixed-point rounding
 *   crates/viprs-ops-spatial/src/resample/shrinkh.rs:220-471   ShrinkH (x86 dispatch)
 *   crates/viprs-ops-spatial/src/resample/shrinkv.rs:122-339   ShrinkV
 *   crates/viprs-ops-spatial/src/resample/thumbnail.rs:229-524 Thumbnail planner
 *   crates/viprs-ops-spatial/src/resample/resize.rs:163-345    Resize planner (downscale)
 *   crates/viprs-ops-spatial/src/convolution/gauss_blur.rs     integer/float Gaussian
 *   crates/viprs-core/src/shared_ops/gauss_kernel.rs           float Gaussian weights
 *   crates/viprs-ops-spatial/src/convolution/convsep.rs        ConvSepH/V
 *   crates/viprs-ops-spatial/src/structural/{premultiply,unpremultiply}.rs
 *   crates/viprs-runtime/src/scheduler/rayon_scheduler/execute.rs:413-849 tile executor
 *   crates/viprs-runtime/src/scheduler/rayon_scheduler/planning.rs:509-601 tiles
 *   crates/viprs-runtime/src/sources/memory.rs:149-235         clamp-to-edge source read
 *
 * Parity pins (tests/test_oracle_golden.py, tests/test_oracle_pins.py):
 *   reduce tables, ReduceH/ReduceV u8  -> libvips fixtures in reduceh_libvips,
 *       reducev_libvips, libvips_parity_resample (test_reduce{h,v}), in-crate
 *       known answers (reduceh.rs:574-610, reducev.rs:510-570), invariants
 *       (f64 phase sums = 1, i16 sums within +-taps of 4096, offset -0.5 at f=2)
 *   ReduceH/ReduceV u16                -> exact integer closed form (i16 table
 *       row times clamped samples, (acc + 2048) >> 12)
 *   ReduceH/ReduceV f32                -> exact rational value of each output's
 *       weighted sum with a rigorous FP64 error bound: bit-exact wherever that
 *       interval decides the f32 (order-independent)
 *   ShrinkH/ShrinkV u8                  -> shrinkh/, shrinkv/ fixtures,
 *       shrink_libvips/rgb_factor4, shrinkh.rs:1479-1506,1784-1901 known answers
 *   ShrinkH/ShrinkV u16 / f32          -> exact closed forms (f64 box sum,
 *       times 1/f, from_f64_clamped)
 *   Gaussian integer masks              -> gauss_blur.rs:56-65 sigma=1.5 constants,
 *       brute-force 2-D convolution on small images; u8 libvips fixtures (+-1)
 *   Gaussian float                      -> gauss_blur_libvips/nonuniform_5x3 and
 *       uniform_field, gauss_blur/fractional_ramp_sigma_1_0 (<= 1 ULP),
 *       libvips_parity_convolution (1e-6), uniform_sigma_1_5 (1e-3); exact
 *       rational closed form per pass (f32 and u16 inputs)
 *   ConvSep                             -> convsep_libvips/gaussian_sigma_1_0
 *       (<= 1 ULP); exact rational closed form with scale and offset (f32, u8,
 *       u16 inputs)
 *   Resize                              -> libvips_parity_resample resize fixtures
 *       (exact)
 *   Premultiply/unpremultiply           -> unpremultiply.rs:239-248 known answer,
 *       exhaustive closed-form check over all 65536 (colour, alpha) pairs
 *   Thumbnail/resize planners           -> thumbnail.rs:531-682 / resize.rs:392-536
 *       shape tests restated in the test file; thumbnail libvips fixtures
 *   Chain executor                      -> strip-height invariance, single-op
 *       equality with clamp-padded closed forms
 * The float summation ORDER (reduce_simd.rs:2621-2694,2824-2888; FMA or not)
 * is restated from the AVX2 source; it decides only the outputs whose exact
 * value lies within a few FP64 ulps of an f32 rounding boundary, which the
 * closed forms leave open (about 1 in 2 x denominator samples for masks with
 * small-denominator weights, far fewer otherwise).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define REDUCE_PHASES 64
#define MAX_REDUCE_TAPS 2000
#define MAX_REDUCE_TAP_RADIUS ((MAX_REDUCE_TAPS - 1) / 2)
#define PLAN_EPS 1e-9

/* ------------------------------------------------------------------ helpers */

/* Rust `f64 as u32` / `as i16`: saturating truncation, NaN -> 0. */
static uint32_t sat_u32(double v) {
    if (!(v == v)) return 0;
    if (v <= 0.0) return 0;
    if (v >= 4294967295.0) return 4294967295u;
    return (uint32_t)v;
}
static int16_t sat_i16(double v) {
    if (!(v == v)) return 0;
    if (v <= -32768.0) return -32768;
    if (v >= 32767.0) return 32767;
    return (int16_t)v;
}
static int64_t sat_i64(double v) {
    if (!(v == v)) return 0;
    if (v <= -9.2233720368547758e18) return INT64_MIN;
    if (v >= 9.2233720368547758e18) return INT64_MAX;
    return (int64_t)v;
}
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

typedef struct {
    int64_t x, y;
    int64_t w, h;
} rgn;

/* ------------------------------------------------------------- kernel.rs */

/* InterpolationKernel::support (kernel.rs:66-75). */
static double kernel_support(int k) {
    switch (k) {
    case ORC_NEAREST: return 0.5;
    case ORC_BILINEAR: return 1.0;
    case ORC_LANCZOS3: return 3.0;
    default: return 2.0; /* Bicubic, CatmullRom, Lanczos2 */
    }
}

/* InterpolationKernel::interpolate (kernel.rs:133-211). */
double orc_kernel_weight(int k, double x) {
    switch (k) {
    case ORC_NEAREST: return fabs(x) < 0.5 ? 1.0 : 0.0;
    case ORC_BILINEAR: {
        double a = fabs(x);
        return a < 1.0 ? 1.0 - a : 0.0;
    }
    case ORC_BICUBIC:
    case ORC_CATMULLROM: {
        double a = fabs(x);
        if (a < 1.0) return fma(2.5 * a, -a, 1.5 * a * a * a) + 1.0;
        if (a < 2.0) return fma(4.0, -a, fma(2.5 * a, a, -0.5 * a * a * a)) + 2.0;
        return 0.0;
    }
    case ORC_LANCZOS2: {
        double a = fabs(x);
        if (a < DBL_EPSILON) return 1.0;
        if (a < 2.0) {
            double px = M_PI * a;
            return 2.0 * sin(px) * sin(px / 2.0) / (px * px);
        }
        return 0.0;
    }
    case ORC_LANCZOS3: {
        double a = fabs(x);
        if (a < DBL_EPSILON) return 1.0;
        if (a < 3.0) {
            double px = M_PI * a;
            return 3.0 * sin(px) * sin(px / 3.0) / (px * px);
        }
        return 0.0;
    }
    default: return 0.0;
    }
}

/* ------------------------------------------------------ reduce_common.rs */

/* validate_reduce_kernel (reduce_common.rs:94-124). */
static int valid_reduce_kernel(int k) {
    return k == ORC_NEAREST || k == ORC_BILINEAR || k == ORC_BICUBIC || k == ORC_LANCZOS2 ||
           k == ORC_LANCZOS3 || k == ORC_CATMULLROM;
}

/* tap_count (reduce_common.rs:257-277). Returns -1 for "None". */
int orc_reduce_taps(int kernel, double factor) {
    if (fabs(factor - 1.0) < DBL_EPSILON) return 1;
    if (kernel == ORC_NEAREST) return 1;
    double r = round(kernel_support(kernel) * factor);
    if (!isfinite(r) || r < 0.0 || r > (double)MAX_REDUCE_TAP_RADIUS) return -1;
    long taps = (long)r * 2 + 1;
    return taps <= MAX_REDUCE_TAPS ? (int)taps : -1;
}

/* build_coefficients + quantize_coefficients (reduce_common.rs:279-316), all
 * 65 phases (ReduceKernel::new, reduce_common.rs:137-164). */
int orc_reduce_tables(int kernel, double factor, double *cf, int16_t *ci) {
    if (!isfinite(factor) || factor < 1.0) return ORC_EINVAL;
    if (!valid_reduce_kernel(kernel)) return ORC_EKERNEL;
    int taps = orc_reduce_taps(kernel, factor);
    if (taps < 0) return ORC_EINVAL;
    for (int p = 0; p <= REDUCE_PHASES; ++p) {
        double *row = cf + (size_t)p * taps;
        if (kernel == ORC_NEAREST) {
            row[0] = 1.0;
        } else {
            double phase = (double)p / (double)REDUCE_PHASES;
            double half = phase + (double)taps / 2.0 - 1.0;
            double scale = 1.0 / factor;
            double sum = 0.0;
            for (int t = 0; t < taps; ++t) {
                double d = ((double)t - half) * scale;
                double wt = orc_kernel_weight(kernel, fabs(d));
                row[t] = wt;
                sum += wt;
            }
            if (sum > 0.0)
                for (int t = 0; t < taps; ++t) row[t] /= sum;
        }
        for (int t = 0; t < taps; ++t) ci[(size_t)p * taps + t] = sat_i16(row[t] * 4096.0);
    }
    return taps;
}

/* ReduceConfig::output_len (crates/viprs-core/src/resample.rs:86-93). */
uint32_t orc_reduce_output_len(double factor, uint32_t in) {
    if (in == 0) return 0;
    return sat_u32(round((double)in / factor));
}

/* ReduceKernel::bind_input_len (reduce_common.rs:171-180). */
double orc_reduce_offset(int kernel, double factor, uint32_t in) {
    (void)kernel;
    if (in == 0 || fabs(factor - 1.0) < DBL_EPSILON) return 0.0;
    uint32_t out = orc_reduce_output_len(factor, in);
    double extra = fma((double)out, factor, -(double)in);
    return (1.0 + extra) / 2.0 - 1.0;
}

typedef struct {
    int kernel;
    double factor;
    int taps;
    int64_t pad;
    double offset;
    double *cf;
    int16_t *ci;
} rk_t;

static int rk_init(rk_t *k, int kernel, double factor, uint32_t input_len) {
    memset(k, 0, sizeof(*k));
    if (!isfinite(factor) || factor < 1.0) return ORC_EINVAL;
    if (!valid_reduce_kernel(kernel)) return ORC_EKERNEL;
    int taps = orc_reduce_taps(kernel, factor);
    if (taps < 0) return ORC_EINVAL;
    k->kernel = kernel;
    k->factor = factor;
    k->taps = taps;
    k->pad = (int64_t)ceil((double)taps / 2.0) - 1;
    k->cf = (double *)malloc(sizeof(double) * (size_t)taps * (REDUCE_PHASES + 1));
    k->ci = (int16_t *)malloc(sizeof(int16_t) * (size_t)taps * (REDUCE_PHASES + 1));
    orc_reduce_tables(kernel, factor, k->cf, k->ci);
    k->offset = orc_reduce_offset(kernel, factor, input_len);
    return ORC_OK;
}

/* ReduceKernel::source_position (reduce_common.rs:183-185). */
static double rk_sp(const rk_t *k, double index) { return fma(index + 0.5, k->factor, -0.5) - k->offset; }

/* plan_for_source_position (reduce_common.rs:127-135). */
static void rk_plan(const rk_t *k, double sp, int64_t *start, int *phase) {
    double fl = floor(sp);
    double ph = clampd(round(fma(sp - fl, (double)REDUCE_PHASES, 0.0)), 0.0, (double)REDUCE_PHASES);
    *phase = (int)ph;
    *start = sat_i64(fl) - k->pad;
}

void orc_reduce_plan(int kernel, double factor, uint32_t input_len, double origin, int accumulate,
                     uint32_t n, int64_t *starts, uint8_t *phases) {
    rk_t k;
    if (rk_init(&k, kernel, factor, input_len) != ORC_OK) return;
    double sx = rk_sp(&k, origin);
    for (uint32_t i = 0; i < n; ++i) {
        int ph;
        double sp = accumulate ? sx : rk_sp(&k, origin + (double)i);
        rk_plan(&k, sp, &starts[i], &ph);
        phases[i] = (uint8_t)ph;
        sx += factor;
    }
    free(k.cf);
    free(k.ci);
}

/* required_input_region_h / _v (reduce_common.rs:218-254). */
static void rk_required(const rk_t *k, int64_t o, int64_t n, int64_t *start, int64_t *len) {
    if (n == 0) {
        *start = o;
        *len = 0;
        return;
    }
    int64_t first = sat_i64(floor(rk_sp(k, (double)o)));
    int64_t last_i = clampi(o + (n - 1), INT32_MIN, INT32_MAX);
    int64_t last = sat_i64(floor(rk_sp(k, (double)last_i)));
    int64_t s = clampi(first - k->pad, INT32_MIN, INT32_MAX);
    int64_t e = clampi(last - k->pad + k->taps - 1, INT32_MIN, INT32_MAX);
    *start = s;
    *len = e - s + 1;
}

/* clamp_axis (reduce_common.rs:318-322). */
static inline int64_t clamp_axis(int64_t abs, int64_t origin, int64_t len) {
    return clampi(abs - origin, 0, len > 0 ? len - 1 : 0);
}

/* Gaussian integer kernel (gauss_blur.rs:80-126; gauss_kernel.rs:13-56). */
int orc_gauss_int_kernel(float sigma_f, int16_t *coeffs, int max_taps, int32_t *scale) {
    double sigma = (double)sigma_f;
    if (sigma < 0.2) {
        if (max_taps < 1) return ORC_EINVAL;
        coeffs[0] = 1;
        *scale = 1;
        return 1;
    }
    double sig2 = 2.0 * sigma * sigma;
    size_t max_x = (size_t)ceil(8.0 * sigma);
    size_t x = 0;
    while (x < max_x) {
        double v = exp(-((double)(x * x)) / sig2);
        if (v < 0.2) break;
        ++x;
    }
    size_t radius = x > 0 ? x - 1 : 0;
    size_t size = 2 * radius + 1;
    if ((int)size > max_taps) return ORC_EINVAL;
    int32_t s = 0;
    for (size_t i = 0; i < size; ++i) {
        int32_t off = (int32_t)i - (int32_t)radius;
        double v = exp(-(double)(off * off) / sig2);
        int32_t tap = (int32_t)nearbyint(20.0 * v);
        coeffs[i] = (int16_t)tap;
        s += tap;
    }
    if (s == 0) {
        coeffs[0] = 1;
        *scale = 1;
        return 1;
    }
    *scale = s;
    return (int)size;
}

/* ------------------------------------------------------------- pipeline */

enum {
    N_SHRINK_H = 1,
    N_SHRINK_V = 2,
    N_REDUCE_H = 3,
    N_REDUCE_V = 4,
    N_GAUSS_H = 5,
    N_GAUSS_V = 6,
    N_CONVSEP_H = 7,
    N_CONVSEP_V = 8,
    N_PREMULT = 9,
    N_UNPREMULT = 10,
    N_COPY = 11,
    N_REDUCE_OP_H = 12, /* ReduceOp's horizontal pass (reduce.rs:104-133) */
    N_REDUCE_OP_V = 13, /* ReduceOp's vertical pass (reduce.rs:136-166) */
    N_EXTRACT = 14,     /* ExtractArea (structural/extract_area.rs:23-60) */
    N_AFFINE = 15,      /* Affine, ExtendMode::Copy, source bounds = its input
                         * (affine/core.rs; builder/state.rs:298-341) */
    N_ZOOM = 16         /* Zoom (structural/zoom.rs:339-400) */
};

typedef struct {
    int type;
    int in_fmt, out_fmt;
    uint32_t bands;
    uint32_t in_w, in_h, out_w, out_h;
    /* shrink */
    uint32_t factor;
    uint32_t source_width;
    /* reduce */
    rk_t rk;
    /* gauss / convsep */
    int ntaps;
    int16_t gi[512];
    int32_t gscale;
    double *gf;
    double cscale, coffset;
    double param;
    /* extract area: offset of the output in the input */
    int64_t ex, ey;
    /* affine: [a, b, c, d], translation, kernel, SmallTile demand */
    double am[4], atx, aty;
    int akern, asmall;
    /* zoom factors */
    uint32_t zx, zy;
} node_t;

struct orc_pipeline {
    uint32_t src_w, src_h, bands;
    int src_fmt;
    uint32_t cur_w, cur_h;
    int cur_fmt;
    int n;
    node_t nodes[32];
    int thin_override; /* the ThinStrip override of large thumbnails (builder/resample.rs:908-917) */
    /* A box pre-shrink of the source (orc_pipeline_box_shrink): the decoded
     * image is orig_w x orig_h, the node chain reads its shrunk copy, whose
     * size is src_w x src_h. */
    uint32_t box_f;
    int box_ceil;
    uint32_t orig_w, orig_h;
};

static size_t fmt_size(int f) {
    return f == ORC_U8 ? 1 : (f == ORC_U16 || f == ORC_I16) ? 2 : f == ORC_F64 ? 8 : 4;
}
static int fmt_is_int(int f) { return f != ORC_F32 && f != ORC_F64; }
static inline double ld(const void *b, int fmt, size_t i);
static void store_shrunk(void *out, int fmt, size_t oi, double v);
/* The integer formats' ranges (band.rs), as doubles and as integers. */
static double fmt_lo(int f) {
    return f == ORC_I16 ? -32768.0 : f == ORC_I32 ? -2147483648.0 : 0.0;
}
static double fmt_hi(int f) {
    switch (f) {
    case ORC_U8: return 255.0;
    case ORC_U16: return 65535.0;
    case ORC_I16: return 32767.0;
    case ORC_U32: return 4294967295.0;
    default: return 2147483647.0;
    }
}

orc_pipeline *orc_pipeline_new(uint32_t w, uint32_t h, uint32_t bands, int fmt) {
    if (bands == 0 || fmt < ORC_U8 || fmt > ORC_F64) return NULL;
    orc_pipeline *p = (orc_pipeline *)calloc(1, sizeof(orc_pipeline));
    p->src_w = p->cur_w = p->orig_w = w;
    p->src_h = p->cur_h = p->orig_h = h;
    p->bands = bands;
    p->src_fmt = p->cur_fmt = fmt;
    return p;
}

void orc_pipeline_free(orc_pipeline *p) {
    if (!p) return;
    for (int i = 0; i < p->n; ++i) {
        free(p->nodes[i].rk.cf);
        free(p->nodes[i].rk.ci);
        free(p->nodes[i].gf);
    }
    free(p);
}

static node_t *push(orc_pipeline *p, int type) {
    if (p->n >= 32) return NULL;
    node_t *nd = &p->nodes[p->n++];
    memset(nd, 0, sizeof(*nd));
    nd->type = type;
    nd->in_fmt = nd->out_fmt = p->cur_fmt;
    nd->bands = p->bands;
    nd->in_w = nd->out_w = p->cur_w;
    nd->in_h = nd->out_h = p->cur_h;
    return nd;
}
static void commit(orc_pipeline *p, node_t *nd) {
    p->cur_w = nd->out_w;
    p->cur_h = nd->out_h;
    p->cur_fmt = nd->out_fmt;
}

/* PipelineBuilder::reduce_h -> ReduceHBridge::new(.., input_w = current width)
 * (builder/resample.rs:62-123; reduceh.rs:40-58,187-206). */
int orc_pipeline_reduce_h(orc_pipeline *p, double factor, int kernel) {
    if (!valid_reduce_kernel(kernel)) return ORC_EKERNEL;
    if (!isfinite(factor) || factor < 1.0) return ORC_EINVAL;
    if (orc_reduce_taps(kernel, factor) < 0) return ORC_EINVAL;
    node_t *nd = push(p, N_REDUCE_H);
    if (!nd) return ORC_EINVAL;
    rk_init(&nd->rk, kernel, factor, p->cur_w);
    nd->out_w = orc_reduce_output_len(factor, p->cur_w);
    nd->param = factor;
    commit(p, nd);
    return ORC_OK;
}

int orc_pipeline_reduce_v(orc_pipeline *p, double factor, int kernel) {
    if (!valid_reduce_kernel(kernel)) return ORC_EKERNEL;
    if (!isfinite(factor) || factor < 1.0) return ORC_EINVAL;
    if (orc_reduce_taps(kernel, factor) < 0) return ORC_EINVAL;
    node_t *nd = push(p, N_REDUCE_V);
    if (!nd) return ORC_EINVAL;
    rk_init(&nd->rk, kernel, factor, p->cur_h);
    nd->out_h = orc_reduce_output_len(factor, p->cur_h);
    nd->param = factor;
    commit(p, nd);
    return ORC_OK;
}

/* PipelineBuilder::reduce (builder/resample.rs:349-382) -> ReduceOp::new
 * (reduce.rs:51-67): validate_reduce_factors over both factors, then the
 * kernel, then both tap limits (reduce_common.rs:26-44,60-67).  The filters
 * are ReduceKernel::new without bind_input_len, so both axes have offset 0.
 * The op is held as two nodes, H then V: the backward walk then gives the V
 * node's input region = required_input_region_v(out) = ReduceOp's scratch
 * region and the H node's input region = required_input_region_h(that),
 * exactly ReduceOp::required_input_region (reduce.rs:181-184,224-250). */
int orc_pipeline_reduce(orc_pipeline *p, double hfactor, double vfactor, int kernel) {
    if (!isfinite(hfactor) || !isfinite(vfactor) || hfactor < 1.0 || vfactor < 1.0) return ORC_EINVAL;
    if (!valid_reduce_kernel(kernel)) return ORC_EKERNEL;
    if (orc_reduce_taps(kernel, hfactor) < 0 || orc_reduce_taps(kernel, vfactor) < 0) return ORC_EINVAL;
    if (p->n + 2 > 32) return ORC_EINVAL;
    node_t *nh = push(p, N_REDUCE_OP_H);
    rk_init(&nh->rk, kernel, hfactor, 0);
    nh->out_w = orc_reduce_output_len(hfactor, p->cur_w);
    nh->param = hfactor;
    commit(p, nh);
    node_t *nv = push(p, N_REDUCE_OP_V);
    rk_init(&nv->rk, kernel, vfactor, 0);
    nv->out_h = orc_reduce_output_len(vfactor, p->cur_h);
    nv->param = vfactor;
    commit(p, nv);
    return ORC_OK;
}

/* PipelineBuilder::extract_area (builder/geometry.rs:40-80): bounds as
 * validate_extract_area_bounds (builder/state.rs:238-270): a non-empty
 * rectangle inside the current image. */
int orc_pipeline_extract_area(orc_pipeline *p, uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    if (w == 0 || h == 0 || (uint64_t)x + w > p->cur_w || (uint64_t)y + h > p->cur_h) return ORC_EINVAL;
    node_t *nd = push(p, N_EXTRACT);
    if (!nd) return ORC_EINVAL;
    nd->ex = x;
    nd->ey = y;
    nd->out_w = w;
    nd->out_h = h;
    commit(p, nd);
    return ORC_OK;
}

/* PipelineBuilder::shrink_h_with_ceil (builder/resample.rs:133-209): the
 * bridge always carries source_width = current width (shrinkh.rs:1366-1383). */
int orc_pipeline_shrink_h(orc_pipeline *p, uint32_t factor, int ceil_mode) {
    if (factor == 0) return ORC_EINVAL;
    node_t *nd = push(p, N_SHRINK_H);
    if (!nd) return ORC_EINVAL;
    nd->factor = factor;
    nd->source_width = p->cur_w;
    nd->out_w = ceil_mode ? (p->cur_w + factor - 1) / factor : p->cur_w / factor;
    nd->param = factor;
    commit(p, nd);
    return ORC_OK;
}

int orc_pipeline_shrink_v(orc_pipeline *p, uint32_t factor, int ceil_mode) {
    if (factor == 0) return ORC_EINVAL;
    node_t *nd = push(p, N_SHRINK_V);
    if (!nd) return ORC_EINVAL;
    nd->factor = factor;
    nd->out_h = ceil_mode ? (p->cur_h + factor - 1) / factor : p->cur_h / factor;
    nd->param = factor;
    commit(p, nd);
    return ORC_OK;
}

/* PipelineBuilder::shrink(h, v) -> ShrinkBridge / Shrink<F> (builder/resample.rs:
 * 219-268; shrink.rs:37-293): ShrinkH (without the full-row borrow) composed with
 * ShrinkV inside one op; the u8 2x2 direct path (shrink.rs:225-293) is the same
 * two pair roundings. */
int orc_pipeline_shrink(orc_pipeline *p, uint32_t hshrink, uint32_t vshrink) {
    if (hshrink == 0 || vshrink == 0) return ORC_EINVAL;
    int rc = orc_pipeline_shrink_h(p, hshrink, 0);
    if (rc != ORC_OK) return rc;
    p->nodes[p->n - 1].source_width = 0;
    return orc_pipeline_shrink_v(p, vshrink, 0);
}

/* PipelineBuilder::gauss_blur (builder/colour.rs:130-197): U8 stays U8, every
 * other format produces F32 (gauss_blur.rs:214-403). */
int orc_pipeline_gauss_blur(orc_pipeline *p, float sigma) {
    int16_t c[512];
    int32_t scale;
    int taps = orc_gauss_int_kernel(sigma, c, 512, &scale);
    if (taps < 0) return ORC_EINVAL;
    for (int pass = 0; pass < 2; ++pass) {
        node_t *nd = push(p, pass == 0 ? N_GAUSS_H : N_GAUSS_V);
        if (!nd) return ORC_EINVAL;
        nd->ntaps = taps;
        memcpy(nd->gi, c, sizeof(int16_t) * (size_t)taps);
        nd->gscale = scale;
        nd->gf = (double *)malloc(sizeof(double) * (size_t)taps);
        /* gaussian_kernel_1d -> normalise_integer_kernel (gauss_kernel.rs:58-65,113-117) */
        for (int i = 0; i < taps; ++i) nd->gf[i] = (double)c[i] / (double)scale;
        nd->out_fmt = (p->cur_fmt == ORC_U8) ? ORC_U8 : ORC_F32;
        nd->param = sigma;
        commit(p, nd);
    }
    return ORC_OK;
}

/* ConvSep::with_mask (convsep.rs:218-233): H uses (scale, offset), V uses
 * (scale, 0). validate_kernel_1d (convolution/common.rs:11-27). */
int orc_pipeline_convsep(orc_pipeline *p, const double *coeffs, uint32_t n, double scale,
                         double offset) {
    if (n == 0 || n % 2 == 0 || n > 511) return ORC_EINVAL;
    for (uint32_t i = 0; i < n; ++i)
        if (!isfinite(coeffs[i])) return ORC_EINVAL;
    if (!isfinite(scale) || scale == 0.0 || !isfinite(offset)) return ORC_EINVAL;
    for (int pass = 0; pass < 2; ++pass) {
        node_t *nd = push(p, pass == 0 ? N_CONVSEP_H : N_CONVSEP_V);
        if (!nd) return ORC_EINVAL;
        nd->ntaps = (int)n;
        nd->gf = (double *)malloc(sizeof(double) * n);
        memcpy(nd->gf, coeffs, sizeof(double) * n);
        nd->cscale = scale;
        nd->coffset = pass == 0 ? offset : 0.0;
        nd->out_fmt = ORC_F32;
        commit(p, nd);
    }
    return ORC_OK;
}

int orc_pipeline_premultiply(orc_pipeline *p) {
    if (p->cur_fmt != ORC_U8) return ORC_EFORMAT;
    node_t *nd = push(p, N_PREMULT);
    if (!nd) return ORC_EINVAL;
    commit(p, nd);
    return ORC_OK;
}

int orc_pipeline_unpremultiply(orc_pipeline *p) {
    if (p->cur_fmt != ORC_U8) return ORC_EFORMAT;
    node_t *nd = push(p, N_UNPREMULT);
    if (!nd) return ORC_EINVAL;
    commit(p, nd);
    return ORC_OK;
}

/* PipelineBuilder::affine_backward_with_extend_and_hint (builder/resample.rs:
 * 478-560) -> AffineBridge::new_with_extend (builder/state.rs:298-341) ->
 * Affine::try_new (affine/interpolation.rs:13-41): finite coefficients and a
 * non-singular matrix; the diagonal matrix [a, 0, 0, d] of the thumbnail,
 * ExtendMode::Copy, source bounds = the whole input. */
int orc_pipeline_affine(orc_pipeline *p, double a, double d, double tx, double ty, uint32_t ow, uint32_t oh,
                        int kernel, int small_tile) {
    const double m[4] = {a, 0.0, 0.0, d};
    for (int i = 0; i < 4; ++i)
        if (!isfinite(m[i])) return ORC_EINVAL;
    if (!isfinite(tx) || !isfinite(ty)) return ORC_EINVAL;
    if (fabs(fma(m[1], -m[2], m[0] * m[3])) <= DBL_EPSILON) return ORC_EINVAL;
    if (kernel < ORC_NEAREST || kernel > ORC_VSQBS) return ORC_EUNSUPPORTED; /* Nohalo is not restated */
    node_t *nd = push(p, N_AFFINE);
    if (!nd) return ORC_EINVAL;
    memcpy(nd->am, m, sizeof(m));
    nd->atx = tx;
    nd->aty = ty;
    nd->akern = kernel;
    nd->asmall = small_tile;
    nd->out_w = ow;
    nd->out_h = oh;
    commit(p, nd);
    return ORC_OK;
}

/* ------------------------------------------------------ thumbnail.rs planner */

enum { T_PREMULT = 1, T_SHRINK_H, T_SHRINK_V, T_REDUCE_H, T_REDUCE_V, T_UNPREMULT, T_EXTRACT, T_AFFINE };

/* Thumbnail::target_dimensions (thumbnail.rs:458-485). */
static void target_dims(int target, uint32_t tw, uint32_t th, uint32_t iw, uint32_t ih,
                        uint32_t *ow, uint32_t *oh) {
    double fw = (double)iw, fh = (double)ih;
    switch (target) {
    case ORC_FITBOX: {
        double hs = (double)(tw > 1 ? tw : 1) / fw;
        double vs = (double)(th > 1 ? th : 1) / fh;
        double s = fmin(fmin(hs, vs), 1.0);
        double a = round(fw * s), b = round(fh * s);
        *ow = sat_u32(a > 1.0 ? a : 1.0);
        *oh = sat_u32(b > 1.0 ? b : 1.0);
        break;
    }
    case ORC_FORCE_EXACT:
        *ow = tw > 1 ? tw : 1;
        *oh = th > 1 ? th : 1;
        break;
    case ORC_WIDTH: {
        uint32_t w = tw > 1 ? tw : 1;
        if (w > iw) w = iw;
        double b = round((fh * (double)w) / fw);
        *ow = w;
        *oh = sat_u32(b > 1.0 ? b : 1.0);
        break;
    }
    default: {
        uint32_t h = th > 1 ? th : 1;
        if (h > ih) h = ih;
        double a = round((fw * (double)h) / fh);
        *oh = h;
        *ow = sat_u32(a > 1.0 ? a : 1.0);
        break;
    }
    }
}

/* choose_integer_shrink (thumbnail.rs:488-511). */
static uint32_t choose_integer_shrink(uint32_t in_len, uint32_t target, double total) {
    if (total <= 1.0 + PLAN_EPS) return 1;
    double f0 = floor(total);
    uint32_t f = sat_u32(f0 > 1.0 ? f0 : 1.0);
    while (f > 1) {
        uint32_t after = in_len / f;
        double residual = total / (double)f;
        uint32_t ar;
        if (residual > 1.0 + PLAN_EPS) {
            double v = round((double)after / residual);
            ar = sat_u32(v > 1.0 ? v : 1.0);
        } else {
            ar = after;
        }
        if (ar >= target) break;
        --f;
    }
    return f;
}

/* Thumbnail::into_pipeline_nodes_without_shrink_hint (thumbnail.rs:229-456),
 * crop = false. Returns node count, or ORC_EUNSUPPORTED when the plan needs an
 * Affine geometry fix-up. */
/* Thumbnail::into_pipeline_nodes_internal (thumbnail.rs:229-436) with the
 * crop option (thumbnail.rs:185-190): FitBox / ForceExact crops keep the box
 * as the target; the geometry is the covering size round(in * max(scale))
 * (at least the target), and a centred ExtractArea ((geometry - target) / 2)
 * follows the reduce nodes when the two differ.  extract[4] receives that
 * rectangle (x, y, width, height) for the T_EXTRACT node, and affine[7] the
 * T_AFFINE node's (a, d, tx, ty, output width, output height, kernel): the
 * 1-pixel-input branch (thumbnail.rs:286-323) and the geometry fix-up after
 * the reduce nodes (thumbnail.rs:396-419), both with the kernel mapped by
 * affine_kernel_for_resize_kernel (thumbnail.rs:31-37). */
static void thumb_affine(double inv_x, double inv_y, uint32_t ow, uint32_t oh, int kernel, double *affine) {
    affine[0] = inv_x;
    affine[1] = inv_y;
    affine[2] = kernel == ORC_NEAREST ? 0.0 : 0.5 * (1.0 - inv_x);
    affine[3] = kernel == ORC_NEAREST ? 0.0 : 0.5 * (1.0 - inv_y);
    affine[4] = ow;
    affine[5] = oh;
    affine[6] = kernel == ORC_NEAREST ? ORC_NEAREST : (kernel == ORC_BILINEAR ? ORC_BILINEAR : ORC_BICUBIC);
}

int orc_thumbnail_plan_crop(uint32_t iw, uint32_t ih, uint32_t bands, int target, uint32_t tw,
                            uint32_t th, int kernel, int crop, int *types, double *params, int max_nodes,
                            uint32_t *out_w, uint32_t *out_h, uint32_t *extract, double *affine) {
    if (bands == 0) return ORC_EINVAL;
    if (iw == 0 || ih == 0) {
        *out_w = *out_h = 0;
        return 0;
    }
    int allow_upscale = target == ORC_FORCE_EXACT;
    int single = iw == 1 && ih == 1;
    uint32_t tgw, tgh;
    if (crop && (target == ORC_FITBOX || target == ORC_FORCE_EXACT)) {
        tgw = tw > 1 ? tw : 1;
        tgh = th > 1 ? th : 1;
    } else {
        target_dims(target, tw, th, iw, ih, &tgw, &tgh);
    }
    if (allow_upscale && !single) {
        if (tgw < 1) tgw = 1;
        if (tgh < 1) tgh = 1;
    } else {
        if (tgw > iw) tgw = iw;
        if (tgh > ih) tgh = ih;
        if (tgw < 1) tgw = 1;
        if (tgh < 1) tgh = 1;
    }
    uint32_t gw = tgw, gh = tgh;
    if (crop) {
        double hsc = (double)(tgw > 1 ? tgw : 1) / (double)iw;
        double vsc = (double)(tgh > 1 ? tgh : 1) / (double)ih;
        double cover = hsc > vsc ? hsc : vsc;
        double cw = round((double)iw * cover), ch = round((double)ih * cover);
        gw = sat_u32(cw > (double)tgw ? cw : (double)tgw);
        gh = sat_u32(ch > (double)tgh ? ch : (double)tgh);
    }
    if (allow_upscale && !single) {
        if (gw < 1) gw = 1;
        if (gh < 1) gh = 1;
    } else {
        if (gw > iw) gw = iw;
        if (gh > ih) gh = ih;
        if (gw < 1) gw = 1;
        if (gh < 1) gh = 1;
    }
    int has_alpha = bands == 2 || bands == 4;
    *out_w = tgw;
    *out_h = tgh;
    int geo[8];
    double gp[8];
    int ng = 0;
    if (iw == 1 || ih == 1) {
        if (single) *out_w = *out_h = 1;
        if (*out_w != iw || *out_h != ih) {
            thumb_affine((double)iw / (double)*out_w, (double)ih / (double)*out_h, *out_w, *out_h, kernel, affine);
            geo[ng] = T_AFFINE;
            gp[ng++] = 0.0;
        }
    } else {
        double hs = (double)iw / (double)gw;
        double vs = (double)ih / (double)gh;
        uint32_t ih_ = choose_integer_shrink(iw, gw, hs);
        uint32_t iv_ = choose_integer_shrink(ih, gh, vs);
        uint32_t asw = iw / (ih_ > 1 ? ih_ : 1);
        uint32_t ash = ih / (iv_ > 1 ? iv_ : 1);
        double rh = hs > 1.0 ? hs / (double)ih_ : 1.0;
        double rv = vs > 1.0 ? vs / (double)iv_ : 1.0;
        uint32_t arw = asw, arh = ash;
        if (rh > 1.0 + PLAN_EPS) {
            double v = round((double)asw / rh);
            arw = sat_u32(v > 1.0 ? v : 1.0);
        }
        if (rv > 1.0 + PLAN_EPS) {
            double v = round((double)ash / rv);
            arh = sat_u32(v > 1.0 ? v : 1.0);
        }
        if (ih_ > 1) { geo[ng] = T_SHRINK_H; gp[ng++] = ih_; }
        if (iv_ > 1) { geo[ng] = T_SHRINK_V; gp[ng++] = iv_; }
        if (rh > 1.0 + PLAN_EPS) { geo[ng] = T_REDUCE_H; gp[ng++] = rh; }
        if (rv > 1.0 + PLAN_EPS) { geo[ng] = T_REDUCE_V; gp[ng++] = rv; }
        double ax = (double)gw / (double)arw, ay = (double)gh / (double)arh;
        if (fabs(ax - 1.0) > PLAN_EPS || fabs(ay - 1.0) > PLAN_EPS) {
            thumb_affine(1.0 / ax, 1.0 / ay, gw, gh, kernel, affine);
            geo[ng] = T_AFFINE;
            gp[ng++] = 0.0;
        }
        if (crop && (gw != tgw || gh != tgh)) { /* thumbnail.rs:421-428 */
            extract[0] = (gw - tgw) / 2;
            extract[1] = (gh - tgh) / 2;
            extract[2] = tgw;
            extract[3] = tgh;
            geo[ng] = T_EXTRACT;
            gp[ng++] = 0.0;
        }
    }
    /* wrap_geometry_nodes (thumbnail.rs:438-456) */
    int n = 0;
    if (has_alpha && ng > 0) {
        if (n >= max_nodes) return ORC_EINVAL;
        types[n] = T_PREMULT;
        params[n++] = 0.0;
    }
    for (int i = 0; i < ng; ++i) {
        if (n >= max_nodes) return ORC_EINVAL;
        types[n] = geo[i];
        params[n++] = gp[i];
    }
    if (has_alpha && ng > 0) {
        if (n >= max_nodes) return ORC_EINVAL;
        types[n] = T_UNPREMULT;
        params[n++] = 0.0;
    }
    return n;
}

/* shrink_on_load_factor (thumbnail.rs:514-524) of the plan's common shrink
 * min(in / geometry) over the axes (thumbnail.rs:326-330); 0 = None, also for
 * 1-pixel inputs (thumbnail.rs:286-323) and empty ones. */
uint32_t orc_shrink_on_load_factor(double common_shrink) {
    return common_shrink >= 16.0 ? 16 : (common_shrink >= 8.0 ? 4 : (common_shrink >= 4.0 ? 2 : 0));
}

uint32_t orc_thumbnail_shrink_on_load(uint32_t iw, uint32_t ih, int target, uint32_t tw, uint32_t th, int crop) {
    if (iw <= 1 || ih <= 1) return 0;
    int allow_upscale = target == ORC_FORCE_EXACT;
    uint32_t tgw, tgh;
    if (crop && (target == ORC_FITBOX || target == ORC_FORCE_EXACT)) {
        tgw = tw > 1 ? tw : 1;
        tgh = th > 1 ? th : 1;
    } else {
        target_dims(target, tw, th, iw, ih, &tgw, &tgh);
    }
    if (!allow_upscale) {
        if (tgw > iw) tgw = iw;
        if (tgh > ih) tgh = ih;
    }
    if (tgw < 1) tgw = 1;
    if (tgh < 1) tgh = 1;
    uint32_t gw = tgw, gh = tgh;
    if (crop) {
        double cover = fmax((double)tgw / (double)iw, (double)tgh / (double)ih);
        double cw = round((double)iw * cover), ch = round((double)ih * cover);
        gw = sat_u32(fmax(cw, (double)tgw));
        gh = sat_u32(fmax(ch, (double)tgh));
        if (!allow_upscale) {
            if (gw > iw) gw = iw;
            if (gh > ih) gh = ih;
        }
    }
    return orc_shrink_on_load_factor(fmin((double)iw / (double)gw, (double)ih / (double)gh));
}

int orc_thumbnail_plan(uint32_t iw, uint32_t ih, uint32_t bands, int target, uint32_t tw,
                       uint32_t th, int kernel, int *types, double *params, int max_nodes,
                       uint32_t *out_w, uint32_t *out_h) {
    uint32_t extract[4];
    double affine[7];
    return orc_thumbnail_plan_crop(iw, ih, bands, target, tw, th, kernel, 0, types, params, max_nodes, out_w,
                                   out_h, extract, affine);
}

/* PipelineBuilder::thumbnail (builder/resample.rs:871-979) for an in-memory
 * source: the shrink-on-load hint is never honoured, so the plan is the
 * "without shrink hint" one. */
int orc_pipeline_thumbnail_crop(orc_pipeline *p, int target, uint32_t tw, uint32_t th, int kernel, int crop) {
    int types[16];
    double params[16];
    uint32_t ow, oh, extract[4] = {0, 0, 0, 0};
    double affine[7] = {0, 0, 0, 0, 0, 0, 0};
    if (p->bands == 0) return ORC_EINVAL;
    if (!valid_reduce_kernel(kernel)) return ORC_EKERNEL;
    const uint32_t iw = p->cur_w, ih = p->cur_h;
    int n = orc_thumbnail_plan_crop(iw, ih, p->bands, target, tw, th, kernel, crop, types, params, 16,
                                    &ow, &oh, extract, affine);
    if (n < 0) return n;
    /* builder/resample.rs:908-917,939-943: inputs above 512 px run as
     * ThinStrip (forced when the thumbnail is the first node); smaller ones
     * give the Affine the SmallTile demand. */
    const int large = iw > 512 || ih > 512;
    if (p->n == 0 && large) p->thin_override = 1;
    for (int i = 0; i < n; ++i) {
        int rc = ORC_OK;
        switch (types[i]) {
        case T_PREMULT: rc = orc_pipeline_premultiply(p); break;
        case T_UNPREMULT: rc = orc_pipeline_unpremultiply(p); break;
        case T_SHRINK_H: rc = orc_pipeline_shrink_h(p, (uint32_t)params[i], 0); break;
        case T_SHRINK_V: rc = orc_pipeline_shrink_v(p, (uint32_t)params[i], 0); break;
        case T_REDUCE_H: rc = orc_pipeline_reduce_h(p, params[i], kernel); break;
        case T_REDUCE_V: rc = orc_pipeline_reduce_v(p, params[i], kernel); break;
        case T_EXTRACT: rc = orc_pipeline_extract_area(p, extract[0], extract[1], extract[2], extract[3]); break;
        case T_AFFINE:
            rc = orc_pipeline_affine(p, affine[0], affine[1], affine[2], affine[3], (uint32_t)affine[4],
                                 (uint32_t)affine[5], (int)affine[6], !large);
            break;
        }
        if (rc != ORC_OK) return rc;
    }
    if (n == 0) {
        node_t *nd = push(p, N_COPY);
        if (!nd) return ORC_EINVAL;
        commit(p, nd);
    }
    return ORC_OK;
}

int orc_pipeline_thumbnail(orc_pipeline *p, int target, uint32_t tw, uint32_t th, int kernel) {
    return orc_pipeline_thumbnail_crop(p, target, tw, th, kernel, 0);
}

/* The source-side box shrinks of the PNG thumbnail path.
 *
 * ceil_mode == 0: software_box_shrink_generic (sources/decoder_source/
 * shrink.rs:47-117), the shrink applied to an eager decoded backing:
 * dst = max(src / f, 1) per axis; window [d*f, min(d*f + f, src)) (cut short
 * only when the source is smaller than f); u8 sums in integers and rounds
 * (sum + n/2) / n with n = the window's pixel count; the other formats sum in
 * f64 in row order (sy outer, sx inner) and store from_f64_clamped(sum / n)
 * (shrinkh.rs:44-128).
 *
 * ceil_mode == 1 (u8 only): decode_png_with_box_shrink_u8 (codecs/src/png/
 * decode_full.rs:320-452), the shrink applied row by row while a PNG file
 * decodes: dst = ceil(src / f); the last column and row take the partial
 * windows, each divided by its own pixel count with the same rounding. */
int orc_box_shrink(const void *src, uint32_t w, uint32_t h, uint32_t bands, int fmt, uint32_t f, int ceil_mode,
                   void *dst, uint32_t *out_w, uint32_t *out_h) {
    if (f == 0 || bands == 0 || fmt < ORC_U8 || fmt > ORC_F64) return ORC_EINVAL;
    if (ceil_mode && fmt != ORC_U8) return ORC_EFORMAT;
    uint32_t dw, dh;
    if (ceil_mode) {
        dw = (w + f - 1) / f;
        dh = (h + f - 1) / f;
    } else {
        dw = w / f > 1 ? w / f : 1;
        dh = h / f > 1 ? h / f : 1;
    }
    if (w == 0 || h == 0) dw = dh = 0;
    if (out_w) *out_w = dw;
    if (out_h) *out_h = dh;
    if (!dst) return ORC_OK;
    const size_t B = bands;
    for (uint32_t dy = 0; dy < dh; ++dy)
        for (uint32_t dx = 0; dx < dw; ++dx) {
            const size_t sy0 = (size_t)dy * f, sx0 = (size_t)dx * f;
            const size_t sy1 = sy0 + f < h ? sy0 + f : h, sx1 = sx0 + f < w ? sx0 + f : w;
            size_t n = (sy1 - sy0) * (sx1 - sx0);
            if (n == 0) n = 1;
            const size_t o = ((size_t)dy * dw + dx) * B;
            for (size_t b = 0; b < B; ++b) {
                if (fmt == ORC_U8) {
                    const uint8_t *s = (const uint8_t *)src;
                    uint32_t sum = 0;
                    for (size_t sy = sy0; sy < sy1; ++sy)
                        for (size_t sx = sx0; sx < sx1; ++sx) sum += s[(sy * w + sx) * B + b];
                    ((uint8_t *)dst)[o + b] = (uint8_t)((sum + (uint32_t)n / 2) / (uint32_t)n);
                } else {
                    double sum = 0.0;
                    for (size_t sy = sy0; sy < sy1; ++sy)
                        for (size_t sx = sx0; sx < sx1; ++sx) sum += ld(src, fmt, (sy * w + sx) * B + b);
                    store_shrunk(dst, fmt, o + b, sum / (double)n);
                }
            }
        }
    return ORC_OK;
}

/* The source replaced by its box-shrunk copy before any node is added
 * (DecoderSource::set_thumbnail_shrink_on_load, sources/decoder_source/
 * source_impl.rs:245-275): the pipeline then sees the shrunk dimensions. */
int orc_pipeline_box_shrink(orc_pipeline *p, uint32_t factor, int ceil_mode) {
    if (!p || factor == 0) return ORC_EINVAL;
    if (p->n != 0 || p->box_f != 0) return ORC_EINVAL;
    uint32_t dw, dh;
    int rc = orc_box_shrink(NULL, p->src_w, p->src_h, p->bands, p->src_fmt, factor, ceil_mode, NULL, &dw, &dh);
    if (rc != ORC_OK) return rc;
    p->box_f = factor;
    p->box_ceil = ceil_mode;
    p->src_w = p->cur_w = dw;
    p->src_h = p->cur_h = dh;
    return ORC_OK;
}

/* PipelineBuilder::thumbnail over a source with the PNG SoftwareBoxShrink
 * pre-shrink (builder/resample.rs:871-906; decoder_source/shrink.rs:5-30,
 * 119-125): when the thumbnail is the first node and the plan's shrink-on-load
 * hint is set, the eager backing is box-shrunk by the hint (2, 4 or 16: all
 * kept by normalize_shrink_factor) and the thumbnail is planned again, without
 * the hint, over the shrunk dimensions. */
int orc_pipeline_thumbnail_shrink_on_load(orc_pipeline *p, int target, uint32_t tw, uint32_t th, int kernel,
                                          int crop) {
    if (!p) return ORC_EINVAL;
    if (p->n == 0 && p->box_f == 0) {
        const uint32_t f = orc_thumbnail_shrink_on_load(p->cur_w, p->cur_h, target, tw, th, crop);
        if (f == 2 || f == 4 || f == 8 || f == 16) {
            const int rc = orc_pipeline_box_shrink(p, f, 0);
            if (rc != ORC_OK) return rc;
        }
    }
    return orc_pipeline_thumbnail_crop(p, target, tw, th, kernel, crop);
}

/* ResizeOp::into_pipeline_nodes (resize.rs:163-284) and PipelineBuilder::resize
 * (builder/resample.rs:637-711). */
int orc_pipeline_resize(orc_pipeline *p, double hscale, double vscale, int kernel) {
    if (p->bands == 0) return ORC_EINVAL;
    if (!(isfinite(hscale) && hscale > 0.0)) hscale = 1.0;
    if (!(isfinite(vscale) && vscale > 0.0)) vscale = 1.0;
    if (kernel == ORC_NOHALO) kernel = ORC_BICUBIC;
    const double gap = 2.0;
    uint32_t iw = p->cur_w, ih = p->cur_h;
    if (iw == 0 || ih == 0) return ORC_EUNSUPPORTED;
    double a = round((double)iw * hscale), b = round((double)ih * vscale);
    uint32_t tw = sat_u32(a > 1.0 ? a : 1.0), th = sat_u32(b > 1.0 ? b : 1.0);
    int rc;
    if (kernel == ORC_LBB) {
        /* resize.rs:176-199: the whole resize is one LBB Affine */
        if (tw != iw || th != ih) {
            const double ax = (double)tw / (double)iw, ay = (double)th / (double)ih;
            rc = orc_pipeline_affine(p, 1.0 / ax, 1.0 / ay, 0.5 * (1.0 - 1.0 / ax), 0.5 * (1.0 - 1.0 / ay), tw, th,
                                     kernel, 1);
            if (rc != ORC_OK) return rc;
        } else {
            node_t *nd = push(p, N_COPY);
            if (!nd) return ORC_EINVAL;
            commit(p, nd);
        }
        return ORC_OK;
    }
    uint32_t cw = iw, ch = ih;
    uint32_t hsh = 1, vsh = 1;
    if (tw < cw && kernel != ORC_NEAREST) {
        double v = floor((double)cw / (double)tw / gap);
        hsh = sat_u32(v > 1.0 ? v : 1.0);
    }
    if (th < ch && kernel != ORC_NEAREST) {
        double v = floor((double)ch / (double)th / gap);
        vsh = sat_u32(v > 1.0 ? v : 1.0);
    }
    int any = 0;
    if (vsh > 1) {
        if ((rc = orc_pipeline_shrink_v(p, vsh, 0)) != ORC_OK) return rc;
        ch /= vsh;
        any = 1;
    }
    if (hsh > 1) {
        if ((rc = orc_pipeline_shrink_h(p, hsh, 0)) != ORC_OK) return rc;
        cw /= hsh;
        any = 1;
    }
    if (th < ch) {
        if ((rc = orc_pipeline_reduce_v(p, (double)ch / (double)th, kernel)) != ORC_OK) return rc;
        ch = th;
        any = 1;
    }
    if (tw < cw) {
        if ((rc = orc_pipeline_reduce_h(p, (double)cw / (double)tw, kernel)) != ORC_OK) return rc;
        cw = tw;
        any = 1;
    }
    /* resize.rs:240-280: integer Zoom of what remains to grow (not for
     * VSQBS, resize.rs:241-243), then the Affine fix-up (SmallTile demand,
     * builder/resample.rs:690-706).  A VSQBS reduce above fails with
     * ORC_EKERNEL (reduce_common.rs:107-113). */
    const int zoom_ok = kernel != ORC_VSQBS;
    const uint32_t zx = zoom_ok && tw > cw ? tw / cw : 1, zy = zoom_ok && th > ch ? th / ch : 1;
    if (zx > 1 || zy > 1) {
        node_t *nd = push(p, N_ZOOM);
        if (!nd) return ORC_EINVAL;
        nd->zx = zx;
        nd->zy = zy;
        nd->out_w = cw * zx;
        nd->out_h = ch * zy;
        commit(p, nd);
        cw *= zx;
        ch *= zy;
        any = 1;
    }
    const double ax = (double)tw / (double)cw, ay = (double)th / (double)ch;
    if (fabs(ax - 1.0) > PLAN_EPS || fabs(ay - 1.0) > PLAN_EPS) {
        const double ox = kernel == ORC_NEAREST ? 0.0 : 0.5 * (1.0 - 1.0 / ax);
        const double oy = kernel == ORC_NEAREST ? 0.0 : 0.5 * (1.0 - 1.0 / ay);
        if ((rc = orc_pipeline_affine(p, 1.0 / ax, 1.0 / ay, ox, oy, tw, th, kernel, 1)) != ORC_OK) return rc;
        any = 1;
    }
    if (!any) {
        node_t *nd = push(p, N_COPY);
        if (!nd) return ORC_EINVAL;
        commit(p, nd);
    }
    return ORC_OK;
}

int orc_pipeline_output(const orc_pipeline *p, uint32_t *w, uint32_t *h, uint32_t *bands,
                        int *fmt) {
    *w = p->cur_w;
    *h = p->cur_h;
    *bands = p->bands;
    *fmt = p->cur_fmt;
    return ORC_OK;
}

int orc_pipeline_node_count(const orc_pipeline *p) { return p->n; }

int orc_pipeline_node_info(const orc_pipeline *p, int i, int *type, double *param) {
    if (i < 0 || i >= p->n) return ORC_EINVAL;
    *type = p->nodes[i].type;
    *param = p->nodes[i].param;
    return ORC_OK;
}

/* ------------------------------------------------- required input regions */

/* InterpolationKernel::affine_padding (kernel.rs:80-117): window_offset
 * and window_size - offset - 1. */
static void affine_pad(int k, int64_t *left, int64_t *right) {
    const int64_t size = k == ORC_NEAREST ? 1 : (k == ORC_BILINEAR ? 2 : (k == ORC_LANCZOS3 ? 6 : 4));
    *left = (k == ORC_NEAREST || k == ORC_BILINEAR) ? 0 : (k == ORC_LANCZOS3 ? 2 : 1);
    *right = size - *left - 1;
}

/* Affine::map_to_input (affine/core.rs:78-84). */
static void affine_map(const node_t *nd, double x, double y, double *mx, double *my) {
    *mx = fma(nd->am[1], y, nd->am[0] * x) + nd->atx;
    *my = fma(nd->am[3], y, nd->am[2] * x) + nd->aty;
}

/* Affine::required_input_axis_bounds with source bounds (0, len)
 * (affine/core.rs:373-405). */
static void affine_axis(int64_t start, int64_t end, uint32_t len, int64_t *o, int64_t *n) {
    if (len == 0) {
        *o = 0;
        *n = 0;
        return;
    }
    const int64_t be = (int64_t)len - 1;
    const int64_t cs = start > 0 ? start : 0, ce = end < be ? end : be;
    if (ce < cs) {
        *o = clampi(end < 0 ? 0 : be + 1, INT32_MIN, INT32_MAX);
        *n = 0;
        return;
    }
    const int64_t rs = clampi(cs, INT32_MIN, INT32_MAX), re = clampi(ce, INT32_MIN, INT32_MAX);
    *o = rs;
    *n = re - rs + 1;
}

static rgn required_input(const node_t *nd, rgn o) {
    rgn r = o;
    switch (nd->type) {
    case N_SHRINK_H: { /* shrinkh.rs:220-235 */
        int64_t need = o.w * (int64_t)nd->factor;
        if (o.x == 0 && nd->source_width != 0 && need <= (int64_t)nd->source_width) {
            r.x = 0;
            r.w = nd->source_width;
        } else {
            r.x = clampi(o.x * (int64_t)nd->factor, INT32_MIN, INT32_MAX);
            r.w = need;
        }
        break;
    }
    case N_SHRINK_V: /* shrinkv.rs:122-129 */
        r.y = clampi(o.y * (int64_t)nd->factor, INT32_MIN, INT32_MAX);
        r.h = o.h * (int64_t)nd->factor;
        break;
    case N_REDUCE_H:
    case N_REDUCE_OP_H: rk_required(&nd->rk, o.x, o.w, &r.x, &r.w); break;
    case N_REDUCE_V:
    case N_REDUCE_OP_V: rk_required(&nd->rk, o.y, o.h, &r.y, &r.h); break;
    case N_EXTRACT: /* extract_area.rs:49-56 */
        r.x = clampi(o.x + nd->ex, INT32_MIN, INT32_MAX);
        r.y = clampi(o.y + nd->ey, INT32_MIN, INT32_MAX);
        break;
    case N_ZOOM: { /* zoom.rs:343-349 */
        const int64_t left = o.x / nd->zx, top = o.y / nd->zy;
        const int64_t right = (o.x + o.w - 1) / nd->zx, bottom = (o.y + o.h - 1) / nd->zy;
        r.x = left;
        r.y = top;
        r.w = right - left + 1;
        r.h = bottom - top + 1;
        break;
    }
    case N_AFFINE: { /* affine/core.rs:344-366,1531-1560: mapped corner pixels, floor, padding */
        const double ox0 = (double)o.x, oy0 = (double)o.y;
        const double ox1 = (double)o.x + (double)(o.w > 0 ? o.w - 1 : 0);
        const double oy1 = (double)o.y + (double)(o.h > 0 ? o.h - 1 : 0);
        double cx[4], cy[4];
        affine_map(nd, ox0, oy0, &cx[0], &cy[0]);
        affine_map(nd, ox1, oy0, &cx[1], &cy[1]);
        affine_map(nd, ox0, oy1, &cx[2], &cy[2]);
        affine_map(nd, ox1, oy1, &cx[3], &cy[3]);
        double minx = fmin(fmin(fmin(cx[0], cx[1]), cx[2]), cx[3]);
        double maxx = fmax(fmax(fmax(cx[0], cx[1]), cx[2]), cx[3]);
        double miny = fmin(fmin(fmin(cy[0], cy[1]), cy[2]), cy[3]);
        double maxy = fmax(fmax(fmax(cy[0], cy[1]), cy[2]), cy[3]);
        int64_t lp, rp;
        affine_pad(nd->akern, &lp, &rp);
        affine_axis((int64_t)floor(minx) - lp, (int64_t)floor(maxx) + rp, nd->in_w, &r.x, &r.w);
        affine_axis((int64_t)floor(miny) - lp, (int64_t)floor(maxy) + rp, nd->in_h, &r.y, &r.h);
        break;
    }
    case N_GAUSS_H:
    case N_CONVSEP_H: { /* gauss_blur.rs:1801-1808; convsep.rs:65-72 */
        int64_t rad = (nd->ntaps - 1) / 2;
        r.x = o.x - rad;
        r.w = o.w + 2 * rad;
        break;
    }
    case N_GAUSS_V:
    case N_CONVSEP_V: {
        int64_t rad = (nd->ntaps - 1) / 2;
        r.y = o.y - rad;
        r.h = o.h + 2 * rad;
        break;
    }
    default: break; /* pixel-local */
    }
    return r;
}

/* --------------------------------------------------------- sample access */

/* ToF64 (viprs-core shared_ops/sample_conv.rs:1-50): every sample exactly. */
static inline double ld(const void *b, int fmt, size_t i) {
    switch (fmt) {
    case ORC_U8: return (double)((const uint8_t *)b)[i];
    case ORC_U16: return (double)((const uint16_t *)b)[i];
    case ORC_I16: return (double)((const int16_t *)b)[i];
    case ORC_U32: return (double)((const uint32_t *)b)[i];
    case ORC_I32: return (double)((const int32_t *)b)[i];
    case ORC_F64: return ((const double *)b)[i];
    default: return (double)((const float *)b)[i];
    }
}
/* ReduceSample::to_i64 (resample/sample_conv.rs:18-62). */
static inline int64_t ldi(const void *b, int fmt, size_t i) {
    switch (fmt) {
    case ORC_U8: return (int64_t)((const uint8_t *)b)[i];
    case ORC_U16: return (int64_t)((const uint16_t *)b)[i];
    case ORC_I16: return (int64_t)((const int16_t *)b)[i];
    case ORC_U32: return (int64_t)((const uint32_t *)b)[i];
    default: return (int64_t)((const int32_t *)b)[i];
    }
}
/* ReduceSample::from_fixed_i64 (resample/sample_conv.rs:18-62): unsigned
 * (v + 2048) >> 12, signed (v - 2048) >> 12 for a negative v (arithmetic
 * shifts), then the type's range. */
static void store_fixed(void *out, int fmt, size_t oi, int64_t v) {
    const int64_t r = (fmt == ORC_I16 || fmt == ORC_I32) && v < 0 ? (v - 2048) >> 12 : (v + 2048) >> 12;
    const int64_t q = clampi(r, (int64_t)fmt_lo(fmt), (int64_t)fmt_hi(fmt));
    switch (fmt) {
    case ORC_U8: ((uint8_t *)out)[oi] = (uint8_t)q; break;
    case ORC_U16: ((uint16_t *)out)[oi] = (uint16_t)q; break;
    case ORC_I16: ((int16_t *)out)[oi] = (int16_t)q; break;
    case ORC_U32: ((uint32_t *)out)[oi] = (uint32_t)q; break;
    default: ((int32_t *)out)[oi] = (int32_t)q; break;
    }
}

/* ------------------------------------------------------- process_region */

static void process_reduce_h(const node_t *nd, const void *in, rgn ir, void *out, rgn orr) {
    const rk_t *k = &nd->rk;
    size_t B = nd->bands, in_w = (size_t)ir.w, out_w = (size_t)orr.w;
    int64_t *starts = (int64_t *)malloc(sizeof(int64_t) * (out_w + 1));
    int *phases = (int *)malloc(sizeof(int) * (out_w + 1));
    if (nd->in_fmt == ORC_U8) {
        /* reduceh.rs:101-120: plan from f64 accumulation starting at sp(region.x) */
        double sx = rk_sp(k, (double)orr.x);
        for (size_t i = 0; i < out_w; ++i) {
            rk_plan(k, sx, &starts[i], &phases[i]);
            sx += k->factor;
        }
    } else {
        /* reduce_simd.rs:63,2565,2642: direct position per column */
        for (size_t i = 0; i < out_w; ++i) rk_plan(k, rk_sp(k, (double)orr.x + (double)i), &starts[i], &phases[i]);
    }
    int T = k->taps;
    for (int64_t y = 0; y < orr.h; ++y) {
        for (size_t xo = 0; xo < out_w; ++xo) {
            for (size_t b = 0; b < B; ++b) {
                size_t oi = ((size_t)y * out_w + xo) * B + b;
                if (fmt_is_int(nd->in_fmt)) {
                    /* reduce_h_u8_scalar_planned / avx2_pixel (reduce_simd.rs:2463-2538);
                     * reduce_h_scalar's fixed-point loop for I16/U32/I32 (:47-75) */
                    const int16_t *w = k->ci + (size_t)phases[xo] * T;
                    int64_t acc = 0;
                    for (int t = 0; t < T; ++t) {
                        size_t tx = (size_t)clamp_axis(starts[xo] + t, ir.x, ir.w);
                        acc += (int64_t)w[t] * ldi(in, nd->in_fmt, ((size_t)y * in_w + tx) * B + b);
                    }
                    store_fixed(out, nd->in_fmt, oi, acc);
                } else if (nd->in_fmt == ORC_F64) {
                    /* reduce_h_scalar's f64 loop (reduce_simd.rs:76-88): mul_add in tap order */
                    const double *w = k->cf + (size_t)phases[xo] * T;
                    double acc = 0.0;
                    for (int t = 0; t < T; ++t) {
                        size_t tx = (size_t)clamp_axis(starts[xo] + t, ir.x, ir.w);
                        acc = fma(ld(in, ORC_F64, ((size_t)y * in_w + tx) * B + b), w[t], acc);
                    }
                    ((double *)out)[oi] = acc;
                } else {
                    const double *w = k->cf + (size_t)phases[xo] * T;
                    double acc = 0.0;
                    int t = 0;
                    if (B == 1) {
                        /* reduce_h_f32_avx2_pixel (reduce_simd.rs:2663-2693) */
                        for (; t + 4 <= T; t += 4) {
                            double s = 0.0;
                            for (int l = 0; l < 4; ++l) {
                                size_t tx = (size_t)clamp_axis(starts[xo] + t + l, ir.x, ir.w);
                                double prod = ld(in, ORC_F32, (size_t)y * in_w + tx) * w[t + l];
                                s = s + prod;
                            }
                            acc += s;
                        }
                    }
                    /* scalar FMA chain (reduce_simd.rs:78-86) / AVX2 tail */
                    for (; t < T; ++t) {
                        size_t tx = (size_t)clamp_axis(starts[xo] + t, ir.x, ir.w);
                        acc = fma(ld(in, ORC_F32, ((size_t)y * in_w + tx) * B + b), w[t], acc);
                    }
                    ((float *)out)[oi] = (float)acc;
                }
            }
        }
    }
    free(starts);
    free(phases);
}

static void process_reduce_v(const node_t *nd, const void *in, rgn ir, void *out, rgn orr) {
    const rk_t *k = &nd->rk;
    size_t B = nd->bands, row_len = (size_t)orr.w * B, stride = (size_t)ir.w * B;
    int T = k->taps;
    /* reduce_v_f32_avx2 vectorises the row prefix in groups of 4 (reduce_simd.rs:2839-2887) */
    size_t vec_len = row_len / 4 * 4;
    for (int64_t yo = 0; yo < orr.h; ++yo) {
        int64_t start;
        int phase;
        /* x86 AVX2: direct source position per row (reduce_simd.rs:2716-2717,2781,2846) */
        rk_plan(k, rk_sp(k, (double)orr.y + (double)yo), &start, &phase);
        for (size_t x = 0; x < row_len; ++x) {
            size_t oi = (size_t)yo * row_len + x;
            if (fmt_is_int(nd->in_fmt)) {
                const int16_t *w = k->ci + (size_t)phase * T;
                int64_t acc = 0;
                for (int t = 0; t < T; ++t) {
                    size_t ty = (size_t)clamp_axis(start + t, ir.y, ir.h);
                    acc += (int64_t)w[t] * ldi(in, nd->in_fmt, ty * stride + x);
                }
                store_fixed(out, nd->in_fmt, oi, acc);
            } else if (nd->in_fmt == ORC_F64) {
                /* reduce_v_scalar's f64 loop (reduce_simd.rs:127-141) */
                const double *w = k->cf + (size_t)phase * T;
                double acc = 0.0;
                for (int t = 0; t < T; ++t) {
                    size_t ty = (size_t)clamp_axis(start + t, ir.y, ir.h);
                    acc = fma(ld(in, ORC_F64, ty * stride + x), w[t], acc);
                }
                ((double *)out)[oi] = acc;
            } else {
                const double *w = k->cf + (size_t)phase * T;
                double acc = 0.0;
                for (int t = 0; t < T; ++t) {
                    size_t ty = (size_t)clamp_axis(start + t, ir.y, ir.h);
                    double xv = ld(in, ORC_F32, ty * stride + x);
                    if (x < vec_len) {
                        double prod = xv * w[t];
                        acc = acc + prod;
                    } else {
                        acc = fma(xv, w[t], acc);
                    }
                }
                ((float *)out)[oi] = (float)acc;
            }
        }
    }
}

static void store_int(void *out, int fmt, size_t oi, double q) {
    switch (fmt) {
    case ORC_U8: ((uint8_t *)out)[oi] = (uint8_t)q; break;
    case ORC_U16: ((uint16_t *)out)[oi] = (uint16_t)q; break;
    case ORC_I16: ((int16_t *)out)[oi] = (int16_t)q; break;
    case ORC_U32: ((uint32_t *)out)[oi] = (uint32_t)q; break;
    default: ((int32_t *)out)[oi] = (int32_t)q; break;
    }
}

/* F::Sample::from_f64 (viprs-core shared_ops/sample_conv.rs:53-99): integers
 * clamp to the type's range, then round half away from zero; f32 is a plain
 * cast, f64 the value. */
static void store_from_f64(void *out, int fmt, size_t oi, double v) {
    if (fmt == ORC_F32)
        ((float *)out)[oi] = (float)v;
    else if (fmt == ORC_F64)
        ((double *)out)[oi] = v;
    else
        store_int(out, fmt, oi, round(clampd(v, fmt_lo(fmt), fmt_hi(fmt))));
}

/* The three sample conversions, one value at a time, for the known-answer
 * tests: from_fixed_i64, FromF64::from_f64, ShrinkSample::from_f64_clamped. */
void orc_convert_fixed(int64_t v, int fmt, void *out) { store_fixed(out, fmt, 0, v); }
void orc_convert_from_f64(double v, int fmt, void *out) { store_from_f64(out, fmt, 0, v); }
void orc_convert_shrunk(double v, int fmt, void *out) { store_shrunk(out, fmt, 0, v); }

/* ReduceOp passes (reduce.rs:104-166): per output position the direct
 * source position (offset 0), taps_for_f64 (reduce_common.rs:188-191), an
 * f64 mul_add chain in tap order over clamp_axis'd input samples, then
 * from_f64.  Same loop for H (vertical = 0) and V. */
static void process_reduce_op(const node_t *nd, const void *in, rgn ir, void *out, rgn orr, int vertical) {
    const rk_t *k = &nd->rk;
    size_t B = nd->bands;
    int T = k->taps;
    for (int64_t y = 0; y < orr.h; ++y) {
        for (int64_t x = 0; x < orr.w; ++x) {
            int64_t start;
            int phase;
            rk_plan(k, rk_sp(k, vertical ? (double)orr.y + (double)y : (double)orr.x + (double)x), &start, &phase);
            const double *w = k->cf + (size_t)phase * T;
            for (size_t b = 0; b < B; ++b) {
                double acc = 0.0;
                for (int t = 0; t < T; ++t) {
                    size_t idx;
                    if (vertical)
                        idx = ((size_t)clamp_axis(start + t, ir.y, ir.h) * (size_t)ir.w + (size_t)x) * B + b;
                    else
                        idx = ((size_t)y * (size_t)ir.w + (size_t)clamp_axis(start + t, ir.x, ir.w)) * B + b;
                    acc = fma(ld(in, nd->in_fmt, idx), w[t], acc);
                }
                store_from_f64(out, nd->out_fmt, ((size_t)y * (size_t)orr.w + (size_t)x) * B + b, acc);
            }
        }
    }
}

/* u8 box-average rounding of ShrinkH on x86 (shrinkh.rs:323-471,473-731). */
static inline uint8_t shrink_h_u8_round(uint32_t sum, uint32_t f, uint32_t bands) {
    if (f == 2) return (uint8_t)((sum + 1) >> 1);
    if (f == 5 && (bands == 1 || bands == 3 || bands == 4)) return (uint8_t)((sum + 2) / 5);
    uint64_t mult = (1ull << 32) / (256ull * f);
    return (uint8_t)(((uint64_t)(sum + f / 2) * mult) >> 24);
}
/* ShrinkV u8 (shrinkv.rs:206-339). */
static inline uint8_t shrink_v_u8_round(uint32_t sum, uint32_t f) {
    if (f == 2) return (uint8_t)((sum + 1) >> 1);
    uint32_t mult = (uint32_t)((1ull << 32) / (256ull * f));
    return (uint8_t)(((uint64_t)(sum + f / 2) * mult) >> 24);
}

static void store_shrunk(void *out, int fmt, size_t oi, double v) {
    /* ShrinkSample::from_f64_clamped (shrinkh.rs:44-128): unsigned round then
     * clamp; signed clamp(v + 0.5) then truncate; f32 cast; f64 as is. */
    if (fmt == ORC_F32)
        ((float *)out)[oi] = (float)v;
    else if (fmt == ORC_F64)
        ((double *)out)[oi] = v;
    else if (fmt == ORC_I16 || fmt == ORC_I32)
        store_int(out, fmt, oi, trunc(clampd(v + 0.5, fmt_lo(fmt), fmt_hi(fmt))));
    else
        store_int(out, fmt, oi, clampd(round(v), fmt_lo(fmt), fmt_hi(fmt)));
}

static void process_shrink_h(const node_t *nd, const void *in, rgn ir, void *out, rgn orr) {
    size_t B = nd->bands, f = nd->factor, in_w = (size_t)ir.w, out_w = (size_t)orr.w;
    double inv = 1.0 / (double)f;
    for (int64_t y = 0; y < orr.h; ++y) {
        for (size_t x = 0; x < out_w; ++x) {
            for (size_t b = 0; b < B; ++b) {
                size_t oi = ((size_t)y * out_w + x) * B + b;
                size_t base = ((size_t)y * in_w + x * f) * B + b;
                if (nd->in_fmt == ORC_U8) {
                    const uint8_t *s = (const uint8_t *)in;
                    if (f == 1) {
                        ((uint8_t *)out)[oi] = s[base];
                        continue;
                    }
                    uint32_t sum = 0;
                    for (size_t j = 0; j < f; ++j) sum += s[base + j * B];
                    ((uint8_t *)out)[oi] = shrink_h_u8_round(sum, (uint32_t)f, (uint32_t)B);
                } else {
                    double sum = 0.0;
                    for (size_t j = 0; j < f; ++j) sum += ld(in, nd->in_fmt, base + j * B);
                    store_shrunk(out, nd->in_fmt, oi, sum * inv);
                }
            }
        }
    }
}

static void process_shrink_v(const node_t *nd, const void *in, rgn ir, void *out, rgn orr) {
    (void)ir;
    size_t B = nd->bands, f = nd->factor, row = (size_t)orr.w * B;
    double inv = 1.0 / (double)f;
    for (int64_t y = 0; y < orr.h; ++y) {
        for (size_t x = 0; x < row; ++x) {
            size_t oi = (size_t)y * row + x;
            if (nd->in_fmt == ORC_U8) {
                const uint8_t *s = (const uint8_t *)in;
                if (f == 1) {
                    ((uint8_t *)out)[oi] = s[oi];
                    continue;
                }
                uint32_t sum = 0;
                for (size_t k = 0; k < f; ++k) sum += s[((size_t)y * f + k) * row + x];
                ((uint8_t *)out)[oi] = shrink_v_u8_round(sum, (uint32_t)f);
            } else {
                double sum = 0.0;
                for (size_t k = 0; k < f; ++k) sum += ld(in, nd->in_fmt, ((size_t)y * f + k) * row + x);
                store_shrunk(out, nd->in_fmt, oi, sum * inv);
            }
        }
    }
}

static void process_conv(const node_t *nd, const void *in, rgn ir, void *out, rgn orr, int vertical) {
    size_t B = nd->bands, in_w = (size_t)ir.w, out_w = (size_t)orr.w;
    int T = nd->ntaps;
    for (int64_t y = 0; y < orr.h; ++y) {
        for (size_t x = 0; x < out_w; ++x) {
            for (size_t b = 0; b < B; ++b) {
                size_t oi = ((size_t)y * out_w + x) * B + b;
#define IDX(t) (vertical ? (((size_t)y + (size_t)(t)) * in_w + x) * B + b : ((size_t)y * in_w + x + (size_t)(t)) * B + b)
                if (nd->type == N_GAUSS_H || nd->type == N_GAUSS_V) {
                    if (nd->in_fmt == ORC_U8) {
                        /* process_{horizontal,vertical}_u8_scalar + clip_u8_fixed (gauss_blur.rs:405-518) */
                        const uint8_t *s = (const uint8_t *)in;
                        int64_t sum = 0;
                        for (int t = 0; t < T; ++t) sum += (int64_t)nd->gi[t] * (int64_t)s[IDX(t)];
                        int64_t v = (sum + (int64_t)(nd->gscale / 2)) / (int64_t)nd->gscale;
                        ((uint8_t *)out)[oi] = (uint8_t)clampi(v, 0, 255);
                    } else {
                        /* process_{horizontal,vertical}_float (gauss_blur.rs:410-464): no FMA */
                        double acc = 0.0;
                        for (int t = 0; t < T; ++t) {
                            double xv = (double)(float)ld(in, nd->in_fmt, IDX(t));
                            double prod = xv * nd->gf[t];
                            acc = acc + prod;
                        }
                        ((float *)out)[oi] = (float)acc;
                    }
                } else {
                    /* ConvSepH/V::process_region (convsep.rs:87-107,184-205) */
                    double acc = 0.0;
                    for (int t = 0; t < T; ++t) acc = fma(ld(in, nd->in_fmt, IDX(t)), nd->gf[t], acc);
                    ((float *)out)[oi] = (float)(acc / nd->cscale + nd->coffset);
                }
#undef IDX
            }
        }
    }
}

static void process_alpha(const node_t *nd, const void *in, void *out, rgn orr) {
    size_t B = nd->bands, npix = (size_t)orr.w * (size_t)orr.h;
    const uint8_t *s = (const uint8_t *)in;
    uint8_t *d = (uint8_t *)out;
    for (size_t i = 0; i < npix; ++i) {
        const uint8_t *px = s + i * B;
        uint8_t *po = d + i * B;
        double a = (double)px[B - 1];
        for (size_t c = 0; c + 1 < B; ++c) {
            double v;
            if (nd->type == N_PREMULT) {
                /* premultiply.rs:103-115 */
                double na = clampd(a, 0.0, 255.0) / 255.0;
                v = clampd(round((double)px[c] * na), 0.0, 255.0);
            } else {
                /* unpremultiply.rs:96-109 */
                v = a == 0.0 ? 0.0 : clampd(round(((double)px[c] * 255.0) / a), 0.0, 255.0);
            }
            po[c] = (uint8_t)v;
        }
        po[B - 1] = px[B - 1];
    }
}

/* Affine::sample_band (affine/core.rs:442-461) under ExtendMode::Copy and
 * source bounds (0, 0, in_w, in_h): coordinates clamp to the image
 * (resolve_source_coord, affine/interpolation.rs:93-125); outside the tile
 * the fill value 0 is read. */
static double affine_sample(const node_t *nd, const void *in, rgn ir, int64_t ix, int64_t iy, int band) {
    const int64_t rx = clampi(ix, 0, (int64_t)nd->in_w - 1), ry = clampi(iy, 0, (int64_t)nd->in_h - 1);
    const int64_t tx = rx - ir.x, ty = ry - ir.y;
    if (tx < 0 || tx >= ir.w || ty < 0 || ty >= ir.h) return 0.0;
    return ld(in, nd->in_fmt, ((size_t)ty * (size_t)ir.w + (size_t)tx) * nd->bands + (size_t)band);
}

/* sample_band_premultiplied (affine/core.rs:463-482); alpha = band 3. */
static double affine_sample_pm(const node_t *nd, const void *in, rgn ir, int64_t ix, int64_t iy, int band,
                               int pm, double amax) {
    const double v = affine_sample(nd, in, ir, ix, iy, band);
    if (!pm || band == 3) return v;
    return v * (affine_sample(nd, in, ir, ix, iy, 3) / amax);
}

/* bilinear_mix (affine/core.rs:197-201). */
static double bilinear_mix(double p00, double p01, double p10, double p11, double fx, double fy) {
    const double top = fma(p01, fx, p00 * (1.0 - fx));
    const double bottom = fma(p11, fx, p10 * (1.0 - fx));
    return top * (1.0 - fy) + bottom * fy;
}

/* One band at (x, y): interp_nearest / interp_bilinear / interp_separable
 * for the cubic and Lanczos kernels (affine/core.rs:485-514,648-697,728-749), and their premultiplied
 * forms (sample_at_premultiplied, affine/core.rs:751-935,937-986) when pm. */
/* interpolate_vsqbs_2d (viprs-core kernel.rs:863-923) over n[row][col]. */
static double vsqbs_eval(double xph, double yph, double n[4][4]) {
    int ax, sx, ay, sy;
    double x0, y0;
    if (xph < 0.5) { ax = 1; sx = 1; x0 = xph; } else { ax = 2; sx = -1; x0 = xph - 1.0; }
    if (yph < 0.5) { ay = 1; sy = 1; y0 = yph; } else { ay = 2; sy = -1; y0 = yph - 1.0; }
    const double x = (double)(2 * sx) * x0 - 0.5, y = (double)(2 * sy) * y0 - 0.5;
    const double cent = 0.75 - x * x, mid = 0.75 - y * y;
    const double left = fma(-0.5, x + cent, 0.5), top = fma(-0.5, y + mid, 0.5);
    const double lc = left + cent, tm = top + mid, cr = 1.0 - left, mb = 1.0 - top;
    const double rite = 1.0 - lc, bot = 1.0 - tm;
    const double k_t_m = lc * top, k_m_l = left * tm, k_m_m = lc + tm, k_m_r = cr * tm + rite;
    const double k_b_m = mb * lc + bot, k_b_r = mb * rite + cr * bot;
    const double k_t_r = top - k_t_m, k_b_l = left - k_m_l;
    const int T = ay - sy, M = ay, Bo = ay + sy, L = ax - sx, C = ax, R = ax + sx;
    return (((k_t_m * n[T][C] + k_m_l * n[M][L]) + (k_m_m * n[M][C] + k_m_r * n[M][R])) +
            ((k_b_m * n[Bo][C] + k_b_r * n[Bo][R]) + (k_t_r * n[T][R] + k_b_l * n[Bo][L]))) *
           0.25;
}

static inline double lmin(double a, double b) { return a <= b ? a : b; }
static inline double lmax(double a, double b) { return a >= b ? a : b; }
static inline double lsgn(double a) { return a >= 0.0 ? 1.0 : -1.0; }

/* One corner's limited slopes and cross derivative of lbbicubic
 * (affine/interpolation.rs:170-378): z its value, lo / hi its local bounds,
 * gx / gy the raw central differences, xy the raw cross difference. */
static void lbb_corner(double z, double lo, double hi, double gx, double gy, double xy, double *dx, double *dy,
                       double *q, double *u, double *v) {
    *u = z - lo;
    *v = hi - z;
    const double lim = 6.0 * lmin(*u, *v);
    *dx = lsgn(gx) * gx <= lim ? gx : lsgn(gx) * lim;
    *dy = lsgn(gy) * gy <= lim ? gy : lsgn(gy) * lim;
    const double as = fabs(6.0 * (*dx + *dy)), ad = fabs(6.0 * (*dx - *dy));
    const double u36 = 36.0 * *u, v36 = 36.0 * *v;
    double t = lmax(xy, as - u36);
    t = lmin(t, v36 - as);
    t = lmax(t, ad - v36);
    *q = lmin(t, u36 - ad);
}

/* lbbicubic (affine/interpolation.rs:170-441) over n[row][col], [1][1] at
 * (floor x, floor y), phases (rx, ry). */
static double lbb_eval(double n[4][4], double rx, double ry) {
    /* ordered central pairs of the two middle rows, and the arm pairs */
    const double a1 = n[1][1] <= n[1][2] ? n[1][1] : n[1][2], b1 = n[1][1] <= n[1][2] ? n[1][2] : n[1][1];
    const double a2 = n[2][1] <= n[2][2] ? n[2][1] : n[2][2], b2 = n[2][1] <= n[2][2] ? n[2][2] : n[2][1];
    const double p3 = n[0][1] <= n[1][0] ? n[0][1] : n[1][0], P3 = n[0][1] <= n[1][0] ? n[1][0] : n[0][1];
    const double p4 = n[0][2] <= n[1][3] ? n[0][2] : n[1][3], P4 = n[0][2] <= n[1][3] ? n[1][3] : n[0][2];
    const double p5 = n[2][0] <= n[3][1] ? n[2][0] : n[3][1], P5 = n[2][0] <= n[3][1] ? n[3][1] : n[2][0];
    const double p6 = n[2][3] <= n[3][2] ? n[2][3] : n[3][2], P6 = n[2][3] <= n[3][2] ? n[3][2] : n[2][3];
    const double lo00 = lmin(lmin(a1, n[2][1]), p3), hi00 = lmax(lmax(b1, n[2][1]), P3);
    const double lo10 = lmin(lmin(a1, n[2][2]), p4), hi10 = lmax(lmax(b1, n[2][2]), P4);
    const double lo01 = lmin(lmin(a2, n[1][1]), p5), hi01 = lmax(lmax(b2, n[1][1]), P5);
    const double lo11 = lmin(lmin(a2, n[1][2]), p6), hi11 = lmax(lmax(b2, n[1][2]), P6);
    const double gx00 = n[1][2] - n[1][0], gx10 = n[1][3] - n[1][1];
    const double gx01 = n[2][2] - n[2][0], gx11 = n[2][3] - n[2][1];
    const double gy00 = n[2][1] - n[0][1], gy10 = n[2][2] - n[0][2];
    const double gy01 = n[3][1] - n[1][1], gy11 = n[3][2] - n[1][2];
    const double xy00 = n[0][0] - n[0][2] + gx01, xy10 = n[0][1] - n[0][3] + gx11;
    const double xy01 = n[3][2] - n[3][0] - gx00, xy11 = n[3][3] - n[3][1] - gx10;
    double dx[4], dy[4], q[4], u, v;
    lbb_corner(n[1][1], lo00, hi00, gx00, gy00, xy00, &dx[0], &dy[0], &q[0], &u, &v);
    lbb_corner(n[1][2], lo10, hi10, gx10, gy10, xy10, &dx[1], &dy[1], &q[1], &u, &v);
    lbb_corner(n[2][1], lo01, hi01, gx01, gy01, xy01, &dx[2], &dy[2], &q[2], &u, &v);
    lbb_corner(n[2][2], lo11, hi11, gx11, gy11, xy11, &dx[3], &dy[3], &q[3], &u, &v);
    /* Hermite basis */
    const double xa = rx, xb = rx - 1.0, ya = ry, yb = ry - 1.0;
    const double ta = (0.5 + xa) + (0.5 + xa), tb = (1.5 - xa) + (1.5 - xa);
    const double sa = (0.5 + ya) + (0.5 + ya), sb = (1.5 - ya) + (1.5 - ya);
    const double xa2 = xa * xa, xb2 = xb * xb, ya2 = ya * ya, yb2 = yb * yb;
    const double e00 = xb2 * yb2, e10 = xa2 * yb2, e01 = xb2 * ya2, e11 = xa2 * ya2;
    const double v1 = (ta * sa) * e00 * n[1][1] + (tb * sa) * e10 * n[1][2] + (ta * sb) * e01 * n[2][1] +
                      (tb * sb) * e11 * n[2][2];
    const double v2 = (xa * sa) * e00 * dx[0] + (xb * sa) * e10 * dx[1] + (xa * sb) * e01 * dx[2] +
                      (xb * sb) * e11 * dx[3] + (ta * ya) * e00 * dy[0] + (tb * ya) * e10 * dy[1] +
                      (ta * yb) * e01 * dy[2] + (tb * yb) * e11 * dy[3];
    const double v3 = (xa * ya) * e00 * q[0] + (xb * ya) * e10 * q[1] + (xa * yb) * e01 * q[2] +
                      (xb * yb) * e11 * q[3];
    return fma(0.25, v3, fma(0.5, v2, v1));
}

/* The two 2-D stencil interpolators alone, for the known-answer tests:
 * n16 is the 4 x 4 neighbourhood row by row. */
double orc_interp_2d(int kernel, const double *n16, double xph, double yph) {
    double n[4][4];
    for (int i = 0; i < 16; ++i) n[i / 4][i % 4] = n16[i];
    return kernel == ORC_VSQBS ? vsqbs_eval(xph, yph, n) : lbb_eval(n, xph, yph);
}

static double affine_band(const node_t *nd, const void *in, rgn ir, double x, double y, int band, int pm,
                          double amax) {
    if (nd->akern == ORC_NEAREST) return affine_sample_pm(nd, in, ir, (int64_t)x, (int64_t)y, band, pm, amax);
    if (nd->akern == ORC_BILINEAR) {
        const int64_t x0 = (int64_t)x, y0 = (int64_t)y; /* truncation, as `as i64` */
        const double fx = x - (double)x0, fy = y - (double)y0;
        return bilinear_mix(affine_sample_pm(nd, in, ir, x0, y0, band, pm, amax),
                            affine_sample_pm(nd, in, ir, x0 + 1, y0, band, pm, amax),
                            affine_sample_pm(nd, in, ir, x0, y0 + 1, band, pm, amax),
                            affine_sample_pm(nd, in, ir, x0 + 1, y0 + 1, band, pm, amax), fx, fy);
    }
    if (nd->akern == ORC_VSQBS || nd->akern == ORC_LBB) {
        /* interp_vsqbs / interp_lbb (affine/core.rs:588-610,699-725) */
        const double fx = floor(x), fy = floor(y);
        const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
        double n[4][4];
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) n[r][c] = affine_sample_pm(nd, in, ir, ix + c - 1, iy + r - 1, band, pm, amax);
        return nd->akern == ORC_VSQBS ? vsqbs_eval(x - fx, y - fy, n) : lbb_eval(n, x - fx, y - fy);
    }
    /* interp_separable<SUPPORT>: SUPPORT 3 for Lanczos3, else 2; taps
     * floor - SUPPORT + 1 .. floor + SUPPORT */
    const int64_t S = nd->akern == ORC_LANCZOS3 ? 3 : 2;
    const int64_t cx = (int64_t)floor(x), cy = (int64_t)floor(y);
    double acc = 0.0, wsum = 0.0;
    for (int64_t ky = 1 - S; ky <= S; ++ky) {
        const int64_t iy = cy + ky;
        const double wy = orc_kernel_weight(nd->akern, fabs((double)iy - y));
        if (wy == 0.0) continue;
        for (int64_t kx = 1 - S; kx <= S; ++kx) {
            const int64_t ix = cx + kx;
            const double wx = orc_kernel_weight(nd->akern, fabs((double)ix - x));
            if (wx == 0.0) continue;
            const double w = wx * wy;
            wsum += w;
            acc = fma(affine_sample_pm(nd, in, ir, ix, iy, band, pm, amax), w, acc);
        }
    }
    return wsum > 0.0 ? acc / wsum : 0.0;
}

/* Affine::process_region, the generic path (affine/core.rs:1593-1632; the
 * fast paths need a constant-fill extend): the source position starts from
 * the region origin and is stepped by one matrix column per pixel and one
 * per row, so it depends on the tile origin.  Four-band images go through
 * sample_pixel_at_with_alpha (affine/core.rs:988-1020): the Affine's own
 * premultiplied flag is never set by the builder. */
static void process_affine(const node_t *nd, const void *in, rgn ir, void *out, rgn orr) {
    const size_t B = nd->bands;
    const int fmt = nd->out_fmt;
    const double amax = fmt_is_int(fmt) ? fmt_hi(fmt) : 1.0; /* Affine::alpha_max (affine/core.rs:101-109) */
    double row_x = fma(nd->am[1], (double)orr.y, nd->am[0] * (double)orr.x) + nd->atx;
    double row_y = fma(nd->am[3], (double)orr.y, nd->am[2] * (double)orr.x) + nd->aty;
    for (int64_t yy = 0; yy < orr.h; ++yy) {
        double x = row_x, y = row_y;
        for (int64_t xx = 0; xx < orr.w; ++xx) {
            const size_t o = ((size_t)yy * (size_t)orr.w + (size_t)xx) * B;
            if (B == 4) {
                const double alpha = clampd(affine_band(nd, in, ir, x, y, 3, 1, amax), 0.0, amax);
                const double na = alpha / amax;
                for (int b = 0; b < 3; ++b) {
                    const double pmv = affine_band(nd, in, ir, x, y, b, 1, amax);
                    store_from_f64(out, fmt, o + (size_t)b, na > 0.0 ? pmv / na : 0.0);
                }
                store_from_f64(out, fmt, o + 3, alpha);
            } else {
                for (size_t b = 0; b < B; ++b)
                    store_from_f64(out, fmt, o + b, affine_band(nd, in, ir, x, y, (int)b, 0, amax));
            }
            x += nd->am[0];
            y += nd->am[2];
        }
        row_x += nd->am[1];
        row_y += nd->am[3];
    }
}

/* Zoom::process_region (zoom.rs:364-400): every output pixel repeats the
 * input pixel (x / xfac, y / yfac). */
static void process_zoom(const node_t *nd, const void *in, rgn ir, void *out, rgn orr) {
    const size_t ps = (size_t)nd->bands * fmt_size(nd->in_fmt);
    for (int64_t y = 0; y < orr.h; ++y)
        for (int64_t x = 0; x < orr.w; ++x) {
            const int64_t sx = (orr.x + x) / nd->zx - ir.x, sy = (orr.y + y) / nd->zy - ir.y;
            memcpy((uint8_t *)out + ((size_t)y * (size_t)orr.w + (size_t)x) * ps,
                   (const uint8_t *)in + ((size_t)sy * (size_t)ir.w + (size_t)sx) * ps, ps);
        }
}

static void process_node(const node_t *nd, const void *in, rgn ir, void *out, rgn orr) {
    switch (nd->type) {
    case N_REDUCE_H: process_reduce_h(nd, in, ir, out, orr); break;
    case N_REDUCE_V: process_reduce_v(nd, in, ir, out, orr); break;
    case N_REDUCE_OP_H: process_reduce_op(nd, in, ir, out, orr, 0); break;
    case N_REDUCE_OP_V: process_reduce_op(nd, in, ir, out, orr, 1); break;
    case N_SHRINK_H: process_shrink_h(nd, in, ir, out, orr); break;
    case N_SHRINK_V: process_shrink_v(nd, in, ir, out, orr); break;
    case N_GAUSS_H:
    case N_CONVSEP_H: process_conv(nd, in, ir, out, orr, 0); break;
    case N_GAUSS_V:
    case N_CONVSEP_V: process_conv(nd, in, ir, out, orr, 1); break;
    case N_PREMULT:
    case N_UNPREMULT: process_alpha(nd, in, out, orr); break;
    case N_AFFINE: process_affine(nd, in, ir, out, orr); break;
    case N_ZOOM: process_zoom(nd, in, ir, out, orr); break;
    default: memcpy(out, in, (size_t)orr.w * (size_t)orr.h * nd->bands * fmt_size(nd->in_fmt)); break;
    }
}

/* MemorySource::read_region with clamp-to-edge (sources/memory.rs:149-235). */
static void read_source(const orc_pipeline *p, const void *src, rgn r, void *dst) {
    size_t ps = (size_t)p->bands * fmt_size(p->src_fmt);
    const uint8_t *s = (const uint8_t *)src;
    uint8_t *d = (uint8_t *)dst;
    for (int64_t row = 0; row < r.h; ++row) {
        int64_t sy = clampi(r.y + row, 0, (int64_t)p->src_h - 1);
        for (int64_t col = 0; col < r.w; ++col) {
            int64_t sx = clampi(r.x + col, 0, (int64_t)p->src_w - 1);
            memcpy(d + ((size_t)row * (size_t)r.w + (size_t)col) * ps,
                   s + ((size_t)sy * p->src_w + (size_t)sx) * ps, ps);
        }
    }
}

/* execute_tile (execute.rs:174-852) for a linear chain: backward region
 * propagation, clamp-to-edge source read, forward op calls. */
static void run_tile(const orc_pipeline *p, const void *src, uint8_t *dst, rgn tile) {
    int n = p->n;
    rgn outr[32], inr[32];
    outr[n - 1] = tile;
    for (int i = n - 1; i >= 0; --i) {
        inr[i] = required_input(&p->nodes[i], outr[i]);
        if (i > 0) outr[i - 1] = inr[i];
    }
    size_t sz = (size_t)inr[0].w * (size_t)inr[0].h * p->bands * fmt_size(p->src_fmt);
    void *cur = malloc(sz ? sz : 1);
    read_source(p, src, inr[0], cur);
    for (int i = 0; i < n; ++i) {
        const node_t *nd = &p->nodes[i];
        size_t osz = (size_t)outr[i].w * (size_t)outr[i].h * nd->bands * fmt_size(nd->out_fmt);
        void *o = malloc(osz ? osz : 1);
        process_node(nd, cur, inr[i], o, outr[i]);
        free(cur);
        cur = o;
    }
    /* sink write: tile rows into the output image */
    size_t ps = (size_t)p->bands * fmt_size(p->cur_fmt);
    size_t row = (size_t)tile.w * ps;
    for (int64_t r = 0; r < tile.h; ++r)
        memcpy(dst + ((size_t)(tile.y + r) * p->cur_w + (size_t)tile.x) * ps, (uint8_t *)cur + (size_t)r * row, row);
    free(cur);
}

typedef struct {
    const orc_pipeline *p;
    const void *src;
    uint8_t *dst;
    uint32_t tw, th, cols;
    uint32_t ntiles;
    int tid, nthreads;
} job_t;

static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    for (uint32_t t = (uint32_t)j->tid; t < j->ntiles; t += (uint32_t)j->nthreads) {
        rgn tile;
        tile.x = (int64_t)(t % j->cols) * j->tw;
        tile.y = (int64_t)(t / j->cols) * j->th;
        tile.w = j->tw;
        tile.h = j->th;
        if (tile.x + tile.w > (int64_t)j->p->cur_w) tile.w = (int64_t)j->p->cur_w - tile.x;
        if (tile.y + tile.h > (int64_t)j->p->cur_h) tile.h = (int64_t)j->p->cur_h - tile.y;
        run_tile(j->p, j->src, j->dst, tile);
    }
    return NULL;
}

/* RayonScheduler::run (runtime.rs:1019-1235) over the tiles of
 * generate_tiles_with_l2_budget (planning.rs:513-601; op/demand.rs:41-85).
 * The pipeline's demand is the ThinStrip override or the largest node demand
 * (arena.rs:411-420): every node here is ThinStrip except an Affine after a
 * small thumbnail, which is SmallTile.  ThinStrip: full-width strips of 16
 * rows (1 row beyond 10000 px); SmallTile: 128 x 128 tiles.
 * strip_height != 0 instead runs full-width strips of that height. */
int orc_pipeline_run(const orc_pipeline *p, const void *src, void *dst, int threads,
                     uint32_t strip_height) {
    if (p->cur_w == 0 || p->cur_h == 0) return ORC_OK;
    if (p->box_f) {
        /* The node chain runs over the box-shrunk source. */
        void *shrunk = malloc((size_t)p->src_w * p->src_h * p->bands * fmt_size(p->src_fmt));
        if (!shrunk) return ORC_EINVAL;
        orc_box_shrink(src, p->orig_w, p->orig_h, p->bands, p->src_fmt, p->box_f, p->box_ceil, shrunk, NULL, NULL);
        orc_pipeline bare = *p;
        bare.box_f = 0;
        const int rc = orc_pipeline_run(&bare, shrunk, dst, threads, strip_height);
        free(shrunk);
        return rc;
    }
    if (p->n == 0) {
        memcpy(dst, src, (size_t)p->src_w * p->src_h * p->bands * fmt_size(p->src_fmt));
        return ORC_OK;
    }
    int small = 0;
    for (int i = 0; i < p->n; ++i)
        if (p->nodes[i].type == N_AFFINE && p->nodes[i].asmall) small = 1;
    if (p->thin_override) small = 0;
    uint32_t tile_w = p->cur_w, tile_h;
    if (strip_height) {
        tile_h = strip_height;
    } else if (small) {
        tile_w = 128;
        tile_h = 128;
    } else {
        tile_h = p->cur_w > 10000 ? 1 : 16;
    }
    if (tile_h > p->cur_h) tile_h = p->cur_h;
    uint32_t cols = (p->cur_w + tile_w - 1) / tile_w;
    uint32_t ntiles = cols * ((p->cur_h + tile_h - 1) / tile_h);
    if (threads < 1) threads = 1;
    if ((uint32_t)threads > ntiles) threads = (int)ntiles;
    pthread_t th[256];
    job_t jobs[256];
    if (threads > 256) threads = 256;
    for (int i = 0; i < threads; ++i) {
        jobs[i] = (job_t){p, src, (uint8_t *)dst, tile_w, tile_h, cols, ntiles, i, threads};
        if (threads > 1)
            pthread_create(&th[i], NULL, worker, &jobs[i]);
        else
            worker(&jobs[i]);
    }
    if (threads > 1)
        for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    return ORC_OK;
}

/* ---------------------------------------------------------- FP64 reference */

typedef struct {
    const float *src;
    double *dst;
    uint32_t w, h, bands;
    rk_t *kh, *kv;
    uint32_t ow, oh;
    double *tmp;
    int tid, nthreads;
    int pass;
} f64job_t;

static void *f64_reduce_worker(void *arg) {
    f64job_t *j = (f64job_t *)arg;
    size_t B = j->bands;
    int T;
    if (j->pass == 0) {
        T = j->kh->taps;
        for (uint32_t y = (uint32_t)j->tid; y < j->h; y += (uint32_t)j->nthreads)
            for (uint32_t x = 0; x < j->ow; ++x) {
                int64_t st;
                int ph;
                rk_plan(j->kh, rk_sp(j->kh, (double)x), &st, &ph);
                const double *wt = j->kh->cf + (size_t)ph * T;
                for (size_t b = 0; b < B; ++b) {
                    double acc = 0.0;
                    for (int t = 0; t < T; ++t) {
                        int64_t sx = clampi(st + t, 0, (int64_t)j->w - 1);
                        acc += wt[t] * (double)j->src[((size_t)y * j->w + (size_t)sx) * B + b];
                    }
                    j->tmp[((size_t)y * j->ow + x) * B + b] = acc;
                }
            }
    } else {
        T = j->kv->taps;
        for (uint32_t y = (uint32_t)j->tid; y < j->oh; y += (uint32_t)j->nthreads) {
            int64_t st;
            int ph;
            rk_plan(j->kv, rk_sp(j->kv, (double)y), &st, &ph);
            const double *wt = j->kv->cf + (size_t)ph * T;
            for (uint32_t x = 0; x < j->ow; ++x)
                for (size_t b = 0; b < B; ++b) {
                    double acc = 0.0;
                    for (int t = 0; t < T; ++t) {
                        int64_t sy = clampi(st + t, 0, (int64_t)j->h - 1);
                        acc += wt[t] * j->tmp[((size_t)sy * j->ow + x) * B + b];
                    }
                    j->dst[((size_t)y * j->ow + x) * B + b] = acc;
                }
        }
    }
    return NULL;
}

/* reduce_h then reduce_v on an f32 image, carried in double: the same
 * 65-phase f64 weight tables, direct per-axis plans and clamp-to-edge borders as
 * the f32 path (reduce_simd.rs:47-144), but with no f32 rounding between the
 * passes or at the output.  This is the FP64 oracle the f32 / bf16 error study
 * is measured against. */
int orc_f64_reduce_hv(const float *src, uint32_t w, uint32_t h, uint32_t bands, double factor,
                      int kernel, double *dst, uint32_t *out_w, uint32_t *out_h, int threads) {
    if (bands == 0 || bands > 4) return ORC_EINVAL;
    rk_t kh, kv;
    int rc;
    if ((rc = rk_init(&kh, kernel, factor, w)) != ORC_OK) return rc;
    if ((rc = rk_init(&kv, kernel, factor, h)) != ORC_OK) return rc;
    uint32_t ow = orc_reduce_output_len(factor, w), oh = orc_reduce_output_len(factor, h);
    *out_w = ow;
    *out_h = oh;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)ow * h * bands);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    for (int pass = 0; pass < 2; ++pass) {
        pthread_t th[256];
        f64job_t jobs[256];
        for (int i = 0; i < threads; ++i) {
            jobs[i] = (f64job_t){src, dst, w, h, bands, &kh, &kv, ow, oh, tmp, i, threads, pass};
            pthread_create(&th[i], NULL, f64_reduce_worker, &jobs[i]);
        }
        for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    }
    free(tmp);
    free(kh.cf);
    free(kh.ci);
    free(kv.cf);
    free(kv.ci);
    return ORC_OK;
}

typedef struct {
    const float *src;
    double *dst, *tmp;
    uint32_t w, h, bands;
    const double *k;
    int taps;
    int tid, nthreads, pass;
} gjob_t;

static void *f64_gauss_worker(void *arg) {
    gjob_t *j = (gjob_t *)arg;
    int r = (j->taps - 1) / 2;
    size_t B = j->bands;
    for (uint32_t y = (uint32_t)j->tid; y < j->h; y += (uint32_t)j->nthreads)
        for (uint32_t x = 0; x < j->w; ++x)
            for (size_t b = 0; b < B; ++b) {
                double acc = 0.0;
                for (int t = 0; t < j->taps; ++t) {
                    if (j->pass == 0) {
                        int64_t sx = clampi((int64_t)x + t - r, 0, (int64_t)j->w - 1);
                        acc += j->k[t] * (double)j->src[((size_t)y * j->w + (size_t)sx) * B + b];
                    } else {
                        int64_t sy = clampi((int64_t)y + t - r, 0, (int64_t)j->h - 1);
                        acc += j->k[t] * j->tmp[((size_t)sy * j->w + x) * B + b];
                    }
                }
                (j->pass == 0 ? j->tmp : j->dst)[((size_t)y * j->w + x) * B + b] = acc;
            }
    return NULL;
}

/* Separable Gaussian with the same normalised integer mask, in double, with
 * clamp-to-edge borders and no intermediate f32 rounding. */
int orc_f64_gauss(const float *src, uint32_t w, uint32_t h, uint32_t bands, float sigma,
                  double *dst, int threads) {
    int16_t c[512];
    int32_t scale;
    int taps = orc_gauss_int_kernel(sigma, c, 512, &scale);
    if (taps < 0) return ORC_EINVAL;
    double k[512];
    for (int i = 0; i < taps; ++i) k[i] = (double)c[i] / (double)scale;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)w * h * bands);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    for (int pass = 0; pass < 2; ++pass) {
        pthread_t th[256];
        gjob_t jobs[256];
        for (int i = 0; i < threads; ++i) {
            jobs[i] = (gjob_t){src, dst, tmp, w, h, bands, k, taps, i, threads, pass};
            pthread_create(&th[i], NULL, f64_gauss_worker, &jobs[i]);
        }
        for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    }
    free(tmp);
    return ORC_OK;
}

/* ------------------------------------------------------- synthetic inputs */

/* Seeded synthetic bytes (test input generator, no reference counterpart):
 * 8-byte word i (little-endian) is splitmix64(seed + (i + 1) *
 * 0x9E3779B97F4A7C15) (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014); the same stream the library's device fill
 * writes, so large device-made inputs can be regenerated here for checking. */
static uint64_t splitmix_word(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct {
    uint8_t *dst;
    size_t n, w0, w1;
    uint64_t seed;
} synth_job;

static void *synth_worker(void *arg) {
    const synth_job *j = (const synth_job *)arg;
    for (size_t w = j->w0; w < j->w1; ++w) {
        const uint64_t z = splitmix_word(j->seed, w);
        const size_t at = w * 8;
        memcpy(j->dst + at, &z, j->n - at < 8 ? j->n - at : 8);
    }
    return NULL;
}

void orc_synthetic_bytes(uint8_t *dst, size_t n, uint64_t seed) {
    enum { MAXT = 16 };
    const size_t words = (n + 7) / 8;
    int t = words < (1u << 20) ? 1 : MAXT;
    pthread_t th[MAXT];
    synth_job jobs[MAXT];
    for (int k = 0; k < t; ++k) {
        jobs[k] = (synth_job){dst, n, words * k / t, words * (k + 1) / t, seed};
        if (t == 1) synth_worker(&jobs[k]);
        else pthread_create(&th[k], NULL, synth_worker, &jobs[k]);
    }
    if (t > 1)
        for (int k = 0; k < t; ++k) pthread_join(th[k], NULL);
}
