/*
 * CPU oracle for the resample + separable-convolution path of viprs.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar C restatement of the reference
 * CPU algorithm (x86-64 + AVX2 dispatch semantics), used by tests/ as the
 * checker, by __graft_entry__.smoke() as the checker, and by bench.py as the
 * timed CPU baseline.  The product library (viprs_amd/) never links, loads or
 * calls anything here.
 *
 * Every function cites the reference file:line range it restates (paths are
 * relative to the viprs repository root).
 *
 * Parity pins (see oracle.c header for the full list): reduce tables and the
 * u8 reduce/shrink kernels are pinned by the reference's committed libvips
 * fixtures and in-crate known answers; the chain executor is pinned by the
 * same fixtures plus tile-height invariance properties.  Functions without
 * such a pin are marked "parity unpinned" in oracle.c.
 */
#ifndef VIPRS_ORACLE_H
#define VIPRS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Band formats (crates/viprs-core/src/format/band.rs:19-35), numbered as the
 * library numbers them. */
enum { ORC_U8 = 0, ORC_U16 = 1, ORC_F32 = 2, ORC_I16 = 3, ORC_U32 = 4, ORC_I32 = 5, ORC_F64 = 6 };

/* InterpolationKernel (crates/viprs-core/src/kernel.rs:22-49). */
enum {
    ORC_NEAREST = 0,
    ORC_BILINEAR = 1,
    ORC_BICUBIC = 2,
    ORC_LANCZOS2 = 3,
    ORC_LANCZOS3 = 4,
    ORC_CATMULLROM = 5,
    ORC_LBB = 6,
    ORC_VSQBS = 7,
    ORC_NOHALO = 8
};

/* Thumbnail targets (crates/viprs-ops-spatial/src/resample/thumbnail.rs:125-145). */
enum { ORC_FITBOX = 0, ORC_FORCE_EXACT = 1, ORC_WIDTH = 2, ORC_HEIGHT = 3 };

/* Status codes. */
enum {
    ORC_OK = 0,
    ORC_EINVAL = -1,      /* BuildError::InvalidReduceParameters & co */
    ORC_EKERNEL = -2,     /* BuildError::InvalidKernel */
    ORC_EFORMAT = -3,     /* unsupported format for this op */
    ORC_EUNSUPPORTED = -4 /* plan needs a node outside the restated path (Affine/crop) */
};

double orc_kernel_weight(int kernel, double x);
int orc_reduce_taps(int kernel, double factor);
int orc_reduce_tables(int kernel, double factor, double *f64_out, int16_t *i16_out);
double orc_reduce_offset(int kernel, double factor, uint32_t input_len);
uint32_t orc_reduce_output_len(double factor, uint32_t input_len);
void orc_reduce_plan(int kernel, double factor, uint32_t input_len, double origin,
                     int accumulate, uint32_t n, int64_t *starts, uint8_t *phases);
int orc_gauss_int_kernel(float sigma, int16_t *coeffs, int max_taps, int32_t *scale);

/* Pipeline (PipelineBuilder restatement + tile scheduler). */
typedef struct orc_pipeline orc_pipeline;
orc_pipeline *orc_pipeline_new(uint32_t width, uint32_t height, uint32_t bands, int format);
void orc_pipeline_free(orc_pipeline *p);
int orc_pipeline_reduce_h(orc_pipeline *p, double factor, int kernel);
int orc_pipeline_reduce_v(orc_pipeline *p, double factor, int kernel);
int orc_pipeline_reduce(orc_pipeline *p, double hfactor, double vfactor, int kernel);
int orc_pipeline_shrink_h(orc_pipeline *p, uint32_t factor, int ceil_mode);
int orc_pipeline_shrink_v(orc_pipeline *p, uint32_t factor, int ceil_mode);
int orc_pipeline_shrink(orc_pipeline *p, uint32_t hshrink, uint32_t vshrink);
int orc_pipeline_gauss_blur(orc_pipeline *p, float sigma);
int orc_pipeline_convsep(orc_pipeline *p, const double *coeffs, uint32_t n, double scale,
                         double offset);
int orc_pipeline_premultiply(orc_pipeline *p);
int orc_pipeline_unpremultiply(orc_pipeline *p);
int orc_pipeline_thumbnail(orc_pipeline *p, int target, uint32_t tw, uint32_t th, int kernel);
int orc_pipeline_thumbnail_crop(orc_pipeline *p, int target, uint32_t tw, uint32_t th, int kernel, int crop);
/* PNG box pre-shrinks: software_box_shrink_generic (ceil_mode 0) and
 * decode_png_with_box_shrink_u8 (ceil_mode 1); dst may be NULL to ask for the
 * output size only. */
int orc_box_shrink(const void *src, uint32_t w, uint32_t h, uint32_t bands, int fmt, uint32_t factor,
                   int ceil_mode, void *dst, uint32_t *out_w, uint32_t *out_h);
int orc_pipeline_box_shrink(orc_pipeline *p, uint32_t factor, int ceil_mode);
int orc_pipeline_thumbnail_shrink_on_load(orc_pipeline *p, int target, uint32_t tw, uint32_t th, int kernel,
                                          int crop);
/* The thumbnail's Affine node: matrix [a, 0, 0, d], translation (tx, ty),
 * ExtendMode::Copy, kernel Nearest / Bilinear / Bicubic; small_tile gives it
 * the SmallTile demand (128 x 128 scheduler tiles) instead of ThinStrip. */
int orc_pipeline_affine(orc_pipeline *p, double a, double d, double tx, double ty, uint32_t width,
                        uint32_t height, int kernel, int small_tile);
int orc_pipeline_extract_area(orc_pipeline *p, uint32_t x, uint32_t y, uint32_t w, uint32_t h);
int orc_pipeline_resize(orc_pipeline *p, double hscale, double vscale, int kernel);
int orc_pipeline_output(const orc_pipeline *p, uint32_t *w, uint32_t *h, uint32_t *bands,
                        int *format);
int orc_pipeline_node_count(const orc_pipeline *p);
int orc_pipeline_node_info(const orc_pipeline *p, int i, int *type, double *param);
int orc_pipeline_run(const orc_pipeline *p, const void *src, void *dst, int threads,
                     uint32_t strip_height);

/* Thumbnail plan only (no execution): node types and parameters. */
int orc_thumbnail_plan(uint32_t in_w, uint32_t in_h, uint32_t bands, int target, uint32_t tw,
                       uint32_t th, int kernel, int *types, double *params, int max_nodes,
                       uint32_t *out_w, uint32_t *out_h);
/* The same with the crop option; extract[4] gets the ExtractArea rectangle,
 * affine[7] the Affine node's (a, d, tx, ty, width, height, kernel). */
/* The shrink-on-load hint of the thumbnail plan (thumbnail.rs:326-330,514-524). */
uint32_t orc_shrink_on_load_factor(double common_shrink);
uint32_t orc_thumbnail_shrink_on_load(uint32_t in_w, uint32_t in_h, int target, uint32_t tw, uint32_t th, int crop);
int orc_thumbnail_plan_crop(uint32_t in_w, uint32_t in_h, uint32_t bands, int target, uint32_t tw,
                            uint32_t th, int kernel, int crop, int *types, double *params, int max_nodes,
                            uint32_t *out_w, uint32_t *out_h, uint32_t *extract, double *affine);

/* FP64 reference for floating-point error studies: the same separable
 * Lanczos/Gaussian math as the f32 path, carried entirely in double with no
 * f32 rounding between passes or at the output. */
int orc_f64_reduce_hv(const float *src, uint32_t w, uint32_t h, uint32_t bands, double factor,
                      int kernel, double *dst, uint32_t *out_w, uint32_t *out_h, int threads);
int orc_f64_gauss(const float *src, uint32_t w, uint32_t h, uint32_t bands, float sigma,
                  double *dst, int threads);

/* Seeded synthetic bytes: splitmix64 words, the device fill's stream. */
void orc_synthetic_bytes(uint8_t *dst, size_t n, uint64_t seed);
/* Sample conversions of one value into *out (a sample of format fmt):
 * ReduceSample::from_fixed_i64, FromF64::from_f64 and
 * ShrinkSample::from_f64_clamped. */
void orc_convert_fixed(int64_t v, int fmt, void *out);
void orc_convert_from_f64(double v, int fmt, void *out);
void orc_convert_shrunk(double v, int fmt, void *out);
/* interpolate_vsqbs_2d (kernel == ORC_VSQBS) or lbbicubic (ORC_LBB) over a
 * 4 x 4 neighbourhood given row by row, [1][1] at the floor of the position,
 * with phases (xph, yph). */
double orc_interp_2d(int kernel, const double *n16, double xph, double yph);

/* Baseline JFIF JPEG of host u8 pixels (jpeg.c): the bytes libjpeg-turbo
 * writes through tj3Compress8 with only quality and subsampling set
 * (jpeg/encode.rs:169-258).  subsampling: 0 auto (4:2:0 below quality 90),
 * 1 off, 2 4:2:0.  *len = the stream's size; out may be NULL. */
int orc_jpeg_encode(const uint8_t *px, uint32_t w, uint32_t h, uint32_t bands, int quality, int subsampling,
                    uint8_t *out, size_t cap, size_t *len);

#ifdef __cplusplus
}
#endif
#endif
