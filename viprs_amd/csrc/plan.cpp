// Host planning: coefficient tables, plans, masks, planners, regions.
//
// The formulas below have to be the reference's, operation for operation,
// because their FP64 rounding decides which integer coefficients and which
// source positions the device uses:
//   kernel weights            crates/viprs-core/src/kernel.rs:133-211
//   tap count / tables        crates/viprs-ops-spatial/src/resample/reduce_common.rs:137-164,257-316
//   offset / source position  reduce_common.rs:171-185
//   plan (start, phase)       reduce_common.rs:127-135
//   required input regions    reduce_common.rs:218-254
//   output length             crates/viprs-core/src/resample.rs:86-93
//   Gaussian integer mask     crates/viprs-ops-spatial/src/convolution/gauss_blur.rs:80-126
//   thumbnail planner         crates/viprs-ops-spatial/src/resample/thumbnail.rs:229-524
//   resize planner            crates/viprs-ops-spatial/src/resample/resize.rs:163-345
// glibc sin/exp/fma are the same functions Rust's f64 methods call on
// x86_64-linux-gnu, so the tables come out bit-identical.
#include "plan.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>

namespace vgpu {

namespace {

constexpr double kPlanEps = 1e-9;  // thumbnail.rs:27, resize.rs:9
constexpr int kMaxTaps = 2000;     // reduce_common.rs:10 (libvips MAX_POINT)

uint32_t to_u32_sat(double v) {
    if (!(v == v) || v <= 0.0) return 0;
    if (v >= 4294967295.0) return 0xffffffffu;
    return (uint32_t)v;
}
int16_t to_i16_sat(double v) {
    if (!(v == v)) return 0;
    if (v <= -32768.0) return INT16_MIN;
    if (v >= 32767.0) return INT16_MAX;
    return (int16_t)v;
}
int64_t to_i64_sat(double v) {
    if (!(v == v)) return 0;
    if (v <= -9.2233720368547758e18) return INT64_MIN;
    if (v >= 9.2233720368547758e18) return INT64_MAX;
    return (int64_t)v;
}
int64_t clamp_i32_range(int64_t v) {
    return std::min<int64_t>(std::max<int64_t>(v, INT32_MIN), INT32_MAX);
}

double support_of(int k) {
    switch (k) {
    case NEAREST: return 0.5;
    case BILINEAR: return 1.0;
    case LANCZOS3: return 3.0;
    default: return 2.0;
    }
}

}  // namespace

double kernel_weight(int k, double x) {
    const double a = std::fabs(x);
    switch (k) {
    case NEAREST: return a < 0.5 ? 1.0 : 0.0;
    case BILINEAR: return a < 1.0 ? 1.0 - a : 0.0;
    case BICUBIC:
    case CATMULLROM:
        // Keys a = -0.5 in mul_add form.
        if (a < 1.0) return std::fma(2.5 * a, -a, 1.5 * a * a * a) + 1.0;
        if (a < 2.0) return std::fma(4.0, -a, std::fma(2.5 * a, a, -0.5 * a * a * a)) + 2.0;
        return 0.0;
    case LANCZOS2:
    case LANCZOS3: {
        const double lobes = (k == LANCZOS3) ? 3.0 : 2.0;
        if (a < DBL_EPSILON) return 1.0;
        if (a >= lobes) return 0.0;
        const double px = M_PI * a;
        return lobes * std::sin(px) * std::sin(px / lobes) / (px * px);
    }
    default: return 0.0;
    }
}

Status validate_reduce(int kernel, double factor) {
    if (kernel == LBB || kernel == VSQBS || kernel == NOHALO) return E_INVALID_KERNEL;
    if (kernel < NEAREST || kernel > CATMULLROM) return E_INVALID_KERNEL;
    if (!std::isfinite(factor) || factor < 1.0) return E_INVALID_ARGUMENT;
    if (reduce_tap_count(kernel, factor) < 0) return E_INVALID_ARGUMENT;
    return OK;
}

int reduce_tap_count(int kernel, double factor) {
    if (std::fabs(factor - 1.0) < DBL_EPSILON || kernel == NEAREST) return 1;
    const double radius = std::round(support_of(kernel) * factor);
    if (!std::isfinite(radius) || radius < 0.0 || radius > (double)((kMaxTaps - 1) / 2)) return -1;
    const long taps = 2 * (long)radius + 1;
    return taps <= kMaxTaps ? (int)taps : -1;
}

uint32_t reduce_output_len(double factor, uint32_t in) {
    return in == 0 ? 0 : to_u32_sat(std::round((double)in / factor));
}

Status build_reduce_axis(int kernel, double factor, uint32_t input_len, ReduceAxis* ax) {
    Status st = validate_reduce(kernel, factor);
    if (st != OK) return st;
    const int taps = reduce_tap_count(kernel, factor);
    ax->kernel = kernel;
    ax->factor = factor;
    ax->taps = taps;
    ax->pad = (int64_t)std::ceil((double)taps / 2.0) - 1;
    ax->wf64.assign((size_t)(kPhases + 1) * taps, 0.0);
    ax->wi16.assign((size_t)(kPhases + 1) * taps, 0);
    const double inv = 1.0 / factor;
    for (int p = 0; p <= kPhases; ++p) {
        double* row = &ax->wf64[(size_t)p * taps];
        if (kernel == NEAREST) {
            row[0] = 1.0;
        } else {
            const double centre = (double)p / (double)kPhases + (double)taps / 2.0 - 1.0;
            double sum = 0.0;
            for (int t = 0; t < taps; ++t) {
                row[t] = kernel_weight(kernel, std::fabs(((double)t - centre) * inv));
                sum += row[t];
            }
            if (sum > 0.0)
                for (int t = 0; t < taps; ++t) row[t] /= sum;
        }
        for (int t = 0; t < taps; ++t) ax->wi16[(size_t)p * taps + t] = to_i16_sat(row[t] * 4096.0);
    }
    // bind_input_len: centre the output grid like libvips.
    if (input_len == 0 || std::fabs(factor - 1.0) < DBL_EPSILON) {
        ax->offset = 0.0;
    } else {
        const double extra = std::fma((double)reduce_output_len(factor, input_len), factor,
                                      -(double)input_len);
        ax->offset = (1.0 + extra) / 2.0 - 1.0;
    }
    return OK;
}

double ReduceAxis::source_position(double index) const {
    return std::fma(index + 0.5, factor, -0.5) - offset;
}

void ReduceAxis::plan(double sp, int64_t* start, int* phase) const {
    const double fl = std::floor(sp);
    double ph = std::round(std::fma(sp - fl, (double)kPhases, 0.0));
    ph = std::min(std::max(ph, 0.0), (double)kPhases);
    *phase = (int)ph;
    *start = to_i64_sat(fl) - pad;
}

void ReduceAxis::required(int64_t origin, int64_t len, int64_t* start, int64_t* n) const {
    if (len <= 0) {
        *start = origin;
        *n = 0;
        return;
    }
    const int64_t first = to_i64_sat(std::floor(source_position((double)origin)));
    const int64_t last_i = clamp_i32_range(origin + len - 1);
    const int64_t last = to_i64_sat(std::floor(source_position((double)last_i)));
    const int64_t s = clamp_i32_range(first - pad);
    const int64_t e = clamp_i32_range(last - pad + taps - 1);
    *start = s;
    *n = e - s + 1;
}

GaussMask build_gauss_mask(float sigma_f) {
    GaussMask m;
    const double sigma = (double)sigma_f;
    if (sigma < 0.2) {
        m.taps = {1};
        m.scale = 1;
        m.weights = {1.0};
        return m;
    }
    const double sig2 = 2.0 * sigma * sigma;
    const size_t limit = (size_t)std::ceil(8.0 * sigma);
    size_t x = 0;
    while (x < limit && std::exp(-((double)(x * x)) / sig2) >= 0.2) ++x;
    const size_t radius = x == 0 ? 0 : x - 1;
    int32_t scale = 0;
    for (size_t i = 0; i < 2 * radius + 1; ++i) {
        const int32_t off = (int32_t)i - (int32_t)radius;
        const int32_t tap = (int32_t)std::nearbyint(20.0 * std::exp(-(double)(off * off) / sig2));
        m.taps.push_back((int16_t)tap);
        scale += tap;
    }
    if (scale == 0) {
        m.taps = {1};
        scale = 1;
    }
    m.scale = scale;
    for (int16_t t : m.taps) m.weights.push_back((double)t / (double)scale);
    return m;
}

// Affine::required_input_axis_bounds with source bounds (0, len)
// (affine/core.rs:373-405).
static void affine_axis(int64_t start, int64_t end, uint32_t len, int64_t* o, int64_t* n) {
    if (len == 0) {
        *o = 0;
        *n = 0;
        return;
    }
    const int64_t be = (int64_t)len - 1;
    const int64_t cs = std::max<int64_t>(start, 0), ce = std::min<int64_t>(end, be);
    if (ce < cs) {
        *o = clamp_i32_range(end < 0 ? 0 : be + 1);
        *n = 0;
        return;
    }
    *o = clamp_i32_range(cs);
    *n = clamp_i32_range(ce) - *o + 1;
}

void affine_positions(const Node& nd, bool columns, int64_t first, int64_t n, int64_t tile, int64_t shift,
                      std::vector<double>* pos) {
    pos->resize((size_t)std::max<int64_t>(n, 0));
    const double step = columns ? nd.am[0] : nd.am[3];
    for (int64_t i = 0; i < n; ++i) {
        const int64_t p = first + i;
        const int64_t t0 = (p - shift) / tile * tile + shift;  // origin of p's tile
        if (i > 0 && (p - shift) % tile != 0) {
            (*pos)[(size_t)i] = (*pos)[(size_t)i - 1] + step;
            continue;
        }
        // row_x = fma(b, oy, a * ox) + tx and row_y = fma(d, oy, c * ox) + ty
        // with b = c = 0: the other coordinate's term is an exact zero.
        double v = columns ? std::fma(nd.am[1], 0.0, nd.am[0] * (double)t0) + nd.atx
                           : std::fma(nd.am[3], (double)t0, nd.am[2] * 0.0) + nd.aty;
        for (int64_t q = t0; q < p; ++q) v += step;
        (*pos)[(size_t)i] = v;
    }
}

Rect Node::required_input(const Rect& o) const {
    Rect r = o;
    switch (kind) {
    case N_SHRINK_H: {
        const int64_t need = o.w * (int64_t)factor;
        if (o.x == 0 && source_width != 0 && need <= (int64_t)source_width) {
            r.x = 0;
            r.w = source_width;  // full-row borrow (shrinkh.rs:220-229)
        } else {
            r.x = clamp_i32_range(o.x * (int64_t)factor);
            r.w = need;
        }
        break;
    }
    case N_SHRINK_V:
        r.y = clamp_i32_range(o.y * (int64_t)factor);
        r.h = o.h * (int64_t)factor;
        break;
    case N_BOX_SHRINK:  // the whole decoded image, whatever is asked of the shrunk one
        r = Rect{0, 0, (int64_t)in_w, (int64_t)in_h};
        break;
    case N_REDUCE_H:
    case N_REDUCE_OP_H: axis.required(o.x, o.w, &r.x, &r.w); break;
    case N_REDUCE_V:
    case N_REDUCE_OP_V: axis.required(o.y, o.h, &r.y, &r.h); break;
    case N_EXTRACT:  // extract_area.rs:49-56
        r.x = clamp_i32_range(o.x + ex);
        r.y = clamp_i32_range(o.y + ey);
        break;
    case N_ZOOM: {  // zoom.rs:343-349
        const int64_t left = o.x / zx, top = o.y / zy;
        const int64_t right = (o.x + o.w - 1) / zx, bottom = (o.y + o.h - 1) / zy;
        r.x = left;
        r.y = top;
        r.w = right - left + 1;
        r.h = bottom - top + 1;
        break;
    }
    case N_AFFINE: {  // affine/core.rs:344-366,373-405,1531-1560
        const double ox0 = (double)o.x, oy0 = (double)o.y;
        const double ox1 = (double)o.x + (double)(o.w > 0 ? o.w - 1 : 0);
        const double oy1 = (double)o.y + (double)(o.h > 0 ? o.h - 1 : 0);
        double cx[4], cy[4];
        const double px[4] = {ox0, ox1, ox0, ox1}, py[4] = {oy0, oy0, oy1, oy1};
        for (int i = 0; i < 4; ++i) affine_map(px[i], py[i], &cx[i], &cy[i]);
        const double minx = std::fmin(std::fmin(std::fmin(cx[0], cx[1]), cx[2]), cx[3]);
        const double maxx = std::fmax(std::fmax(std::fmax(cx[0], cx[1]), cx[2]), cx[3]);
        const double miny = std::fmin(std::fmin(std::fmin(cy[0], cy[1]), cy[2]), cy[3]);
        const double maxy = std::fmax(std::fmax(std::fmax(cy[0], cy[1]), cy[2]), cy[3]);
        // InterpolationKernel::affine_padding (kernel.rs:80-117)
        const int64_t size = akern == NEAREST ? 1 : (akern == BILINEAR ? 2 : (akern == LANCZOS3 ? 6 : 4));
        const int64_t lp = (akern == NEAREST || akern == BILINEAR) ? 0 : (akern == LANCZOS3 ? 2 : 1);
        const int64_t rp = size - lp - 1;
        affine_axis((int64_t)std::floor(minx) - lp, (int64_t)std::floor(maxx) + rp, in_w, &r.x, &r.w);
        affine_axis((int64_t)std::floor(miny) - lp, (int64_t)std::floor(maxy) + rp, in_h, &r.y, &r.h);
        break;
    }
    case N_GAUSS_H:
    case N_CONVSEP_H:
        r.x = o.x - radius();
        r.w = o.w + 2 * radius();
        break;
    case N_GAUSS_V:
    case N_CONVSEP_V:
        r.y = o.y - radius();
        r.h = o.h + 2 * radius();
        break;
    default: break;
    }
    return r;
}

namespace {

void thumb_target_dims(int target, uint32_t tw, uint32_t th, uint32_t iw, uint32_t ih,
                       uint32_t* ow, uint32_t* oh) {
    const double fw = (double)iw, fh = (double)ih;
    switch (target) {
    case FITBOX: {
        const double s = std::min(std::min((double)std::max(tw, 1u) / fw,
                                           (double)std::max(th, 1u) / fh),
                                  1.0);
        *ow = to_u32_sat(std::max(std::round(fw * s), 1.0));
        *oh = to_u32_sat(std::max(std::round(fh * s), 1.0));
        break;
    }
    case FORCE_EXACT:
        *ow = std::max(tw, 1u);
        *oh = std::max(th, 1u);
        break;
    case WIDTH: {
        const uint32_t w = std::min(std::max(tw, 1u), iw);
        *ow = w;
        *oh = to_u32_sat(std::max(std::round((fh * (double)w) / fw), 1.0));
        break;
    }
    default: {
        const uint32_t h = std::min(std::max(th, 1u), ih);
        *oh = h;
        *ow = to_u32_sat(std::max(std::round((fw * (double)h) / fh), 1.0));
        break;
    }
    }
}

uint32_t after_reduce(uint32_t after_shrink, double residual) {
    if (residual > 1.0 + kPlanEps)
        return to_u32_sat(std::max(std::round((double)after_shrink / residual), 1.0));
    return after_shrink;
}

uint32_t integer_shrink(uint32_t in_len, uint32_t target, double total) {
    if (total <= 1.0 + kPlanEps) return 1;
    uint32_t f = to_u32_sat(std::max(std::floor(total), 1.0));
    for (; f > 1; --f)
        if (after_reduce(in_len / f, total / (double)f) >= target) break;
    return f;
}

}  // namespace

// The thumbnail's Affine step: matrix [inv_x, 0, 0, inv_y], translation
// 0.5 (1 - inv) except for Nearest, kernel mapped by
// affine_kernel_for_resize_kernel (thumbnail.rs:31-37).
static PlanStep thumb_affine(double inv_x, double inv_y, uint32_t ow, uint32_t oh, int kernel) {
    PlanStep s{N_AFFINE, 0.0};
    s.a = inv_x;
    s.d = inv_y;
    s.tx = kernel == NEAREST ? 0.0 : 0.5 * (1.0 - inv_x);
    s.ty = kernel == NEAREST ? 0.0 : 0.5 * (1.0 - inv_y);
    s.ew = ow;
    s.eh = oh;
    s.akern = kernel == NEAREST ? NEAREST : (kernel == BILINEAR ? BILINEAR : BICUBIC);
    return s;
}

Status plan_thumbnail(uint32_t iw, uint32_t ih, uint32_t bands, int target, uint32_t tw,
                      uint32_t th, std::vector<PlanStep>* steps, uint32_t* out_w,
                      uint32_t* out_h, bool crop, int kernel, uint32_t* shrink_on_load) {
    steps->clear();
    if (shrink_on_load) *shrink_on_load = 0;
    if (bands == 0) return E_INVALID_ARGUMENT;
    if (target < FITBOX || target > HEIGHT) return E_INVALID_ARGUMENT;
    if (iw == 0 || ih == 0) {
        *out_w = *out_h = 0;
        return OK;
    }
    const bool upscale = target == FORCE_EXACT;
    const bool single = iw == 1 && ih == 1;
    auto clamp_dims = [&](uint32_t* w, uint32_t* h) {
        if (!(upscale && !single)) {
            *w = std::max(std::min(*w, iw), 1u);
            *h = std::max(std::min(*h, ih), 1u);
        }
    };
    // Target (the output) and geometry (what shrink + reduce produce): equal
    // without crop; with crop a FitBox / ForceExact box is the target and the
    // geometry covers it (thumbnail.rs:246-285).
    uint32_t tgw, tgh;
    if (crop && (target == FITBOX || target == FORCE_EXACT)) {
        tgw = std::max(tw, 1u);
        tgh = std::max(th, 1u);
    } else {
        thumb_target_dims(target, tw, th, iw, ih, &tgw, &tgh);
    }
    clamp_dims(&tgw, &tgh);
    uint32_t gw = tgw, gh = tgh;
    if (crop) {
        const double cover = std::max((double)std::max(tgw, 1u) / (double)iw, (double)std::max(tgh, 1u) / (double)ih);
        gw = to_u32_sat(std::max(std::round((double)iw * cover), (double)tgw));
        gh = to_u32_sat(std::max(std::round((double)ih * cover), (double)tgh));
    }
    clamp_dims(&gw, &gh);
    std::vector<PlanStep> geo;
    if (iw == 1 || ih == 1) {
        if (single) tgw = tgh = 1;
        gw = tgw;
        gh = tgh;
        if (tgw != iw || tgh != ih)  // thumbnail.rs:292-316
            geo.push_back(thumb_affine((double)iw / (double)tgw, (double)ih / (double)tgh, tgw, tgh, kernel));
    } else {
        const double hs = (double)iw / (double)gw, vs = (double)ih / (double)gh;
        if (shrink_on_load) {  // shrink_on_load_factor (thumbnail.rs:326-330,514-524)
            const double common = std::min(hs, vs);
            *shrink_on_load = common >= 16.0 ? 16 : (common >= 8.0 ? 4 : (common >= 4.0 ? 2 : 0));
        }
        const uint32_t sh = integer_shrink(iw, gw, hs), sv = integer_shrink(ih, gh, vs);
        const double rh = hs > 1.0 ? hs / (double)sh : 1.0;
        const double rv = vs > 1.0 ? vs / (double)sv : 1.0;
        const uint32_t arw = after_reduce(iw / std::max(sh, 1u), rh);
        const uint32_t arh = after_reduce(ih / std::max(sv, 1u), rv);
        if (sh > 1) geo.push_back({N_SHRINK_H, (double)sh});
        if (sv > 1) geo.push_back({N_SHRINK_V, (double)sv});
        if (rh > 1.0 + kPlanEps) geo.push_back({N_REDUCE_H, rh});
        if (rv > 1.0 + kPlanEps) geo.push_back({N_REDUCE_V, rv});
        const double ax = (double)gw / (double)arw, ay = (double)gh / (double)arh;
        if (std::fabs(ax - 1.0) > kPlanEps || std::fabs(ay - 1.0) > kPlanEps)  // thumbnail.rs:396-419
            geo.push_back(thumb_affine(1.0 / ax, 1.0 / ay, gw, gh, kernel));
        if (crop && (gw != tgw || gh != tgh)) {  // thumbnail.rs:421-428
            PlanStep e{N_EXTRACT, 0.0};
            e.ex = (gw - tgw) / 2;
            e.ey = (gh - tgh) / 2;
            e.ew = tgw;
            e.eh = tgh;
            geo.push_back(e);
        }
    }
    const bool alpha = bands == 2 || bands == 4;
    if (alpha && !geo.empty()) steps->push_back({N_PREMULT, 0.0});
    steps->insert(steps->end(), geo.begin(), geo.end());
    if (alpha && !geo.empty()) steps->push_back({N_UNPREMULT, 0.0});
    *out_w = tgw;
    *out_h = tgh;
    return OK;
}

Status plan_resize(uint32_t iw, uint32_t ih, double hscale, double vscale, int kernel,
                   std::vector<PlanStep>* steps, int* eff_kernel) {
    steps->clear();
    if (!(std::isfinite(hscale) && hscale > 0.0)) hscale = 1.0;
    if (!(std::isfinite(vscale) && vscale > 0.0)) vscale = 1.0;
    if (kernel == NOHALO) kernel = BICUBIC;
    if (kernel < NEAREST || kernel > NOHALO) return E_INVALID_KERNEL;
    *eff_kernel = kernel;
    if (iw == 0 || ih == 0) return E_UNSUPPORTED;
    const uint32_t tw = to_u32_sat(std::max(std::round((double)iw * hscale), 1.0));
    const uint32_t th = to_u32_sat(std::max(std::round((double)ih * vscale), 1.0));
    auto affine = [&](double ax, double ay) {
        PlanStep s{N_AFFINE, 0.0};
        s.a = 1.0 / ax;
        s.d = 1.0 / ay;
        s.tx = kernel == NEAREST ? 0.0 : 0.5 * (1.0 - 1.0 / ax);
        s.ty = kernel == NEAREST ? 0.0 : 0.5 * (1.0 - 1.0 / ay);
        s.ew = tw;
        s.eh = th;
        s.akern = kernel;
        steps->push_back(s);
    };
    if (kernel == LBB) {
        // resize.rs:176-199: LBB exists only as the 2-D affine interpolator,
        // one Affine over the whole image.
        if (tw != iw || th != ih) affine((double)tw / (double)iw, (double)th / (double)ih);
        return OK;
    }
    auto preshrink = [&](uint32_t in, uint32_t t) -> uint32_t {
        if (t == 0 || in <= t || kernel == NEAREST) return 1;
        return to_u32_sat(std::max(std::floor((double)in / (double)t / 2.0), 1.0));
    };
    uint32_t cw = iw, ch = ih;
    const uint32_t hs = tw < cw ? preshrink(cw, tw) : 1;
    const uint32_t vs = th < ch ? preshrink(ch, th) : 1;
    if (vs > 1) {
        steps->push_back({N_SHRINK_V, (double)vs});
        ch /= vs;
    }
    if (hs > 1) {
        steps->push_back({N_SHRINK_H, (double)hs});
        cw /= hs;
    }
    if (th < ch) {
        steps->push_back({N_REDUCE_V, (double)ch / (double)th});
        ch = th;
    }
    if (tw < cw) {
        steps->push_back({N_REDUCE_H, (double)cw / (double)tw});
        cw = tw;
    }
    // resize.rs:240-280: integer Zoom of what remains to grow (never for
    // VSQBS, whose 4 x 4 stencil the Zoom would bypass), then the Affine
    // fix-up with the resize kernel.  A VSQBS reduce step is refused by the
    // reduce builder (reduce_common.rs:107-113), so a VSQBS downscale fails
    // as in the reference.
    const bool can_zoom = kernel != VSQBS;
    const uint32_t zx = can_zoom && tw > cw ? tw / cw : 1, zy = can_zoom && th > ch ? th / ch : 1;
    if (zx > 1 || zy > 1) {
        PlanStep z{N_ZOOM, 0.0};
        z.zx = zx;
        z.zy = zy;
        steps->push_back(z);
        cw *= zx;
        ch *= zy;
    }
    const double ax = (double)tw / (double)cw, ay = (double)th / (double)ch;
    if (std::fabs(ax - 1.0) > kPlanEps || std::fabs(ay - 1.0) > kPlanEps) affine(ax, ay);
    return OK;
}

}  // namespace vgpu
