// Device region executor, buffer pool and the C ABI.
//
// A pipeline is the reference's linear node chain (PipelineBuilder,
// crates/viprs-runtime/src/pipeline/builder/*.rs).  Running it restates
// execute_tile (crates/viprs-runtime/src/scheduler/rayon_scheduler/
// execute.rs:413-849) once over the whole output instead of per 16-row strip:
// the backward region walk gives every node the rectangle of Z^2 the
// reference computes it on (rows unioned over strips, columns identical), the
// source is read with clamp-to-edge, and each intermediate is evaluated on
// that rectangle, negative and past-the-edge coordinates included.  Adjacent
// nodes are fused into one launch where the intermediate can stay on chip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "runtime.hpp"

using namespace vgpu;

namespace vgpu {

Rect whole(uint32_t w, uint32_t h) { return Rect{0, 0, (int64_t)w, (int64_t)h}; }

InView view_of(const void* ptr, const Rect& buf, const Rect& clip) {
    InView v;
    v.ptr = ptr;
    v.bx = (int32_t)buf.x;
    v.by = (int32_t)buf.y;
    v.bw = (int32_t)buf.w;
    v.bh = (int32_t)buf.h;
    v.cx0 = (int32_t)clip.x;
    v.cy0 = (int32_t)clip.y;
    v.cx1 = (int32_t)(clip.x + std::max<int64_t>(clip.w, 1) - 1);
    v.cy1 = (int32_t)(clip.y + std::max<int64_t>(clip.h, 1) - 1);
    return v;
}

OutView outview_of(void* ptr, const Rect& r) {
    OutView o;
    o.ptr = ptr;
    o.x = (int32_t)r.x;
    o.y = (int32_t)r.y;
    o.w = (int32_t)r.w;
    o.h = (int32_t)r.h;
    return o;
}

namespace {

// Per-axis (start, phase) plan for a reduce node producing `n` outputs from
// `origin`.  ReduceH on u8 accumulates the f64 source position column by
// column from the region origin (reduceh.rs:112-119); every other case
// evaluates the position directly (reduce_simd.rs:63,111,2716-2717).
void build_plan(const ReduceAxis& ax, int64_t origin, int64_t n, bool accumulate,
                std::vector<int32_t>* starts, std::vector<uint8_t>* phases) {
    starts->resize((size_t)std::max<int64_t>(n, 1));
    phases->resize((size_t)std::max<int64_t>(n, 1));
    double sx = ax.source_position((double)origin);
    for (int64_t i = 0; i < n; ++i) {
        int64_t st;
        int ph;
        const double sp = accumulate ? sx : ax.source_position((double)origin + (double)i);
        ax.plan(sp, &st, &ph);
        (*starts)[(size_t)i] = (int32_t)std::min<int64_t>(std::max<int64_t>(st, INT32_MIN), INT32_MAX);
        (*phases)[(size_t)i] = (uint8_t)ph;
        sx += ax.factor;
    }
}

// The identity axis a fused reduce stage uses for a missing ReduceH or
// ReduceV: factor 1, one tap of weight 1.0 (4096 in fixed point), so the
// staged intermediate is an exact copy of its input.
const ReduceAxis& identity_axis() {
    static const ReduceAxis ax = [] {
        ReduceAxis a;
        build_reduce_axis(NEAREST, 1.0, 0, &a);
        return a;
    }();
    return ax;
}

// Buffer-relative clamped column / row, as the device's col_of / row_of.
int64_t clamp_rel(int64_t x, int64_t c0, int64_t cn, int64_t b0, int64_t bn) {
    x = std::min(std::max(x, c0), c0 + std::max<int64_t>(cn, 1) - 1);
    x = std::min(std::max(x, b0), b0 + std::max<int64_t>(bn, 1) - 1);
    return x - b0;
}

// One pass of the MFMA kernel as the host folds it: for output index i (a
// column of the pass, or a row) the taps start at absolute position start[i]
// and weigh wrow(i)[0 .. taps).
struct SepAxis {
    const std::vector<int32_t>* start = nullptr;
    const std::vector<uint8_t>* phase = nullptr;  // null: one weight row
    const int16_t* wtab = nullptr;
    int taps = 1;
    // 1 when start[i] = start[0] + i (and no phases): a window clear of the
    // clamps then folds to the same pattern as any other at the same
    // alignment, which the planner reuses without building its key.
    int unit_step = 0;
    const int16_t* row(size_t i) const { return wtab + (phase ? (size_t)(*phase)[i] * taps : 0); }
};

// Lay out folded weights W[j][k] (16 x ksteps*64) as MFMA operand bytes: lane
// j + 16 q holds K = 64 s + 16 q .. + 15 of step s; one int8 limb, or the
// {hi, lo} limbs of w = 128 a + c (c in [-64, 63]); then 16 int32 seeds
// 128 * sum_k W[j][k] + bias.  Only row j's window lo[j] .. hi[j] can be
// nonzero; the rest of the operand bytes stay zero.
void emit_operands(const std::vector<int32_t>& w, int ksteps, int limbs, int32_t bias, const int64_t (&lo)[16],
                   const int64_t (&hi)[16], std::vector<unsigned char>* out) {
    const size_t ops = (size_t)ksteps * limbs * 64 * 16;
    out->assign(ops + 16 * sizeof(int32_t), 0);
    unsigned char* o = out->data();
    const int kw = 64 * ksteps;
    int32_t seed[16];
    for (int j = 0; j < 16; ++j) {
        int64_t sum = 0;
        for (int64_t k = lo[j]; k <= hi[j]; ++k) {
            const int32_t v = w[(size_t)(j * kw + k)];
            sum += v;
            const int s = (int)(k >> 6), lane = j + 16 * (int)((k >> 4) & 3), e = (int)(k & 15);
            if (limbs == 1) {
                o[((size_t)s * 64 + lane) * 16 + e] = (unsigned char)(v & 255);
            } else {
                const int32_t c = ((v + 64) & 127) - 64;
                const int32_t a = (v - c) >> 7;
                o[((size_t)(2 * s) * 64 + lane) * 16 + e] = (unsigned char)(a & 255);
                o[((size_t)(2 * s + 1) * 64 + lane) * 16 + e] = (unsigned char)(c & 255);
            }
        }
        seed[j] = (int32_t)(128 * sum + bias);
    }
    std::memcpy(o + ops, seed, sizeof(seed));
}

// Zero the touched K windows of a folded-weight buffer (rows of kw entries).
template <typename T>
void clear_rows(std::vector<T>* w, int kw, const int64_t (&lo)[16], const int64_t (&hi)[16]) {
    for (int n = 0; n < 16; ++n)
        for (int64_t k = lo[n]; k <= hi[n]; ++k) (*w)[(size_t)(n * kw + k)] = T(0);
}

// Identical operand patterns are stored once.  A pattern is named by a key
// that determines its folded weights: for each of its 16 rows (or samples),
// the unclamped position of tap 0, the clamped positions of the first and the
// last tap (clamping is monotone, so these fix where every tap lands) and the
// weight row, all relative to the window start.  Only a new key folds and lays
// out its operand bytes.
struct PatternKey {
    int64_t k[16 * 4];
    void empty(int n) {
        k[4 * n] = INT64_MIN;
        k[4 * n + 1] = 0;
        k[4 * n + 2] = -1;
        k[4 * n + 3] = -1;
    }
    void set(int n, int64_t start, int64_t lo, int64_t hi, int64_t phase) {
        k[4 * n] = start;
        k[4 * n + 1] = lo;
        k[4 * n + 2] = hi;
        k[4 * n + 3] = phase;
    }
};

struct PatternSet {
    std::vector<std::vector<unsigned char>> pats;
    std::vector<PatternKey> keys;
    std::unordered_map<uint64_t, std::vector<int>> by_hash;
    template <typename Emit>
    int add(const PatternKey& key, Emit emit) {
        uint64_t h = 1469598103934665603ull;
        for (int64_t v : key.k) {
            h = (h ^ (uint64_t)v) * 1099511628211ull;
            h ^= h >> 29;
        }
        std::vector<int>& ids = by_hash[h];
        for (int id : ids)
            if (std::memcmp(keys[(size_t)id].k, key.k, sizeof(key.k)) == 0) return id;
        const int id = (int)pats.size();
        pats.emplace_back();
        emit(&pats.back());
        keys.push_back(key);
        ids.push_back(id);
        return id;
    }
    size_t put(Staging* st, int* bytes) const {
        *bytes = pats.empty() ? 0 : (int)pats[0].size();
        const size_t off = st->put(nullptr, pats.size() * (size_t)*bytes);
        unsigned char* dst = st->bytes.data() + off;
        for (const auto& p : pats) {
            std::memcpy(dst, p.data(), p.size());
            dst += p.size();
        }
        return off;
    }
};

// Largest sum |w| over the weight rows an axis uses.
int64_t axis_abs_sum(const SepAxis& ax, size_t n) {
    int64_t worst = 0;
    for (size_t i = 0; i < n; ++i) {
        int64_t sum = 0;
        for (int t = 0; t < ax.taps; ++t) sum += std::abs((int)ax.row(i)[t]);
        worst = std::max(worst, sum);
    }
    return worst;
}

// Plan the int8-limb MFMA job of a u8 H -> V pair over output rectangle `out`
// (H input clip `clip`, input buffer `buf`, V input rows `vclip`): K steps of
// the H span of each 16-sample strip and of the V window of each 16 output
// rows, limbs, ring size and task split, and the V operand patterns (and, for
// family 1, the H patterns).  Returns false when a window is too wide.
bool plan_sep(int family, const SepAxis& hax, const SepAxis& vax, int32_t hbias, int32_t vbias, uint32_t bands,
              const Rect& out, const Rect& clip, const Rect& buf, const Rect& vclip, const StageOptions& opt,
              Staging* st, Stage* stg) {
    SepJob* j = &stg->sep;
    const int64_t nsamp = out.w * (int64_t)bands;
    if (nsamp <= 0 || out.h <= 0 || bands < 1 || bands > 4) return false;
    // The kernel addresses its input with 32-bit byte offsets.
    if (buf.w * buf.h * (int64_t)bands >= ((int64_t)1 << 31)) return false;
    const std::vector<int32_t>& hs = *hax.start;
    const std::vector<int32_t>& vs = *vax.start;
    auto hcol = [&](int64_t x) { return clamp_rel(x, clip.x, clip.w, buf.x, buf.w); };
    const int64_t ntiles = (out.h + 15) / 16, nstrips = (nsamp + 15) / 16;
    int64_t span = 0;
    for (int64_t jx = 0; jx < nstrips; ++jx) {
        const int64_t s0 = 16 * jx, s1 = std::min<int64_t>(s0 + 15, nsamp - 1);
        const int64_t c0 = hcol(hs[(size_t)(s0 / bands)]);
        const int64_t c1 = hcol(hs[(size_t)(s1 / bands)] + hax.taps - 1);
        span = std::max(span, (c1 - c0 + 1) * (int64_t)bands);
    }
    const int64_t v0 = vclip.y, v1 = vclip.y + std::max<int64_t>(vclip.h, 1) - 1;
    auto vrow = [&](int64_t r) { return std::min(std::max(r, v0), v1); };
    int64_t window = 0;
    for (int64_t ti = 0; ti < ntiles; ++ti) {
        const int64_t y0 = 16 * ti, y1 = std::min<int64_t>(y0 + 15, out.h - 1);
        const int64_t lo = vrow(vs[(size_t)y0]), hi = vrow(vs[(size_t)y1] + vax.taps - 1);
        window = std::max(window, hi - (lo & ~(int64_t)15) + 1);
    }
    const int limbs = axis_abs_sum(hax, (size_t)out.w) <= 127 && axis_abs_sum(vax, (size_t)out.h) <= 127 ? 1 : 2;
    const int kh = (int)((span + 63) / 64), kv = (int)((window + 63) / 64);
    const int khmax = family == 0 ? 2 : (limbs == 1 ? 6 : 3);
    // kv <= 7 keeps the ring (16 x 512 bytes) and four wavefronts' LDS under 64 KiB.
    if (kh < 1 || kh > khmax || kv < 1 || kv > 7) return false;
    if (family == 0 && limbs != 2) return false;
    j->family = family;
    j->limbs = limbs;
    j->bands = (int)bands;
    j->kh = kh;
    j->kv = kv;
    int ring = 128;
    while (ring < 64 * kv + 16) ring *= 2;
    j->ring = ring;
    j->gx = (int)nstrips;
    int rpt = opt.sep_rows_per_task;
    if (rpt <= 0) {
        // One image alone: long runs while about 2048 wavefronts remain.
        rpt = 128;
        while (rpt > 16 && (int64_t)j->gx * ((out.h + rpt - 1) / rpt) < 2048) rpt /= 2;
    }
    j->rows_per_task = rpt;
    j->gy = (int)((out.h + rpt - 1) / rpt);

    // V operand patterns per 16 output rows, with each tile's (base16, hi).
    PatternSet vp;
    std::vector<int32_t> vtile((size_t)ntiles * 3);
    std::vector<int32_t> w;
    int vmemo[16];  // unit-step axis: pattern of a full, unclamped tile by its offset in base16
    std::fill(vmemo, vmemo + 16, -1);
    for (int64_t ti = 0; ti < ntiles; ++ti) {
        const int64_t y0 = 16 * ti, nrow = std::min<int64_t>(16, out.h - y0);
        const int64_t lo = vrow(vs[(size_t)y0]), hi = vrow(vs[(size_t)(y0 + nrow - 1)] + vax.taps - 1);
        const int64_t base16 = lo & ~(int64_t)15;
        if (ti == 0) w.assign((size_t)16 * 64 * kv, 0);
        vtile[(size_t)(3 * ti + 1)] = (int32_t)base16;
        vtile[(size_t)(3 * ti + 2)] = (int32_t)hi;
        const bool clear = vax.unit_step && !vax.phase && nrow == 16 && lo == vs[(size_t)y0] &&
                           hi == vs[(size_t)(y0 + 15)] + vax.taps - 1;
        if (clear && vmemo[lo - base16] >= 0) {
            vtile[(size_t)(3 * ti)] = vmemo[lo - base16];
            continue;
        }
        PatternKey key;
        for (int m = 0; m < 16; ++m) {
            if (m >= nrow) {
                key.empty(m);
                continue;
            }
            const int64_t st0 = vs[(size_t)(y0 + m)];
            key.set(m, st0 - base16, vrow(st0) - base16, vrow(st0 + vax.taps - 1) - base16,
                    vax.phase ? (*vax.phase)[(size_t)(y0 + m)] : 0);
        }
        vtile[(size_t)(3 * ti)] = vp.add(key, [&](std::vector<unsigned char>* b) {
            int64_t klo[16], khi[16];
            for (int m = 0; m < 16; ++m) {
                klo[m] = key.k[4 * m + 1];
                khi[m] = key.k[4 * m + 2];
            }
            for (int64_t m = 0; m < nrow; ++m) {
                const int16_t* row = vax.row((size_t)(y0 + m));
                for (int t = 0; t < vax.taps; ++t)
                    w[(size_t)(m * 64 * kv + (vrow(vs[(size_t)(y0 + m)] + t) - base16))] += row[t];
            }
            emit_operands(w, kv, limbs, vbias, klo, khi, b);
            clear_rows(&w, 64 * kv, klo, khi);
        });
        if (clear) vmemo[lo - base16] = vtile[(size_t)(3 * ti)];
    }
    stg->v_tab = vp.put(st, &j->vtab_bytes);
    stg->v_tile = st->put(vtile.data(), vtile.size() * sizeof(int32_t));

    // Block-tile form: four strips x R output rows per block with its input
    // rows, intermediate columns and V operands staged in LDS (H operands
    // from host-folded patterns for both families).  R: opt.sep_rows_per_task
    // if set, else the largest of 64 / 32 / 16 that still gives 1024 blocks
    // (16 otherwise); the form is used when that R needs at most 48 KiB.
    j->tile = 0;
    if (opt.sep_form == 0 && kh <= (limbs == 1 ? 6 : 3)) {
        std::vector<int64_t> span0s((size_t)nstrips);
        for (int64_t jx = 0; jx < nstrips; ++jx) span0s[(size_t)jx] = hcol(hs[(size_t)(16 * jx / bands)]);
        const int64_t gxb = (nstrips + 3) / 4;
        int64_t rsb = 0;
        for (int64_t bx = 0; bx < gxb; ++bx) {
            const int64_t f = span0s[(size_t)(4 * bx)], l = span0s[(size_t)std::min(4 * bx + 3, nstrips - 1)];
            rsb = std::max(rsb, (l - f) * (int64_t)bands + 64 * kh + 8);
        }
        rsb = (rsb + 15) / 16 * 16;
        struct Geo {
            int64_t rows = 0, rs = 0, lds = 0, blocks = 0;
        };
        auto geo = [&](int64_t R) {
            Geo g;
            const int64_t gyb = (out.h + R - 1) / R;
            for (int64_t by = 0; by < gyb; ++by) {
                const int64_t t0 = by * R / 16, t1 = (std::min(out.h, by * R + R) - 1) / 16;
                const int64_t r_lo = vtile[(size_t)(3 * t0 + 1)], r_hi = vtile[(size_t)(3 * t1 + 2)];
                g.rows = std::max(g.rows, r_hi - r_lo + 1);
                g.rs = std::max(g.rs, vtile[(size_t)(3 * t1 + 1)] - r_lo + 64 * (int64_t)kv);
            }
            g.rs = (g.rs + 15) / 16 * 16;
            if ((g.rs / 16) % 2 == 0) g.rs += 16;  // odd 16-byte stride: sample columns spread over the banks
            g.lds = g.rows * rsb + 64 * g.rs + (R / 16) * (int64_t)kv * limbs * 1024;
            g.blocks = gxb * gyb;
            return g;
        };
        int64_t R = 0;
        Geo g;
        if (opt.sep_rows_per_task > 0) {
            for (int64_t r = opt.sep_rows_per_task & ~15; r >= 16 && !R; r = (r / 2) & ~15) {
                g = geo(r);
                if (g.lds <= 48 * 1024) R = r;
            }
        } else {
            for (int64_t r : {64, 32, 16}) {
                const Geo c = geo(r);
                if (c.lds > 48 * 1024) continue;
                R = r;
                g = c;
                if (c.blocks >= 1024) break;
            }
        }
        if (R > 0) {
            j->tile = 1;
            j->rows_per_task = (int)R;
            j->gy = (int)((out.h + R - 1) / R);
            j->t_row_stride = (int)rsb;
            j->t_rows = (int)g.rows;
            j->t_inter_stride = (int)g.rs;
            j->t_lds = (int)g.lds;
            const uint64_t chunks = (uint64_t)rsb / 16;
            j->t_chunk_magic = (uint32_t)(((1ull << 32) + chunks - 1) / chunks);
        }
    }

    if (family == 1 || j->tile) {
        // H operand patterns per 16-sample strip, with each strip's first
        // span pixel.
        PatternSet hp;
        std::vector<int32_t> hstrip((size_t)nstrips * 2);
        int hmemo[4] = {-1, -1, -1, -1};  // unit-step axis: full, unclamped strip by s0 % bands
        for (int64_t jx = 0; jx < nstrips; ++jx) {
            const int64_t s0 = 16 * jx;
            const int64_t span0 = hcol(hs[(size_t)(s0 / bands)]);
            if (jx == 0) w.assign((size_t)16 * 64 * kh, 0);
            hstrip[(size_t)(2 * jx + 1)] = (int32_t)span0;
            const int64_t xl = (s0 + 15) / bands;
            const bool clear = hax.unit_step && !hax.phase && s0 + 15 < nsamp &&
                               span0 == hs[(size_t)(s0 / bands)] - buf.x &&
                               hcol(hs[(size_t)xl] + hax.taps - 1) == hs[(size_t)xl] + hax.taps - 1 - buf.x;
            if (clear && hmemo[s0 % bands] >= 0) {
                hstrip[(size_t)(2 * jx)] = hmemo[s0 % bands];
                continue;
            }
            PatternKey key;
            for (int n = 0; n < 16; ++n) {
                if (s0 + n >= nsamp) {
                    key.empty(n);
                    continue;
                }
                const int64_t smp = s0 + n, x = smp / bands, b = smp - x * bands;
                const int64_t st0 = hs[(size_t)x];
                key.set(n, (st0 - span0) * bands + b, (hcol(st0) - span0) * bands + b,
                        (hcol(st0 + hax.taps - 1) - span0) * bands + b, hax.phase ? (*hax.phase)[(size_t)x] : 0);
            }
            hstrip[(size_t)(2 * jx)] = hp.add(key, [&](std::vector<unsigned char>* bytes) {
                int64_t klo[16], khi[16];
                for (int n = 0; n < 16; ++n) {
                    klo[n] = key.k[4 * n + 1];
                    khi[n] = key.k[4 * n + 2];
                }
                for (int n = 0; n < 16 && s0 + n < nsamp; ++n) {
                    const int64_t smp = s0 + n, x = smp / bands, b = smp - x * bands;
                    const int16_t* row = hax.row((size_t)x);
                    for (int t = 0; t < hax.taps; ++t)
                        w[(size_t)(n * 64 * kh + (hcol(hs[(size_t)x] + t) - span0) * bands + b)] += row[t];
                }
                emit_operands(w, kh, limbs, hbias, klo, khi, bytes);
                clear_rows(&w, 64 * kh, klo, khi);
            });
            if (clear) hmemo[s0 % bands] = hstrip[(size_t)(2 * jx)];
        }
        stg->h_tab = hp.put(st, &j->htab_bytes);
        stg->h_strip = st->put(hstrip.data(), hstrip.size() * sizeof(int32_t));
    }
    if (family == 0) j->h_bias = hbias;
    stg->kind = ST_SEP;
    return true;
}

// Tap tables of an Affine stage (node i): the source position of every
// output column and row as the reference's tile walk computes it, then its
// AffineTap.  The scheduler's tiles cover the pipeline's output (the last
// rectangle of out_r): 128 x 128 for SmallTile, else full-width strips of 16
// rows (1 row beyond 10000 px) (op/demand.rs:41-85); the ExtractArea nodes
// after the Affine place that output at their summed offset.
void stage_affine_params(const std::vector<Node>& nodes, int i, const std::vector<Rect>& out_r, Staging* st,
                         Stage* s) {
    const Node& nd = nodes[(size_t)i];
    const Rect& fin = out_r.back();
    int64_t sx = 0, sy = 0;
    for (size_t k = (size_t)i + 1; k < nodes.size(); ++k)
        if (nodes[k].kind == N_EXTRACT) {
            sx += nodes[k].ex;
            sy += nodes[k].ey;
        }
    const int64_t tw = nd.asmall ? 128 : std::max<int64_t>(fin.w, 1);
    const int64_t th = std::max<int64_t>(std::min<int64_t>(nd.asmall ? 128 : (fin.w > 10000 ? 1 : 16), fin.h), 1);
    auto taps = [&](bool columns, int64_t first, int64_t n, int64_t tile, int64_t shift) {
        std::vector<double> pos;
        affine_positions(nd, columns, first, n, tile, shift, &pos);
        std::vector<AffineTap> t(pos.size());
        for (size_t k = 0; k < pos.size(); ++k) {
            const double x = pos[k];
            AffineTap& a = t[k];
            if (nd.akern == LBB || nd.akern == VSQBS) {
                // the 4 x 4 stencil from floor(x) - 1; phase x - floor(x)
                const double c = std::floor(x);
                a.i0 = (int32_t)((int64_t)c - 1);
                a.f = x - c;
            } else if (nd.akern != NEAREST && nd.akern != BILINEAR) {
                // interp_separable<S>: taps floor(x) - S + 1 .. floor(x) + S
                const int S = nd.akern == LANCZOS3 ? 3 : 2;
                const int64_t c = (int64_t)std::floor(x);
                a.i0 = (int32_t)(c - S + 1);
                for (int q = 0; q < 2 * S; ++q)
                    a.w[q] = kernel_weight(nd.akern, std::fabs((double)(c - S + 1 + q) - x));
            } else {  // `x as i64` truncates toward zero
                const int64_t c = (int64_t)x;
                a.i0 = (int32_t)c;
                a.f = x - (double)c;
            }
        }
        return t;
    };
    const std::vector<AffineTap> xs = taps(true, s->out.x, s->out.w, tw, sx);
    const std::vector<AffineTap> ys = taps(false, s->out.y, s->out.h, th, sy);
    s->aff_x = st->put(xs.data(), xs.size() * sizeof(AffineTap));
    s->aff_y = st->put(ys.data(), ys.size() * sizeof(AffineTap));
}

void stage_reduce_params(const Node& nd, const Rect& out, Staging* st, Stage* s) {
    std::vector<int32_t> starts;
    std::vector<uint8_t> phases;
    if (nd.kind == N_REDUCE_H || nd.kind == N_REDUCE_OP_H)
        build_plan(nd.axis, out.x, out.w, nd.kind == N_REDUCE_H && nd.in_fmt == U8, &starts, &phases);
    else
        build_plan(nd.axis, out.y, out.h, false, &starts, &phases);
    s->plan_start = st->put(starts.data(), starts.size() * sizeof(int32_t));
    s->plan_phase = st->put(phases.data(), phases.size());
    s->w16 = st->put(nd.axis.wi16.data(), nd.axis.wi16.size() * sizeof(int16_t));
    s->w64 = st->put(nd.axis.wf64.data(), nd.axis.wf64.size() * sizeof(double));
    s->taps = nd.axis.taps;
}

// Widest span of source columns the H taps of one tx-column tile read.
int64_t reduce_span(const std::vector<int32_t>& hs, int htaps, int64_t outw, int tx) {
    int64_t span = 1;
    for (int64_t x0 = 0; x0 < outw; x0 += tx) {
        const int64_t x1 = std::min<int64_t>(x0 + tx, outw) - 1;
        span = std::max<int64_t>(span, (int64_t)hs[(size_t)x1] + htaps - hs[(size_t)x0]);
    }
    return span;
}

// Tile of the fused VALU reduce kernels.  Candidates, largest first: take the
// first whose staged H rows plus two staged source rows fit 64 KiB of LDS and
// that still gives >= opt.reduce_min_blocks blocks; otherwise the fitting
// candidate with the most blocks.  The search starts at candidate
// opt.reduce_first_tile.
bool pick_reduce_tile(const std::vector<int32_t>& vs, int vtaps, const std::vector<int32_t>& hs, int htaps,
                      const Rect& out, uint32_t bands, size_t esz, bool stage, const StageOptions& opt, int* tx,
                      int* ty, int* nr) {
    static const int cand[][2] = {{64, 16}, {32, 16}, {32, 8}, {16, 8}, {16, 4}, {8, 4}, {8, 2}, {8, 1}, {4, 1}, {1, 1}};
    int64_t best_blocks = -1;
    const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
    for (int ci = std::min(std::max(opt.reduce_first_tile, 0), ncand - 1); ci < ncand; ++ci) {
        const int* c = cand[ci];
        int rows = 1;
        for (int64_t y0 = 0; y0 < out.h; y0 += c[1]) {
            const int64_t y1 = std::min<int64_t>(y0 + c[1], out.h) - 1;
            rows = std::max<int>(rows, vs[(size_t)y1] + vtaps - vs[(size_t)y0]);
        }
        // With staged source rows (stage): the H rows plus one source row
        // per row lane of the block (a chunk of 256 / tx rows) within 40 KiB;
        // else the H rows within 64 KiB.
        const size_t hbuf = ((size_t)rows * c[0] * bands * esz + 15) & ~(size_t)15;
        const size_t srow = (size_t)reduce_span(hs, htaps, out.w, c[0]) * bands * esz;
        if (stage ? hbuf + (size_t)(256 / c[0]) * srow > 40 * 1024 : hbuf > 64 * 1024) continue;
        const int64_t blocks = ((out.w + c[0] - 1) / c[0]) * ((out.h + c[1] - 1) / c[1]);
        if (blocks > best_blocks) {
            best_blocks = blocks;
            *tx = c[0];
            *ty = c[1];
            *nr = rows;
        }
        if (blocks >= opt.reduce_min_blocks) break;
    }
    return best_blocks >= 0;
}

// H source staging of a planned fused VALU reduce job (f32 and ReduceOp;
// integer fused reduces read the source directly, hr = 0): the widest tile
// span and the rows per staged chunk, rb (1, 2 or 4) rows per row lane of
// the block, as many as keep the tile's LDS within 40 KiB (pick_reduce_tile
// guarantees one).
void set_reduce_staging(ReduceJob* j, const std::vector<int32_t>& hs, int htaps, int64_t outw, size_t esz) {
    if (esz < 4 && !j->op) {
        j->span_max = 1;
        j->hr = 0;
        j->rb = 1;
        return;
    }
    j->span_max = (int)reduce_span(hs, htaps, outw, j->tx);
    const size_t hbuf = ((size_t)j->nr_max * j->tx * j->bands * esz + 15) & ~(size_t)15;
    const size_t srow = (size_t)j->span_max * j->bands * esz;
    const int lanes = std::max(1, 256 / std::max(j->tx, 1));
    int rb = 4;
    while (rb > 1 && hbuf + (size_t)(rb * lanes) * srow > 40 * 1024) rb /= 2;
    j->rb = rb;
    j->hr = rb * lanes;
    if (esz < 4) {
        // ReduceOp on integers sums one row per thread iteration: stage as
        // many rows per chunk as fit (fewer chunks, fewer barriers).
        const size_t fit = hbuf < 40 * 1024 ? (40 * 1024 - hbuf) / srow : 0;
        j->hr = (int)std::max<size_t>((size_t)lanes, std::min<size_t>(fit, (size_t)j->nr_max));
    }
}

// Streaming form of an f32 reduce pair: the widest strip (at most 256 / B
// output columns, at least 128 samples) whose H span stays within 1024
// staged samples a row, and runs of output rows sized so that the blocks fill
// the chip once (each run re-reads the taps_v - stride halo rows).  Leaves
// job.stream = 0 where no strip fits.
void plan_reduce_stream(ReduceJob* j, const std::vector<int32_t>& hs, const std::vector<int32_t>& vs, int htaps,
                        int vtaps, const Rect& out) {
    j->stream = 0;
    const int B = j->bands;
    if (B < 1 || B > 4 || out.w < 1 || out.h < 1) return;
    int tx = (int)std::min<int64_t>(256 / B, out.w);
    int64_t span = 0;
    for (; tx >= 1; --tx) {
        span = reduce_span(hs, htaps, out.w, tx);
        if (span * B <= 1024) break;
    }
    if (tx < 1 || ((int64_t)tx * B < 128 && tx < out.w)) return;
    // Output rows whose clamped V windows hold a row at once (the kernel
    // keeps one accumulator per open output: 8 or 16).
    const int64_t c0 = j->vclip0, c1 = j->vclip1;
    int open = 0;
    for (int64_t y = 0, y1 = 0; y < out.h; ++y) {
        const int64_t lo = std::min(std::max<int64_t>(vs[(size_t)y], c0), c1);
        while (y1 < y && std::min(std::max<int64_t>((int64_t)vs[(size_t)y1] + vtaps - 1, c0), c1) < lo) ++y1;
        open = std::max(open, (int)(y - y1 + 1));
    }
    if (open > 16) return;
    ReduceJob t = *j;
    t.st_tx = tx;
    t.st_span = (int)span;
    t.st_open = open;
    t.st_gx = (int)((out.w + tx - 1) / tx);
    // One round of resident blocks (LDS-limited, at most 4 a CU) where the
    // runs stay within 128 rows: a second, partly filled round would idle
    // most of the chip for a whole block lifetime.
    t.stream = 0;
    for (int64_t per_cu = 4; per_cu >= 1 && !t.stream; --per_cu) {
        const int64_t gy = std::max<int64_t>(1, 256 * per_cu / t.st_gx);
        t.st_run = (int)std::min<int64_t>(128, (out.h + gy - 1) / gy);
        // Intermediate rows of the widest run, plus one.
        int64_t rows = 1;
        for (int64_t y0 = 0; y0 < out.h; y0 += t.st_run) {
            const int64_t y1 = std::min<int64_t>(y0 + t.st_run, out.h) - 1;
            const int64_t lo = std::min(std::max<int64_t>(vs[(size_t)y0], c0), c1);
            const int64_t hi = std::min(std::max<int64_t>((int64_t)vs[(size_t)y1] + vtaps - 1, c0), c1);
            rows = std::max(rows, hi - lo + 1);
        }
        t.st_pad = (int)rows + 1;
        const size_t lds = reduce_stream_lds(t, 0);
        if (lds <= 64 * 1024 && lds * (size_t)per_cu <= 160 * 1024) t.stream = 1;
    }
    if (!t.stream) return;
    t.st_gy = (int)((out.h + t.st_run - 1) / t.st_run);
    *j = t;
}

// f32 copy of a weight table.
size_t put_f32(const std::vector<double>& w, Staging* st) {
    std::vector<float> f(w.size());
    for (size_t i = 0; i < w.size(); ++i) f[i] = (float)w[i];
    return st->put(f.data(), f.size() * sizeof(float));
}

// Float pair on the FP16 hi/lo MFMA kernel: fold the f64 weights of each
// 16-sample strip and each 16 output rows (clamped taps added), split every
// folded weight w (as f32) into FP16 hi = f16(w), lo = f16(w - hi), and lay
// them out as operands of K steps of 32 (lane j + 16 q holds K = 32 s + 8 q
// .. + 7), deduplicated.  Returns false when a window does not fit.
struct F16Axis {
    const std::vector<int32_t>* start = nullptr;
    const std::vector<uint8_t>* phase = nullptr;  // null: one weight row
    const double* wtab = nullptr;
    int taps = 1;
    const double* row(size_t i) const { return wtab + (phase ? (size_t)(*phase)[i] * taps : 0); }
};

void emit_f16_operands(const std::vector<double>& w, int ksteps, int exp2, const int64_t (&lo)[16],
                       const int64_t (&hi)[16], std::vector<unsigned char>* out) {
    out->assign((size_t)ksteps * 2 * 64 * 16, 0);
    unsigned char* o = out->data();
    const int kw = 32 * ksteps;
    for (int j = 0; j < 16; ++j)
        for (int64_t k = lo[j]; k <= hi[j]; ++k) {
            const float f = std::ldexp((float)w[(size_t)(j * kw + k)], exp2);
            const _Float16 h = (_Float16)f;
            const _Float16 l = (_Float16)(f - (float)h);
            const int s = (int)(k >> 5), lane = j + 16 * (int)((k >> 3) & 3), e = (int)(k & 7);
            std::memcpy(o + ((size_t)(2 * s) * 64 + lane) * 16 + 2 * e, &h, 2);
            std::memcpy(o + ((size_t)(2 * s + 1) * 64 + lane) * 16 + 2 * e, &l, 2);
        }
}

bool plan_f16(const F16Axis& hax, const F16Axis& vax, uint32_t bands, const Rect& out, const Rect& clip,
              const Rect& buf, const Rect& vclip, const StageOptions& opt, Staging* st, Stage* stg) {
    F16Job* j = &stg->f16;
    const int64_t nsamp = out.w * (int64_t)bands;
    if (nsamp <= 0 || out.h <= 0 || bands < 1 || bands > 4) return false;
    if (buf.w * buf.h * (int64_t)bands >= ((int64_t)1 << 31)) return false;  // 32-bit element offsets
    const std::vector<int32_t>& hs = *hax.start;
    const std::vector<int32_t>& vs = *vax.start;
    auto hcol = [&](int64_t x) { return clamp_rel(x, clip.x, clip.w, buf.x, buf.w); };
    const int64_t v0 = vclip.y, v1 = vclip.y + std::max<int64_t>(vclip.h, 1) - 1;
    auto vrow = [&](int64_t r) { return std::min(std::max(r, v0), v1); };
    const int64_t ntiles = (out.h + 15) / 16, nstrips = (nsamp + 15) / 16;
    int64_t span = 0, window = 0;
    for (int64_t jx = 0; jx < nstrips; ++jx) {
        const int64_t s0 = 16 * jx, s1 = std::min<int64_t>(s0 + 15, nsamp - 1);
        span = std::max(span, (hcol(hs[(size_t)(s1 / bands)] + hax.taps - 1) - hcol(hs[(size_t)(s0 / bands)]) + 1) *
                                  (int64_t)bands);
    }
    for (int64_t ti = 0; ti < ntiles; ++ti) {
        const int64_t y0 = 16 * ti, y1 = std::min<int64_t>(y0 + 15, out.h - 1);
        const int64_t lo = vrow(vs[(size_t)y0]), hi = vrow(vs[(size_t)y1] + vax.taps - 1);
        window = std::max(window, hi - (lo & ~(int64_t)7) + 1);
    }
    const int kh = (int)((span + 31) / 32), kv = (int)((window + 31) / 32);
    if (kh < 1 || kh > 4 || kv < 1 || kv > 2) return false;
    j->bands = (int)bands;
    j->kh = kh;
    j->kv = kv;
    j->ring = 128;
    j->gx = (int)nstrips;
    int rpt = opt.sep_rows_per_task & ~15;  // whole 16-row tiles per task
    if (rpt <= 0) {
        rpt = 128;
        while (rpt > 16 && (int64_t)j->gx * ((out.h + rpt - 1) / rpt) < 2048) rpt /= 2;
        // Longer tasks recompute fewer halo rows (a task's first tile needs
        // the V window's rows above it): lengthen them while about three
        // wavefronts per SIMD are left.  4096^2x3 sigma 8 (29 taps): 128 ->
        // 1024 rows, 322 -> 277 us; the 2.5x reduce keeps 128.
        while (rpt < 1024 && (int64_t)j->gx * ((out.h + 2 * rpt - 1) / (2 * rpt)) >= 3072) rpt *= 2;
    }
    j->rows_per_task = rpt;
    j->gy = (int)((out.h + rpt - 1) / rpt);
    // Weight scales: the folded sums are bounded by the sum of |w| of a row,
    // so max|w| * 2^exp < 2^14 holds for every folded weight.
    auto axis_exp = [](const F16Axis& a, size_t rows) {
        double m = 0;
        for (size_t i = 0; i < rows; ++i) {
            double t = 0;
            for (int k = 0; k < a.taps; ++k) t += std::fabs(a.row(i)[k]);
            m = std::max(m, t);
        }
        int e = 0;
        std::frexp(m, &e);
        return 14 - e;
    };
    j->h_exp = axis_exp(hax, (size_t)out.w);
    j->v_exp = axis_exp(vax, (size_t)out.h);
    std::vector<double> w;
    PatternSet vp, hp;
    std::vector<int32_t> vtile((size_t)ntiles * 3), hstrip((size_t)nstrips * 2);
    for (int64_t ti = 0; ti < ntiles; ++ti) {
        const int64_t y0 = 16 * ti, nrow = std::min<int64_t>(16, out.h - y0);
        const int64_t lo = vrow(vs[(size_t)y0]), hi = vrow(vs[(size_t)(y0 + nrow - 1)] + vax.taps - 1);
        const int64_t base8 = lo & ~(int64_t)7;
        if (ti == 0) w.assign((size_t)16 * 32 * kv, 0.0);
        PatternKey key;
        for (int m = 0; m < 16; ++m) {
            if (m >= nrow) {
                key.empty(m);
                continue;
            }
            const int64_t st0 = vs[(size_t)(y0 + m)];
            key.set(m, st0 - base8, vrow(st0) - base8, vrow(st0 + vax.taps - 1) - base8,
                    vax.phase ? (*vax.phase)[(size_t)(y0 + m)] : 0);
        }
        vtile[(size_t)(3 * ti)] = vp.add(key, [&](std::vector<unsigned char>* b) {
            int64_t klo[16], khi[16];
            for (int m = 0; m < 16; ++m) {
                klo[m] = key.k[4 * m + 1];
                khi[m] = key.k[4 * m + 2];
            }
            for (int64_t m = 0; m < nrow; ++m) {
                const double* row = vax.row((size_t)(y0 + m));
                for (int t = 0; t < vax.taps; ++t)
                    w[(size_t)(m * 32 * kv + (vrow(vs[(size_t)(y0 + m)] + t) - base8))] += row[t];
            }
            emit_f16_operands(w, kv, j->v_exp, klo, khi, b);
            clear_rows(&w, 32 * kv, klo, khi);
        });
        vtile[(size_t)(3 * ti + 1)] = (int32_t)base8;
        vtile[(size_t)(3 * ti + 2)] = (int32_t)hi;
    }
    for (int64_t jx = 0; jx < nstrips; ++jx) {
        const int64_t s0 = 16 * jx;
        const int64_t span0 = hcol(hs[(size_t)(s0 / bands)]);
        if (jx == 0) w.assign((size_t)16 * 32 * kh, 0.0);
        PatternKey key;
        for (int n = 0; n < 16; ++n) {
            if (s0 + n >= nsamp) {
                key.empty(n);
                continue;
            }
            const int64_t smp = s0 + n, x = smp / bands, b = smp - x * bands;
            const int64_t st0 = hs[(size_t)x];
            key.set(n, (st0 - span0) * bands + b, (hcol(st0) - span0) * bands + b,
                    (hcol(st0 + hax.taps - 1) - span0) * bands + b, hax.phase ? (*hax.phase)[(size_t)x] : 0);
        }
        hstrip[(size_t)(2 * jx)] = hp.add(key, [&](std::vector<unsigned char>* bytes) {
            int64_t klo[16], khi[16];
            for (int n = 0; n < 16; ++n) {
                klo[n] = key.k[4 * n + 1];
                khi[n] = key.k[4 * n + 2];
            }
            for (int n = 0; n < 16 && s0 + n < nsamp; ++n) {
                const int64_t smp = s0 + n, x = smp / bands, b = smp - x * bands;
                const double* row = hax.row((size_t)x);
                for (int t = 0; t < hax.taps; ++t)
                    w[(size_t)(n * 32 * kh + (hcol(hs[(size_t)x] + t) - span0) * bands + b)] += row[t];
            }
            emit_f16_operands(w, kh, j->h_exp, klo, khi, bytes);
            clear_rows(&w, 32 * kh, klo, khi);
        });
        hstrip[(size_t)(2 * jx + 1)] = (int32_t)span0;
    }
    stg->v_tab = vp.put(st, &j->vtab_bytes);
    stg->v_tile = st->put(vtile.data(), vtile.size() * sizeof(int32_t));
    stg->h_tab = hp.put(st, &j->htab_bytes);
    stg->h_strip = st->put(hstrip.data(), hstrip.size() * sizeof(int32_t));
    stg->kind = ST_SEP_F16;
    return true;
}

// Fused reduce stage: H plan over the intermediate's columns (`mid`), V plan
// over the output rows, and a tile shape whose staged rows fit 64 KiB of LDS.
// `accumulate_h` selects the u8 ReduceH plan; an identity axis stands in for
// a missing pass.  Returns false when no tile fits (very large factors).
bool stage_reduce_hv_params(const ReduceAxis& hax, bool accumulate_h, const ReduceAxis& vax,
                            int fmt, uint32_t bands, const Rect& mid, const Rect& out,
                            const Rect& vclip, const Rect& clip, const Rect& buf,
                            const StageOptions& opt, Staging* st, Stage* s, bool reduce_op = false) {
    if (!base_format(fmt)) return false;  // I16, U32, I32, F64: the per-node kernels
    std::vector<int32_t> hs, vs;
    std::vector<uint8_t> hp, vp;
    build_plan(hax, mid.x, mid.w, accumulate_h, &hs, &hp);
    build_plan(vax, out.y, out.h, false, &vs, &vp);
    s->kind = ST_REDUCE_HV;  // replaced below by ST_SEP, ST_RED_F32 or ST_SEP_F16
    if (fmt == U8 && opt.integer_mfma && !reduce_op) {
        const SepAxis ha{&hs, &hp, hax.wi16.data(), hax.taps}, va{&vs, &vp, vax.wi16.data(), vax.taps};
        plan_sep(0, ha, va, 2048, 2048, bands, out, clip, buf, vclip, opt, st, s);  // ST_SEP on success
    }
    int tx = 0, ty = 0, nr = 0;
    const bool tile_ok = pick_reduce_tile(vs, vax.taps, hs, hax.taps, out, bands, format_size(fmt),
                                          fmt == F32 || reduce_op, opt, &tx, &ty, &nr);
    if (!tile_ok) {
        if (s->kind != ST_SEP) return false;
        tx = ty = nr = 1;  // the VALU tile is not used
    }
    s->plan_start = st->put(hs.data(), hs.size() * sizeof(int32_t));
    s->plan_phase = st->put(hp.data(), hp.size());
    s->w16 = st->put(hax.wi16.data(), hax.wi16.size() * sizeof(int16_t));
    s->w64 = st->put(hax.wf64.data(), hax.wf64.size() * sizeof(double));
    s->taps = hax.taps;
    s->v_start = st->put(vs.data(), vs.size() * sizeof(int32_t));
    s->v_phase = st->put(vp.data(), vp.size());
    s->v_w16 = st->put(vax.wi16.data(), vax.wi16.size() * sizeof(int16_t));
    s->v_w64 = st->put(vax.wf64.data(), vax.wf64.size() * sizeof(double));
    s->v_taps = vax.taps;
    ReduceJob& j = s->rjob;
    j.bands = (int)bands;
    j.fmt = fmt;
    j.tx = tx;
    j.ty = ty;
    j.nr_max = nr;
    j.vclip0 = (int)vclip.y;
    j.vclip1 = (int)(vclip.y + std::max<int64_t>(vclip.h, 1) - 1);
    j.gx = (int)((out.w + tx - 1) / tx);
    j.gy = (int)((out.h + ty - 1) / ty);
    j.op = reduce_op ? 1 : 0;
    set_reduce_staging(&j, hs, hax.taps, out.w, format_size(fmt));
    // int32 sums are exact when max over phases of sum |w| times the largest
    // sample stays below 2^31.
    const int64_t maxv = fmt == U8 ? 255 : 65535;
    int64_t worst = 0;
    for (const ReduceAxis* ax : {&hax, &vax})
        for (int p = 0; p <= kPhases; ++p) {
            int64_t sabs = 0;
            for (int t = 0; t < ax->taps; ++t) sabs += std::abs((int64_t)ax->wi16[(size_t)p * ax->taps + t]);
            worst = std::max(worst, sabs * maxv);
        }
    j.wide_acc = worst >= (int64_t)INT32_MAX - 4096;
    j.op = reduce_op ? 1 : 0;
    if (fmt == F32 && !reduce_op && opt.reduce_form == 0) {
        j.ph.taps = hax.taps;  // the planner's LDS estimate reads the tap counts
        j.pv.taps = vax.taps;
        plan_reduce_stream(&j, hs, vs, hax.taps, vax.taps, out);
    }
    if (fmt == F32 && opt.float_mode >= 1 && tile_ok && !reduce_op) {
        s->w32 = put_f32(hax.wf64, st);
        s->v_w32 = put_f32(vax.wf64, st);
        s->kind = ST_RED_F32;
        if (opt.float_mode == 2) {
            const F16Axis ha{&hs, &hp, hax.wf64.data(), hax.taps}, va{&vs, &vp, vax.wf64.data(), vax.taps};
            plan_f16(ha, va, bands, out, clip, buf, vclip, opt, st, s);  // ST_SEP_F16 on success
        }
    }
    return true;
}

// GaussBlurH -> GaussBlurV on u8 as one MFMA stage: every output column /
// row starts its taps `radius` before itself, one phase, the integer taps.
bool stage_gauss_sep_params(const Node& gh, const Node& gv, const Rect& mid, const Rect& out,
                            const Rect& vclip, const Rect& clip, const Rect& buf,
                            const StageOptions& opt, Staging* st, Stage* s) {
    const int taps = (int)gh.itaps.size();
    const int rad = (taps - 1) / 2;
    // mul_hi by ceil(2^32 / scale) is an exact division for sums below
    // 256 * scale when scale < 4096.
    if (taps < 1 || gv.itaps != gh.itaps || gh.iscale != gv.iscale || gh.iscale < 1 || gh.iscale >= 4096)
        return false;
    std::vector<int32_t> hs((size_t)std::max<int64_t>(mid.w, 1)), vs((size_t)std::max<int64_t>(out.h, 1));
    for (int64_t x = 0; x < mid.w; ++x) hs[(size_t)x] = (int32_t)(mid.x + x - rad);
    for (int64_t y = 0; y < out.h; ++y) vs[(size_t)y] = (int32_t)(out.y + y - rad);
    const SepAxis ha{&hs, nullptr, gh.itaps.data(), taps, 1}, va{&vs, nullptr, gv.itaps.data(), taps, 1};
    const int32_t bias = gh.iscale / 2;
    if (!plan_sep(1, ha, va, bias, bias, gh.bands, out, clip, buf, vclip, opt, st, s)) return false;
    SepJob& j = s->sep;
    if (gh.iscale == 1) {
        j.h_mul = j.v_mul = 1;
        j.h_hi = j.v_hi = 0;
    } else {
        const uint32_t mul = (uint32_t)((((uint64_t)1 << 32) + (uint64_t)gh.iscale - 1) / (uint64_t)gh.iscale);
        j.h_mul = j.v_mul = mul;
        j.h_hi = j.v_hi = 1;
    }
    j.h_shift = j.v_shift = 0;
    return true;
}

// GaussBlurH -> GaussBlurV on 3- or 4-band u8 as one VALU stage (packed
// 16-bit sums, the integer taps, (sum + scale / 2) / scale per pass).
bool stage_gauss_u8_params(const Node& gh, const Node& gv, const Rect& out, const StageOptions& opt, Stage* s) {
    const int taps = (int)gh.itaps.size();
    if (opt.gauss_rows < 0) return false;
    if (gv.itaps != gh.itaps || gh.iscale != gv.iscale || !gauss_u8_supported((int)gh.bands, taps, gh.iscale))
        return false;
    GaussJob& j = s->gauss;
    j.bands = (int)gh.bands;
    j.radius = (taps - 1) / 2;
    // 32-row tiles pay off once the mask is 5 taps or wider (fewer halo rows
    // per output row); 16 for 3 taps (measured on 4096x4096 images).
    j.th = opt.gauss_rows ? opt.gauss_rows : (j.radius >= 2 ? 32 : 16);
    for (int t = 0; t < taps; ++t) {
        const uint32_t w = (uint32_t)(uint16_t)gh.itaps[(size_t)t];
        j.taps2[t] = w | (w << 16);
    }
    const uint32_t half = (uint32_t)(gh.iscale / 2);
    j.half2 = half | (half << 16);
    j.mul = gh.iscale == 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)gh.iscale - 1) / (uint64_t)gh.iscale);
    j.gx = (int)((out.w + 63) / 64);
    // Blocks walk down their column of tiles (the next tile's loads in flight
    // while one is computed); enough blocks for about 8 a CU.
    const int64_t rows = (out.h + j.th - 1) / j.th;
    const int64_t want = std::max<int64_t>(1, (int64_t)256 * 8 / std::max(j.gx, 1));
    j.tiles_per_block = (int)std::max<int64_t>(1, (rows + want - 1) / want);
    j.gy = (int)((rows + j.tiles_per_block - 1) / j.tiles_per_block);
    s->kind = ST_GAUSS_U8;
    return true;
}

// GaussBlurH -> GaussBlurV on f32 images in plain FP32: the float kernel
// taps / scale (gauss_blur.rs G1) as f32 weights, one phase.
bool stage_gauss_float_params(const Node& gh, const Node& gv, const Rect& mid, const Rect& out, const Rect& vclip,
                              const Rect& clip, const Rect& buf, const StageOptions& opt, Staging* st, Stage* s) {
    const int taps = (int)gh.conv.size();
    const int rad = (taps - 1) / 2;
    if (taps < 1 || gv.conv != gh.conv) return false;
    std::vector<int32_t> hs((size_t)std::max<int64_t>(mid.w, 1)), vs((size_t)std::max<int64_t>(out.h, 1));
    for (int64_t x = 0; x < mid.w; ++x) hs[(size_t)x] = (int32_t)(mid.x + x - rad);
    for (int64_t y = 0; y < out.h; ++y) vs[(size_t)y] = (int32_t)(out.y + y - rad);
    int tx, ty, nr;
    if (!pick_reduce_tile(vs, taps, hs, taps, out, gh.bands, sizeof(float), true, opt, &tx, &ty, &nr)) return false;
    const std::vector<uint8_t> hp(hs.size(), 0), vp(vs.size(), 0);
    s->plan_start = st->put(hs.data(), hs.size() * sizeof(int32_t));
    s->plan_phase = st->put(hp.data(), hp.size());
    s->w32 = put_f32(gh.conv, st);
    s->taps = taps;
    s->v_start = st->put(vs.data(), vs.size() * sizeof(int32_t));
    s->v_phase = st->put(vp.data(), vp.size());
    s->v_w32 = put_f32(gv.conv, st);
    s->v_taps = taps;
    ReduceJob& j = s->rjob;
    j.bands = (int)gh.bands;
    j.fmt = F32;
    j.tx = tx;
    j.ty = ty;
    j.nr_max = nr;
    j.vclip0 = (int)vclip.y;
    j.vclip1 = (int)(vclip.y + std::max<int64_t>(vclip.h, 1) - 1);
    j.gx = (int)((out.w + tx - 1) / tx);
    j.gy = (int)((out.h + ty - 1) / ty);
    set_reduce_staging(&j, hs, taps, out.w, sizeof(float));
    s->kind = ST_RED_F32;
    if (opt.float_mode == 2) {
        const F16Axis ha{&hs, nullptr, gh.conv.data(), taps}, va{&vs, nullptr, gv.conv.data(), taps};
        plan_f16(ha, va, gh.bands, out, clip, buf, vclip, opt, st, s);  // ST_SEP_F16 on success
    }
    return true;
}

// Box rounding of ShrinkH (x86 dispatch, shrinkh.rs:381-397) and ShrinkV
// (shrinkv.rs:206-339) as (bias, mul, shift, mul_hi): pair rounding at f = 2,
// exact /5 for ShrinkH f = 5 with 1, 3 or 4 bands, else the multiply-shift
// ((sum + f/2) * floor(2^24 / f)) >> 24.
void shrink_rounding(int f, int bands, bool horizontal, int32_t* bias, uint32_t* mul, int* shift, int* hi) {
    if (f == 2) {
        *bias = 1, *mul = 1, *shift = 1, *hi = 0;
    } else if (horizontal && f == 5 && bands != 2 && bands <= 4) {
        *bias = 2, *mul = 0xCCCCCCCDu, *shift = 2, *hi = 1;  // floor(x / 5) = mul_hi(x, ceil(2^34 / 5)) >> 2
    } else {
        *bias = f / 2, *mul = (uint32_t)((1u << 24) / (uint32_t)f), *shift = 24, *hi = 0;
    }
}

// [Premultiply] -> ShrinkH(fh) -> ShrinkV(fv) -> [Unpremultiply] on u8 as one
// MFMA stage: box taps of weight 1, fh = 1 / fv = 1 for a missing pass.
bool stage_shrink_sep_params(int fh, int fv, uint32_t bands, const Rect& mid, const Rect& out, const Rect& vclip,
                             const Rect& clip, const Rect& buf, const StageOptions& opt, Staging* st, Stage* s) {
    if (fh < 1 || fv < 1 || fh > 127 || fv > 127) return false;  // folded counts fit int8
    std::vector<int32_t> hs((size_t)std::max<int64_t>(mid.w, 1)), vs((size_t)std::max<int64_t>(out.h, 1));
    for (int64_t x = 0; x < mid.w; ++x) hs[(size_t)x] = (int32_t)((mid.x + x) * fh);
    for (int64_t y = 0; y < out.h; ++y) vs[(size_t)y] = (int32_t)((out.y + y) * fv);
    const std::vector<int16_t> hones((size_t)fh, 1), vones((size_t)fv, 1);
    const SepAxis ha{&hs, nullptr, hones.data(), fh}, va{&vs, nullptr, vones.data(), fv};
    int32_t hb, vb;
    uint32_t hm, vm;
    int hsh, vsh, hhi, vhi;
    shrink_rounding(fh, (int)bands, true, &hb, &hm, &hsh, &hhi);
    shrink_rounding(fv, (int)bands, false, &vb, &vm, &vsh, &vhi);
    if (!plan_sep(1, ha, va, hb, vb, bands, out, clip, buf, vclip, opt, st, s)) return false;
    SepJob& j = s->sep;
    j.h_mul = hm, j.h_shift = hsh, j.h_hi = hhi;
    j.v_mul = vm, j.v_shift = vsh, j.v_hi = vhi;
    return true;
}

// The fused float pair: both masks odd and of one length, the strip and its
// H ring within the kernel's limits; rows per block run chosen so that the
// launch has about 1000 blocks or more (each run recomputes 2r halo H rows).
bool stage_conv_pair_params(const Node& nh, const Node& nv, int kind, const Rect& out, const StageOptions& opt,
                            Staging* st, Stage* s) {
    (void)opt;
    const size_t n = nh.conv.size();
    if (n == 0 || n % 2 == 0 || nv.conv.size() != n) return false;
    const int rad = (int)(n / 2);
    const int B = (int)nh.bands;
    if (!conv_pair_supported(B, rad)) return false;
    ConvPairJob& j = s->cpair;
    j.bands = B;
    j.tw = kConvPairThreads / B;
    j.rad = rad;
    const int64_t strips = (out.w + j.tw - 1) / j.tw;
    int run = 256;
    while (run > 32 && strips * ((out.h + run - 1) / run) < 1024) run /= 2;
    j.run = run;
    j.hscale = nh.cscale;
    j.hoffset = nh.coffset;
    j.vscale = nv.cscale;
    j.voffset = nv.coffset;
    s->kind = ST_CONV_PAIR;
    s->cpair_kind = kind;
    s->cp_wh = st->put(nh.conv.data(), n * sizeof(double));
    s->cp_wv = st->put(nv.conv.data(), n * sizeof(double));
    std::vector<float> h32(n), v32(n);
    for (size_t t = 0; t < n; ++t) {
        h32[t] = (float)nh.conv[t];
        v32[t] = (float)nv.conv[t];
    }
    s->cp_wh32 = st->put(h32.data(), n * sizeof(float));
    s->cp_wv32 = st->put(v32.data(), n * sizeof(float));
    return true;
}

void stage_conv_params(const Node& nd, Staging* st, Stage* s) {
    s->itaps = st->put(nd.itaps.data(), nd.itaps.size() * sizeof(int16_t));
    s->wts = st->put(nd.conv.data(), nd.conv.size() * sizeof(double));
    s->taps = (int)nd.conv.size();
}

}  // namespace

ShrinkJob shrink_job_of(const Stage& s, const Node& first, const InView& in, const OutView& out) {
    ShrinkJob j;
    j.in = in;
    j.out = out;
    j.bands = (int)first.bands;
    j.fh = s.fh;
    j.fv = s.fv;
    j.premult = s.premult;
    j.unpremult = s.unpremult;
    return j;
}

ReduceJob reduce_job_of(const Stage& s, const InView& in, const OutView& out, const unsigned char* blob) {
    ReduceJob j = s.rjob;
    j.in = in;
    j.out = out;
    j.ph.start = reinterpret_cast<const int32_t*>(blob + s.plan_start);
    j.ph.phase = reinterpret_cast<const uint8_t*>(blob + s.plan_phase);
    j.ph.wi16 = reinterpret_cast<const int16_t*>(blob + s.w16);
    j.ph.wf64 = reinterpret_cast<const double*>(blob + s.w64);
    j.ph.taps = s.taps;
    j.pv.start = reinterpret_cast<const int32_t*>(blob + s.v_start);
    j.pv.phase = reinterpret_cast<const uint8_t*>(blob + s.v_phase);
    j.pv.wi16 = reinterpret_cast<const int16_t*>(blob + s.v_w16);
    j.pv.wf64 = reinterpret_cast<const double*>(blob + s.v_w64);
    j.pv.taps = s.v_taps;
    j.ph.wf32 = reinterpret_cast<const float*>(blob + s.w32);
    j.pv.wf32 = reinterpret_cast<const float*>(blob + s.v_w32);
    j.premult = s.premult;
    j.unpremult = s.unpremult;
    return j;
}

SepJob sep_job_of(const Stage& s, const InView& in, const OutView& out, const unsigned char* blob) {
    SepJob j = s.sep;
    j.in = in;
    j.out = out;
    j.ph.start = reinterpret_cast<const int32_t*>(blob + s.plan_start);
    j.ph.phase = reinterpret_cast<const uint8_t*>(blob + s.plan_phase);
    j.ph.wi16 = reinterpret_cast<const int16_t*>(blob + s.w16);
    j.ph.taps = s.taps;
    j.vtab = blob + s.v_tab;
    j.vtile = reinterpret_cast<const int32_t*>(blob + s.v_tile);
    j.htab = blob + s.h_tab;
    j.hstrip = reinterpret_cast<const int32_t*>(blob + s.h_strip);
    j.premult = s.premult;
    j.unpremult = s.unpremult;
    return j;
}

// Launch one stage reading `in` and writing `dst` (dense over s.out).
int launch_stage(const std::vector<Node>& nodes, const Stage& s, const InView& in, void* dst,
                 const unsigned char* blob, hipStream_t stream, const StageOptions& opt) {
    const Node& nd = nodes[(size_t)s.first];
    const OutView out = outview_of(dst, s.out);
    const int B = (int)nd.bands;
    hipError_t e = hipSuccess;
    switch (s.kind) {
    case ST_SHRINK_HV_U8: {
        ShrinkJob j = shrink_job_of(s, nd, in, out);
        if (plan_shrink_job(&j, opt.shrink_row_split, opt.shrink_run)) {
            e = launch_shrink_job(j, stream);
        } else {
            e = launch_shrink_hv_u8(in, out, B, s.fh, s.fv, s.premult, stream);
            if (e == hipSuccess && s.unpremult)  // pixel-local, in place
                e = launch_alpha_u8(view_of(dst, s.out, s.out), out, B, ALPHA_UNPREMULT, stream);
        }
        break;
    }
    case ST_SHRINK:
        e = launch_shrink_generic(in, out, B, nd.in_fmt, (int)nd.factor, nd.kind == N_SHRINK_V, stream);
        break;
    case ST_BOX: {
        // The u8 windows that lie inside the source on the box-sum form of
        // the shrink kernel; the partial windows of the last column and row
        // (and every other format) one thread per output pixel.
        const int f = (int)nd.factor, sw = (int)nd.in_w, sh = (int)nd.in_h;
        const int dw = (int)nd.out_w, dh = (int)nd.out_h;
        const int fw = std::min(dw, sw / f), fhh = std::min(dh, sh / f);
        bool fast = false;
        if (nd.in_fmt == U8 && B <= 4 && fw > 0 && fhh > 0) {
            ShrinkJob j;
            j.in = in;
            j.out = OutView{dst, 0, 0, fw, fhh};
            j.ostride = dw;
            j.bands = B;
            j.fh = j.fv = f;
            if (plan_box_shrink_job(&j, opt.shrink_row_split)) {
                fast = true;
                e = launch_shrink_job(j, stream);
                if (e == hipSuccess && dw > fw)
                    e = launch_box_generic(in.ptr, sw, sh, B, U8, f, dst, dw, fw, 0, dw - fw, fhh, stream);
                if (e == hipSuccess && dh > fhh)
                    e = launch_box_generic(in.ptr, sw, sh, B, U8, f, dst, dw, 0, fhh, dw, dh - fhh, stream);
            }
        }
        if (!fast) e = launch_box_generic(in.ptr, sw, sh, B, nd.in_fmt, f, dst, dw, 0, 0, dw, dh, stream);
        break;
    }
    case ST_REDUCE_H:
    case ST_REDUCE_V: {
        DevPlan p;
        p.start = reinterpret_cast<const int32_t*>(blob + s.plan_start);
        p.phase = reinterpret_cast<const uint8_t*>(blob + s.plan_phase);
        p.wi16 = reinterpret_cast<const int16_t*>(blob + s.w16);
        p.wf64 = reinterpret_cast<const double*>(blob + s.w64);
        p.taps = s.taps;
        e = s.kind == ST_REDUCE_H ? launch_reduce_h(in, out, B, nd.in_fmt, p, s.reduce_op, stream)
                                  : launch_reduce_v(in, out, B, nd.in_fmt, p, s.reduce_op, stream);
        break;
    }
    case ST_REDUCE_HV: e = launch_reduce_job(reduce_job_of(s, in, out, blob), stream); break;
    case ST_SEP: e = launch_sep_job(sep_job_of(s, in, out, blob), stream); break;
    case ST_RED_F32: e = launch_red_f32_job(reduce_job_of(s, in, out, blob), stream); break;
    case ST_GAUSS_U8: {
        GaussJob j = s.gauss;
        j.in = in;
        j.out = out;
        e = launch_gauss_u8(j, stream);
        break;
    }
    case ST_SEP_F16: {
        F16Job j = s.f16;
        j.in = in;
        j.out = out;
        j.htab = blob + s.h_tab;
        j.hstrip = reinterpret_cast<const int32_t*>(blob + s.h_strip);
        j.vtab = blob + s.v_tab;
        j.vtile = reinterpret_cast<const int32_t*>(blob + s.v_tile);
        e = launch_f16_job(j, stream);
        break;
    }
    case ST_CONV: {
        const bool gauss = nd.kind == N_GAUSS_H || nd.kind == N_GAUSS_V;
        const bool vert = nd.kind == N_GAUSS_V || nd.kind == N_CONVSEP_V;
        e = launch_conv(in, out, B, nd.in_fmt, nd.out_fmt, gauss ? 0 : 1, vert ? 1 : 0,
                        reinterpret_cast<const int16_t*>(blob + s.itaps), nd.iscale,
                        reinterpret_cast<const double*>(blob + s.wts), s.taps, nd.cscale,
                        nd.coffset, stream);
        break;
    }
    case ST_CONV_PAIR: {
        ConvPairJob j = s.cpair;
        j.in = in;
        j.out = out;
        j.wh = reinterpret_cast<const double*>(blob + s.cp_wh);
        j.wv = reinterpret_cast<const double*>(blob + s.cp_wv);
        j.wh32 = reinterpret_cast<const float*>(blob + s.cp_wh32);
        j.wv32 = reinterpret_cast<const float*>(blob + s.cp_wv32);
        e = launch_conv_pair(j, s.in_fmt, s.cpair_kind, stream);
        break;
    }
    case ST_ALPHA:
        e = launch_alpha_u8(in, out, B, nd.kind == N_PREMULT ? ALPHA_PREMULT : ALPHA_UNPREMULT, stream);
        break;
    case ST_PLACE: {  // dst is the whole destination: rows of place_w pixels
        const size_t ps = (size_t)B * format_size(s.out_fmt);
        const size_t row = (size_t)s.out.w * ps;
        unsigned char* d = static_cast<unsigned char*>(dst) + ((size_t)s.out.y * (size_t)s.place_w + (size_t)s.out.x) * ps;
        const unsigned char* sp = static_cast<const unsigned char*>(in.ptr) +
                                  ((size_t)(s.out.y - in.by) * (size_t)in.bw + (size_t)(s.out.x - in.bx)) * ps;
        e = hipMemcpy2DAsync(d, (size_t)s.place_w * ps, sp, (size_t)in.bw * ps, row, (size_t)s.out.h,
                             hipMemcpyDeviceToDevice, stream);
        break;
    }
    case ST_ZOOM: e = launch_zoom(in, out, B, nd.in_fmt, (int)nd.zx, (int)nd.zy, stream); break;
    case ST_AFFINE: {
        const Node& an = nodes[(size_t)s.aff_node];
        AffineJob j;
        j.in = in;
        j.out = out;
        j.bands = B;
        j.fmt = an.in_fmt;
        if (s.aff_node > s.first && nodes[(size_t)s.aff_node - 1].kind == N_ZOOM) {
            j.zx = (int32_t)nodes[(size_t)s.aff_node - 1].zx;
            j.zy = (int32_t)nodes[(size_t)s.aff_node - 1].zy;
        }
        j.premult = s.premult;
        j.unpremult = s.unpremult;
        j.kernel = an.akern == NEAREST    ? AFF_NEAREST
                   : an.akern == BILINEAR ? AFF_BILINEAR
                   : an.akern == LANCZOS3 ? AFF_SEP6
                   : an.akern == VSQBS    ? AFF_VSQBS
                   : an.akern == LBB      ? AFF_LBB
                                          : AFF_SEP4;
        j.in_w = (int32_t)an.in_w;
        j.in_h = (int32_t)an.in_h;
        j.xs = reinterpret_cast<const AffineTap*>(blob + s.aff_x);
        j.ys = reinterpret_cast<const AffineTap*>(blob + s.aff_y);
        e = launch_affine(j, stream);
        break;
    }
    case ST_COPY: {
        // ExtractArea reads its input at the output's coordinates plus its offset.
        OutView o = out;
        if (nd.kind == N_EXTRACT) {
            o.x += nd.ex;
            o.y += nd.ey;
        }
        e = launch_copy(in, o, B, nd.in_fmt, stream);
        break;
    }
    }
    return from_hip(e);
}

// Group a pipeline's nodes into fused stages over the regions of the backward
// walk; every stage's plan data goes into `st`.
bool affine_tail_supported(const std::vector<Node>& nodes) {
    bool seen = false;
    for (const Node& nd : nodes) {
        if (seen && nd.kind != N_EXTRACT && nd.kind != N_UNPREMULT) return false;
        if (nd.kind == N_AFFINE) seen = true;
    }
    return true;
}

std::vector<Stage> plan_stages(const std::vector<Node>& nodes, const std::vector<Rect>& in_r,
                               const std::vector<Rect>& out_r, Staging* st, const StageOptions& opt,
                               int first_node) {
    std::vector<Stage> stages;
    const int n = (int)nodes.size();
    auto kind_at = [&](int k) { return k < n ? nodes[(size_t)k].kind : N_COPY; };
    for (int i = first_node; i < n;) {
        const Node& nd = nodes[(size_t)i];
        Stage s;
        s.first = i;
        s.clip = in_r[(size_t)i];
        s.in_fmt = nd.in_fmt;
        const bool small_bands = nd.bands <= 4;
        const bool u8small = nd.in_fmt == U8 && small_bands;
        const bool alpha_ok = nd.in_fmt == U8 && (nd.bands == 2 || nd.bands == 4);
        // The stage's input buffer: the source, or the previous stage's output.
        const Rect buf = i == 0 ? whole(nodes[0].in_w, nodes[0].in_h) : out_r[(size_t)i - 1];
        bool done = false;
        if (u8small && (nd.kind == N_SHRINK_H || nd.kind == N_SHRINK_V ||
                        (alpha_ok && nd.kind == N_PREMULT &&
                         (kind_at(i + 1) == N_SHRINK_H || kind_at(i + 1) == N_SHRINK_V)))) {
            // [Premultiply] -> ShrinkH -> [ShrinkV] -> [Unpremultiply], or a lone ShrinkV.
            int j = i, jh = -1, jv = -1;
            s.kind = ST_SHRINK_HV_U8;
            if (nodes[(size_t)j].kind == N_PREMULT) {
                s.premult = 1;
                ++j;
            }
            if (nodes[(size_t)j].kind == N_SHRINK_H) {
                jh = j;
                s.fh = (int)nodes[(size_t)j].factor;
                if (kind_at(j + 1) == N_SHRINK_V) {
                    jv = ++j;
                    s.fv = (int)nodes[(size_t)j].factor;
                }
            } else {
                jv = j;
                s.fv = (int)nodes[(size_t)j].factor;
            }
            const int jlast = j;
            if (alpha_ok && kind_at(j + 1) == N_UNPREMULT) {
                s.unpremult = 1;
                ++j;
            }
            s.last = j;
            done = true;
            if (opt.integer_mfma && opt.mfma_shrink) {
                // The box pair on the MFMA kernel (ST_SEP) when its windows fit.
                const Rect mid = jh >= 0 ? out_r[(size_t)jh] : in_r[(size_t)jv];
                const Rect vclip = jv >= 0 ? in_r[(size_t)jv] : out_r[(size_t)jh];
                stage_shrink_sep_params(s.fh, s.fv, nd.bands, mid, out_r[(size_t)jlast], vclip, s.clip, buf, opt,
                                        st, &s);
            }
        } else if (small_bands && (nd.kind == N_REDUCE_H || nd.kind == N_REDUCE_V ||
                                   (alpha_ok && nd.kind == N_PREMULT &&
                                    (kind_at(i + 1) == N_REDUCE_H || kind_at(i + 1) == N_REDUCE_V)))) {
            // [Premultiply] -> ReduceH -> ReduceV, a lone ReduceH or a lone
            // ReduceV, -> [Unpremultiply].  Premultiply is pixel-local, so its
            // region equals the first reduce node's input region.
            const int r0 = nd.kind == N_PREMULT ? i + 1 : i;
            const Node& rn = nodes[(size_t)r0];
            const bool has_h = rn.kind == N_REDUCE_H;
            const int vi = has_h ? (kind_at(r0 + 1) == N_REDUCE_V ? r0 + 1 : -1) : r0;
            const int last = vi >= 0 ? vi : r0;
            const ReduceAxis& hax = has_h ? rn.axis : identity_axis();
            const ReduceAxis& vax = vi >= 0 ? nodes[(size_t)vi].axis : identity_axis();
            const Rect mid = has_h ? out_r[(size_t)r0] : in_r[(size_t)r0];
            const Rect out = out_r[(size_t)last];
            const Rect vclip = vi >= 0 ? in_r[(size_t)vi] : out_r[(size_t)r0];
            if (stage_reduce_hv_params(hax, has_h && rn.in_fmt == U8, vax, rn.in_fmt, rn.bands, mid, out,
                                       vclip, s.clip, buf, opt, st, &s)) {
                s.premult = nd.kind == N_PREMULT ? 1 : 0;
                s.last = last;
                if (alpha_ok && kind_at(last + 1) == N_UNPREMULT) {
                    s.unpremult = 1;
                    s.last = last + 1;
                }
                done = true;
            }
        }
        if (!done && small_bands && nd.kind == N_REDUCE_OP_H && kind_at(i + 1) == N_REDUCE_OP_V) {
            // ReduceOp: both f64 passes on one tile, the H results kept in
            // LDS in the sample type as the reference's scratch holds them.
            if (stage_reduce_hv_params(nd.axis, false, nodes[(size_t)i + 1].axis, nd.in_fmt, nd.bands,
                                       out_r[(size_t)i], out_r[(size_t)i + 1], in_r[(size_t)i + 1], s.clip, buf,
                                       opt, st, &s, true)) {
                s.last = i + 1;
                done = true;
            }
        }
        if (!done && u8small && nd.kind == N_GAUSS_H && kind_at(i + 1) == N_GAUSS_V && nd.out_fmt == U8 &&
            nodes[(size_t)i + 1].out_fmt == U8 && stage_gauss_u8_params(nd, nodes[(size_t)i + 1], out_r[(size_t)i + 1], opt, &s)) {
            // GaussBlurH -> GaussBlurV (u8 -> u8, 3 or 4 bands) on VALU tiles.
            s.last = i + 1;
            done = true;
        }
        if (!done && u8small && opt.integer_mfma && nd.kind == N_GAUSS_H && kind_at(i + 1) == N_GAUSS_V &&
            nd.out_fmt == U8 && nodes[(size_t)i + 1].out_fmt == U8) {
            // GaussBlurH -> GaussBlurV (u8 -> u8) on the MFMA kernel.
            if (stage_gauss_sep_params(nd, nodes[(size_t)i + 1], out_r[(size_t)i], out_r[(size_t)i + 1],
                                       in_r[(size_t)i + 1], s.clip, buf, opt, st, &s)) {
                s.last = i + 1;
                done = true;
            }
        }
        if (!done && small_bands && i + 1 < n) {
            // GaussBlurH -> GaussBlurV on the float path (reference order, or
            // plain FP32 in float mode 1) or ConvSepH -> ConvSepV: one fused
            // strip kernel.
            const Node& nv = nodes[(size_t)i + 1];
            int kind = -1;
            if (nd.kind == N_GAUSS_H && nv.kind == N_GAUSS_V && nd.in_fmt != U8 && nd.out_fmt == F32 &&
                nv.out_fmt == F32 && opt.float_mode <= 1)
                kind = opt.float_mode == 1 ? 2 : 0;
            if (nd.kind == N_CONVSEP_H && nv.kind == N_CONVSEP_V && base_format(nd.in_fmt)) kind = 1;
            if (kind >= 0 && stage_conv_pair_params(nd, nv, kind, out_r[(size_t)i + 1], opt, st, &s)) {
                s.last = i + 1;
                done = true;
            }
        }
        if (!done && small_bands && opt.float_mode >= 1 && nd.kind == N_GAUSS_H && kind_at(i + 1) == N_GAUSS_V &&
            nd.in_fmt == F32 && nodes[(size_t)i + 1].out_fmt == F32) {
            // GaussBlurH -> GaussBlurV on f32 in a fast float mode.
            if (stage_gauss_float_params(nd, nodes[(size_t)i + 1], out_r[(size_t)i], out_r[(size_t)i + 1],
                                         in_r[(size_t)i + 1], s.clip, buf, opt, st, &s)) {
                s.last = i + 1;
                done = true;
            }
        }
        if (!done) {
            // [Premultiply] -> [Zoom] -> Affine -> [Unpremultiply]: the Zoom
            // folds into the Affine's sample reads, the alpha steps (u8, 2 or 4
            // bands) into its loads and stores.
            int j = i;
            const bool pm = alpha_ok && nd.kind == N_PREMULT;
            if (pm) ++j;
            if (kind_at(j) == N_ZOOM) ++j;
            if (kind_at(j) == N_AFFINE) {
                s.kind = ST_AFFINE;
                s.premult = pm ? 1 : 0;
                s.aff_node = j;
                s.last = j;
                if (alpha_ok && kind_at(j + 1) == N_UNPREMULT) {
                    s.unpremult = 1;
                    s.last = j + 1;
                }
                done = true;
            }
        }
        if (!done) {
            s.last = i;
            switch (nd.kind) {
            case N_SHRINK_H:
            case N_SHRINK_V: s.kind = ST_SHRINK; break;
            case N_REDUCE_H: s.kind = ST_REDUCE_H; break;
            case N_REDUCE_V: s.kind = ST_REDUCE_V; break;
            case N_REDUCE_OP_H:
                s.kind = ST_REDUCE_H;
                s.reduce_op = 1;
                break;
            case N_REDUCE_OP_V:
                s.kind = ST_REDUCE_V;
                s.reduce_op = 1;
                break;
            case N_GAUSS_H:
            case N_GAUSS_V:
            case N_CONVSEP_H:
            case N_CONVSEP_V: s.kind = ST_CONV; break;
            case N_PREMULT:
            case N_UNPREMULT: s.kind = ST_ALPHA; break;
            case N_AFFINE: s.kind = ST_AFFINE; break;
            case N_ZOOM: s.kind = ST_ZOOM; break;
            case N_BOX_SHRINK: s.kind = ST_BOX; break;
            default: s.kind = ST_COPY; break;
            }
        }
        s.out = out_r[(size_t)s.last];
        s.out_fmt = nodes[(size_t)s.last].out_fmt;
        if (s.kind == ST_AFFINE) stage_affine_params(nodes, s.aff_node, out_r, st, &s);
        if (s.kind == ST_REDUCE_H || s.kind == ST_REDUCE_V) stage_reduce_params(nd, s.out, st, &s);
        if (s.kind == ST_CONV) stage_conv_params(nd, st, &s);
        stages.push_back(s);
        i = s.last + 1;
    }
    return stages;
}

static void backward_regions_from(const std::vector<Node>& nodes, const Rect& fin, std::vector<Rect>* in_r,
                                  std::vector<Rect>* out_r) {
    const int n = (int)nodes.size();
    in_r->assign((size_t)n, Rect{});
    out_r->assign((size_t)n, Rect{});
    if (n == 0) return;
    (*out_r)[(size_t)n - 1] = fin;
    for (int i = n - 1; i >= 0; --i) {
        (*in_r)[(size_t)i] = nodes[(size_t)i].required_input((*out_r)[(size_t)i]);
        if (i > 0) (*out_r)[(size_t)i - 1] = (*in_r)[(size_t)i];
    }
    if (nodes[0].kind == N_BOX_SHRINK && n > 1) {
        // The pre-shrunk image is the chain's source: computed whole, read
        // with clamp-to-edge over whatever region node 1 asks for.
        (*out_r)[0] = whole(nodes[0].out_w, nodes[0].out_h);
    }
}

void backward_regions(const std::vector<Node>& nodes, uint32_t w, uint32_t h, std::vector<Rect>* in_r,
                      std::vector<Rect>* out_r) {
    backward_regions_from(nodes, whole(w, h), in_r, out_r);
}

std::vector<Stage> plan_pipeline(const std::vector<Node>& nodes, uint32_t w, uint32_t h, Staging* st,
                                 const StageOptions& opt) {
    bool small = false, u8_reduce_h = false;
    for (const Node& nd : nodes) {
        if (nd.kind == N_AFFINE && nd.asmall) small = true;
        if (nd.kind == N_REDUCE_H && nd.in_fmt == U8) u8_reduce_h = true;
    }
    std::vector<Rect> in_r, out_r;
    // A box pre-shrink (node 0) is one stage over the whole image; the other
    // nodes read its output as their source.
    const bool box = !nodes.empty() && nodes[0].kind == N_BOX_SHRINK;
    std::vector<Stage> all;
    if (box) {
        Stage b;
        b.kind = ST_BOX;
        b.first = b.last = 0;
        b.clip = whole(nodes[0].in_w, nodes[0].in_h);
        b.out = whole(nodes[0].out_w, nodes[0].out_h);
        b.in_fmt = b.out_fmt = nodes[0].in_fmt;
        all.push_back(b);
    }
    if (!(small && u8_reduce_h) || nodes.empty()) {
        backward_regions(nodes, w, h, &in_r, &out_r);
        std::vector<Stage> rest = plan_stages(nodes, in_r, out_r, st, opt, box ? 1 : 0);
        all.insert(all.end(), rest.begin(), rest.end());
        return all;
    }
    for (int64_t x0 = 0; x0 < (int64_t)w; x0 += 128) {
        const Rect fin{x0, 0, std::min<int64_t>(128, (int64_t)w - x0), (int64_t)h};
        backward_regions_from(nodes, fin, &in_r, &out_r);
        std::vector<Stage> run = plan_stages(nodes, in_r, out_r, st, opt, box ? 1 : 0);
        if (run.empty()) continue;
        run.front().from_source = true;
        run.front().src_stage = box ? 0 : -1;
        Stage pl;
        pl.kind = ST_PLACE;
        pl.first = pl.last = (int)nodes.size() - 1;
        pl.clip = fin;
        pl.out = fin;
        pl.in_fmt = pl.out_fmt = nodes.back().out_fmt;
        pl.to_dst = true;
        pl.place_w = w;
        run.push_back(pl);
        all.insert(all.end(), run.begin(), run.end());
    }
    return all;
}

}  // namespace vgpu

int vgpu_pipeline::plan() {
    if (planned) return VGPU_OK;
    if (!affine_tail_supported(nodes)) return VGPU_E_UNSUPPORTED;
    const int n = (int)nodes.size();
    if (n == 0 || cur_w == 0 || cur_h == 0) {
        planned = true;
        return VGPU_OK;
    }
    // Backward region walk (execute.rs:413-480) from the full output (or
    // from each tile column, see plan_pipeline).
    Staging st;
    stages = plan_pipeline(nodes, cur_w, cur_h, &st, opt);
    host_blob = std::move(st.bytes);
    planned = true;
    return VGPU_OK;
}

int vgpu_pipeline::compile() {
    if (compiled) return VGPU_OK;
    int rc = plan();
    if (rc != VGPU_OK) return rc;
    if (!host_blob.empty()) {
        blob.ctx = ctx;
        blob.ptr = ctx->pool.acquire(host_blob.size());
        if (!blob.ptr) return VGPU_E_NO_MEMORY;
        hipError_t e = hipMemcpyAsync(blob.ptr, host_blob.data(), host_blob.size(), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return VGPU_E_DEVICE;
    }
    for (size_t k = 0; k + 1 < stages.size(); ++k) {
        const Stage& s = stages[k];
        if (stage_writes_dst(stages, k)) {
            inter.push_back(nullptr);
            continue;
        }
        const size_t bytes = (size_t)s.out.w * (size_t)s.out.h * bands * format_size(s.out_fmt);
        void* p = ctx->pool.acquire(bytes);
        if (!p) return VGPU_E_NO_MEMORY;
        inter.push_back(p);
    }
    compiled = true;
    return VGPU_OK;
}

// ------------------------------------------------- fixed scratch (no-alloc)

// The no-alloc run's plan tables and intermediates.  In host code `scratch`
// names a host-side stand-in, not device memory: only the address
// hipGetSymbolAddress returns for the current device may be passed to kernels
// or copies.
static __device__ __align__(256) unsigned char scratch[VGPU_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_by_device;  // per device id, looked up once

// This device's copy of `scratch`; the device must be current.
int scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if ((size_t)device >= g_scratch_by_device.size()) g_scratch_by_device.resize((size_t)device + 1, nullptr);
    if (!g_scratch_by_device[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return VGPU_E_DEVICE;
        }
        g_scratch_by_device[device] = ptr;
    }
    *out = g_scratch_by_device[device];
    return VGPU_OK;
}
}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* vgpu_status_string(int s) {
    switch (s) {
    case VGPU_OK: return "ok";
    case VGPU_E_INVALID_ARGUMENT: return "invalid argument";
    case VGPU_E_INVALID_KERNEL: return "invalid interpolation kernel for this op";
    case VGPU_E_UNSUPPORTED_FORMAT: return "unsupported band format";
    case VGPU_E_UNSUPPORTED: return "plan needs a node outside the GPU path";
    case VGPU_E_BUFFER_SIZE: return "buffer length does not match region";
    case VGPU_E_DEVICE: return "HIP device error";
    case VGPU_E_NO_MEMORY: return "out of device memory";
    case VGPU_E_CANCELLED: return "cancelled";
    case VGPU_E_ADMISSION: return "request exceeds the device memory budget";
    case VGPU_E_NOT_READY: return "request not finished";
    case VGPU_E_SCRATCH_LIMIT: return "plan tables and intermediates exceed the fixed scratch array";
    default: return "unknown status";
    }
}

int vgpu_context_create(int device, vgpu_context** out) {
    if (!out) return VGPU_E_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        (void)hipGetLastError();
        return VGPU_E_DEVICE;
    }
    if (hipSetDevice(device) != hipSuccess) return VGPU_E_DEVICE;
    void* scratch_ptr = nullptr;
    int rc = scratch_address(device, &scratch_ptr);
    if (rc != VGPU_OK) return rc;
    auto* c = new (std::nothrow) vgpu_context();
    if (!c) return VGPU_E_NO_MEMORY;
    c->device = device;
    c->scratch = scratch_ptr;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return VGPU_E_DEVICE;
    }
    c->stream = c->own;
    *out = c;
    return VGPU_OK;
}

void vgpu_context_destroy(vgpu_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->pool.trim(true);
    (void)hipStreamDestroy(c->own);
    delete c;
}

int vgpu_context_set_stream(vgpu_context* c, void* s) {
    if (!c) return VGPU_E_INVALID_ARGUMENT;
    c->stream = s ? static_cast<hipStream_t>(s) : c->own;
    return VGPU_OK;
}

void* vgpu_context_stream(vgpu_context* c) { return c ? (void*)c->stream : nullptr; }

int vgpu_synchronize(vgpu_context* c) {
    if (!c) return VGPU_E_INVALID_ARGUMENT;
    if (hipStreamSynchronize(c->stream) != hipSuccess) c->sticky = VGPU_E_DEVICE;
    return c->sticky;
}

size_t vgpu_context_pool_bytes(const vgpu_context* c) { return c ? c->pool.held() : 0; }

int vgpu_context_trim(vgpu_context* c) {
    if (!c) return VGPU_E_INVALID_ARGUMENT;
    c->pool.trim(false);
    return VGPU_OK;
}

int vgpu_buffer_alloc(vgpu_context* c, size_t bytes, void** out) {
    if (!c || !out) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    *out = c->pool.acquire(bytes);
    return *out ? VGPU_OK : VGPU_E_NO_MEMORY;
}

int vgpu_buffer_free(vgpu_context* c, void* ptr) {
    if (!c) return VGPU_E_INVALID_ARGUMENT;
    if (!ptr) return VGPU_OK;
    // Queued work may still read the buffer; the pool hands it out again only
    // after the stream has drained.
    if (hipStreamSynchronize(c->stream) != hipSuccess) c->sticky = VGPU_E_DEVICE;
    c->pool.release(ptr);
    return VGPU_OK;
}

int vgpu_memcpy_h2d(vgpu_context* c, void* dst, const void* src, size_t bytes) {
    if (!c || (!dst && bytes) || (!src && bytes)) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    if (bytes == 0) return VGPU_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return from_hip(e);
}

int vgpu_memcpy_d2h(vgpu_context* c, void* dst, const void* src, size_t bytes) {
    if (!c || (!dst && bytes) || (!src && bytes)) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    if (bytes == 0) return VGPU_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return from_hip(e);
}

int vgpu_fill_random_u8(vgpu_context* c, void* dst, size_t bytes, uint64_t seed) {
    if (!c || (!dst && bytes)) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    return from_hip(launch_fill_random(dst, bytes, seed, c->stream));
}

int vgpu_measure_stream(vgpu_context* c, size_t bytes, int iterations, float* read_ms, float* copy_ms) {
    if (!c || bytes < 16 || iterations < 1 || !read_ms || !copy_ms) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    bytes &= ~(size_t)15;
    constexpr int kBlocks = 256 * 16;  // a persistent grid of 16 blocks per CU
    unsigned char* a = static_cast<unsigned char*>(c->pool.acquire(bytes));
    unsigned char* b = static_cast<unsigned char*>(c->pool.acquire(bytes));
    uint32_t* sink = static_cast<uint32_t*>(c->pool.acquire(kBlocks * sizeof(uint32_t)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = VGPU_OK;
    if (!a || !b || !sink) rc = VGPU_E_NO_MEMORY;
    else if (!create_event_pair(&e0, &e1)) rc = VGPU_E_DEVICE;
    auto time = [&](auto launch, float* ms) {
        if (rc != VGPU_OK) return;
        hipError_t e = launch();  // warm
        e = e == hipSuccess ? hipEventRecord(e0, c->stream) : e;
        for (int i = 0; i < iterations && e == hipSuccess; ++i) e = launch();
        e = e == hipSuccess ? hipEventRecord(e1, c->stream) : e;
        e = e == hipSuccess ? hipEventSynchronize(e1) : e;
        e = e == hipSuccess ? hipEventElapsedTime(ms, e0, e1) : e;
        if (e != hipSuccess) rc = VGPU_E_DEVICE;
        else *ms /= (float)iterations;
    };
    if (rc == VGPU_OK) rc = from_hip(launch_fill_random(a, bytes, 1, c->stream));
    time([&]() { return launch_stream_read(a, bytes, sink, kBlocks, c->stream); }, read_ms);
    time([&]() { return hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, c->stream); }, copy_ms);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c->stream);
    c->pool.release(a);
    c->pool.release(b);
    c->pool.release(sink);
    if (rc == VGPU_E_DEVICE) c->sticky = rc;
    return rc;
}

int vgpu_pipeline_new(vgpu_context* ctx, uint32_t w, uint32_t h, uint32_t bands, int fmt,
                      vgpu_pipeline** out) {
    if (!ctx || !out) return VGPU_E_INVALID_ARGUMENT;
    *out = nullptr;
    if (bands == 0) return VGPU_E_INVALID_ARGUMENT;
    if (fmt < U8 || fmt > F64) return VGPU_E_UNSUPPORTED_FORMAT;
    if (w > INT32_MAX / 4 || h > INT32_MAX / 4) return VGPU_E_INVALID_ARGUMENT;
    auto* p = new (std::nothrow) vgpu_pipeline();
    if (!p) return VGPU_E_NO_MEMORY;
    p->ctx = ctx;
    p->src_w = p->cur_w = w;
    p->src_h = p->cur_h = h;
    p->bands = bands;
    p->src_fmt = p->cur_fmt = fmt;
    *out = p;
    return VGPU_OK;
}

void vgpu_pipeline_destroy(vgpu_pipeline* p) { delete p; }

int vgpu_pipeline_reduce_h(vgpu_pipeline* p, double factor, int kernel) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    ReduceAxis ax;
    int st = build_reduce_axis(kernel, factor, p->cur_w, &ax);
    if (st != OK) return st;
    Node& nd = p->push(N_REDUCE_H);
    nd.axis = std::move(ax);
    nd.out_w = reduce_output_len(factor, p->cur_w);
    p->commit();
    return VGPU_OK;
}

int vgpu_pipeline_reduce_v(vgpu_pipeline* p, double factor, int kernel) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    ReduceAxis ax;
    int st = build_reduce_axis(kernel, factor, p->cur_h, &ax);
    if (st != OK) return st;
    Node& nd = p->push(N_REDUCE_V);
    nd.axis = std::move(ax);
    nd.out_h = reduce_output_len(factor, p->cur_h);
    p->commit();
    return VGPU_OK;
}

// PipelineBuilder::extract_area (builder/geometry.rs:40-80): a non-empty
// rectangle inside the current image (validate_extract_area_bounds,
// builder/state.rs:238-270).
int vgpu_pipeline_extract_area(vgpu_pipeline* p, uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (w == 0 || h == 0 || (uint64_t)x + w > p->cur_w || (uint64_t)y + h > p->cur_h) return VGPU_E_INVALID_ARGUMENT;
    Node& nd = p->push(N_EXTRACT);
    nd.ex = (int32_t)x;
    nd.ey = (int32_t)y;
    nd.out_w = w;
    nd.out_h = h;
    p->commit();
    return VGPU_OK;
}

// PipelineBuilder::reduce (builder/resample.rs:349-382) -> ReduceOp::new
// (reduce.rs:51-67): both factors validated first, then the kernel, then each
// axis's tap limit; the filters are never bound to the input length.
int vgpu_pipeline_reduce(vgpu_pipeline* p, double hfactor, double vfactor, int kernel) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (!std::isfinite(hfactor) || !std::isfinite(vfactor) || hfactor < 1.0 || vfactor < 1.0)
        return VGPU_E_INVALID_ARGUMENT;
    ReduceAxis hax, vax;
    int st = build_reduce_axis(kernel, hfactor, 0, &hax);
    if (st == OK) st = build_reduce_axis(kernel, vfactor, 0, &vax);
    if (st != OK) return st;
    Node& nh = p->push(N_REDUCE_OP_H);
    nh.axis = std::move(hax);
    nh.out_w = reduce_output_len(hfactor, p->cur_w);
    p->commit();
    Node& nv = p->push(N_REDUCE_OP_V);
    nv.axis = std::move(vax);
    nv.out_h = reduce_output_len(vfactor, p->cur_h);
    p->commit();
    return VGPU_OK;
}

int vgpu_pipeline_shrink_h(vgpu_pipeline* p, uint32_t factor, int ceil_mode) {
    if (!p || factor == 0) return VGPU_E_INVALID_ARGUMENT;
    const uint32_t w = p->cur_w;
    Node& nd = p->push(N_SHRINK_H);
    nd.factor = factor;
    nd.source_width = w;  // the builder always passes Some(current width)
    nd.out_w = ceil_mode ? (w + factor - 1) / factor : w / factor;
    p->commit();
    return VGPU_OK;
}

int vgpu_pipeline_shrink_v(vgpu_pipeline* p, uint32_t factor, int ceil_mode) {
    if (!p || factor == 0) return VGPU_E_INVALID_ARGUMENT;
    const uint32_t h = p->cur_h;
    Node& nd = p->push(N_SHRINK_V);
    nd.factor = factor;
    nd.out_h = ceil_mode ? (h + factor - 1) / factor : h / factor;
    p->commit();
    return VGPU_OK;
}

int vgpu_pipeline_shrink(vgpu_pipeline* p, uint32_t hshrink, uint32_t vshrink) {
    if (!p || hshrink == 0 || vshrink == 0) return VGPU_E_INVALID_ARGUMENT;
    // Shrink<F> = ShrinkH (no full-row borrow) then ShrinkV inside one op
    // (shrink.rs:37-293); the two nodes fuse into one launch.
    int rc = vgpu_pipeline_shrink_h(p, hshrink, 0);
    if (rc != VGPU_OK) return rc;
    p->nodes.back().source_width = 0;
    return vgpu_pipeline_shrink_v(p, vshrink, 0);
}

int vgpu_pipeline_box_shrink(vgpu_pipeline* p, uint32_t factor, int ceil_mode) {
    if (!p || factor == 0 || factor > (uint32_t)INT32_MAX) return VGPU_E_INVALID_ARGUMENT;
    if (!p->nodes.empty()) return VGPU_E_INVALID_ARGUMENT;  // it replaces the source
    if (ceil_mode && p->cur_fmt != U8) return VGPU_E_UNSUPPORTED_FORMAT;
    const uint32_t w = p->cur_w, h = p->cur_h;
    Node& nd = p->push(N_BOX_SHRINK);
    nd.factor = factor;
    nd.box_ceil = ceil_mode ? 1 : 0;
    if (w == 0 || h == 0) {
        nd.out_w = nd.out_h = 0;
    } else if (ceil_mode) {  // div_ceil (decode_full.rs:343-344)
        nd.out_w = (uint32_t)(((uint64_t)w + factor - 1) / factor);
        nd.out_h = (uint32_t)(((uint64_t)h + factor - 1) / factor);
    } else {  // (src / f).max(1) (shrink.rs:58-59)
        nd.out_w = std::max(w / factor, 1u);
        nd.out_h = std::max(h / factor, 1u);
    }
    p->commit();
    return VGPU_OK;
}

int vgpu_pipeline_gauss_blur(vgpu_pipeline* p, float sigma) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (!(sigma == sigma)) return VGPU_E_INVALID_ARGUMENT;
    GaussMask m = build_gauss_mask(sigma);
    for (int pass = 0; pass < 2; ++pass) {
        Node& nd = p->push(pass == 0 ? N_GAUSS_H : N_GAUSS_V);
        nd.itaps = m.taps;
        nd.iscale = m.scale;
        nd.conv = m.weights;
        nd.out_fmt = p->cur_fmt == U8 ? U8 : F32;
        p->commit();
    }
    return VGPU_OK;
}

int vgpu_pipeline_convsep(vgpu_pipeline* p, const double* coeffs, uint32_t n, double scale,
                          double offset) {
    if (!p || !coeffs || n == 0 || n % 2 == 0) return VGPU_E_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i)
        if (!std::isfinite(coeffs[i])) return VGPU_E_INVALID_ARGUMENT;
    if (!std::isfinite(scale) || scale == 0.0 || !std::isfinite(offset)) return VGPU_E_INVALID_ARGUMENT;
    for (int pass = 0; pass < 2; ++pass) {
        Node& nd = p->push(pass == 0 ? N_CONVSEP_H : N_CONVSEP_V);
        nd.conv.assign(coeffs, coeffs + n);
        nd.itaps.assign(n, 0);
        nd.cscale = scale;
        nd.coffset = pass == 0 ? offset : 0.0;
        nd.out_fmt = F32;
        p->commit();
    }
    return VGPU_OK;
}

int vgpu_pipeline_premultiply(vgpu_pipeline* p) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (p->cur_fmt != U8) return VGPU_E_UNSUPPORTED_FORMAT;
    if (p->bands < 2) return VGPU_E_INVALID_ARGUMENT;
    p->push(N_PREMULT);
    p->commit();
    return VGPU_OK;
}

int vgpu_pipeline_unpremultiply(vgpu_pipeline* p) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (p->cur_fmt != U8) return VGPU_E_UNSUPPORTED_FORMAT;
    if (p->bands < 2) return VGPU_E_INVALID_ARGUMENT;
    p->push(N_UNPREMULT);
    p->commit();
    return VGPU_OK;
}

// ThumbnailNode::Affine / ResizeNode::Affine ->
// PipelineBuilder::affine_backward_with_extend_and_hint (builder/resample.rs:
// 478-560, 690-706, 932-955) -> AffineBridge::new_with_extend
// (builder/state.rs:298-341) -> Affine::try_new (affine/interpolation.rs:
// 13-41): finite coefficients, a non-singular matrix, ExtendMode::Copy over
// the whole input.  small: the SmallTile demand (resize, and a thumbnail of
// an input no larger than 512 x 512).
static int push_affine(vgpu_pipeline* p, const PlanStep& s, bool small) {
    const double m[4] = {s.a, 0.0, 0.0, s.d};
    for (double v : m)
        if (!std::isfinite(v)) return VGPU_E_INVALID_ARGUMENT;
    if (!std::isfinite(s.tx) || !std::isfinite(s.ty)) return VGPU_E_INVALID_ARGUMENT;
    if (std::fabs(std::fma(m[1], -m[2], m[0] * m[3])) <= DBL_EPSILON) return VGPU_E_INVALID_ARGUMENT;
    if (s.ew == 0 || s.eh == 0) return VGPU_E_INVALID_ARGUMENT;
    Node& nd = p->push(N_AFFINE);
    for (int k = 0; k < 4; ++k) nd.am[k] = m[k];
    nd.atx = s.tx;
    nd.aty = s.ty;
    nd.akern = s.akern;
    nd.asmall = small;
    nd.out_w = s.ew;
    nd.out_h = s.eh;
    p->commit();
    return VGPU_OK;
}

static int apply_steps(vgpu_pipeline* p, const std::vector<PlanStep>& steps, int kernel, bool affine_small = false) {
    for (const PlanStep& s : steps) {
        int rc = VGPU_OK;
        switch (s.kind) {
        case N_PREMULT: rc = vgpu_pipeline_premultiply(p); break;
        case N_UNPREMULT: rc = vgpu_pipeline_unpremultiply(p); break;
        case N_SHRINK_H: rc = vgpu_pipeline_shrink_h(p, (uint32_t)s.param, 0); break;
        case N_SHRINK_V: rc = vgpu_pipeline_shrink_v(p, (uint32_t)s.param, 0); break;
        case N_REDUCE_H: rc = vgpu_pipeline_reduce_h(p, s.param, kernel); break;
        case N_REDUCE_V: rc = vgpu_pipeline_reduce_v(p, s.param, kernel); break;
        case N_EXTRACT: rc = vgpu_pipeline_extract_area(p, s.ex, s.ey, s.ew, s.eh); break;
        case N_AFFINE: rc = push_affine(p, s, affine_small); break;
        case N_ZOOM: {  // ResizeNode::Zoom -> PipelineBuilder::zoom (builder/resample.rs:683-689)
            Node& nd = p->push(N_ZOOM);
            nd.zx = s.zx;
            nd.zy = s.zy;
            nd.out_w = p->cur_w * s.zx;
            nd.out_h = p->cur_h * s.zy;
            p->commit();
            break;
        }
        default: rc = VGPU_E_UNSUPPORTED; break;
        }
        if (rc != VGPU_OK) return rc;
    }
    if (steps.empty()) {
        p->push(N_COPY);
        p->commit();
    }
    return VGPU_OK;
}

int vgpu_thumbnail_shrink_on_load_factor(uint32_t width, uint32_t height, uint32_t bands, int target, uint32_t tw,
                                         uint32_t th, int crop, uint32_t* factor) {
    if (!factor) return VGPU_E_INVALID_ARGUMENT;
    std::vector<PlanStep> steps;
    uint32_t ow = 0, oh = 0;
    const int rc = plan_thumbnail(width, height, bands, target, tw, th, &steps, &ow, &oh, crop != 0, LANCZOS3, factor);
    return rc == OK ? VGPU_OK : rc;
}

int vgpu_pipeline_thumbnail(vgpu_pipeline* p, int target, uint32_t w, uint32_t h, int kernel) {
    return vgpu_pipeline_thumbnail_crop(p, target, w, h, kernel, 0);
}

int vgpu_pipeline_thumbnail_crop(vgpu_pipeline* p, int target, uint32_t w, uint32_t h, int kernel, int crop) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (kernel == LBB || kernel == VSQBS || kernel == NOHALO || kernel < NEAREST || kernel > CATMULLROM)
        return VGPU_E_INVALID_KERNEL;
    std::vector<PlanStep> steps;
    uint32_t ow = 0, oh = 0;
    int rc = plan_thumbnail(p->cur_w, p->cur_h, p->bands, target, w, h, &steps, &ow, &oh, crop != 0, kernel);
    if (rc != OK) return rc;
    if (p->cur_fmt != U8)
        for (const PlanStep& s : steps)
            if (s.kind == N_PREMULT) return VGPU_E_UNSUPPORTED_FORMAT;
    // builder/resample.rs:908-917,939-943: the Affine of a thumbnail of an
    // input up to 512 x 512 asks for SmallTile tiles.
    const bool small = p->cur_w <= 512 && p->cur_h <= 512;
    const size_t n0 = p->nodes.size();
    const uint32_t w0 = p->cur_w, h0 = p->cur_h;
    const int f0 = p->cur_fmt;
    rc = apply_steps(p, steps, kernel, small);
    if (rc == VGPU_OK && !affine_tail_supported(p->nodes)) rc = VGPU_E_UNSUPPORTED;
    if (rc != VGPU_OK) {  // leave the pipeline as it was
        p->release();
        p->nodes.resize(n0);
        p->cur_w = w0;
        p->cur_h = h0;
        p->cur_fmt = f0;
    }
    return rc;
}

int vgpu_pipeline_thumbnail_shrink_on_load(vgpu_pipeline* p, int target, uint32_t w, uint32_t h, int kernel,
                                           int crop) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (kernel == LBB || kernel == VSQBS || kernel == NOHALO || kernel < NEAREST || kernel > CATMULLROM)
        return VGPU_E_INVALID_KERNEL;
    uint32_t f = 0;
    if (p->nodes.empty()) {  // the hint is only taken by the source (builder/resample.rs:878-881)
        std::vector<PlanStep> steps;
        uint32_t ow = 0, oh = 0;
        const int rc = plan_thumbnail(p->cur_w, p->cur_h, p->bands, target, w, h, &steps, &ow, &oh, crop != 0,
                                      kernel, &f);
        if (rc != OK) return rc;
    }
    // normalize_shrink_factor (shrink.rs:119-125) keeps 2, 4, 8 and 16.
    const bool box = f == 2 || f == 4 || f == 8 || f == 16;
    if (box) {
        const int rc = vgpu_pipeline_box_shrink(p, f, 0);
        if (rc != VGPU_OK) return rc;
    }
    const int rc = vgpu_pipeline_thumbnail_crop(p, target, w, h, kernel, crop);
    if (rc != VGPU_OK && box) {  // leave the pipeline as it was
        p->release();
        p->nodes.clear();
        p->cur_w = p->src_w;
        p->cur_h = p->src_h;
        p->cur_fmt = p->src_fmt;
    }
    return rc;
}

int vgpu_pipeline_resize(vgpu_pipeline* p, double hscale, double vscale, int kernel) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    std::vector<PlanStep> steps;
    int eff = kernel;
    int rc = plan_resize(p->cur_w, p->cur_h, hscale, vscale, kernel, &steps, &eff);
    if (rc != OK) return rc;
    const size_t n0 = p->nodes.size();
    const uint32_t w0 = p->cur_w, h0 = p->cur_h;
    const int f0 = p->cur_fmt;
    // builder/resample.rs:690-706: resize's Affine always asks for SmallTile tiles.
    rc = apply_steps(p, steps, eff, true);
    if (rc == VGPU_OK && !affine_tail_supported(p->nodes)) rc = VGPU_E_UNSUPPORTED;
    if (rc != VGPU_OK) {  // leave the pipeline as it was
        p->release();
        p->nodes.resize(n0);
        p->cur_w = w0;
        p->cur_h = h0;
        p->cur_fmt = f0;
    }
    return rc;
}

int vgpu_pipeline_output_info(const vgpu_pipeline* p, uint32_t* w, uint32_t* h, uint32_t* b,
                              int* f) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (w) *w = p->cur_w;
    if (h) *h = p->cur_h;
    if (b) *b = p->bands;
    if (f) *f = p->cur_fmt;
    return VGPU_OK;
}

int vgpu_pipeline_set_integer_mfma(vgpu_pipeline* p, int enable) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(p->ctx->device);
    if (hipStreamSynchronize(p->ctx->stream) != hipSuccess) return p->ctx->sticky = VGPU_E_DEVICE;
    p->release();
    p->opt.integer_mfma = (enable & 1) != 0;
    p->opt.mfma_shrink = (enable & 2) != 0;
    return VGPU_OK;
}

int vgpu_pipeline_set_option(vgpu_pipeline* p, int option, int value) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(p->ctx->device);
    if (hipStreamSynchronize(p->ctx->stream) != hipSuccess) return p->ctx->sticky = VGPU_E_DEVICE;
    switch (option) {
    case VGPU_OPT_SEP_ROWS_PER_TASK:
        if (value < 0 || value % 16) return VGPU_E_INVALID_ARGUMENT;
        p->release();
        p->opt.sep_rows_per_task = value;
        return VGPU_OK;
    case VGPU_OPT_SHRINK_RUN:
        if (value < 0 || value > 64) return VGPU_E_INVALID_ARGUMENT;
        p->release();
        p->opt.shrink_run = value;
        return VGPU_OK;
    case VGPU_OPT_GAUSS_ROWS:
        if (value != -1 && value != 0 && value != 16 && value != 32) return VGPU_E_INVALID_ARGUMENT;
        p->release();
        p->opt.gauss_rows = value;
        return VGPU_OK;
    case VGPU_OPT_SEP_FORM:
        if (value != 0 && value != 1) return VGPU_E_INVALID_ARGUMENT;
        p->release();
        p->opt.sep_form = value;
        return VGPU_OK;
    case VGPU_OPT_REDUCE_FORM:
        if (value != 0 && value != 1) return VGPU_E_INVALID_ARGUMENT;
        p->release();
        p->opt.reduce_form = value;
        return VGPU_OK;
    default: return VGPU_E_INVALID_ARGUMENT;
    }
}

int vgpu_pipeline_set_float_mode(vgpu_pipeline* p, int mode) {
    if (!p || mode < 0 || mode > 2) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(p->ctx->device);
    if (hipStreamSynchronize(p->ctx->stream) != hipSuccess) return p->ctx->sticky = VGPU_E_DEVICE;
    p->release();
    p->opt.float_mode = mode;
    return VGPU_OK;
}

namespace {
const char* stage_kernel_name(const Stage& s) {
    switch (s.kind) {
    case ST_SEP: return "sep_mfma_i8";
    case ST_GAUSS_U8: return "gauss_u8";
    case ST_RED_F32: return "reduce_f32";
    case ST_SEP_F16: return "sep_mfma_f16x3";
    case ST_SHRINK_HV_U8: return "shrink_hv";
    case ST_SHRINK: return "shrink";
    case ST_REDUCE_H: return "reduce_h";
    case ST_REDUCE_V: return "reduce_v";
    case ST_REDUCE_HV: return "reduce_hv";
    case ST_CONV: return "conv";
    case ST_ALPHA: return "alpha";
    case ST_COPY: return "copy";
    case ST_AFFINE: return "affine";
    case ST_ZOOM: return "zoom";
    case ST_PLACE: return "place";
    case ST_BOX: return "box_shrink";
    case ST_CONV_PAIR: return s.cpair_kind == 2 ? "conv_pair_f32" : "conv_pair_f64";
    }
    return "unknown";
}
}  // namespace

int vgpu_pipeline_stage_kernel(vgpu_pipeline* p, int index, char* name, size_t cap) {
    if (!p || !name || cap == 0) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(p->ctx->device);
    const int rc = p->compile();
    if (rc != VGPU_OK) return rc;
    if (index < 0 || (size_t)index >= p->stages.size()) return VGPU_E_INVALID_ARGUMENT;
    std::snprintf(name, cap, "%s", stage_kernel_name(p->stages[(size_t)index]));
    return VGPU_OK;
}

int vgpu_pipeline_stage_count(vgpu_pipeline* p) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(p->ctx->device);
    int rc = p->compile();
    return rc != VGPU_OK ? rc : (int)p->stages.size();
}

struct vgpu_cancel_token {
    std::atomic<bool> cancelled{false};
};

int vgpu_cancel_token_new(vgpu_cancel_token** out) {
    if (!out) return VGPU_E_INVALID_ARGUMENT;
    *out = new (std::nothrow) vgpu_cancel_token();
    return *out ? VGPU_OK : VGPU_E_NO_MEMORY;
}
void vgpu_cancel_token_destroy(vgpu_cancel_token* t) { delete t; }
void vgpu_cancel_token_cancel(vgpu_cancel_token* t) {
    if (t) t->cancelled.store(true, std::memory_order_relaxed);
}
int vgpu_cancel_token_is_cancelled(const vgpu_cancel_token* t) {
    return t && t->cancelled.load(std::memory_order_relaxed) ? 1 : 0;
}

int vgpu_pipeline_run(vgpu_pipeline* p, const void* src, size_t src_bytes, void* dst,
                      size_t dst_bytes) {
    return vgpu_pipeline_run_cancellable(p, src, src_bytes, dst, dst_bytes, nullptr);
}

int vgpu_pipeline_run_cancellable(vgpu_pipeline* p, const void* src, size_t src_bytes, void* dst,
                                  size_t dst_bytes, const vgpu_cancel_token* token) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    if (vgpu_cancel_token_is_cancelled(token)) return VGPU_E_CANCELLED;
    vgpu_context* ctx = p->ctx;
    if (ctx->sticky != VGPU_OK) return ctx->sticky;
    const size_t in_bytes = (size_t)p->src_w * p->src_h * p->bands * format_size(p->src_fmt);
    const size_t out_bytes = (size_t)p->cur_w * p->cur_h * p->bands * format_size(p->cur_fmt);
    if (src_bytes != in_bytes || dst_bytes != out_bytes) return VGPU_E_BUFFER_SIZE;
    if (out_bytes == 0) return VGPU_OK;
    if (!src || !dst) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    int rc = p->compile();
    if (rc != VGPU_OK) return rc;
    if (p->stages.empty()) {
        return from_hip(hipMemcpyAsync(dst, src, out_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
    const unsigned char* blob = static_cast<const unsigned char*>(p->blob.ptr);
    for (size_t k = 0; k < p->stages.size(); ++k) {
        // The scheduler polls the token between tiles (runtime.rs:1237); the
        // units here are the fused stage launches (and tile-column runs).
        if (vgpu_cancel_token_is_cancelled(token)) return VGPU_E_CANCELLED;
        const Stage& s = p->stages[k];
        const int si = stage_input(p->stages, k);
        const InView in = si < 0 ? view_of(src, whole(p->src_w, p->src_h), s.clip)
                                 : view_of(p->inter[(size_t)si], p->stages[(size_t)si].out, s.clip);
        void* o = stage_writes_dst(p->stages, k) ? dst : p->inter[k];
        rc = launch_stage(p->nodes, s, in, o, blob, ctx->stream, p->opt);
        if (rc != VGPU_OK) {
            ctx->sticky = rc;
            return rc;
        }
    }
    return VGPU_OK;
}

int vgpu_pipeline_run_noalloc(vgpu_pipeline* p, const void* src, size_t src_bytes, void* dst,
                              size_t dst_bytes) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = p->ctx;
    if (ctx->sticky != VGPU_OK) return ctx->sticky;
    const size_t in_bytes = (size_t)p->src_w * p->src_h * p->bands * format_size(p->src_fmt);
    const size_t out_bytes = (size_t)p->cur_w * p->cur_h * p->bands * format_size(p->cur_fmt);
    if (src_bytes != in_bytes || dst_bytes != out_bytes) return VGPU_E_BUFFER_SIZE;
    if (out_bytes == 0) return VGPU_OK;
    if (!src || !dst) return VGPU_E_INVALID_ARGUMENT;
    if (!ctx->scratch) return VGPU_E_DEVICE;
    (void)hipSetDevice(ctx->device);
    int rc = p->plan();  // host memory only
    if (rc != VGPU_OK) return rc;
    if (p->stages.empty()) {
        return from_hip(hipMemcpyAsync(dst, src, out_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
    // Layout in the scratch array: plan tables, then each intermediate.
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t used = up(p->host_blob.size());
    std::vector<size_t> off(p->stages.size(), 0);
    for (size_t k = 0; k < p->stages.size(); ++k) {
        if (stage_writes_dst(p->stages, k)) continue;
        const Stage& s = p->stages[k];
        off[k] = used;
        used += up((size_t)s.out.w * (size_t)s.out.h * p->bands * format_size(s.out_fmt));
        if (used > VGPU_NOALLOC_SCRATCH_BYTES) return VGPU_E_SCRATCH_LIMIT;
    }
    if (used > VGPU_NOALLOC_SCRATCH_BYTES) return VGPU_E_SCRATCH_LIMIT;
    unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
    if (!p->host_blob.empty()) {
        rc = from_hip(hipMemcpyAsync(base, p->host_blob.data(), p->host_blob.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
        if (rc != VGPU_OK) return ctx->sticky = rc;
    }
    for (size_t k = 0; k < p->stages.size(); ++k) {
        const Stage& s = p->stages[k];
        const int si = stage_input(p->stages, k);
        const InView in = si < 0 ? view_of(src, whole(p->src_w, p->src_h), s.clip)
                                 : view_of(base + off[(size_t)si], p->stages[(size_t)si].out, s.clip);
        void* o = stage_writes_dst(p->stages, k) ? dst : static_cast<void*>(base + off[k]);
        rc = launch_stage(p->nodes, s, in, o, base, ctx->stream, p->opt);
        if (rc != VGPU_OK) return ctx->sticky = rc;
    }
    return VGPU_OK;
}

int vgpu_pipeline_run_profiled(vgpu_pipeline* p, const void* src, size_t src_bytes, void* dst,
                               size_t dst_bytes, int iterations, float* stage_ms, int max_stages) {
    if (!p || iterations < 1 || !stage_ms) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    int rc = p->compile();
    if (rc != VGPU_OK) return rc;
    const int ns = (int)p->stages.size();
    if (ns == 0) return VGPU_E_UNSUPPORTED;
    if (max_stages < ns) return VGPU_E_BUFFER_SIZE;
    const size_t in_bytes = (size_t)p->src_w * p->src_h * p->bands * format_size(p->src_fmt);
    const size_t out_bytes = (size_t)p->cur_w * p->cur_h * p->bands * format_size(p->cur_fmt);
    if (src_bytes != in_bytes || dst_bytes != out_bytes) return VGPU_E_BUFFER_SIZE;
    std::vector<hipEvent_t> ev((size_t)ns + 1);
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return VGPU_E_DEVICE;
    std::vector<double> acc((size_t)ns, 0.0);
    const unsigned char* blob = static_cast<const unsigned char*>(p->blob.ptr);
    for (int it = 0; it < iterations && rc == VGPU_OK; ++it) {
        (void)hipEventRecord(ev[0], ctx->stream);
        for (int k = 0; k < ns && rc == VGPU_OK; ++k) {
            const Stage& s = p->stages[(size_t)k];
            const int si = stage_input(p->stages, (size_t)k);
            const InView in = si < 0 ? view_of(src, whole(p->src_w, p->src_h), s.clip)
                                     : view_of(p->inter[(size_t)si], p->stages[(size_t)si].out, s.clip);
            void* o = stage_writes_dst(p->stages, (size_t)k) ? dst : p->inter[(size_t)k];
            rc = launch_stage(p->nodes, s, in, o, blob, ctx->stream, p->opt);
            (void)hipEventRecord(ev[(size_t)k + 1], ctx->stream);
        }
        if (hipEventSynchronize(ev[(size_t)ns]) != hipSuccess) rc = VGPU_E_DEVICE;
        for (int k = 0; k < ns && rc == VGPU_OK; ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[(size_t)k], ev[(size_t)k + 1]);
            acc[(size_t)k] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < ns; ++k) stage_ms[k] = (float)(acc[(size_t)k] / iterations);
    return rc == VGPU_OK ? ns : rc;
}

int vgpu_pipeline_time_stage(vgpu_pipeline* p, const void* src, size_t src_bytes, void* dst, size_t dst_bytes,
                             int stage, int iterations, float* mean_ms) {
    return vgpu_pipeline_time_stage_sources(p, &src, 1, src_bytes, dst, dst_bytes, stage, iterations, mean_ms);
}

int vgpu_pipeline_time_stage_sources(vgpu_pipeline* p, const void* const* srcs, int nsrc, size_t src_bytes,
                                     void* dst, size_t dst_bytes, int stage, int iterations, float* mean_ms) {
    if (!p || !srcs || nsrc < 1 || iterations < 1 || !mean_ms) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    int rc = vgpu_pipeline_run(p, srcs[0], src_bytes, dst, dst_bytes);  // compiles, fills the intermediates
    if (rc != VGPU_OK) return rc;
    const int ns = (int)p->stages.size();
    if (stage < 0 || stage >= ns) return VGPU_E_INVALID_ARGUMENT;
    for (int i = 0; i < nsrc; ++i)
        if (!srcs[i]) return VGPU_E_INVALID_ARGUMENT;
    // Input view of the stage: the source (cycled over srcs) or the output
    // of the stage it reads.
    const int si = stage_input(p->stages, (size_t)stage);
    const Stage& s = p->stages[(size_t)stage];
    void* o = stage_writes_dst(p->stages, (size_t)stage) ? dst : p->inter[(size_t)stage];
    const unsigned char* blob = static_cast<const unsigned char*>(p->blob.ptr);
    auto view = [&](int i) {
        if (si < 0) return view_of(srcs[i % nsrc], whole(p->src_w, p->src_h), s.clip);
        return view_of(p->inter[(size_t)si], p->stages[(size_t)si].out, s.clip);
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!create_event_pair(&e0, &e1)) return VGPU_E_DEVICE;
    for (int i = 0; i < 3 && rc == VGPU_OK; ++i) rc = launch_stage(p->nodes, s, view(i), o, blob, ctx->stream, p->opt);
    (void)hipEventRecord(e0, ctx->stream);
    for (int i = 0; i < iterations && rc == VGPU_OK; ++i)
        rc = launch_stage(p->nodes, s, view(i), o, blob, ctx->stream, p->opt);
    (void)hipEventRecord(e1, ctx->stream);
    float ms = 0.f;
    if (rc == VGPU_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = VGPU_E_DEVICE;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *mean_ms = ms / (float)iterations;
    return rc;
}

int vgpu_pipeline_benchmark(vgpu_pipeline* p, const void* src, size_t src_bytes, void* dst,
                            size_t dst_bytes, int warmup, int iterations, float* total_ms,
                            float* step_ms) {
    return vgpu_pipeline_benchmark_sources(p, &src, 1, src_bytes, dst, dst_bytes, warmup, iterations, total_ms,
                                           step_ms);
}

int vgpu_pipeline_benchmark_sources(vgpu_pipeline* p, const void* const* srcs, int nsrc, size_t src_bytes,
                                    void* dst, size_t dst_bytes, int warmup, int iterations, float* total_ms,
                                    float* step_ms) {
    if (!p || !srcs || nsrc < 1 || iterations < 1 || !total_ms) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = p->ctx;
    int rc = VGPU_OK;
    auto step = [&](int i) { return vgpu_pipeline_run(p, srcs[i % nsrc], src_bytes, dst, dst_bytes); };
    for (int i = 0; i < warmup && rc == VGPU_OK; ++i) rc = step(i);
    if (rc != VGPU_OK) return rc;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return ctx->sticky = VGPU_E_DEVICE;
    hipEvent_t a, b;
    if (!create_event_pair(&a, &b)) return VGPU_E_DEVICE;
    // Back-to-back steps: whole-job throughput.
    (void)hipEventRecord(a, ctx->stream);
    for (int i = 0; i < iterations && rc == VGPU_OK; ++i) rc = step(i);
    (void)hipEventRecord(b, ctx->stream);
    if (rc == VGPU_OK && hipEventSynchronize(b) != hipSuccess) rc = VGPU_E_DEVICE;
    if (rc == VGPU_OK) (void)hipEventElapsedTime(total_ms, a, b);
    // Per-step latency: one event pair per step, steps still back to back.
    if (rc == VGPU_OK && step_ms) {
        std::vector<hipEvent_t> ev((size_t)iterations + 1);
        for (auto& e : ev) (void)hipEventCreate(&e);
        (void)hipEventRecord(ev[0], ctx->stream);
        for (int i = 0; i < iterations && rc == VGPU_OK; ++i) {
            rc = step(i);
            (void)hipEventRecord(ev[(size_t)i + 1], ctx->stream);
        }
        if (rc == VGPU_OK && hipEventSynchronize(ev[(size_t)iterations]) != hipSuccess) rc = VGPU_E_DEVICE;
        for (int i = 0; i < iterations && rc == VGPU_OK; ++i)
            (void)hipEventElapsedTime(&step_ms[i], ev[(size_t)i], ev[(size_t)i + 1]);
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

int vgpu_pipeline_run_host(vgpu_pipeline* p, const void* src, size_t src_bytes, void* dst,
                           size_t dst_bytes) {
    if (!p) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    void* dsrc = ctx->pool.acquire(src_bytes);
    void* ddst = ctx->pool.acquire(dst_bytes);
    if (!dsrc || !ddst) {
        ctx->pool.release(dsrc);
        ctx->pool.release(ddst);
        return VGPU_E_NO_MEMORY;
    }
    int rc = from_hip(hipMemcpyAsync(dsrc, src, src_bytes, hipMemcpyHostToDevice, ctx->stream));
    if (rc == VGPU_OK) rc = vgpu_pipeline_run(p, dsrc, src_bytes, ddst, dst_bytes);
    if (rc == VGPU_OK) rc = from_hip(hipMemcpyAsync(dst, ddst, dst_bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (rc == VGPU_OK) rc = vgpu_synchronize(ctx);
    else (void)hipStreamSynchronize(ctx->stream);
    ctx->pool.release(dsrc);
    ctx->pool.release(ddst);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------- op level

// A device buffer that queued work may still read: returned to the pool once
// the event recorded after that work has completed.
struct Retired {
    void* ptr = nullptr;
    hipEvent_t done = nullptr;
    std::vector<unsigned char> host;  // host source of a copy still in flight
};

struct vgpu_op {
    vgpu_context* ctx = nullptr;
    Node node;
    Node vnode;          // ReduceOp: the vertical pass (node holds the horizontal one)
    bool pair = false;   // ReduceOp
    // Serialises calls on one op: the blob, the scratch and the staging below
    // are per-op state (the reference gives each rayon worker its own State;
    // here the state is device memory shared by the op's calls).
    std::mutex mu;
    DeviceBlob blob;
    std::vector<unsigned char> blob_host;  // bytes last uploaded into blob
    Rect blob_key;  // output region the uploaded plan was built for (region-dependent plans)
    bool have_blob = false;
    std::vector<Retired> retired;
    DeviceBlob scratch;  // ReduceOp's intermediate when its passes run as two kernels
    size_t scratch_bytes = 0;
    ~vgpu_op() {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (Retired& r : retired) {
            ctx->pool.release(r.ptr);
            if (r.done) (void)hipEventDestroy(r.done);
        }
    }
};

// Return retired buffers whose readers have finished to the pool.
static void sweep_retired(vgpu_op* op) {
    auto& v = op->retired;
    for (size_t i = 0; i < v.size();) {
        if (!v[i].done || hipEventQuery(v[i].done) == hipSuccess) {
            op->ctx->pool.release(v[i].ptr);
            if (v[i].done) (void)hipEventDestroy(v[i].done);
            v[i] = std::move(v.back());
            v.pop_back();
        } else {
            ++i;
        }
    }
    (void)hipGetLastError();
}

// Plans of ReduceH / ReduceV / ReduceOp depend on the output region (tile
// origins decide the accumulated u8 positions and the tap windows); the
// others (Gaussian, ConvSep, shrink, alpha) do not.
static bool plan_depends_on_region(const vgpu_op* op) {
    const NodeKind k = op->node.kind;
    return op->pair || k == N_REDUCE_H || k == N_REDUCE_V;
}

static int make_op(vgpu_context* ctx, Node nd, vgpu_op** out) {
    if (!ctx || !out) return VGPU_E_INVALID_ARGUMENT;
    auto* op = new (std::nothrow) vgpu_op();
    if (!op) return VGPU_E_NO_MEMORY;
    op->ctx = ctx;
    op->node = std::move(nd);
    op->blob.ctx = ctx;
    *out = op;
    return VGPU_OK;
}

// Uploads an op's plan tables unless the blob already holds them (the same
// bytes; for region-dependent plans the same output region).  The new blob
// replaces the old one only after its allocation and copy have been
// enqueued; the old one is retired until the work already queued on the
// stream has finished with it, so no call drains the stream.
static int upload_op_blob(vgpu_op* op, Staging& st, const Rect& out_r) {
    if (st.bytes.empty()) return VGPU_OK;
    vgpu_context* ctx = op->ctx;
    sweep_retired(op);
    if (op->have_blob) {
        const bool same_region = op->blob_key.x == out_r.x && op->blob_key.y == out_r.y &&
                                 op->blob_key.w == out_r.w && op->blob_key.h == out_r.h;
        if ((!plan_depends_on_region(op) || same_region) && op->blob_host == st.bytes) return VGPU_OK;
    }
    void* fresh = ctx->pool.acquire(st.bytes.size());
    if (!fresh) return VGPU_E_NO_MEMORY;
    std::vector<unsigned char> host = std::move(st.bytes);
    if (hipMemcpyAsync(fresh, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        ctx->pool.release(fresh);
        return ctx->sticky = VGPU_E_DEVICE;
    }
    if (op->blob.ptr) {
        Retired r;
        r.ptr = op->blob.ptr;
        r.host = std::move(op->blob_host);
        if (hipEventCreateWithFlags(&r.done, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(r.done, ctx->stream) != hipSuccess) {
            (void)hipGetLastError();
            if (r.done) (void)hipEventDestroy(r.done);
            r.done = nullptr;
            if (hipStreamSynchronize(ctx->stream) != hipSuccess) return ctx->sticky = VGPU_E_DEVICE;
        }
        op->retired.push_back(std::move(r));
        op->blob.ptr = nullptr;
    }
    op->blob.ptr = fresh;
    op->blob_host = std::move(host);
    op->blob_key = out_r;
    op->have_blob = true;
    return VGPU_OK;
}

// ReduceOp::process_region (reduce.rs:224-250): the H pass fills the scratch
// region required_input_region_v(out), clamping columns to the input tile;
// its rows are the scratch rows themselves, so the input tile must carry
// exactly those rows.  One fused stage for <= 4 bands, else two kernels
// through a device scratch owned by the op.
static int process_reduce_op(vgpu_op* op, const void* input, void* output, const Rect& in_r, const Rect& out_r) {
    vgpu_context* ctx = op->ctx;
    const Node& nh = op->node;
    const Node& nv = op->vnode;
    const Rect mid = nv.required_input(out_r);
    if (in_r.y != mid.y || in_r.h != mid.h) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    std::vector<Node> nodes{nh, nv};
    Staging st;
    Stage s;
    s.first = 0;
    s.clip = in_r;
    s.in_fmt = s.out_fmt = nh.in_fmt;
    const InView in = view_of(input, in_r, in_r);
    if (nh.bands <= 4 && stage_reduce_hv_params(nh.axis, false, nv.axis, nh.in_fmt, nh.bands, mid, out_r, mid, in_r,
                                                in_r, StageOptions(), &st, &s, true)) {
        s.last = 1;
        s.out = out_r;
        int rc = upload_op_blob(op, st, out_r);
        if (rc == VGPU_OK)
            rc = launch_stage(nodes, s, in, output, static_cast<const unsigned char*>(op->blob.ptr), ctx->stream);
        if (rc != VGPU_OK) ctx->sticky = rc;
        return rc;
    }
    Stage sv = s;
    s.last = 0;
    s.kind = ST_REDUCE_H;
    s.reduce_op = 1;
    s.out = mid;
    stage_reduce_params(nh, mid, &st, &s);
    sv.first = sv.last = 1;
    sv.kind = ST_REDUCE_V;
    sv.reduce_op = 1;
    sv.clip = mid;
    sv.out = out_r;
    stage_reduce_params(nv, out_r, &st, &sv);
    int rc = upload_op_blob(op, st, out_r);
    if (rc != VGPU_OK) return rc;
    const size_t bytes = (size_t)mid.w * mid.h * nh.bands * format_size(nh.in_fmt);
    if (bytes > op->scratch_bytes) {
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return ctx->sticky = VGPU_E_DEVICE;
        op->scratch.reset();
        op->scratch.ctx = ctx;
        op->scratch.ptr = ctx->pool.acquire(bytes);
        op->scratch_bytes = op->scratch.ptr ? bytes : 0;
        if (!op->scratch.ptr) return VGPU_E_NO_MEMORY;
    }
    const unsigned char* blob = static_cast<const unsigned char*>(op->blob.ptr);
    rc = launch_stage(nodes, s, in, op->scratch.ptr, blob, ctx->stream);
    if (rc == VGPU_OK) rc = launch_stage(nodes, sv, view_of(op->scratch.ptr, mid, mid), output, blob, ctx->stream);
    if (rc != VGPU_OK) ctx->sticky = rc;
    return rc;
}

static Node base_node(NodeKind k, uint32_t bands, int fmt) {
    Node nd;
    nd.kind = k;
    nd.bands = bands;
    nd.in_fmt = nd.out_fmt = fmt;
    return nd;
}

static bool valid_fmt(int f) { return f >= U8 && f <= F64; }

extern "C" {

int vgpu_op_reduce_h_new(vgpu_context* ctx, double factor, int kernel, uint32_t bands,
                         uint32_t input_w, int fmt, vgpu_op** out) {
    if (bands == 0) return VGPU_E_INVALID_ARGUMENT;
    if (!valid_fmt(fmt)) return VGPU_E_UNSUPPORTED_FORMAT;
    Node nd = base_node(N_REDUCE_H, bands, fmt);
    int st = build_reduce_axis(kernel, factor, input_w, &nd.axis);
    if (st != OK) return st;
    return make_op(ctx, std::move(nd), out);
}

int vgpu_op_reduce_v_new(vgpu_context* ctx, double factor, int kernel, uint32_t bands,
                         uint32_t input_h, int fmt, vgpu_op** out) {
    if (bands == 0) return VGPU_E_INVALID_ARGUMENT;
    if (!valid_fmt(fmt)) return VGPU_E_UNSUPPORTED_FORMAT;
    Node nd = base_node(N_REDUCE_V, bands, fmt);
    int st = build_reduce_axis(kernel, factor, input_h, &nd.axis);
    if (st != OK) return st;
    return make_op(ctx, std::move(nd), out);
}

int vgpu_op_reduce_new(vgpu_context* ctx, double hfactor, double vfactor, int kernel, uint32_t bands, int fmt,
                       vgpu_op** out) {
    if (bands == 0) return VGPU_E_INVALID_ARGUMENT;
    if (!valid_fmt(fmt)) return VGPU_E_UNSUPPORTED_FORMAT;
    if (!std::isfinite(hfactor) || !std::isfinite(vfactor) || hfactor < 1.0 || vfactor < 1.0)
        return VGPU_E_INVALID_ARGUMENT;
    Node nh = base_node(N_REDUCE_OP_H, bands, fmt), nv = base_node(N_REDUCE_OP_V, bands, fmt);
    int st = build_reduce_axis(kernel, hfactor, 0, &nh.axis);
    if (st == OK) st = build_reduce_axis(kernel, vfactor, 0, &nv.axis);
    if (st != OK) return st;
    st = make_op(ctx, std::move(nh), out);
    if (st != VGPU_OK) return st;
    (*out)->vnode = std::move(nv);
    (*out)->pair = true;
    return VGPU_OK;
}

int vgpu_op_shrink_h_new(vgpu_context* ctx, uint32_t factor, int ceil_mode, uint32_t bands,
                         uint32_t source_width, int fmt, vgpu_op** out) {
    (void)ceil_mode;
    if (factor == 0 || bands == 0) return VGPU_E_INVALID_ARGUMENT;
    if (!valid_fmt(fmt)) return VGPU_E_UNSUPPORTED_FORMAT;
    Node nd = base_node(N_SHRINK_H, bands, fmt);
    nd.factor = factor;
    nd.source_width = source_width;
    return make_op(ctx, std::move(nd), out);
}

int vgpu_op_shrink_v_new(vgpu_context* ctx, uint32_t factor, int ceil_mode, uint32_t bands,
                         int fmt, vgpu_op** out) {
    (void)ceil_mode;
    if (factor == 0 || bands == 0) return VGPU_E_INVALID_ARGUMENT;
    if (!valid_fmt(fmt)) return VGPU_E_UNSUPPORTED_FORMAT;
    Node nd = base_node(N_SHRINK_V, bands, fmt);
    nd.factor = factor;
    return make_op(ctx, std::move(nd), out);
}

static int gauss_op(vgpu_context* ctx, NodeKind k, float sigma, uint32_t bands, int fmt, vgpu_op** out) {
    if (bands == 0 || !(sigma == sigma)) return VGPU_E_INVALID_ARGUMENT;
    if (!valid_fmt(fmt)) return VGPU_E_UNSUPPORTED_FORMAT;
    Node nd = base_node(k, bands, fmt);
    GaussMask m = build_gauss_mask(sigma);
    nd.itaps = m.taps;
    nd.iscale = m.scale;
    nd.conv = m.weights;
    nd.out_fmt = fmt == U8 ? U8 : F32;
    return make_op(ctx, std::move(nd), out);
}

int vgpu_op_gauss_blur_h_new(vgpu_context* ctx, float sigma, uint32_t bands, int fmt, vgpu_op** out) {
    return gauss_op(ctx, N_GAUSS_H, sigma, bands, fmt, out);
}
int vgpu_op_gauss_blur_v_new(vgpu_context* ctx, float sigma, uint32_t bands, int fmt, vgpu_op** out) {
    return gauss_op(ctx, N_GAUSS_V, sigma, bands, fmt, out);
}

// ConvSepH::with_mask / ConvSepV::with_mask (convolution/convsep.rs:23-52,
// 120-149): an odd-length finite mask with its scale and offset; any input
// format, F32 output.
static int convsep_op(vgpu_context* ctx, NodeKind k, const double* coeffs, uint32_t n, double scale, double offset,
                      uint32_t bands, int fmt, vgpu_op** out) {
    if (bands == 0 || !coeffs || n == 0 || n % 2 == 0) return VGPU_E_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i)
        if (!std::isfinite(coeffs[i])) return VGPU_E_INVALID_ARGUMENT;
    if (!std::isfinite(scale) || scale == 0.0 || !std::isfinite(offset)) return VGPU_E_INVALID_ARGUMENT;
    if (!valid_fmt(fmt)) return VGPU_E_UNSUPPORTED_FORMAT;
    Node nd = base_node(k, bands, fmt);
    nd.conv.assign(coeffs, coeffs + n);
    nd.itaps.assign(n, 0);
    nd.cscale = scale;
    nd.coffset = offset;
    nd.out_fmt = F32;
    return make_op(ctx, std::move(nd), out);
}

int vgpu_op_convsep_h_new(vgpu_context* ctx, const double* coeffs, uint32_t n, double scale, double offset,
                          uint32_t bands, int fmt, vgpu_op** out) {
    return convsep_op(ctx, N_CONVSEP_H, coeffs, n, scale, offset, bands, fmt, out);
}
int vgpu_op_convsep_v_new(vgpu_context* ctx, const double* coeffs, uint32_t n, double scale, double offset,
                          uint32_t bands, int fmt, vgpu_op** out) {
    return convsep_op(ctx, N_CONVSEP_V, coeffs, n, scale, offset, bands, fmt, out);
}

int vgpu_op_premultiply_new(vgpu_context* ctx, uint32_t bands, int fmt, vgpu_op** out) {
    if (bands < 2) return VGPU_E_INVALID_ARGUMENT;
    if (fmt != U8) return VGPU_E_UNSUPPORTED_FORMAT;
    return make_op(ctx, base_node(N_PREMULT, bands, fmt), out);
}
int vgpu_op_unpremultiply_new(vgpu_context* ctx, uint32_t bands, int fmt, vgpu_op** out) {
    if (bands < 2) return VGPU_E_INVALID_ARGUMENT;
    if (fmt != U8) return VGPU_E_UNSUPPORTED_FORMAT;
    return make_op(ctx, base_node(N_UNPREMULT, bands, fmt), out);
}

void vgpu_op_destroy(vgpu_op* op) { delete op; }

static Rect rect_of(vgpu_region r) { return Rect{r.x, r.y, (int64_t)r.width, (int64_t)r.height}; }

int vgpu_op_required_input_region(const vgpu_op* op, vgpu_region o, vgpu_region* in) {
    if (!op || !in) return VGPU_E_INVALID_ARGUMENT;
    // ReduceOp::required_input_region = h(v(out)) (reduce.rs:181-184).
    Rect r = op->node.required_input(op->pair ? op->vnode.required_input(rect_of(o)) : rect_of(o));
    if (r.w < 0 || r.h < 0 || r.w > UINT32_MAX || r.h > UINT32_MAX) return VGPU_E_INVALID_ARGUMENT;
    in->x = (int32_t)r.x;
    in->y = (int32_t)r.y;
    in->width = (uint32_t)r.w;
    in->height = (uint32_t)r.h;
    return VGPU_OK;
}

int vgpu_op_output_size(const vgpu_op* op, uint32_t iw, uint32_t ih, uint32_t* ow, uint32_t* oh) {
    if (!op || !ow || !oh) return VGPU_E_INVALID_ARGUMENT;
    const Node& nd = op->node;
    *ow = iw;
    *oh = ih;
    switch (nd.kind) {
    case N_REDUCE_H: *ow = reduce_output_len(nd.axis.factor, iw); break;
    case N_REDUCE_V: *oh = reduce_output_len(nd.axis.factor, ih); break;
    case N_REDUCE_OP_H:  // reduce.rs:257-263
        *ow = reduce_output_len(nd.axis.factor, iw);
        *oh = reduce_output_len(op->vnode.axis.factor, ih);
        break;
    case N_SHRINK_H: *ow = iw / nd.factor; break;
    case N_SHRINK_V: *oh = ih / nd.factor; break;
    default: break;
    }
    return VGPU_OK;
}

int vgpu_op_output_format(const vgpu_op* op, int* f) {
    if (!op || !f) return VGPU_E_INVALID_ARGUMENT;
    *f = op->node.out_fmt;
    return VGPU_OK;
}

}  // extern "C"

// Checks the byte lengths against the regions (OperationBridge casts the
// slices and writes nothing on a mismatch, bridge.rs:208-236; here that is
// VGPU_E_BUFFER_SIZE).
static int check_region_lengths(const vgpu_op* op, size_t in_len, size_t out_len, const Rect& in_r,
                                const Rect& out_r) {
    const Node& nd = op->node;
    if (in_len != (size_t)in_r.w * in_r.h * nd.bands * format_size(nd.in_fmt) ||
        out_len != (size_t)out_r.w * out_r.h * nd.bands * format_size(nd.out_fmt))
        return VGPU_E_BUFFER_SIZE;
    return VGPU_OK;
}

// Whether kernels on this context's device may dereference `p`: device or
// managed memory (host memory, pinned or not, is rejected: the op-level
// device entry takes device buffers, the host entry takes host buffers).
static bool device_pointer(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

static int process_region_locked(vgpu_op* op, const void* input, void* output, const Rect& in_r,
                                 const Rect& out_r);

extern "C" {

int vgpu_op_process_region(vgpu_op* op, const void* input, size_t in_len, void* output,
                           size_t out_len, vgpu_region ir, vgpu_region orr) {
    if (!op) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = op->ctx;
    if (ctx->sticky != VGPU_OK) return ctx->sticky;
    const Rect in_r = rect_of(ir), out_r = rect_of(orr);
    int rc = check_region_lengths(op, in_len, out_len, in_r, out_r);
    if (rc != VGPU_OK) return rc;
    if (out_r.empty()) return VGPU_OK;
    if (!input || !output) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (!device_pointer(input) || !device_pointer(output)) return VGPU_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(op->mu);
    return process_region_locked(op, input, output, in_r, out_r);
}

int vgpu_op_process_region_host(vgpu_op* op, const void* input, size_t in_len, void* output,
                                size_t out_len, vgpu_region ir, vgpu_region orr) {
    if (!op) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = op->ctx;
    if (ctx->sticky != VGPU_OK) return ctx->sticky;
    const Rect in_r = rect_of(ir), out_r = rect_of(orr);
    int rc = check_region_lengths(op, in_len, out_len, in_r, out_r);
    if (rc != VGPU_OK) return rc;
    if (out_r.empty()) return VGPU_OK;
    if (!input || !output) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    std::lock_guard<std::mutex> g(op->mu);
    // Staging through pooled device buffers: H2D, the device op, D2H, then
    // wait, so the call returns with `output` written, as dyn_process_region.
    void* d_in = ctx->pool.acquire(in_len);
    void* d_out = ctx->pool.acquire(out_len);
    if (!d_in || !d_out) {
        ctx->pool.release(d_in);
        ctx->pool.release(d_out);
        return VGPU_E_NO_MEMORY;
    }
    rc = from_hip(hipMemcpyAsync(d_in, input, in_len, hipMemcpyHostToDevice, ctx->stream));
    if (rc == VGPU_OK) rc = process_region_locked(op, d_in, d_out, in_r, out_r);
    if (rc == VGPU_OK) rc = from_hip(hipMemcpyAsync(output, d_out, out_len, hipMemcpyDeviceToHost, ctx->stream));
    const int sync = from_hip(hipStreamSynchronize(ctx->stream));
    if (rc == VGPU_OK) rc = sync;
    if (sync == VGPU_OK) {
        ctx->pool.release(d_in);
        ctx->pool.release(d_out);
    }  // else the buffers may still be in use by a faulted stream: leak them
    if (rc == VGPU_E_DEVICE) ctx->sticky = rc;
    return rc;
}

}  // extern "C"

static int process_region_locked(vgpu_op* op, const void* input, void* output, const Rect& in_r,
                                 const Rect& out_r) {
    vgpu_context* ctx = op->ctx;
    const Node& nd = op->node;
    if (op->pair) return process_reduce_op(op, input, output, in_r, out_r);
    if (nd.kind != N_REDUCE_H && nd.kind != N_REDUCE_V) {
        // These ops index their input relative to the tile origin the
        // scheduler derived from required_input_region; insist on it.
        const Rect need = nd.required_input(out_r);
        if (need.x != in_r.x || need.y != in_r.y || need.w != in_r.w || need.h != in_r.h)
            return VGPU_E_INVALID_ARGUMENT;
    }
    Stage s;
    s.first = s.last = 0;
    s.clip = in_r;
    s.out = out_r;
    s.in_fmt = nd.in_fmt;
    s.out_fmt = nd.out_fmt;
    Staging st;
    switch (nd.kind) {
    case N_SHRINK_H:
    case N_SHRINK_V:
        if (nd.in_fmt == U8 && nd.bands <= 4) {
            s.kind = ST_SHRINK_HV_U8;
            (nd.kind == N_SHRINK_H ? s.fh : s.fv) = (int)nd.factor;
        } else {
            s.kind = ST_SHRINK;
        }
        break;
    case N_REDUCE_H: s.kind = ST_REDUCE_H; stage_reduce_params(nd, out_r, &st, &s); break;
    case N_REDUCE_V: s.kind = ST_REDUCE_V; stage_reduce_params(nd, out_r, &st, &s); break;
    case N_GAUSS_H:
    case N_GAUSS_V:
    case N_CONVSEP_H:
    case N_CONVSEP_V: s.kind = ST_CONV; stage_conv_params(nd, &st, &s); break;
    case N_PREMULT:
    case N_UNPREMULT: s.kind = ST_ALPHA; break;
    default: s.kind = ST_COPY; break;
    }
    int rc = upload_op_blob(op, st, out_r);
    if (rc != VGPU_OK) return rc;
    std::vector<Node> one{nd};
    rc = launch_stage(one, s, view_of(input, in_r, in_r), output,
                      static_cast<const unsigned char*>(op->blob.ptr), ctx->stream);
    if (rc != VGPU_OK) ctx->sticky = rc;
    return rc;
}
