// Internal declarations shared by the pipeline executor (runtime.cpp) and
// the batch executor (batch.cpp): context, buffer pool, fused stages.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/viprs_amd.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace vgpu {

// Size-bucketed cache of device allocations: hipMalloc only on a miss, never
// on the steady-state run path.
class BufferPool {
  public:
    ~BufferPool() { trim(true); }
    void* acquire(size_t bytes) {
        bytes = round_up(bytes);
        std::lock_guard<std::mutex> g(mu_);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            free_.erase(it);
            return p;
        }
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        sizes_[p] = bytes;
        held_ += bytes;
        return p;
    }
    void release(void* p) {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        auto it = sizes_.find(p);
        if (it != sizes_.end()) free_.emplace(it->second, p);
    }
    void trim(bool all) {
        std::lock_guard<std::mutex> g(mu_);
        for (auto& kv : free_) {
            (void)hipFree(kv.second);
            held_ -= sizes_[kv.second];
            sizes_.erase(kv.second);
        }
        free_.clear();
        if (all) {
            for (auto& kv : sizes_) (void)hipFree(kv.first);
            sizes_.clear();
            held_ = 0;
        }
    }
    size_t held() const { return held_; }

  private:
    static size_t round_up(size_t b) { return (std::max<size_t>(b, 1) + 4095) & ~(size_t)4095; }
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    std::map<void*, size_t> sizes_;
    size_t held_ = 0;
};

inline int from_hip(hipError_t e) { return e == hipSuccess ? VGPU_OK : VGPU_E_DEVICE; }

// Two timing events, or none: the first is destroyed if the second fails.
inline bool create_event_pair(hipEvent_t* a, hipEvent_t* b) {
    if (hipEventCreate(a) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (hipEventCreate(b) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipEventDestroy(*a);
        return false;
    }
    return true;
}

}  // namespace vgpu

struct vgpu_context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    vgpu::BufferPool pool;
    int sticky = VGPU_OK;
    // Device address of this device's copy of the library's fixed scratch
    // array (vgpu_pipeline_run_noalloc), looked up once at context creation.
    void* scratch = nullptr;
};

namespace vgpu {

struct DeviceBlob {
    void* ptr = nullptr;
    vgpu_context* ctx = nullptr;
    ~DeviceBlob() { reset(); }
    void reset() {
        if (ptr && ctx) ctx->pool.release(ptr);
        ptr = nullptr;
    }
};

// Host staging for per-node plans and tables, uploaded as one blob.
struct Staging {
    std::vector<unsigned char> bytes;
    // Appends n bytes (16-byte aligned) and returns their offset; p == nullptr
    // reserves them for the caller to fill.
    size_t put(const void* p, size_t n) {
        size_t off = (bytes.size() + 15) & ~(size_t)15;
        bytes.resize(off + n);
        if (n && p) std::memcpy(bytes.data() + off, p, n);
        return off;
    }
};

enum StageKind {
    ST_SEP,      // fused u8 H -> V pair on MFMA (reduce, Gaussian or box)
    ST_RED_F32,  // fused f32 H -> V pair in plain FP32 (reduce or Gaussian)
    ST_SEP_F16,  // fused f32 H -> V pair, FP16 hi/lo split on MFMA
    ST_GAUSS_U8, // fused u8 GaussBlurH -> GaussBlurV, 3 or 4 bands, VALU tiles
    ST_SHRINK_HV_U8,
    ST_SHRINK,
    ST_REDUCE_H,
    ST_REDUCE_V,
    ST_REDUCE_HV,
    ST_CONV,
    ST_ALPHA,
    ST_COPY,
    ST_AFFINE,   // the thumbnail / resize Affine tail
    ST_ZOOM,
    ST_PLACE,    // copy a tile-column run's output into its columns of the destination
    ST_CONV_PAIR, // fused float GaussBlurH -> GaussBlurV or ConvSepH -> ConvSepV
    ST_BOX,       // the PNG box pre-shrink of the source
};

struct Stage {
    StageKind kind = ST_COPY;
    int first = 0, last = 0;  // node index range [first, last]
    Rect clip;                // reference tile rectangle of the first node's input
    Rect out;                 // output rectangle of the last node
    int in_fmt = U8, out_fmt = U8;
    int fh = 1, fv = 1, premult = 0, unpremult = 0;
    size_t plan_start = 0, plan_phase = 0, w16 = 0, w64 = 0;  // blob offsets
    size_t itaps = 0, wts = 0;
    int taps = 1;
    int reduce_op = 0;  // ST_REDUCE_H / ST_REDUCE_V: ReduceOp's pass arithmetic
    // fused reduce: the V plan and the job template (pointers set at launch)
    size_t v_start = 0, v_phase = 0, v_w16 = 0, v_w64 = 0;
    int v_taps = 1;
    ReduceJob rjob;
    SepJob sep;  // ST_SEP: job template (pointers set at launch)
    // ST_SEP blob offsets: V operand patterns, per-tile (pattern, base16, hi),
    // and (family 1) H operand patterns, per-strip (pattern, first pixel).
    size_t v_tab = 0, v_tile = 0, h_tab = 0, h_strip = 0;
    size_t w32 = 0, v_w32 = 0;  // ST_RED_F32: f32 weight tables
    size_t aff_x = 0, aff_y = 0;  // ST_AFFINE: per-column / per-row AffineTap tables
    int aff_node = -1;            // ST_AFFINE: index of the Affine node
    // Tile-column runs: the stage reads the pipeline's source (or, after a
    // box pre-shrink, src_stage's output) instead of the previous stage's
    // output / writes the destination (ST_PLACE, whose destination rows are
    // place_w pixels wide).
    bool from_source = false, to_dst = false;
    int src_stage = -1;
    int64_t place_w = 0;
    ConvPairJob cpair;          // ST_CONV_PAIR: job template (views and weights set at launch)
    int cpair_kind = 0;         // 0 Gaussian FP64, 1 ConvSep FP64, 2 Gaussian FP32
    size_t cp_wh = 0, cp_wv = 0, cp_wh32 = 0, cp_wv32 = 0;  // ST_CONV_PAIR blob offsets
    F16Job f16;                 // ST_SEP_F16: job template (table offsets below)
    GaussJob gauss;             // ST_GAUSS_U8: job template (views set at launch)
};

// Stage-planning choices that depend on how the stage is launched: a single
// image wants enough blocks to fill the chip from one stage; a batch gets its
// parallelism from many images and prefers the largest tiles.
struct StageOptions {
    int64_t reduce_min_blocks = 1024;  // stop shrinking reduce tiles once this many blocks exist
    int reduce_first_tile = 0;         // largest reduce tile candidate considered (0 = 64x16)
    bool integer_mfma = true;          // u8 reduce / Gaussian pairs on the int8-limb MFMA kernel
    bool mfma_shrink = false;          // u8 box pre-shrink on it too (slower than the VALU kernel)
    int sep_rows_per_task = 0;         // output rows per MFMA wavefront task (0 = by chip fill)
    // f32 images: 0 = the reference's FP64 accumulation order (bit-exact),
    // 1 = plain FP32, 2 = FP16 hi/lo split on MFMA (falls back to 1).
    int float_mode = 0;
    bool shrink_row_split = true;      // split a shrink row's source rows over wavefronts
    int shrink_run = 0;                // shrink output rows per wavefront run (0 = planner's choice)
    int gauss_rows = 0;                // u8 3/4-band Gaussian pair: VALU tile rows (0 = auto, 16, 32), -1 = MFMA kernel
    int sep_form = 0;                  // int8 MFMA pairs: 0 = block-tile form where it fits, 1 = wavefront tasks
    int reduce_form = 0;               // f32 reduce pairs: 0 = streaming strips where they fit, 1 = tiles
};

Rect whole(uint32_t w, uint32_t h);
InView view_of(const void* ptr, const Rect& buf, const Rect& clip);
OutView outview_of(void* ptr, const Rect& r);
void backward_regions(const std::vector<Node>& nodes, uint32_t w, uint32_t h, std::vector<Rect>* in_r,
                      std::vector<Rect>* out_r);
// Whether the device reproduces a chain holding an Affine tail: the nodes
// after it may only move regions (ExtractArea, Unpremultiply).
bool affine_tail_supported(const std::vector<Node>& nodes);
// Stages of the whole pipeline (w x h output).  When an Affine runs on
// SmallTile (128 x 128) tiles after a u8 ReduceH, whose fixed-point source
// positions accumulate from each tile's left edge, every 128-column tile
// column is planned as its own run from the source (its stages computing
// exactly the regions the reference's tiles of that column compute), closed
// by an ST_PLACE stage.
std::vector<Stage> plan_pipeline(const std::vector<Node>& nodes, uint32_t w, uint32_t h, Staging* st,
                                 const StageOptions& opt);
// The input of stage k: -1 for the pipeline's source, else the index of the
// stage whose output it reads (the previous one, or the box pre-shrink for a
// tile-column run that starts after it).
inline int stage_input(const std::vector<Stage>& stages, size_t k) {
    if (k == 0) return -1;
    if (stages[k].from_source) return stages[k].src_stage;
    return (int)k - 1;
}
inline bool stage_writes_dst(const std::vector<Stage>& stages, size_t k) {
    return k + 1 == stages.size() || stages[k].to_dst;
}
// Stages of nodes [first_node, n) over the regions of the backward walk.
std::vector<Stage> plan_stages(const std::vector<Node>& nodes, const std::vector<Rect>& in_r,
                               const std::vector<Rect>& out_r, Staging* st,
                               const StageOptions& opt = StageOptions(), int first_node = 0);
ShrinkJob shrink_job_of(const Stage& s, const Node& first, const InView& in, const OutView& out);
ReduceJob reduce_job_of(const Stage& s, const InView& in, const OutView& out, const unsigned char* blob);
SepJob sep_job_of(const Stage& s, const InView& in, const OutView& out, const unsigned char* blob);
int launch_stage(const std::vector<Node>& nodes, const Stage& s, const InView& in, void* dst,
                 const unsigned char* blob, hipStream_t stream, const StageOptions& opt = StageOptions());

}  // namespace vgpu

struct vgpu_pipeline {
    vgpu_context* ctx = nullptr;
    uint32_t src_w = 0, src_h = 0, bands = 0;
    int src_fmt = vgpu::U8;
    uint32_t cur_w = 0, cur_h = 0;
    int cur_fmt = vgpu::U8;
    std::vector<vgpu::Node> nodes;

    bool planned = false;   // stages and host_blob built (host only)
    bool compiled = false;  // plus the blob uploaded and intermediates allocated
    vgpu::StageOptions opt;
    std::vector<vgpu::Stage> stages;
    std::vector<unsigned char> host_blob;  // plan tables as uploaded
    vgpu::DeviceBlob blob;
    std::vector<void*> inter;  // stage outputs except the last

    ~vgpu_pipeline() { release(); }
    void release() {
        for (void* p : inter) ctx->pool.release(p);
        inter.clear();
        blob.reset();
        stages.clear();
        host_blob.clear();
        planned = false;
        compiled = false;
    }
    vgpu::Node& push(vgpu::NodeKind k) {
        release();
        vgpu::Node nd;
        nd.kind = k;
        nd.in_fmt = nd.out_fmt = cur_fmt;
        nd.bands = bands;
        nd.in_w = nd.out_w = cur_w;
        nd.in_h = nd.out_h = cur_h;
        nodes.push_back(nd);
        return nodes.back();
    }
    void commit() {
        cur_w = nodes.back().out_w;
        cur_h = nodes.back().out_h;
        cur_fmt = nodes.back().out_fmt;
    }
    int plan();     // host-side planning only: stages and host_blob
    int compile();  // plan(), then the blob and intermediates in pooled HBM
};
