// Batch executor: many independent images, each with its own plan, run as a
// handful of persistent launches over device task lists instead of one launch
// sequence per image.
//
// Every image is planned exactly as a single pipeline would plan it (the
// reference's thumbnail(), crates/viprs-ops-spatial/src/resample/
// thumbnail.rs:229-524, one CompiledPipeline per image), so each output is the
// one the image would get alone.  Stage k of every image forms "level" k; a
// level runs as one persistent launch per kernel family (fused shrink, fused
// reduce) whose blocks walk a task list of (job, bx, by) entries covering all
// images of that level.  Images whose plan needs another stage kind run their
// stages one launch at a time after the batched levels.
//
// Level-0 jobs read the caller's sources and last-stage jobs write the
// caller's outputs through two device slot tables refreshed per run (only
// when the pointers change), so the launch sequence is fixed for the life of
// the batch and is captured once into a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <numeric>
#include <tuple>
#include <vector>

#include "runtime.hpp"

using namespace vgpu;

namespace {

struct BatchImage {
    uint32_t w = 0, h = 0, bands = 0;
    uint32_t ow = 0, oh = 0;
    std::vector<Node> nodes;
    std::vector<Stage> stages;
    std::vector<size_t> inter_off;  // arena offsets of every stage output but the last
    bool batched = true;
};

// One persistent launch over the tasks of one (level, kernel family).
struct Launch {
    int family = 0;  // 0: fused shrink, 1: fused reduce (VALU), 2: MFMA separable pair, 3: its block-tile form
    int variant = 0; // kernel variant within the family
    size_t jobs_off = 0, tasks_off = 0;
    int ntasks = 0;
    size_t lds = 0;
    uint64_t alg_bytes = 0;  // input buffers read + outputs written, summed over jobs
    int chain = -1;          // the group of its images (one stream each), or -1 for mixed
    int wait_for = -1;       // launch (on another chain) whose outputs it reads, or -1
    bool signal = false;     // a later launch waits for this one
    int after = -1;          // interleaved layout: the launch whose outputs it reads (planning only)
    int images = 0;          // images its jobs belong to
};

// Four launch chains, as many as the device's hardware queues, each on its
// own stream (chain 0 on the caller's).
constexpr int kGroups = 4;
constexpr int kChains = kGroups;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

struct vgpu_batch {
    vgpu_context* ctx = nullptr;
    std::vector<BatchImage> images;
    StageOptions opt;
    DeviceBlob blob;           // plans and tables of every image
    void* arena = nullptr;     // stage intermediates
    size_t arena_bytes = 0;
    void* tables = nullptr;    // jobs, tasks and the two slot tables
    size_t slots_off = 0;      // srcs[count] then dsts[count] inside `tables`
    void* host_slots = nullptr;  // pinned staging of the slot tables
    hipEvent_t slots_read = nullptr;
    std::vector<const void*> cur_srcs;
    std::vector<void*> cur_dsts;
    bool slots_valid = false;
    std::vector<Launch> launches;
    std::vector<int> fallback;  // images run stage by stage
    int n_shrink_tasks = 0, n_reduce_tasks = 0;
    int reduce_chain = 2;     // chain layout: 0 level by level, 1 a reduce chain, 2 interleaved (see plan())
    int grid_cap = 0;         // persistent grid size of the shrink launches; 0 = one block per task
    int reduce_grid_cap = 0;  // the same for the reduce and MFMA separable launches
    int use_graph = 1;
    hipGraphExec_t graph = nullptr;
    // Launches that serve images of one band count and size group form a
    // chain; chains run on their own streams, forked from and joined back into
    // the context stream.
    hipStream_t side[kChains] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kChains] = {};
    std::vector<hipEvent_t> done_ev;  // per launch that a launch on another chain waits for
    unsigned chains = 0;  // bit c: some launch runs on chain c

    ~vgpu_batch() {
        release();
        for (int c = 0; c < kChains; ++c) {
            if (side[c]) (void)hipStreamDestroy(side[c]);
            if (join_ev[c]) (void)hipEventDestroy(join_ev[c]);
        }
        if (fork_ev) (void)hipEventDestroy(fork_ev);
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    void release() {
        drop_graph();
        if (ctx) {
            ctx->pool.release(arena);
            ctx->pool.release(tables);
        }
        arena = tables = nullptr;
        blob.reset();
        if (host_slots) (void)hipHostFree(host_slots);
        host_slots = nullptr;
        if (slots_read) (void)hipEventDestroy(slots_read);
        slots_read = nullptr;
        slots_valid = false;
        launches.clear();
        fallback.clear();
        for (hipEvent_t e : done_ev)
            if (e) (void)hipEventDestroy(e);
        done_ev.clear();
    }
    int plan();
    int launch_all(hipStream_t s, hipEvent_t* marks = nullptr);
    int check_buffers(const void* const* srcs, const size_t* src_bytes, void* const* dsts,
                      const size_t* dst_bytes) const;
    int upload_slots(const void* const* srcs, void* const* dsts, hipStream_t s);
};

int vgpu_batch::check_buffers(const void* const* srcs, const size_t* src_bytes, void* const* dsts,
                              const size_t* dst_bytes) const {
    if (!srcs || !dsts || !src_bytes || !dst_bytes) return VGPU_E_INVALID_ARGUMENT;
    for (size_t i = 0; i < images.size(); ++i) {
        const BatchImage& im = images[i];
        if (src_bytes[i] != (size_t)im.w * im.h * im.bands || dst_bytes[i] != (size_t)im.ow * im.oh * im.bands)
            return VGPU_E_BUFFER_SIZE;
        // The shrink kernel's 16-byte row loads need 16-byte aligned sources.
        if (!srcs[i] || !dsts[i] || ((uintptr_t)srcs[i] & 15)) return VGPU_E_INVALID_ARGUMENT;
    }
    return VGPU_OK;
}

// Refresh the device slot tables when the caller's pointers differ from the
// last run's; enqueued on `s` ahead of the launches that read them.
int vgpu_batch::upload_slots(const void* const* srcs, void* const* dsts, hipStream_t s) {
    const size_t count = images.size();
    bool same = slots_valid;
    for (size_t i = 0; same && i < count; ++i) same = cur_srcs[i] == srcs[i] && cur_dsts[i] == dsts[i];
    if (same) return VGPU_OK;
    // The staging may still feed the previous upload: wait for it first.
    if (slots_valid && hipEventSynchronize(slots_read) != hipSuccess) return VGPU_E_DEVICE;
    void** hs = static_cast<void**>(host_slots);
    for (size_t i = 0; i < count; ++i) {
        hs[i] = const_cast<void*>(srcs[i]);
        hs[count + i] = dsts[i];
    }
    unsigned char* t = static_cast<unsigned char*>(tables);
    if (hipMemcpyAsync(t + slots_off, hs, 2 * count * sizeof(void*), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventRecord(slots_read, s) != hipSuccess)
        return VGPU_E_DEVICE;
    cur_srcs.assign(srcs, srcs + count);
    cur_dsts.assign(dsts, dsts + count);
    slots_valid = true;
    return VGPU_OK;
}

// Plan every image's stages with the batch options, lay out intermediates,
// build the job arrays and task lists and upload them.
int vgpu_batch::plan() {
    release();
    const size_t count = images.size();
    Staging st;
    size_t arena_need = 0;
    int levels = 0;
    for (BatchImage& im : images) {
        im.stages.clear();
        im.inter_off.clear();
        im.batched = true;
        if (im.nodes.empty()) continue;
        im.stages = plan_pipeline(im.nodes, im.ow, im.oh, &st, opt);
        for (size_t k = 0; k + 1 < im.stages.size(); ++k) {
            const Stage& s = im.stages[k];
            im.inter_off.push_back(arena_need);
            if (stage_writes_dst(im.stages, k)) continue;  // a tile column's ST_PLACE writes the output
            arena_need = align_up(arena_need + (size_t)s.out.w * s.out.h * im.bands * format_size(s.out_fmt), 256);
        }
        for (const Stage& s : im.stages)
            if (s.kind != ST_SHRINK_HV_U8 && s.kind != ST_REDUCE_HV && s.kind != ST_SEP) im.batched = false;
        levels = std::max(levels, (int)im.stages.size());
    }
    if (!st.bytes.empty()) {
        blob.ctx = ctx;
        blob.ptr = ctx->pool.acquire(st.bytes.size());
        if (!blob.ptr) return VGPU_E_NO_MEMORY;
        if (hipMemcpy(blob.ptr, st.bytes.data(), st.bytes.size(), hipMemcpyHostToDevice) != hipSuccess)
            return VGPU_E_DEVICE;
    }
    arena_bytes = arena_need;
    if (arena_need) {
        arena = ctx->pool.acquire(arena_need);
        if (!arena) return VGPU_E_NO_MEMORY;
    }
    const unsigned char* dblob = static_cast<const unsigned char*>(blob.ptr);
    unsigned char* darena = static_cast<unsigned char*>(arena);

    // Images in decreasing input size, so the biggest tasks start first.
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (uint64_t)images[(size_t)a].w * images[(size_t)a].h * images[(size_t)a].bands >
               (uint64_t)images[(size_t)b].w * images[(size_t)b].h * images[(size_t)b].bands;
    });

    // Shrink groups: images dealt, largest first, to the group that moves
    // the fewest bytes so far; each group's box shrinks form a chain of
    // launches on its own stream.  With a reduce chain (an option) the
    // reduce stages of every level after the first go to the last stream
    // instead, one launch per (kernel variant, producing launch), each
    // waiting only for the shrink launch that wrote its inputs: the shrink
    // chains never wait for a reduce.  (Measured: the hardware dispatches
    // those reduce launches only once the shrink launches, whose persistent
    // grids hold every CU slot, have drained, so the option is off unless
    // the tuner finds it faster.)
    const int shrink_groups = reduce_chain == 1 ? kGroups - 1 : kGroups;
    const int rchain = reduce_chain == 1 ? kGroups - 1 : -1;
    // Interleaved (reduce_chain == 2): each reduce launch follows, on the
    // same stream, the shrink launch that wrote its inputs, and a chain's
    // shrink launches run those with the most images first, so the reduce
    // work of all but a chain's last shrink launch runs while the other
    // chains still stream their shrinks.
    const bool interleave = reduce_chain == 2;
    std::vector<int> group_of(count, 0);
    {
        uint64_t load[kGroups] = {};
        for (int idx : order) {
            const BatchImage& im = images[(size_t)idx];
            uint64_t bytes = 0;
            for (size_t k = 0; k < im.stages.size(); ++k) {
                const Rect in_rect = k == 0 ? whole(im.w, im.h) : im.stages[k - 1].out;
                bytes += ((uint64_t)in_rect.w * in_rect.h + (uint64_t)im.stages[k].out.w * im.stages[k].out.h) *
                         im.bands;
            }
            int g = 0;
            for (int c = 1; c < shrink_groups; ++c)
                if (load[c] < load[g]) g = c;
            group_of[(size_t)idx] = g;
            load[g] += bytes;
        }
    }
    // The launch that ran each image's previous level (-1: none yet).
    std::vector<int> producer(count, -1);
    Staging tbl;
    n_shrink_tasks = n_reduce_tasks = 0;
    // Jobs of one future launch: kernel family and variant, its chain, and
    // the launch it waits for (-1: none).
    struct Bucket {
        std::vector<ShrinkJob> sjobs;
        std::vector<ReduceJob> rjobs;
        std::vector<SepJob> pjobs;
        std::vector<BatchTask> tasks;
        std::vector<int> images;
        size_t lds = 0;
        uint64_t bytes = 0;
    };
    for (int lv = 0; lv < levels; ++lv) {
        // key: (order, family, variant, chain, wait_for); std::map keeps the
        // launches of a level in producer order, then family and variant.
        std::map<std::tuple<int, int, int, int, int>, Bucket> buckets;
        for (size_t rank = 0; rank < order.size(); ++rank) {
            const int idx = order[rank];
            const int grp = group_of[(size_t)idx];
            BatchImage& im = images[(size_t)idx];
            if (!im.batched || lv >= (int)im.stages.size()) continue;
            const Stage& s = im.stages[(size_t)lv];
            const bool first = lv == 0, last = lv + 1 == (int)im.stages.size();
            const Rect in_rect = first ? whole(im.w, im.h) : im.stages[(size_t)lv - 1].out;
            const void* in_ptr = first ? nullptr : darena + im.inter_off[(size_t)lv - 1];
            void* out_ptr = last ? nullptr : darena + im.inter_off[(size_t)lv];
            const InView in = view_of(in_ptr, in_rect, s.clip);
            const OutView out = outview_of(out_ptr, s.out);
            const uint64_t io_bytes = ((uint64_t)in_rect.w * in_rect.h + (uint64_t)s.out.w * s.out.h) * im.bands;
            // Reduce stages go to the reduce chain (after the producing
            // launch) or stay on the image's group.
            const int wait = producer[(size_t)idx];
            const int rkey_chain = rchain >= 0 ? rchain : grp;
            const int rkey_wait = rchain >= 0 ? wait : -1;
            const int rkey_order = rchain >= 0 || interleave ? wait + 1 : 0;
            if (s.kind == ST_SHRINK_HV_U8) {
                ShrinkJob j = shrink_job_of(s, im.nodes[(size_t)s.first], in, out);
                j.src_slot = first ? idx : -1;
                j.dst_slot = last ? idx : -1;
                if (!plan_shrink_job(&j, opt.shrink_row_split, opt.shrink_run)) {
                    im.batched = false;
                    continue;
                }
                Bucket& bk = buckets[std::make_tuple(0, 0, shrink_job_variant(j), grp, -1)];
                const int jid = (int)bk.sjobs.size();
                bk.sjobs.push_back(j);
                bk.bytes += io_bytes;
                bk.lds = shrink_job_lds();
                bk.images.push_back(idx);
                for (int by = 0; by < j.gy; ++by)
                    for (int bx = 0; bx < j.gx; ++bx) bk.tasks.push_back(BatchTask{jid, bx, by});
            } else if (s.kind == ST_SEP) {
                SepJob j = sep_job_of(s, in, out, dblob);
                j.src_slot = first ? idx : -1;
                j.dst_slot = last ? idx : -1;
                // Family 3: the block-tile form (one task per block of four
                // strips x R rows); family 2: one task per wavefront.
                const int fam = j.tile ? 3 : 2;
                Bucket& bk = buckets[std::make_tuple(rkey_order, fam, sep_job_variant(j), rkey_chain, rkey_wait)];
                const int jid = (int)bk.pjobs.size();
                bk.pjobs.push_back(j);
                bk.bytes += io_bytes;
                bk.images.push_back(idx);
                if (j.tile) {
                    bk.lds = std::max(bk.lds, (size_t)j.t_lds);  // per block
                    const int gxb = (j.gx + 3) / 4;
                    for (int by = 0; by < j.gy; ++by)
                        for (int bx = 0; bx < gxb; ++bx) bk.tasks.push_back(BatchTask{jid, bx, by});
                } else {
                    bk.lds = std::max(bk.lds, sep_job_wave_lds(j));  // per wavefront
                    for (int ty = 0; ty < j.gy; ++ty)
                        for (int jx = 0; jx < j.gx; ++jx) bk.tasks.push_back(BatchTask{jid, jx, ty});
                }
            } else {
                ReduceJob j = reduce_job_of(s, in, out, dblob);
                j.src_slot = first ? idx : -1;
                j.dst_slot = last ? idx : -1;
                // The VALU reduce takes every band count in one launch, after
                // every chain (chain -1 joins them).
                Bucket& bk = buckets[std::make_tuple(1 << 30, 1, 0, -1, -1)];
                const int jid = (int)bk.rjobs.size();
                bk.rjobs.push_back(j);
                bk.bytes += io_bytes;
                bk.lds = std::max(bk.lds, reduce_job_lds(j));
                bk.images.push_back(idx);
                for (int by = 0; by < j.gy; ++by)
                    for (int bx = 0; bx < j.gx; ++bx) bk.tasks.push_back(BatchTask{jid, bx, by});
            }
        }
        for (auto& kv : buckets) {
            Bucket& bk = kv.second;
            if (bk.tasks.empty()) continue;
            Launch l;
            l.family = std::get<1>(kv.first);
            l.variant = std::get<2>(kv.first);
            l.chain = std::get<3>(kv.first);
            l.wait_for = std::get<4>(kv.first);
            if (l.family == 0) l.jobs_off = tbl.put(bk.sjobs.data(), bk.sjobs.size() * sizeof(ShrinkJob));
            else if (l.family == 1) l.jobs_off = tbl.put(bk.rjobs.data(), bk.rjobs.size() * sizeof(ReduceJob));
            else l.jobs_off = tbl.put(bk.pjobs.data(), bk.pjobs.size() * sizeof(SepJob));
            l.tasks_off = tbl.put(bk.tasks.data(), bk.tasks.size() * sizeof(BatchTask));
            l.ntasks = (int)bk.tasks.size();
            l.lds = bk.lds;
            l.alg_bytes = bk.bytes;
            l.after = interleave && l.chain >= 0 ? std::get<0>(kv.first) - 1 : -1;
            l.images = (int)bk.images.size();
            const int li = (int)launches.size();
            for (int idx : bk.images) producer[(size_t)idx] = li;
            if (l.wait_for >= 0) launches[(size_t)l.wait_for].signal = true;
            launches.push_back(l);
            (l.family == 0 ? n_shrink_tasks : n_reduce_tasks) += l.ntasks;
        }
    }
    if (interleave) {
        // Launch order: level-0 launches by chain, most images first, each
        // followed by the launches that read its outputs (recursively).  No
        // launch of this layout waits on an event, so only the order within
        // a chain matters.
        std::vector<std::vector<int>> children(launches.size());
        std::vector<int> roots;
        for (int li = 0; li < (int)launches.size(); ++li) {
            const int a = launches[(size_t)li].after;
            if (a >= 0) children[(size_t)a].push_back(li);
            else roots.push_back(li);
        }
        // Launches over every chain (chain -1: joins them) stay last, in
        // their planned order.
        std::stable_sort(roots.begin(), roots.end(), [&](int a, int b) {
            const Launch& x = launches[(size_t)a];
            const Launch& y = launches[(size_t)b];
            if ((x.chain < 0) != (y.chain < 0)) return y.chain < 0;
            if (x.chain < 0) return false;
            if (x.chain != y.chain) return x.chain < y.chain;
            if ((x.family == 0) != (y.family == 0)) return x.family != 0;  // lone reduces (no shrink) first
            return x.images > y.images;
        });
        std::vector<Launch> ordered;
        std::vector<int> stack;
        for (auto it = roots.rbegin(); it != roots.rend(); ++it) stack.push_back(*it);
        while (!stack.empty()) {
            const int li = stack.back();
            stack.pop_back();
            ordered.push_back(launches[(size_t)li]);
            for (auto it = children[(size_t)li].rbegin(); it != children[(size_t)li].rend(); ++it) stack.push_back(*it);
        }
        // Issue the chains' launches round robin (each chain keeps its own
        // order): the captured graph then starts every chain at once.
        std::vector<std::vector<Launch>> per(kChains);
        std::vector<Launch> joins;
        for (const Launch& l : ordered) (l.chain >= 0 ? per[(size_t)l.chain] : joins).push_back(l);
        launches.clear();
        for (size_t k = 0;; ++k) {
            bool any = false;
            for (int c = 0; c < kChains; ++c)
                if (k < per[(size_t)c].size()) {
                    launches.push_back(per[(size_t)c][k]);
                    any = true;
                }
            if (!any) break;
        }
        for (const Launch& l : joins) launches.push_back(l);
    }
    chains = 0;
    for (const Launch& l : launches)
        if (l.chain >= 0) chains |= 1u << l.chain;
    for (int c = 0; c < kChains; ++c) {
        if (c == 0 || !(chains & (1u << c)) || side[c]) continue;  // chain 0 uses the caller's stream
        // Default priority, as the context stream.
        if (hipStreamCreateWithFlags(&side[c], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&join_ev[c], hipEventDisableTiming) != hipSuccess)
            return VGPU_E_DEVICE;
    }
    if (!fork_ev && hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming) != hipSuccess) return VGPU_E_DEVICE;
    done_ev.assign(launches.size(), nullptr);
    for (size_t li = 0; li < launches.size(); ++li)
        if (launches[li].signal && hipEventCreateWithFlags(&done_ev[li], hipEventDisableTiming) != hipSuccess) {
            done_ev[li] = nullptr;
            return VGPU_E_DEVICE;
        }
    // An image whose shrink job could not be planned runs stage by stage
    // after the batched levels (any earlier-level job it already has only
    // writes its own intermediate, which the stage-by-stage run rewrites).
    for (size_t i = 0; i < count; ++i)
        if (!images[i].batched || images[i].stages.empty()) fallback.push_back((int)i);

    slots_off = tbl.put(nullptr, 0);
    slots_off = align_up(slots_off, 16);
    tbl.bytes.resize(slots_off + 2 * count * sizeof(void*), 0);
    tables = ctx->pool.acquire(tbl.bytes.size());
    if (!tables) return VGPU_E_NO_MEMORY;
    if (hipMemcpy(tables, tbl.bytes.data(), tbl.bytes.size(), hipMemcpyHostToDevice) != hipSuccess)
        return VGPU_E_DEVICE;
    if (hipHostMalloc(&host_slots, std::max<size_t>(2 * count * sizeof(void*), 16), hipHostMallocDefault) !=
        hipSuccess) {
        host_slots = nullptr;
        return VGPU_E_NO_MEMORY;
    }
    if (hipEventCreateWithFlags(&slots_read, hipEventDisableTiming) != hipSuccess) {
        slots_read = nullptr;
        return VGPU_E_DEVICE;
    }
    return VGPU_OK;
}

int vgpu_batch::launch_all(hipStream_t s, hipEvent_t* marks) {
    unsigned char* t = static_cast<unsigned char*>(tables);
    const void* const* srcs = reinterpret_cast<const void* const*>(t + slots_off);
    void* const* dsts = reinterpret_cast<void* const*>(t + slots_off + images.size() * sizeof(void*));
    // Per-launch timing (marks) keeps every launch on `s`, in order.
    const bool par = !marks && __builtin_popcount(chains) >= 2;
    auto fork = [&]() -> bool {
        if (hipEventRecord(fork_ev, s) != hipSuccess) return false;
        for (int c = 1; c < kChains; ++c)
            if ((chains & (1u << c)) && hipStreamWaitEvent(side[c], fork_ev, 0) != hipSuccess) return false;
        return true;
    };
    auto join = [&]() -> bool {
        for (int c = 1; c < kChains; ++c)
            if ((chains & (1u << c)) &&
                (hipEventRecord(join_ev[c], side[c]) != hipSuccess || hipStreamWaitEvent(s, join_ev[c], 0) != hipSuccess))
                return false;
        return true;
    };
    if (par && !fork()) return VGPU_E_DEVICE;
    for (size_t li = 0; li < launches.size(); ++li) {
        const Launch& l = launches[li];
        hipStream_t ls = s;
        if (par) {
            if (l.chain >= 0) {
                ls = l.chain == 0 ? s : side[l.chain];  // chain 0 on the caller's stream
                if (l.wait_for >= 0 && launches[(size_t)l.wait_for].chain != l.chain &&
                    hipStreamWaitEvent(ls, done_ev[(size_t)l.wait_for], 0) != hipSuccess)
                    return VGPU_E_DEVICE;
            } else if (!join()) {  // a launch over mixed band counts waits for every chain
                return VGPU_E_DEVICE;
            }
        }
        // A block takes one task, or (MFMA wavefront family) one task per wavefront.
        const int blocks = l.family == 2 ? (l.ntasks + 3) / 4 : l.ntasks;
        const int cap = l.family == 0 ? grid_cap : reduce_grid_cap;
        const int grid = cap > 0 ? std::min(blocks, cap) : blocks;
        const BatchTask* tasks = reinterpret_cast<const BatchTask*>(t + l.tasks_off);
        hipError_t e;
        if (l.family == 3)
            e = launch_sep_tile_batch(l.variant, reinterpret_cast<const SepJob*>(t + l.jobs_off), tasks, l.ntasks,
                                      srcs, dsts, grid, l.lds, ls);
        else if (l.family == 2)
            e = launch_sep_batch(l.variant, reinterpret_cast<const SepJob*>(t + l.jobs_off), tasks, l.ntasks, srcs,
                                 dsts, grid, l.lds, ls);
        else if (l.family == 0)
            e = launch_shrink_batch(l.variant, reinterpret_cast<const ShrinkJob*>(t + l.jobs_off), tasks, l.ntasks,
                                    srcs, dsts, grid, ls);
        else
            e = launch_reduce_batch(reinterpret_cast<const ReduceJob*>(t + l.jobs_off), tasks, l.ntasks, srcs, dsts,
                                    grid, l.lds, ls);
        if (e != hipSuccess) return VGPU_E_DEVICE;
        if (par && l.signal && l.chain >= 0 && hipEventRecord(done_ev[li], ls) != hipSuccess) return VGPU_E_DEVICE;
        if (par && l.chain < 0 && !fork()) return VGPU_E_DEVICE;
        if (marks && hipEventRecord(marks[li + 1], s) != hipSuccess) return VGPU_E_DEVICE;
    }
    if (par && !join()) return VGPU_E_DEVICE;
    return VGPU_OK;
}

namespace {

int run_fallback(vgpu_batch* b, int idx, const void* src, void* dst, hipStream_t s) {
    BatchImage& im = b->images[(size_t)idx];
    const size_t out_bytes = (size_t)im.ow * im.oh * im.bands;
    if (im.stages.empty()) return from_hip(hipMemcpyAsync(dst, src, out_bytes, hipMemcpyDeviceToDevice, s));
    const unsigned char* dblob = static_cast<const unsigned char*>(b->blob.ptr);
    unsigned char* darena = static_cast<unsigned char*>(b->arena);
    for (size_t k = 0; k < im.stages.size(); ++k) {
        const Stage& st = im.stages[k];
        const int si = stage_input(im.stages, k);
        const InView in = si < 0 ? view_of(src, whole(im.w, im.h), st.clip)
                                 : view_of(darena + im.inter_off[(size_t)si], im.stages[(size_t)si].out, st.clip);
        void* o = stage_writes_dst(im.stages, k) ? dst : darena + im.inter_off[k];
        const int rc = launch_stage(im.nodes, st, in, o, dblob, s);
        if (rc != VGPU_OK) return rc;
    }
    return VGPU_OK;
}

}  // namespace

extern "C" {

int vgpu_batch_thumbnail_new(vgpu_context* ctx, uint32_t count, const uint32_t* widths,
                             const uint32_t* heights, const uint32_t* bands, int target,
                             uint32_t width, uint32_t height, int kernel, vgpu_batch** out) {
    if (!ctx || !out || (count && (!widths || !heights || !bands))) return VGPU_E_INVALID_ARGUMENT;
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    std::unique_ptr<vgpu_batch> b(new (std::nothrow) vgpu_batch());
    if (!b) return VGPU_E_NO_MEMORY;
    b->ctx = ctx;
    b->opt.reduce_min_blocks = 1;  // many images: the largest tiles that fit
    b->opt.shrink_row_split = false;
    b->opt.sep_rows_per_task = 64;
    b->images.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        if (bands[i] < 1 || bands[i] > 4 || widths[i] == 0 || heights[i] == 0) return VGPU_E_INVALID_ARGUMENT;
        vgpu_pipeline* p = nullptr;
        int rc = vgpu_pipeline_new(ctx, widths[i], heights[i], bands[i], U8, &p);
        if (rc != VGPU_OK) return rc;
        rc = vgpu_pipeline_thumbnail(p, target, width, height, kernel);
        if (rc != VGPU_OK) {
            vgpu_pipeline_destroy(p);
            return rc;
        }
        BatchImage& im = b->images[i];
        im.w = widths[i];
        im.h = heights[i];
        im.bands = bands[i];
        im.ow = p->cur_w;
        im.oh = p->cur_h;
        im.nodes = p->nodes;
        vgpu_pipeline_destroy(p);
    }
    const int rc = b->plan();
    if (rc != VGPU_OK) return rc;
    *out = b.release();
    return VGPU_OK;
}

void vgpu_batch_destroy(vgpu_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

int vgpu_batch_size(const vgpu_batch* b) { return b ? (int)b->images.size() : VGPU_E_INVALID_ARGUMENT; }

int vgpu_batch_output_info(const vgpu_batch* b, uint32_t index, uint32_t* width, uint32_t* height,
                           uint32_t* bands) {
    if (!b || index >= b->images.size()) return VGPU_E_INVALID_ARGUMENT;
    const BatchImage& im = b->images[index];
    if (width) *width = im.ow;
    if (height) *height = im.oh;
    if (bands) *bands = im.bands;
    return VGPU_OK;
}

int vgpu_batch_info(const vgpu_batch* b, vgpu_batch_stats* st) {
    if (!b || !st) return VGPU_E_INVALID_ARGUMENT;
    std::memset(st, 0, sizeof(*st));
    st->images = (uint32_t)b->images.size();
    st->launches = (uint32_t)b->launches.size();
    st->shrink_tasks = (uint32_t)b->n_shrink_tasks;
    st->reduce_tasks = (uint32_t)b->n_reduce_tasks;
    st->fallback_images = (uint32_t)b->fallback.size();
    st->intermediate_bytes = b->arena_bytes;
    for (const BatchImage& im : b->images) {
        st->input_bytes += (uint64_t)im.w * im.h * im.bands;
        st->output_bytes += (uint64_t)im.ow * im.oh * im.bands;
    }
    st->grid_cap = b->grid_cap;
    st->reduce_grid_cap = b->reduce_grid_cap;
    st->use_graph = b->use_graph;
    st->reduce_first_tile = b->opt.reduce_first_tile;
    st->reduce_chain = b->reduce_chain;
    return VGPU_OK;
}

int vgpu_batch_configure(vgpu_batch* b, int grid_cap, int reduce_grid_cap, int use_graph, int reduce_first_tile) {
    if (!b || grid_cap < 0 || reduce_grid_cap < 0 || reduce_first_tile < 0) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(b->ctx->device);
    if (hipStreamSynchronize(b->ctx->stream) != hipSuccess) return b->ctx->sticky = VGPU_E_DEVICE;
    b->grid_cap = grid_cap;
    b->reduce_grid_cap = reduce_grid_cap;
    b->use_graph = use_graph ? 1 : 0;
    b->drop_graph();
    if (reduce_first_tile != b->opt.reduce_first_tile) {
        b->opt.reduce_first_tile = reduce_first_tile;
        return b->plan();
    }
    return VGPU_OK;
}

int vgpu_batch_set_option(vgpu_batch* b, int option, int value) {
    if (!b) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(b->ctx->device);
    if (hipStreamSynchronize(b->ctx->stream) != hipSuccess) return b->ctx->sticky = VGPU_E_DEVICE;
    switch (option) {
    case VGPU_OPT_SEP_ROWS_PER_TASK:
        if (value < 0 || value % 16) return VGPU_E_INVALID_ARGUMENT;
        b->opt.sep_rows_per_task = value;
        break;
    case VGPU_OPT_SHRINK_RUN:
        if (value < 0 || value > 64) return VGPU_E_INVALID_ARGUMENT;
        b->opt.shrink_run = value;
        break;
    case VGPU_OPT_GAUSS_ROWS:
        if (value != -1 && value != 0 && value != 16 && value != 32) return VGPU_E_INVALID_ARGUMENT;
        b->opt.gauss_rows = value;
        break;
    case VGPU_OPT_BATCH_REDUCE_CHAIN:
        if (value < 0 || value > 2) return VGPU_E_INVALID_ARGUMENT;
        b->reduce_chain = value;
        break;
    case VGPU_OPT_SEP_FORM:
        if (value != 0 && value != 1) return VGPU_E_INVALID_ARGUMENT;
        b->opt.sep_form = value;
        break;
    default: return VGPU_E_INVALID_ARGUMENT;
    }
    b->drop_graph();
    return b->plan();
}

int vgpu_batch_set_integer_mfma(vgpu_batch* b, int enable) {
    if (!b) return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(b->ctx->device);
    if (hipStreamSynchronize(b->ctx->stream) != hipSuccess) return b->ctx->sticky = VGPU_E_DEVICE;
    b->opt.integer_mfma = (enable & 1) != 0;
    b->opt.mfma_shrink = (enable & 2) != 0;
    return b->plan();
}

int vgpu_batch_run(vgpu_batch* b, const void* const* srcs, const size_t* src_bytes, void* const* dsts,
                   const size_t* dst_bytes) {
    if (!b) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = b->ctx;
    if (ctx->sticky != VGPU_OK) return ctx->sticky;
    const size_t count = b->images.size();
    if (count == 0) return VGPU_OK;
    int rc = b->check_buffers(srcs, src_bytes, dsts, dst_bytes);
    if (rc != VGPU_OK) return rc;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    rc = b->upload_slots(srcs, dsts, s);
    if (rc != VGPU_OK) return ctx->sticky = rc;
    if (b->use_graph && !b->launches.empty()) {
        if (!b->graph) {
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return ctx->sticky = VGPU_E_DEVICE;
            rc = b->launch_all(s);
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (rc != VGPU_OK || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return ctx->sticky = VGPU_E_DEVICE;
            }
            const hipError_t ie = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                b->graph = nullptr;
                return ctx->sticky = VGPU_E_DEVICE;
            }
        }
        if (hipGraphLaunch(b->graph, s) != hipSuccess) return ctx->sticky = VGPU_E_DEVICE;
    } else {
        rc = b->launch_all(s);
    }
    for (int idx : b->fallback) {
        if (rc != VGPU_OK) break;
        rc = run_fallback(b, idx, srcs[(size_t)idx], dsts[(size_t)idx], s);
    }
    if (rc != VGPU_OK) ctx->sticky = rc;
    return rc;
}

int vgpu_batch_run_profiled(vgpu_batch* b, const void* const* srcs, const size_t* src_bytes, void* const* dsts,
                            const size_t* dst_bytes, int iterations, float* launch_ms, uint64_t* launch_bytes,
                            int* launch_family, int max_launches) {
    if (!b || iterations < 1 || !launch_ms) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = b->ctx;
    if (ctx->sticky != VGPU_OK) return ctx->sticky;
    const int nl = (int)b->launches.size();
    if (max_launches < nl) return VGPU_E_BUFFER_SIZE;
    int rc = b->check_buffers(srcs, src_bytes, dsts, dst_bytes);
    if (rc != VGPU_OK) return rc;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    rc = b->upload_slots(srcs, dsts, s);
    if (rc != VGPU_OK) return ctx->sticky = rc;
    std::vector<hipEvent_t> ev((size_t)nl + 1);
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return VGPU_E_DEVICE;
    std::vector<double> acc((size_t)nl, 0.0);
    for (int it = 0; it < iterations && rc == VGPU_OK; ++it) {
        (void)hipEventRecord(ev[0], s);
        rc = b->launch_all(s, ev.data());
        if (rc == VGPU_OK && hipEventSynchronize(ev[(size_t)nl]) != hipSuccess) rc = VGPU_E_DEVICE;
        for (int k = 0; k < nl && rc == VGPU_OK; ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[(size_t)k], ev[(size_t)k + 1]);
            acc[(size_t)k] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < nl; ++k) {
        launch_ms[k] = (float)(acc[(size_t)k] / iterations);
        if (launch_bytes) launch_bytes[k] = b->launches[(size_t)k].alg_bytes;
        if (launch_family) launch_family[k] = b->launches[(size_t)k].family;
    }
    if (rc != VGPU_OK) return ctx->sticky = rc;
    return nl;
}

int vgpu_batch_benchmark(vgpu_batch* b, const void* const* srcs, const size_t* src_bytes, void* const* dsts,
                         const size_t* dst_bytes, int warmup, int iterations, float* total_ms, float* step_ms) {
    if (!b || iterations < 1 || !total_ms) return VGPU_E_INVALID_ARGUMENT;
    vgpu_context* ctx = b->ctx;
    int rc = VGPU_OK;
    for (int i = 0; i < warmup && rc == VGPU_OK; ++i) rc = vgpu_batch_run(b, srcs, src_bytes, dsts, dst_bytes);
    if (rc != VGPU_OK) return rc;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return ctx->sticky = VGPU_E_DEVICE;
    hipEvent_t a, e;
    if (!create_event_pair(&a, &e)) return VGPU_E_DEVICE;
    (void)hipEventRecord(a, ctx->stream);
    for (int i = 0; i < iterations && rc == VGPU_OK; ++i) rc = vgpu_batch_run(b, srcs, src_bytes, dsts, dst_bytes);
    (void)hipEventRecord(e, ctx->stream);
    if (rc == VGPU_OK && hipEventSynchronize(e) != hipSuccess) rc = VGPU_E_DEVICE;
    if (rc == VGPU_OK) (void)hipEventElapsedTime(total_ms, a, e);
    if (rc == VGPU_OK && step_ms) {
        std::vector<hipEvent_t> ev((size_t)iterations + 1);
        for (auto& x : ev) (void)hipEventCreate(&x);
        (void)hipEventRecord(ev[0], ctx->stream);
        for (int i = 0; i < iterations && rc == VGPU_OK; ++i) {
            rc = vgpu_batch_run(b, srcs, src_bytes, dsts, dst_bytes);
            (void)hipEventRecord(ev[(size_t)i + 1], ctx->stream);
        }
        if (rc == VGPU_OK && hipEventSynchronize(ev[(size_t)iterations]) != hipSuccess) rc = VGPU_E_DEVICE;
        for (int i = 0; i < iterations && rc == VGPU_OK; ++i)
            (void)hipEventElapsedTime(&step_ms[i], ev[(size_t)i], ev[(size_t)i + 1]);
        for (auto& x : ev) (void)hipEventDestroy(x);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(e);
    return rc;
}

int vgpu_batch_tune(vgpu_batch* b, const void* const* srcs, const size_t* src_bytes, void* const* dsts,
                    const size_t* dst_bytes, int iterations, float* best_ms) {
    if (!b || iterations < 1) return VGPU_E_INVALID_ARGUMENT;
    // Candidates: reduce chain on / off x reduce tile start x persistent grid
    // size of the shrink launches x that of the reduce launches (one block
    // per task, or 4/8/16 blocks per CU walking the list), all launched from
    // a graph.  The VALU reduce tile only matters when the batch has VALU
    // reduce launches.
    static const int shares[] = {2, 0};  // chain layouts: interleaved, level by level
    static const int tiles[] = {0, 1, 2, 3};
    static const int sgrids[] = {0, 256 * 8, 256 * 16};
    static const int rgrids[] = {0, 256 * 4, 256 * 8, 256 * 16};
    float best = 0.f;
    int best_tile = b->opt.reduce_first_tile, best_grid = b->grid_cap, best_rgrid = b->reduce_grid_cap;
    int best_share = b->reduce_chain;
    bool have = false;
    for (int sh : shares) {
        if (sh != b->reduce_chain) {
            (void)hipStreamSynchronize(b->ctx->stream);
            b->reduce_chain = sh;
            b->drop_graph();
            const int rc = b->plan();
            if (rc != VGPU_OK) return rc;
        }
        bool valu_reduce = false;
        for (const Launch& l : b->launches) valu_reduce = valu_reduce || l.family == 1;
        for (int t : tiles) {
            if (!valu_reduce && t != tiles[0]) break;
            for (int g : sgrids) {
                for (int rg : rgrids) {
                    int rc = vgpu_batch_configure(b, g, rg, 1, t);
                    if (rc != VGPU_OK) return rc;
                    float ms = 0.f;
                    rc = vgpu_batch_benchmark(b, srcs, src_bytes, dsts, dst_bytes, 1, iterations, &ms, nullptr);
                    if (rc != VGPU_OK) return rc;
                    if (!have || ms < best) {
                        have = true;
                        best = ms;
                        best_tile = t;
                        best_grid = g;
                        best_rgrid = rg;
                        best_share = sh;
                    }
                }
            }
        }
    }
    if (best_share != b->reduce_chain) {
        (void)hipStreamSynchronize(b->ctx->stream);
        b->reduce_chain = best_share;
        b->drop_graph();
        const int rc = b->plan();
        if (rc != VGPU_OK) return rc;
    }
    const int rc = vgpu_batch_configure(b, best_grid, best_rgrid, 1, best_tile);
    if (best_ms) *best_ms = best / (float)iterations;
    return rc;
}

}  // extern "C"
