// Thumbnail request loop: continuous batching over a device-memory arena.
//
// The reference serves many images by running several pipelines at once
// under a concurrency cap (ExecutionSemaphore, crates/viprs-runtime/src/
// limits.rs:290-360; RayonScheduler::with_max_concurrent_pipelines) and
// rejects requests over its resource limits (ResourceLimits::validate_output,
// limits.rs:238-275).  Here the cap is a number of request slots plus an HBM
// budget: every request holds one contiguous chunk of a preallocated device
// arena (input copy when the pixels come from the host, plan tables, fused
// stage intermediates, output) from admission until its output is released,
// so admission is the arena allocation succeeding.
//
// Threads:
//   - planner threads build each request's fused stage plan on the host
//     (coefficient tables, MFMA operand patterns) and stage the plan tables,
//     and a pageable host input, in pinned memory;
//   - one scheduler thread admits planned requests in arrival order while a
//     slot and arena space are free, enqueues the upload and the stage
//     launches on one of several HIP streams (requests overlap on the
//     device), and retires requests whose completion event has fired,
//     freeing their slot and, when asked to, their memory.
// Requests that arrive while others run are admitted as soon as slots and
// memory free up; nothing waits for a batch to drain.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime.hpp"

using namespace vgpu;

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr size_t kAlign = 256;
size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation, with coalescing.
class Arena {
  public:
    int init(size_t bytes) {
        cap_ = align_up(std::max<size_t>(bytes, kAlign));
        if (hipMalloc(&base_, cap_) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            return VGPU_E_NO_MEMORY;
        }
        free_[0] = cap_;
        return VGPU_OK;
    }
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    unsigned char* at(size_t off) const { return static_cast<unsigned char*>(base_) + off; }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

  private:
    void* base_ = nullptr;
    size_t cap_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

// Pinned host buffers by power-of-two size class (hipHostMalloc only on a miss).
class PinnedPool {
  public:
    ~PinnedPool() {
        for (auto& kv : free_) (void)hipHostFree(kv.second);
    }
    void* acquire(size_t bytes, size_t* cls) {
        size_t c = 4096;
        while (c < bytes) c <<= 1;
        *cls = c;
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = free_.find(c);
            if (it != free_.end()) {
                void* p = it->second;
                free_.erase(it);
                return p;
            }
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return p;
    }
    void release(void* p, size_t cls) {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        free_.emplace(cls, p);
    }

  private:
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
};

enum JobState { J_PLANNING, J_READY, J_RUNNING, J_DONE };

struct Job {
    uint64_t id = 0;
    vgpu_request req{};
    int auto_release = 0;
    // plan
    std::unique_ptr<vgpu_pipeline> pipe;
    std::vector<Stage> stages;
    size_t blob_bytes = 0, in_bytes = 0, out_bytes = 0;
    std::vector<size_t> inter_bytes;
    uint32_t ow = 0, oh = 0, ob = 0;
    int ofmt = U8;
    size_t chunk = 0;   // arena bytes held from admission to release
    void* pin_blob = nullptr;
    size_t pin_blob_cls = 0;
    void* pin_in = nullptr;  // staged copy of a pageable host input
    size_t pin_in_cls = 0;
    // device
    size_t off = 0;
    bool holds_chunk = false;
    int stream = 0;
    hipEvent_t done = nullptr;
    // state
    JobState state = J_PLANNING;
    int status = VGPU_OK;
    bool cancelled = false;  // vgpu_server_cancel: finish with VGPU_E_CANCELLED
    int64_t t_arrival = 0, t_planned = 0, t_admit = 0, t_done = 0;
};

}  // namespace

struct vgpu_server {
    vgpu_context* ctx = nullptr;
    vgpu_server_config cfg{};
    Arena arena;
    PinnedPool pinned;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> free_events;

    std::mutex mu;
    std::condition_variable cv_plan, cv_sched, cv_done;
    uint64_t next_id = 1;
    std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
    std::deque<std::shared_ptr<Job>> to_plan;   // submitted, not planned
    std::deque<std::shared_ptr<Job>> arrival;   // planned or planning, not admitted (arrival order)
    std::vector<std::shared_ptr<Job>> running;
    std::vector<std::thread> planners;
    std::thread scheduler;
    bool stop = false;
    int sticky = VGPU_OK;
    // statistics
    uint64_t completed = 0, rejected = 0, peak_inflight = 0;
    size_t next_stream = 0;

    void plan_job(Job& j);
    bool launch_job(Job& j);
    void finish_job(Job& j);
    void free_chunk(Job& j);
    void planner_loop(int index);
    void scheduler_loop();
};

void vgpu_server::plan_job(Job& j) {
    const vgpu_request& r = j.req;
    vgpu_pipeline* p = nullptr;
    int rc = vgpu_pipeline_new(ctx, r.width, r.height, r.bands, r.format, &p);
    if (rc == VGPU_OK) j.pipe.reset(p);
    if (rc == VGPU_OK && r.thumbnail) rc = vgpu_pipeline_thumbnail(p, r.target, r.target_w, r.target_h, r.kernel);
    if (rc == VGPU_OK && r.blur_sigma > 0.f) rc = vgpu_pipeline_gauss_blur(p, r.blur_sigma);
    if (rc == VGPU_OK) rc = vgpu_pipeline_output_info(p, &j.ow, &j.oh, &j.ob, &j.ofmt);
    if (rc == VGPU_OK && !affine_tail_supported(p->nodes)) rc = VGPU_E_UNSUPPORTED;
    if (rc != VGPU_OK) {
        j.status = rc;
        return;
    }
    j.in_bytes = (size_t)r.width * r.height * r.bands * format_size(r.format);
    j.out_bytes = (size_t)j.ow * j.oh * j.ob * format_size(j.ofmt);
    Staging st;
    if (!p->nodes.empty() && j.out_bytes) {
        j.stages = plan_pipeline(p->nodes, p->cur_w, p->cur_h, &st, p->opt);
        for (size_t k = 0; k + 1 < j.stages.size(); ++k) {
            const Stage& s = j.stages[k];
            // Stages that write the output (a tile column's ST_PLACE) need no space.
            j.inter_bytes.push_back(stage_writes_dst(j.stages, k)
                                        ? 0
                                        : (size_t)s.out.w * (size_t)s.out.h * r.bands * format_size(s.out_fmt));
        }
    }
    j.blob_bytes = st.bytes.size();
    const bool copy_in = !r.src_on_device;
    j.chunk = (copy_in ? align_up(j.in_bytes) : 0) + align_up(j.blob_bytes) + align_up(std::max<size_t>(j.out_bytes, 1));
    for (size_t b : j.inter_bytes) j.chunk += align_up(b);
    if (j.chunk > arena.capacity()) {
        j.status = VGPU_E_ADMISSION;
        return;
    }
    if (j.blob_bytes) {
        j.pin_blob = pinned.acquire(j.blob_bytes, &j.pin_blob_cls);
        if (!j.pin_blob) {
            j.status = VGPU_E_NO_MEMORY;
            return;
        }
        std::memcpy(j.pin_blob, st.bytes.data(), j.blob_bytes);
    }
    if (copy_in && r.src_pinned == 0 && j.in_bytes) {
        j.pin_in = pinned.acquire(j.in_bytes, &j.pin_in_cls);
        if (!j.pin_in) {
            j.status = VGPU_E_NO_MEMORY;
            return;
        }
        std::memcpy(j.pin_in, r.src, j.in_bytes);
    }
}

// Idle threads keep polling this long after their last piece of work before
// they block: a thread woken through a condition variable costs tens of
// microseconds of request latency (the scheduler and two planner threads
// poll; the other planners block at once).
constexpr int64_t kSpinNs = 10000000;

void vgpu_server::planner_loop(int index) {
    (void)hipSetDevice(ctx->device);  // pinned staging allocations
    int64_t last = now_ns();
    for (;;) {
        std::shared_ptr<Job> j;
        {
            std::unique_lock<std::mutex> g(mu);
            for (;;) {
                if (stop && to_plan.empty()) return;
                if (!to_plan.empty()) break;
                if (index < 2 && now_ns() - last < kSpinNs) {
                    g.unlock();
                    std::this_thread::yield();
                    g.lock();
                } else {
                    cv_plan.wait(g);
                }
            }
            j = to_plan.front();
            to_plan.pop_front();
        }
        bool skip;
        {
            std::lock_guard<std::mutex> g(mu);
            skip = j->cancelled;
        }
        if (!skip) plan_job(*j);
        {
            std::lock_guard<std::mutex> g(mu);
            if (j->cancelled) j->status = VGPU_E_CANCELLED;
            j->t_planned = now_ns();
            j->state = J_READY;
        }
        cv_sched.notify_one();
        last = now_ns();
    }
}

// Enqueue one admitted request on its stream: upload tables (and input),
// fused stages in order, completion event.
bool vgpu_server::launch_job(Job& j) {
    hipStream_t s = streams[(size_t)j.stream];
    unsigned char* base = arena.at(j.off);
    size_t o = 0;
    const void* src = j.req.src;
    if (!j.req.src_on_device) {
        void* d_in = base + o;
        o += align_up(j.in_bytes);
        const void* host = j.pin_in ? j.pin_in : j.req.src;
        if (j.in_bytes && hipMemcpyAsync(d_in, host, j.in_bytes, hipMemcpyHostToDevice, s) != hipSuccess) return false;
        src = d_in;
    }
    unsigned char* blob = base + o;
    o += align_up(j.blob_bytes);
    if (j.blob_bytes && hipMemcpyAsync(blob, j.pin_blob, j.blob_bytes, hipMemcpyHostToDevice, s) != hipSuccess)
        return false;
    std::vector<void*> inter;
    for (size_t b : j.inter_bytes) {
        inter.push_back(base + o);
        o += align_up(b);
    }
    void* out = base + o;
    if (j.out_bytes) {
        if (j.stages.empty()) {
            if (hipMemcpyAsync(out, src, j.out_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) return false;
        } else {
            for (size_t k = 0; k < j.stages.size(); ++k) {
                const Stage& st = j.stages[k];
                const int si = stage_input(j.stages, k);
                const InView in = si < 0 ? view_of(src, whole(j.req.width, j.req.height), st.clip)
                                         : view_of(inter[(size_t)si], j.stages[(size_t)si].out, st.clip);
                void* dst = stage_writes_dst(j.stages, k) ? out : inter[k];
                if (launch_stage(j.pipe->nodes, st, in, dst, blob, s, j.pipe->opt) != VGPU_OK) return false;
            }
        }
    }
    return hipEventRecord(j.done, s) == hipSuccess;
}

void vgpu_server::free_chunk(Job& j) {
    if (j.holds_chunk) {
        arena.release(j.off, j.chunk);
        j.holds_chunk = false;
    }
}

// Completion bookkeeping (scheduler thread, lock held).
void vgpu_server::finish_job(Job& j) {
    j.t_done = now_ns();
    j.state = J_DONE;
    if (j.cancelled && j.status == VGPU_OK) j.status = VGPU_E_CANCELLED;
    ++completed;
    pinned.release(j.pin_blob, j.pin_blob_cls);
    pinned.release(j.pin_in, j.pin_in_cls);
    j.pin_blob = j.pin_in = nullptr;
    if (j.done) {
        free_events.push_back(j.done);
        j.done = nullptr;
    }
    if (j.auto_release) free_chunk(j);
}

void vgpu_server::scheduler_loop() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> g(mu);
    int64_t last = now_ns();
    for (;;) {
        // Retire finished requests.
        bool retired = false;
        for (size_t i = 0; i < running.size();) {
            Job& j = *running[i];
            const hipError_t q = hipEventQuery(j.done);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            if (q != hipSuccess) {
                j.status = VGPU_E_DEVICE;
                sticky = VGPU_E_DEVICE;
            }
            finish_job(j);
            running[i] = running.back();
            running.pop_back();
            retired = true;
        }
        if (retired) {
            cv_done.notify_all();
            last = now_ns();
        }
        // Admit planned requests in arrival order while slots and memory allow.
        bool blocked = false;
        while (!arrival.empty() && !blocked) {
            Job& j = *arrival.front();
            if (j.state == J_PLANNING) break;  // keep arrival order
            if (j.cancelled && j.status == VGPU_OK) j.status = VGPU_E_CANCELLED;
            if (j.status != VGPU_OK || sticky != VGPU_OK) {
                if (j.status == VGPU_OK) j.status = sticky;
                if (j.status == VGPU_E_ADMISSION) ++rejected;
                j.t_done = now_ns();
                j.state = J_DONE;
                pinned.release(j.pin_blob, j.pin_blob_cls);
                pinned.release(j.pin_in, j.pin_in_cls);
                j.pin_blob = j.pin_in = nullptr;
                arrival.pop_front();
                cv_done.notify_all();
                continue;
            }
            if (running.size() >= cfg.max_inflight || !arena.alloc(j.chunk, &j.off)) {
                blocked = true;
                break;
            }
            j.holds_chunk = true;
            if (free_events.empty()) {
                hipEvent_t e = nullptr;
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                    (void)hipGetLastError();
                    free_chunk(j);
                    j.status = VGPU_E_DEVICE;
                    continue;
                }
                free_events.push_back(e);
            }
            j.done = free_events.back();
            free_events.pop_back();
            j.stream = (int)(next_stream++ % streams.size());
            j.t_admit = now_ns();
            j.state = J_RUNNING;
            if (!launch_job(j)) {
                (void)hipGetLastError();
                j.status = VGPU_E_DEVICE;
                sticky = VGPU_E_DEVICE;
                (void)hipStreamSynchronize(streams[(size_t)j.stream]);
                free_chunk(j);
                finish_job(j);
                arrival.pop_front();
                cv_done.notify_all();
                continue;
            }
            running.push_back(arrival.front());
            peak_inflight = std::max<uint64_t>(peak_inflight, running.size());
            arrival.pop_front();
            last = now_ns();
        }
        if (stop && arrival.empty() && running.empty() && to_plan.empty()) return;
        if (running.empty() && now_ns() - last >= kSpinNs) {
            // Idle for a while: sleep until a request is planned or released.
            cv_sched.wait(g);
            last = now_ns();
        } else {
            // Requests on the device, or recently: poll without holding the lock long.
            g.unlock();
            std::this_thread::yield();
            g.lock();
        }
    }
}

namespace {
std::shared_ptr<Job> find_job(vgpu_server* s, uint64_t id) {
    auto it = s->jobs.find(id);
    return it == s->jobs.end() ? nullptr : it->second;
}

void fill_result(const Job& j, vgpu_request_result* r) {
    r->status = j.status;
    r->width = j.ow;
    r->height = j.oh;
    r->bands = j.ob;
    r->format = j.ofmt;
    r->out_bytes = j.out_bytes;
    r->arena_bytes = j.chunk;
    r->plan_ns = j.t_planned ? j.t_planned - j.t_arrival : 0;
    r->queue_ns = j.t_admit ? j.t_admit - j.t_arrival : 0;
    r->run_ns = j.t_admit ? j.t_done - j.t_admit : 0;
    r->latency_ns = j.t_done - j.t_arrival;
}
}  // namespace

extern "C" {

int vgpu_server_create(vgpu_context* ctx, const vgpu_server_config* cfg, vgpu_server** out) {
    if (!ctx || !cfg || !out) return VGPU_E_INVALID_ARGUMENT;
    *out = nullptr;
    if (cfg->hbm_budget == 0 || cfg->max_inflight == 0 || cfg->streams == 0 || cfg->streams > 16 ||
        cfg->planner_threads == 0 || cfg->planner_threads > 64)
        return VGPU_E_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    auto* s = new (std::nothrow) vgpu_server();
    if (!s) return VGPU_E_NO_MEMORY;
    s->ctx = ctx;
    s->cfg = *cfg;
    int rc = s->arena.init(cfg->hbm_budget);
    if (rc != VGPU_OK) {
        delete s;
        return rc;
    }
    for (uint32_t i = 0; i < cfg->streams; ++i) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
            delete s;
            return VGPU_E_DEVICE;
        }
        s->streams.push_back(st);
    }
    for (uint32_t i = 0; i < cfg->planner_threads; ++i) s->planners.emplace_back([s, i] { s->planner_loop((int)i); });
    s->scheduler = std::thread([s] { s->scheduler_loop(); });
    *out = s;
    return VGPU_OK;
}

void vgpu_server_destroy(vgpu_server* s) {
    if (!s) return;
    {
        // Outputs nobody will read any more: free them now, and those of
        // requests still running when they finish, so queued requests can
        // still be admitted and drained.
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        for (auto& kv : s->jobs) {
            kv.second->auto_release = 1;
            if (kv.second->state == J_DONE) s->free_chunk(*kv.second);
        }
    }
    s->cv_plan.notify_all();
    s->cv_sched.notify_all();
    for (auto& t : s->planners) t.join();
    s->cv_sched.notify_all();
    s->scheduler.join();
    (void)hipSetDevice(s->ctx->device);
    for (hipStream_t x : s->streams) {
        (void)hipStreamSynchronize(x);
        (void)hipStreamDestroy(x);
    }
    for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
    for (auto& kv : s->jobs) {
        Job& j = *kv.second;
        if (j.done) (void)hipEventDestroy(j.done);
        s->pinned.release(j.pin_blob, j.pin_blob_cls);
        s->pinned.release(j.pin_in, j.pin_in_cls);
    }
    delete s;
}

int vgpu_server_submit(vgpu_server* s, const vgpu_request* r, int auto_release, uint64_t* id) {
    if (!s || !r || !id || !r->src) return VGPU_E_INVALID_ARGUMENT;
    if (r->width == 0 || r->height == 0 || r->bands == 0 || r->format < U8 || r->format > F32)
        return VGPU_E_INVALID_ARGUMENT;
    auto j = std::make_shared<Job>();
    j->req = *r;
    j->auto_release = auto_release;
    j->t_arrival = r->arrival_ns ? r->arrival_ns : now_ns();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return VGPU_E_INVALID_ARGUMENT;
        j->id = s->next_id++;
        *id = j->id;
        s->jobs[j->id] = j;
        s->to_plan.push_back(j);
        s->arrival.push_back(j);
    }
    s->cv_plan.notify_one();
    return VGPU_OK;
}

int vgpu_server_wait(vgpu_server* s, uint64_t id, int64_t timeout_us, vgpu_request_result* res) {
    if (!s || !res) return VGPU_E_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> g(s->mu);
    std::shared_ptr<Job> j = find_job(s, id);
    if (!j) return VGPU_E_INVALID_ARGUMENT;
    auto ready = [&] { return j->state == J_DONE; };
    if (timeout_us < 0) {
        s->cv_done.wait(g, ready);
    } else if (!s->cv_done.wait_for(g, std::chrono::microseconds(timeout_us), ready)) {
        return VGPU_E_NOT_READY;
    }
    fill_result(*j, res);
    return VGPU_OK;
}

int vgpu_server_output(vgpu_server* s, uint64_t id, const void** dev_ptr, size_t* bytes) {
    if (!s || !dev_ptr || !bytes) return VGPU_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    std::shared_ptr<Job> j = find_job(s, id);
    if (!j || j->state != J_DONE) return VGPU_E_INVALID_ARGUMENT;
    if (j->status != VGPU_OK) return j->status;
    if (!j->holds_chunk) return VGPU_E_INVALID_ARGUMENT;  // released
    size_t o = (j->req.src_on_device ? 0 : align_up(j->in_bytes)) + align_up(j->blob_bytes);
    for (size_t b : j->inter_bytes) o += align_up(b);
    *dev_ptr = s->arena.at(j->off + o);
    *bytes = j->out_bytes;
    return VGPU_OK;
}

int vgpu_server_fetch(vgpu_server* s, uint64_t id, void* host_dst, size_t bytes) {
    const void* p = nullptr;
    size_t n = 0;
    int rc = vgpu_server_output(s, id, &p, &n);
    if (rc != VGPU_OK) return rc;
    if (!host_dst || bytes != n) return VGPU_E_BUFFER_SIZE;
    (void)hipSetDevice(s->ctx->device);
    return from_hip(hipMemcpy(host_dst, p, n, hipMemcpyDeviceToHost));
}

int vgpu_server_cancel(vgpu_server* s, uint64_t id) {
    if (!s) return VGPU_E_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> g(s->mu);
        std::shared_ptr<Job> j = find_job(s, id);
        if (!j || j->state == J_DONE) return VGPU_E_INVALID_ARGUMENT;
        j->cancelled = true;
    }
    s->cv_sched.notify_one();
    return VGPU_OK;
}

int vgpu_server_release(vgpu_server* s, uint64_t id) {
    if (!s) return VGPU_E_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> g(s->mu);
        std::shared_ptr<Job> j = find_job(s, id);
        if (!j || j->state != J_DONE) return VGPU_E_INVALID_ARGUMENT;
        s->free_chunk(*j);
        s->jobs.erase(id);
    }
    s->cv_sched.notify_one();
    return VGPU_OK;
}

int vgpu_server_stats(vgpu_server* s, vgpu_server_counters* st) {
    if (!s || !st) return VGPU_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    st->completed = s->completed;
    st->rejected = s->rejected;
    st->inflight = s->running.size();
    st->peak_inflight = s->peak_inflight;
    st->queued = s->arrival.size();
    st->arena_bytes = s->arena.capacity();
    st->arena_used = s->arena.used();
    st->arena_peak = s->arena.peak();
    return s->sticky;
}

// Open-loop trace: request i is submitted at start + arrival_offset_ns[i]
// (busy-waiting the last stretch), outputs are released on completion, and
// latency_ns[i] = completion time - scheduled arrival time.
int vgpu_server_replay(vgpu_server* s, const vgpu_request* reqs, uint32_t n, const int64_t* arrival_offset_ns,
                       int64_t* latency_ns, int32_t* status, int64_t* wall_ns, int64_t* plan_ns,
                       int64_t* queue_ns) {
    if (!s || (!reqs && n) || (!arrival_offset_ns && n) || (!latency_ns && n) || (!status && n))
        return VGPU_E_INVALID_ARGUMENT;
    std::vector<uint64_t> ids(n);
    const int64_t t0 = now_ns();
    for (uint32_t i = 0; i < n; ++i) {
        const int64_t due = t0 + arrival_offset_ns[i];
        for (;;) {
            const int64_t left = due - now_ns();
            if (left <= 0) break;
            if (left > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(left - 100000));
        }
        vgpu_request r = reqs[i];
        r.arrival_ns = due;
        int rc = vgpu_server_submit(s, &r, 1, &ids[i]);
        if (rc != VGPU_OK) return rc;
    }
    int64_t last = t0;
    for (uint32_t i = 0; i < n; ++i) {
        vgpu_request_result res;
        int rc = vgpu_server_wait(s, ids[i], -1, &res);
        if (rc != VGPU_OK) return rc;
        latency_ns[i] = res.latency_ns;
        status[i] = res.status;
        if (plan_ns) plan_ns[i] = res.plan_ns;
        if (queue_ns) queue_ns[i] = res.queue_ns;
        last = std::max(last, (int64_t)(arrival_offset_ns[i] + t0 + res.latency_ns));
        std::lock_guard<std::mutex> g(s->mu);
        s->jobs.erase(ids[i]);
    }
    if (wall_ns) *wall_ns = last - t0;
    return VGPU_OK;
}

int vgpu_host_alloc(size_t bytes, void** out) {
    if (!out || bytes == 0) return VGPU_E_INVALID_ARGUMENT;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return VGPU_E_NO_MEMORY;
    }
    return VGPU_OK;
}

void vgpu_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

}  // extern "C"
