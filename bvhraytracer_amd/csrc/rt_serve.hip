// Serving loop: resident scenes, admission by device-memory footprint,
// continuous batching into batched persistent launches (rt_batch_trace).
//
// Threads: callers submit/wait; one dispatcher thread owns all HIP work
// (H2D of admitted rays, launches, D2H of results, retiring finished
// launches).  Device memory for request rays and results comes from one
// pool of hbm_cap_bytes (first-fit free list); a request is admitted when
// its region fits, strictly in arrival order.  The host copies between the
// callers' buffers and the pinned mirror of the pool run without the server
// mutex, so submit and wait never wait behind a large request's copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host_common.h"

using namespace rt;
using Clock = std::chrono::steady_clock;

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

enum ReqState { kQueued = 0, kStaging, kAdmitted, kLaunched, kRetiring, kDone, kFailed, kCancelled };

struct Request {
    uint64_t id = 0;
    uint32_t scene = 0, n = 0, flags = 0;
    const rt_ray* rays = nullptr;
    void* out = nullptr;
    bool device = false;  // rays and out in device memory: no pool region, no host copies
    size_t rayBytes = 0, outBytes = 0, bytes = 0;
    size_t offset = 0;  // pool offset (rays at offset, results at offset + rayBytes)
    int state = kQueued;
    int error = RT_OK;
    Clock::time_point tSubmit, tDone;
};

// first-fit allocator over [0, size)
struct Pool {
    size_t size = 0;
    std::map<size_t, size_t> freeList;  // offset -> length
    void init(size_t s) {
        size = s;
        freeList.clear();
        freeList[0] = s;
    }
    bool alloc(size_t len, size_t& off) {
        for (auto it = freeList.begin(); it != freeList.end(); ++it) {
            if (it->second >= len) {
                off = it->first;
                const size_t rest = it->second - len;
                freeList.erase(it);
                if (rest) freeList[off + len] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t len) {
        auto it = freeList.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != freeList.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            freeList.erase(nx);
        }
        if (it != freeList.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                freeList.erase(it);
            }
        }
    }
};

// Host copies between the callers' buffers and the pinned pool, spread over
// worker threads in chunks: one thread moves on the order of 10 GB/s, and a
// saturated server moves 40 bytes per ray in and up to 24 out at over a
// billion rays per second.
class CopyPool {
public:
    struct Job {
        void* dst;
        const void* src;
        size_t len;
    };
    void start(unsigned n) {
        for (unsigned i = 0; i < n; ++i) threads_.emplace_back([this] { work(); });
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
        threads_.clear();
    }
    // Performs every copy and returns when all are done (the calling thread
    // takes chunks too).  Called by one thread at a time.
    void run(const std::vector<Job>& jobs) {
        size_t total = 0;
        for (const Job& j : jobs) total += j.len;
        if (threads_.empty() || total < kSmall) {
            for (const Job& j : jobs) std::memcpy(j.dst, j.src, j.len);
            return;
        }
        const size_t chunk = std::max(kMinChunk, total / (4 * (threads_.size() + 1)));
        {
            std::lock_guard<std::mutex> lk(mu_);
            chunks_.clear();
            for (const Job& j : jobs)
                for (size_t o = 0; o < j.len; o += chunk)
                    chunks_.push_back({static_cast<char*>(j.dst) + o, static_cast<const char*>(j.src) + o,
                                       std::min(chunk, j.len - o)});
            next_ = 0;
            pending_ = chunks_.size();
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> lk(mu_);
        doneCv_.wait(lk, [this] { return pending_ == 0; });
    }

private:
    static constexpr size_t kSmall = 256 << 10, kMinChunk = 1 << 20;
    // takes chunks until none is left
    void drain() {
        while (true) {
            Job c;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (next_ >= chunks_.size()) return;
                c = chunks_[next_++];
            }
            std::memcpy(c.dst, c.src, c.len);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) doneCv_.notify_all();
        }
    }
    void work() {
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return quit_ || next_ < chunks_.size(); });
                if (quit_) return;
            }
            drain();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, doneCv_;
    std::vector<Job> chunks_;
    size_t next_ = 0, pending_ = 0;
    bool quit_ = false;
    std::vector<std::thread> threads_;
};

struct Slot {
    hipStream_t stream = nullptr;
    rt_batch_t batch = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Request*> reqs;
    bool busy = false;
};

}  // namespace

struct rt_server {
    rt_server_config cfg{};
    int device = 0;
    std::vector<rt_scene_t> scenes;
    char* dPool = nullptr;  // device pool
    char* hPool = nullptr;  // pinned mirror (staging)
    Pool pool;
    size_t admittedBytes = 0, peakBytes = 0;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv;      // wakes the dispatcher
    std::condition_variable doneCv;  // wakes waiters
    std::deque<Request*> queue;      // submitted, not admitted
    std::deque<Request*> admitted;   // admitted, not launched
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t nextId = 1;
    bool stop = false;
    std::thread worker;
    CopyPool copies;  // used by the dispatcher only

    // stats since the last reset
    std::vector<double> lat;
    uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0, cancelled = 0;
    uint64_t raysDone = 0, launches = 0, launchedReqs = 0;
    bool haveFirst = false;
    Clock::time_point tFirst, tLast;
};

namespace {

// A finished launch's requests, with the launch's status.
struct Retired {
    Request* r;
    int err;
};

// Frees a slot whose launch ended and moves its requests to `out`; their
// results still sit in the pinned pool (called with mu held).
void detach(Slot& sl, int err, std::vector<Retired>& out) {
    for (Request* r : sl.reqs) {
        r->state = kRetiring;
        out.push_back({r, err});
    }
    sl.reqs.clear();
    sl.busy = false;
}

// Completes detached requests after their results were copied out (called
// with mu held).
void complete(rt_server* sv, std::vector<Retired>& done) {
    const auto now = Clock::now();
    for (const Retired& d : done) {
        Request* r = d.r;
        r->tDone = now;
        r->state = d.err == RT_OK ? kDone : kFailed;
        r->error = d.err;
        if (!r->device) {
            sv->pool.release(r->offset, r->bytes);
            sv->admittedBytes -= r->bytes;
        }
        if (d.err == RT_OK) {
            ++sv->completed;
            sv->raysDone += r->n;
            sv->lat.push_back(std::chrono::duration<double, std::milli>(now - r->tSubmit).count());
            sv->tLast = now;
        } else {
            ++sv->failed;
        }
    }
    done.clear();
}

int launch(rt_server* sv, Slot& sl, std::vector<Request*>& batch) {
    // called without mu held; the requests are owned by the dispatcher
    std::vector<rt_batch_item> items;
    items.reserve(batch.size());
    for (Request* r : batch) {
        rt_batch_item it{};
        it.scene = sv->scenes[r->scene];
        if (r->device) {
            it.rays = r->rays;
            it.out = r->out;
        } else {
            char* d = sv->dPool + r->offset;
            if (hipMemcpyAsync(d, sv->hPool + r->offset, r->rayBytes, hipMemcpyHostToDevice, sl.stream) != hipSuccess)
                return set_error(RT_ERR_HIP, "rt_server: ray upload failed");
            it.rays = reinterpret_cast<const rt_ray*>(d);
            it.out = d + r->rayBytes;
        }
        it.counts = nullptr;
        it.n = r->n;
        it.flags = r->flags & RT_ANY;
        items.push_back(it);
    }
    const uint32_t fl = RT_NO_SYNC | (sv->cfg.use_graph ? RT_GRAPH : 0u);
    int rc = rt_batch_trace(sl.batch, items.data(), (uint32_t)items.size(), fl, sl.stream);
    if (rc != RT_OK) return rc;
    for (Request* r : batch) {
        if (r->device) continue;
        if (hipMemcpyAsync(sv->hPool + r->offset + r->rayBytes, sv->dPool + r->offset + r->rayBytes, r->outBytes,
                           hipMemcpyDeviceToHost, sl.stream) != hipSuccess)
            return set_error(RT_ERR_HIP, "rt_server: result download failed");
    }
    if (hipEventRecord(sl.done, sl.stream) != hipSuccess) return set_error(RT_ERR_HIP, "rt_server: event record failed");
    return RT_OK;
}

void dispatcher(rt_server* sv) {
    (void)hipSetDevice(sv->device);
    std::vector<Retired> finished;
    std::vector<Request*> staging;
    std::vector<CopyPool::Job> jobs;
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        bool progress = false;
        // 1. detach finished launches (their slots are free again)
        for (Slot& sl : sv->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            detach(sl, q == hipSuccess ? RT_OK : RT_ERR_HIP, finished);
            progress = true;
        }
        // 2. admit in arrival order while the footprint fits
        while (!sv->queue.empty()) {
            Request* r = sv->queue.front();
            if (r->device) {
                // the caller's device buffers: nothing to allocate or copy
                sv->queue.pop_front();
                r->state = kStaging;
                staging.push_back(r);
                progress = true;
                continue;
            }
            size_t off = 0;
            if (!sv->pool.alloc(r->bytes, off)) break;
            sv->queue.pop_front();
            r->offset = off;
            r->state = kStaging;
            sv->admittedBytes += r->bytes;
            sv->peakBytes = std::max(sv->peakBytes, sv->admittedBytes);
            staging.push_back(r);
            progress = true;
        }
        // copies between the callers' buffers and the pinned pool, without
        // the lock: results out of finished launches, rays of new admissions
        // into their regions (the dispatcher alone touches these requests
        // until they are completed or queued for a launch)
        if (!finished.empty() || !staging.empty()) {
            lk.unlock();
            jobs.clear();
            for (const Retired& d : finished)
                if (d.err == RT_OK && !d.r->device)
                    jobs.push_back({d.r->out, sv->hPool + d.r->offset + d.r->rayBytes, d.r->outBytes});
            for (Request* r : staging)
                if (!r->device) jobs.push_back({sv->hPool + r->offset, r->rays, (size_t)r->n * sizeof(rt_ray)});
            sv->copies.run(jobs);
            lk.lock();
            if (!finished.empty()) {
                complete(sv, finished);
                sv->doneCv.notify_all();
            }
            for (Request* r : staging) {
                r->state = kAdmitted;
                sv->admitted.push_back(r);
            }
            staging.clear();
        }
        // 3. coalesce everything admitted into a launch on a free slot
        for (Slot& sl : sv->slots) {
            if (sv->admitted.empty()) break;
            if (sl.busy) continue;
            std::vector<Request*> batch;
            while (!sv->admitted.empty() && batch.size() < sv->cfg.max_items) {
                Request* r = sv->admitted.front();
                sv->admitted.pop_front();
                r->state = kLaunched;
                batch.push_back(r);
            }
            sl.busy = true;
            sl.reqs = batch;
            ++sv->launches;
            sv->launchedReqs += batch.size();
            lk.unlock();
            const int rc = launch(sv, sl, batch);
            lk.lock();
            if (rc != RT_OK) {
                (void)hipStreamSynchronize(sl.stream);
                detach(sl, rc, finished);
                complete(sv, finished);
                sv->doneCv.notify_all();
            }
            progress = true;
        }
        if (sv->stop && sv->queue.empty() && sv->admitted.empty()) {
            bool busy = false;
            for (Slot& sl : sv->slots) busy |= sl.busy;
            if (!busy) break;
        }
        if (!progress) {
            bool busy = false;
            for (Slot& sl : sv->slots) busy |= sl.busy;
            // poll launches quickly; sleep longer when idle
            sv->cv.wait_for(lk, busy ? std::chrono::microseconds(20) : std::chrono::milliseconds(2));
        }
    }
}

void destroy(rt_server* sv) {
    if (sv->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->stop = true;
        }
        sv->cv.notify_all();
        sv->worker.join();
    }
    sv->copies.stop();
    for (Slot& sl : sv->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.batch) rt_batch_release(sl.batch);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (sv->dPool) (void)hipFree(sv->dPool);
    if (sv->hPool) (void)hipHostFree(sv->hPool);
    for (auto& kv : sv->reqs) delete kv.second;
    delete sv;
}

}  // namespace

extern "C" int rt_server_create(const rt_server_config* cfg, rt_server_t* out) {
    if (!out) return set_error(RT_ERR_ARG, "rt_server_create: null output");
    rt_server* sv = new (std::nothrow) rt_server;
    if (!sv) return set_error(RT_ERR_NOMEM, "rt_server_create: out of host memory");
    if (cfg) sv->cfg = *cfg;
    if (!sv->cfg.hbm_cap_bytes) sv->cfg.hbm_cap_bytes = 4ull << 30;
    if (!sv->cfg.slots) sv->cfg.slots = 2;
    if (!sv->cfg.max_items) sv->cfg.max_items = 512;
    if (!sv->cfg.copy_threads) sv->cfg.copy_threads = 8;
    if (sv->cfg.copy_threads > 64) sv->cfg.copy_threads = 64;
    sv->cfg.hbm_cap_bytes = align_up(sv->cfg.hbm_cap_bytes);
    (void)hipGetDevice(&sv->device);
    if (hipMalloc(reinterpret_cast<void**>(&sv->dPool), sv->cfg.hbm_cap_bytes) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sv->hPool), sv->cfg.hbm_cap_bytes, hipHostMallocDefault) != hipSuccess) {
        destroy(sv);
        return set_error(RT_ERR_NOMEM, "rt_server_create: cannot allocate the request pool");
    }
    // Touch both pools once now: the first writes to fresh pinned pages and
    // the first copies into fresh device pages are slow, and a request
    // should not pay for them (measured: a new server served about 150
    // Mrays/s until its pool had been used once, about 870 after).
    {
        const size_t cap = sv->cfg.hbm_cap_bytes;
        const unsigned nt = std::max(1u, sv->cfg.copy_threads);
        std::vector<std::thread> fill;
        for (unsigned i = 0; i < nt; ++i)
            fill.emplace_back([sv, cap, nt, i] {
                const size_t a = cap / nt * i, b = (i + 1 == nt) ? cap : cap / nt * (i + 1);
                std::memset(sv->hPool + a, 0, b - a);
            });
        for (auto& t : fill) t.join();
        if (hipMemset(sv->dPool, 0, cap) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            destroy(sv);
            return set_error(RT_ERR_HIP, "rt_server_create: cannot initialise the request pool");
        }
    }
    sv->pool.init(sv->cfg.hbm_cap_bytes);
    sv->slots.resize(sv->cfg.slots);
    for (Slot& sl : sv->slots) {
        int rc;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            destroy(sv);
            return set_error(RT_ERR_HIP, "rt_server_create: stream/event creation failed");
        }
        if ((rc = rt_batch_create(sv->cfg.max_items, &sl.batch)) != RT_OK ||
            (rc = rt_batch_set_launch(sl.batch, sv->cfg.block_size, sv->cfg.stack_window)) != RT_OK) {
            destroy(sv);
            return rc;
        }
    }
    // the dispatcher thread copies too, so copy_threads - 1 helpers
    sv->copies.start(sv->cfg.copy_threads - 1);
    sv->worker = std::thread(dispatcher, sv);
    *out = sv;
    return RT_OK;
}

extern "C" int rt_server_add_scene(rt_server_t sv, rt_scene_t scene, uint32_t* scene_id) {
    if (!sv || !scene || !scene_id) return set_error(RT_ERR_ARG, "rt_server_add_scene: null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->scenes.push_back(scene);
    *scene_id = (uint32_t)sv->scenes.size() - 1;
    return RT_OK;
}

extern "C" int rt_server_submit(rt_server_t sv, uint32_t scene_id, const rt_ray* rays, uint32_t n, uint32_t flags,
                                void* out, uint64_t* request_id) {
    if (!sv || !request_id) return set_error(RT_ERR_ARG, "rt_server_submit: null argument");
    if (n == 0 || !rays || !out) return set_error(RT_ERR_ARG, "rt_server_submit: empty request");
    if (flags & ~(RT_ANY | RT_DEVICE_PTRS))
        return set_error(RT_ERR_UNSUPPORTED, "rt_server_submit: RT_CLOSEST or RT_ANY, optionally RT_DEVICE_PTRS");
    Request* r = new (std::nothrow) Request;
    if (!r) return set_error(RT_ERR_NOMEM, "rt_server_submit: out of host memory");
    r->scene = scene_id;
    r->n = n;
    r->flags = flags;
    r->rays = rays;
    r->out = out;
    r->device = (flags & RT_DEVICE_PTRS) != 0;
    r->rayBytes = align_up((size_t)n * sizeof(rt_ray));
    r->outBytes = (flags & RT_ANY) ? (size_t)n : (size_t)n * sizeof(rt_hit);
    r->bytes = r->device ? 0 : r->rayBytes + align_up(r->outBytes);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (scene_id >= sv->scenes.size()) {
        delete r;
        return set_error(RT_ERR_ARG, "rt_server_submit: unknown scene id");
    }
    if (r->bytes > sv->cfg.hbm_cap_bytes) {
        delete r;
        ++sv->rejected;
        return set_error(RT_ERR_BUSY, "rt_server_submit: request exceeds the admission cap");
    }
    r->id = sv->nextId++;
    r->tSubmit = Clock::now();
    if (!sv->haveFirst) {
        sv->haveFirst = true;
        sv->tFirst = r->tSubmit;
    }
    sv->reqs[r->id] = r;
    sv->queue.push_back(r);
    ++sv->submitted;
    *request_id = r->id;
    sv->cv.notify_all();
    return RT_OK;
}

extern "C" int rt_server_wait(rt_server_t sv, uint64_t id, double timeout_s) {
    if (!sv) return set_error(RT_ERR_ARG, "rt_server_wait: null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return set_error(RT_ERR_ARG, "rt_server_wait: unknown request id");
    Request* r = it->second;
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                              std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s));
    while (r->state != kDone && r->state != kFailed && r->state != kCancelled) {
        if (sv->doneCv.wait_until(lk, deadline) == std::cv_status::timeout &&
            r->state != kDone && r->state != kFailed && r->state != kCancelled)
            return set_error(RT_ERR_BUSY, "rt_server_wait: timeout");
    }
    const int st = r->state, err = r->error;
    sv->reqs.erase(it);
    delete r;
    if (st == kCancelled) return set_error(RT_ERR_ARG, "rt_server_wait: request was cancelled");
    if (st == kFailed) return set_error(err ? err : RT_ERR_HIP, "rt_server_wait: the request's launch failed");
    return RT_OK;
}

extern "C" int rt_server_cancel(rt_server_t sv, uint64_t id) {
    if (!sv) return set_error(RT_ERR_ARG, "rt_server_cancel: null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return set_error(RT_ERR_ARG, "rt_server_cancel: unknown request id");
    Request* r = it->second;
    auto q = std::find(sv->queue.begin(), sv->queue.end(), r);
    if (q != sv->queue.end()) {
        sv->queue.erase(q);
    } else {
        auto a = std::find(sv->admitted.begin(), sv->admitted.end(), r);
        if (a == sv->admitted.end()) return set_error(RT_ERR_BUSY, "rt_server_cancel: request already being staged or launched");
        sv->admitted.erase(a);
        if (!r->device) {
            sv->pool.release(r->offset, r->bytes);
            sv->admittedBytes -= r->bytes;
        }
    }
    r->state = kCancelled;
    ++sv->cancelled;
    sv->doneCv.notify_all();
    return RT_OK;
}

extern "C" int rt_server_drain(rt_server_t sv) {
    if (!sv) return set_error(RT_ERR_ARG, "rt_server_drain: null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        bool busy = !sv->queue.empty() || !sv->admitted.empty();
        for (Slot& sl : sv->slots) busy |= sl.busy;
        if (!busy) return RT_OK;
        sv->doneCv.wait_for(lk, std::chrono::milliseconds(5));
    }
}

extern "C" int rt_server_get_stats(rt_server_t sv, rt_server_stats* o, int reset) {
    if (!sv || !o) return set_error(RT_ERR_ARG, "rt_server_get_stats: null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    std::memset(o, 0, sizeof(*o));
    o->submitted = sv->submitted;
    o->completed = sv->completed;
    o->failed = sv->failed;
    o->rejected = sv->rejected;
    o->cancelled = sv->cancelled;
    o->rays_completed = sv->raysDone;
    o->launches = sv->launches;
    o->peak_admitted_bytes = sv->peakBytes;
    o->mean_requests_per_launch = sv->launches ? (double)sv->launchedReqs / sv->launches : 0.0;
    if (!sv->lat.empty()) {
        std::vector<double> v = sv->lat;
        std::sort(v.begin(), v.end());
        auto q = [&](double f) { return v[std::min(v.size() - 1, (size_t)(f * (v.size() - 1) + 0.5))]; };
        o->p50_ms = q(0.50);
        o->p99_ms = q(0.99);
        o->max_ms = v.back();
        double s = 0;
        for (double x : v) s += x;
        o->mean_ms = s / v.size();
    }
    if (sv->haveFirst && sv->completed) {
        o->wall_s = std::chrono::duration<double>(sv->tLast - sv->tFirst).count();
        if (o->wall_s > 0) o->mrays_per_s = sv->raysDone / o->wall_s / 1e6;
    }
    if (reset) {
        sv->lat.clear();
        sv->submitted = sv->completed = sv->failed = sv->rejected = sv->cancelled = 0;
        sv->raysDone = sv->launches = sv->launchedReqs = 0;
        sv->haveFirst = false;
        sv->peakBytes = sv->admittedBytes;
    }
    return RT_OK;
}

extern "C" int rt_server_destroy(rt_server_t sv) {
    if (!sv) return RT_OK;
    destroy(sv);
    return RT_OK;
}
