// Native continuous-batching loop for translation requests on one GPU
// (sfmm_server_* in solidfmm_gpu.h).  The reference has no counterpart: its
// multithreaded_translator (translations.hpp:47-87) serialises whole
// batches with a mutex.
//
// A request is one item list of one operator plus the cells it reads and
// writes.  The server keeps the multipole and local pools resident and runs
// requests concurrently on `slots` HIP streams:
//   arrival      sfmm_server_submit copies the request and queues it, in
//                arrival order, and hands it to a preparation thread;
//   preparation  sfmm_plan_prepare (validation and grouping on the host)
//                on one of `workers` threads, while the GPU runs others;
//   admission    the oldest waiting request starts when it is prepared, a
//                slot is free, its device memory fits in the arena (first
//                fit, 256-byte blocks), and no in-flight request writes a
//                cell it reads or reads or writes a cell it writes.  A
//                blocked head blocks the queue, so every cell sees its
//                updates in arrival order and the pools end as if the
//                requests had run one by one;
//   execution    sfmm_plan_bind into the arena and sfmm_plan_execute on the
//                slot's stream, between two events (no allocation, no wait);
//   completion   a scheduler thread polls the events, releases the slot,
//                the arena range and the hazards, and records the request's
//                completion time on the GPU clock.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include <sys/resource.h>
#include <sys/syscall.h>
#include <unistd.h>

#include "../../include/solidfmm_gpu.h"

// Library-internal (capi.cpp): sfmm_plan_execute without branch streams.
extern "C" __attribute__((visibility("hidden"))) int sfmm_plan_execute_chain(sfmm_plan *plan, const double *src_pool,
                                                                           size_t src_len, double *dst_pool,
                                                                           size_t dst_len, void *stream);

namespace {

constexpr size_t kAlign = 256;
constexpr int kPreparerNice = 5;  // niceness of the preparation threads

struct Request {
    uint64_t id = 0;
    int op = 0;
    std::vector<sfmm_item> items;
    std::vector<uint32_t> src_cells, dst_cells;  // sorted, unique
    double arrival_ms = 0;                       // host clock since the server's start
    double admitted_ms = -1;
    double done_ms = -1;  // GPU clock since the server's start
    double noticed_ms = -1;  // host clock: when the loop saw the completion
    double gpu_ms = -1;
    int rc = SFMM_OK;
    bool prepared = false;
    bool finished = false;
    sfmm_plan *plan = nullptr;
    int slot = -1;
    size_t arena_off = 0, arena_len = 0;
    bool staged = false;  // tables copied into the pinned pool by the preparer
    size_t host_off = 0, host_len = 0;
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
};

// First-fit allocator over [0, size).
class Arena {
public:
    explicit Arena(size_t size) : size_(size) { free_[0] = size; }
    bool alloc(size_t want, size_t &off, size_t &len)
    {
        len = std::max(kAlign, (want + kAlign - 1) / kAlign * kAlign);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            off = it->first;
            const size_t rest = it->second - len;
            free_.erase(it);
            if (rest) free_[off + len] = rest;
            used_ += len;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t len)
    {
        used_ -= len;
        auto it = free_.emplace(off, len).first;
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
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t size() const { return size_; }

private:
    size_t size_, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

// pools read and written per operator: 0 multipole, 1 local
int src_pool_of(int op) { return op == SFMM_L2L ? 1 : 0; }
int dst_pool_of(int op) { return op == SFMM_M2M ? 0 : 1; }

}  // namespace

struct sfmm_server {
    sfmm_context *ctx = nullptr;
    double *pool[2] = {nullptr, nullptr};
    size_t pool_len[2] = {0, 0};
    size_t ncells = 0;
    int device = 0;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> slot_start, slot_end;  // per slot, reused by its requests
    std::vector<int> free_slots;
    void *arena_mem = nullptr;
    unsigned char *host_pool = nullptr;  // pinned: plan tables pass through it
    std::unique_ptr<Arena> arena, host_arena;
    std::vector<int32_t> readers[2], writers[2];
    hipEvent_t ref = nullptr;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work;  // scheduler: something to do
    std::condition_variable cv_prep;  // preparers: queue not empty
    std::condition_variable cv_done;  // drain: a request finished
    std::deque<Request *> waiting;    // arrival order
    std::deque<Request *> to_prepare;
    std::deque<Request *> future;     // submitted ahead of their arrival time, by arrival
    std::vector<Request *> inflight;
    std::vector<std::unique_ptr<Request>> all;
    size_t unfinished = 0;
    bool stop = false;
    size_t max_inflight = 0;
    uint64_t blocked[4] = {0, 0, 0, 0};  // prepare, slot, arena, hazard
    std::vector<std::thread> preparers;
    std::thread scheduler;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    bool conflict(const Request &r) const
    {
        const int sp = src_pool_of(r.op), dp = dst_pool_of(r.op);
        for (uint32_t c : r.dst_cells)
            if (writers[dp][c] || readers[dp][c]) return true;
        for (uint32_t c : r.src_cells)
            if (writers[sp][c]) return true;
        return false;
    }
    void hazards(const Request &r, int k)
    {
        const int sp = src_pool_of(r.op), dp = dst_pool_of(r.op);
        for (uint32_t c : r.src_cells) readers[sp][c] += k;
        for (uint32_t c : r.dst_cells) writers[dp][c] += k;
    }

    void finish_locked(Request &r)
    {
        float ms = 0;
        r.noticed_ms = now_ms();
        if (hipEventElapsedTime(&ms, ref, r.ev_end) == hipSuccess) r.done_ms = ms;
        if (hipEventElapsedTime(&ms, r.ev_start, r.ev_end) == hipSuccess) r.gpu_ms = ms;
        hazards(r, -1);
        arena->release(r.arena_off, r.arena_len);
        if (r.staged) host_arena->release(r.host_off, r.host_len);
        r.staged = false;
        free_slots.push_back(r.slot);
        sfmm_plan_destroy(r.plan);
        r.plan = nullptr;
        r.ev_start = r.ev_end = nullptr;
        r.finished = true;
        --unfinished;
        cv_done.notify_all();
    }

    // drop a request that cannot run (preparation or launch failed)
    void fail_locked(Request &r, int rc)
    {
        r.rc = rc;
        r.finished = true;
        if (r.staged) host_arena->release(r.host_off, r.host_len);
        r.staged = false;
        if (r.plan) sfmm_plan_destroy(r.plan);
        r.plan = nullptr;
        --unfinished;
        cv_done.notify_all();
    }

    bool launch_locked(Request &r)
    {
        const hipStream_t s = streams[size_t(r.slot)];
        const int sp = src_pool_of(r.op), dp = dst_pool_of(r.op);
        r.ev_start = slot_start[size_t(r.slot)];
        r.ev_end = slot_end[size_t(r.slot)];
        if (hipEventRecord(r.ev_start, s) != hipSuccess) return false;
        void *mem = static_cast<unsigned char *>(arena_mem) + r.arena_off;
        // tables staged by the preparer: only the asynchronous copy is left
        // (the pinned range stays reserved until completion); otherwise
        // stage them now, or copy from the plan's pageable memory
        if (!r.staged) {
            const size_t hb = sfmm_plan_host_bytes(r.plan);
            size_t off = 0, len = 0;
            if (hb && host_arena->alloc(hb, off, len)) {
                if (sfmm_plan_copy_tables(r.plan, host_pool + off, len) != SFMM_OK) {
                    host_arena->release(off, len);
                    return false;
                }
                r.staged = true;
                r.host_off = off;
                r.host_len = len;
            }
        }
        const int rc = r.staged ? sfmm_plan_bind_pinned(r.plan, mem, r.arena_len, host_pool + r.host_off, s)
                                : sfmm_plan_bind(r.plan, mem, r.arena_len, s);
        if (rc != SFMM_OK) return false;
        if (sfmm_plan_execute_chain(r.plan, pool[sp], pool_len[sp], pool[dp], pool_len[dp], s) != SFMM_OK)
            return false;
        return hipEventRecord(r.ev_end, s) == hipSuccess;
    }

    void schedule()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lock(mu);
        for (;;) {
            bool progressed = false;
            // requests whose arrival time has come join the queue
            const double now = now_ms();
            while (!future.empty() && future.front()->arrival_ms <= now) {
                Request *r = future.front();
                future.pop_front();
                waiting.push_back(r);
                to_prepare.push_back(r);
                cv_prep.notify_one();
                progressed = true;
            }
            for (size_t i = 0; i < inflight.size();) {
                Request &r = *inflight[i];
                if (hipEventQuery(r.ev_end) == hipSuccess) {
                    finish_locked(r);
                    inflight[i] = inflight.back();
                    inflight.pop_back();
                    progressed = true;
                } else {
                    ++i;
                }
            }
            while (!waiting.empty()) {
                Request &r = *waiting.front();
                if (!r.prepared) {
                    ++blocked[0];
                    break;
                }
                if (r.rc != SFMM_OK) {  // failed in preparation: drop it
                    waiting.pop_front();
                    fail_locked(r, r.rc);
                    progressed = true;
                    continue;
                }
                if (free_slots.empty()) {
                    ++blocked[1];
                    break;
                }
                if (conflict(r)) {
                    ++blocked[3];
                    break;
                }
                size_t off = 0, len = 0;
                if (!arena->alloc(sfmm_plan_device_bytes(r.plan), off, len)) {
                    if (inflight.empty()) {  // can never fit
                        waiting.pop_front();
                        fail_locked(r, SFMM_ERR_ALLOC);
                        progressed = true;
                        continue;
                    }
                    ++blocked[2];
                    break;
                }
                waiting.pop_front();
                r.slot = free_slots.back();
                free_slots.pop_back();
                r.arena_off = off;
                r.arena_len = len;
                r.admitted_ms = now_ms();
                hazards(r, +1);
                if (!launch_locked(r)) {
                    (void)hipStreamSynchronize(streams[size_t(r.slot)]);
                    hazards(r, -1);
                    arena->release(off, len);
                    free_slots.push_back(r.slot);
                    r.ev_start = r.ev_end = nullptr;
                    fail_locked(r, SFMM_ERR_HIP);
                } else {
                    inflight.push_back(&r);
                }
                progressed = true;
            }
            max_inflight = std::max(max_inflight, inflight.size());
            if (stop && waiting.empty() && inflight.empty() && future.empty()) return;
            if (progressed) continue;
            if (!inflight.empty()) {
                // kernels are running: poll again shortly.  The pause lasts
                // about the timer slack (50 us); polling without a pause, or
                // with a 1 us slack, measured slower, with a longer latency
                // tail: the event queries contend with the launches in the
                // HIP runtime
                lock.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(5));
                lock.lock();
            } else if (!future.empty()) {
                const double wait_ms = future.front()->arrival_ms - now_ms();
                if (wait_ms > 0)
                    cv_work.wait_for(lock, std::chrono::duration<double, std::milli>(std::min(wait_ms, 1.0)));
            } else {
                cv_work.wait(lock);
            }
        }
    }

    void prepare_loop()
    {
        // below the scheduler: when the preparers fill every core, a woken
        // scheduler otherwise waits up to a scheduler tick (milliseconds)
        // to notice completions and admit requests
        (void)setpriority(PRIO_PROCESS, static_cast<id_t>(syscall(SYS_gettid)), kPreparerNice);
        std::unique_lock<std::mutex> lock(mu);
        for (;;) {
            cv_prep.wait(lock, [&] { return stop || !to_prepare.empty(); });
            if (to_prepare.empty()) return;  // stopping
            Request *r = to_prepare.front();
            to_prepare.pop_front();
            lock.unlock();
            sfmm_plan *plan = nullptr;
            const int rc = sfmm_plan_prepare(ctx, r->op, r->items.data(), r->items.size(), &plan);
            // stage the tables in pinned memory here, off the scheduler
            size_t off = 0, len = 0;
            bool staged = false;
            const size_t hb = rc == SFMM_OK ? sfmm_plan_host_bytes(plan) : 0;
            if (hb) {
                lock.lock();
                staged = host_arena->alloc(hb, off, len);
                lock.unlock();
                if (staged && sfmm_plan_copy_tables(plan, host_pool + off, len) != SFMM_OK) {
                    lock.lock();
                    host_arena->release(off, len);
                    lock.unlock();
                    staged = false;
                }
            }
            lock.lock();
            r->plan = plan;
            r->rc = rc;
            r->staged = staged;
            r->host_off = off;
            r->host_len = len;
            r->prepared = true;
            cv_work.notify_one();
        }
    }
};

extern "C" {

int sfmm_server_create(sfmm_context *ctx, double *m_pool, size_t m_len, double *l_pool, size_t l_len, size_t ncells,
                       int slots, size_t arena_bytes, int workers, sfmm_server **out)
{
    if (!ctx || !out || slots < 1 || slots > 64 || arena_bytes < kAlign || workers < 1 || workers > 64)
        return SFMM_ERR_LOGIC;
    *out = nullptr;
    auto sv = std::unique_ptr<sfmm_server>(new (std::nothrow) sfmm_server);
    if (!sv) return SFMM_ERR_ALLOC;
    sv->ctx = ctx;
    sv->pool[0] = m_pool;
    sv->pool[1] = l_pool;
    sv->pool_len[0] = m_len;
    sv->pool_len[1] = l_len;
    sv->ncells = ncells;
    for (int p = 0; p < 2; ++p) {
        sv->readers[p].assign(ncells, 0);
        sv->writers[p].assign(ncells, 0);
    }
    sv->device = sfmm_device(ctx);
    hipError_t e = hipSetDevice(sv->device);
    for (int s = 0; s < slots && e == hipSuccess; ++s) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) {
            sv->streams.push_back(st);
            sv->free_slots.push_back(s);
            hipEvent_t a = nullptr, b = nullptr;
            e = hipEventCreate(&a);
            if (e == hipSuccess) sv->slot_start.push_back(a);
            if (e == hipSuccess) e = hipEventCreate(&b);
            if (e == hipSuccess) sv->slot_end.push_back(b);
        }
    }
    if (e == hipSuccess) e = hipMalloc(&sv->arena_mem, arena_bytes);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&sv->host_pool), arena_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&sv->ref);
    if (e == hipSuccess) e = hipEventRecord(sv->ref, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(sv->ref);
    if (e != hipSuccess) {
        for (hipStream_t st : sv->streams) (void)hipStreamDestroy(st);
        for (hipEvent_t ev : sv->slot_start) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : sv->slot_end) (void)hipEventDestroy(ev);
        if (sv->arena_mem) (void)hipFree(sv->arena_mem);
        if (sv->host_pool) (void)hipHostFree(sv->host_pool);
        if (sv->ref) (void)hipEventDestroy(sv->ref);
        return e == hipErrorOutOfMemory ? SFMM_ERR_ALLOC : SFMM_ERR_HIP;
    }
    sv->t0 = std::chrono::steady_clock::now();
    sv->arena.reset(new Arena(arena_bytes));
    sv->host_arena.reset(new Arena(arena_bytes));
    sfmm_server *raw = sv.get();
    for (int w = 0; w < workers; ++w) raw->preparers.emplace_back([raw] { raw->prepare_loop(); });
    raw->scheduler = std::thread([raw] { raw->schedule(); });
    *out = sv.release();
    return SFMM_OK;
}

int sfmm_server_submit(sfmm_server *sv, int op, const sfmm_item *items, size_t n, const uint32_t *src_cells,
                       size_t n_src, const uint32_t *dst_cells, size_t n_dst, uint64_t *rid)
{
    return sfmm_server_submit_at(sv, op, items, n, src_cells, n_src, dst_cells, n_dst, -1.0, rid);
}

int sfmm_server_submit_at(sfmm_server *sv, int op, const sfmm_item *items, size_t n, const uint32_t *src_cells,
                          size_t n_src, const uint32_t *dst_cells, size_t n_dst, double arrival_ms, uint64_t *rid)
{
    if (!sv || !rid || op < SFMM_M2M || op > SFMM_L2L) return SFMM_ERR_LOGIC;
    if ((n && !items) || (n_src && !src_cells) || (n_dst && !dst_cells)) return SFMM_ERR_LOGIC;
    auto r = std::unique_ptr<Request>(new (std::nothrow) Request);
    if (!r) return SFMM_ERR_ALLOC;
    r->op = op;
    r->items.assign(items, items + n);
    r->src_cells.assign(src_cells, src_cells + n_src);
    r->dst_cells.assign(dst_cells, dst_cells + n_dst);
    for (auto *v : {&r->src_cells, &r->dst_cells}) {
        std::sort(v->begin(), v->end());
        v->erase(std::unique(v->begin(), v->end()), v->end());
        if (!v->empty() && v->back() >= sv->ncells) return SFMM_ERR_LOGIC;
    }
    std::lock_guard<std::mutex> lock(sv->mu);
    if (sv->stop) return SFMM_ERR_LOGIC;
    const double now = sv->now_ms();
    r->id = sv->all.size();
    r->arrival_ms = arrival_ms > now ? arrival_ms : now;
    *rid = r->id;
    Request *p = r.get();
    sv->all.push_back(std::move(r));
    ++sv->unfinished;
    if (arrival_ms > now) {
        // held until its arrival time; the queue stays in arrival order
        auto pos = std::upper_bound(sv->future.begin(), sv->future.end(), p,
                                    [](const Request *a, const Request *b) { return a->arrival_ms < b->arrival_ms; });
        sv->future.insert(pos, p);
    } else {
        sv->waiting.push_back(p);
        sv->to_prepare.push_back(p);
        sv->cv_prep.notify_one();
    }
    sv->cv_work.notify_one();
    return SFMM_OK;
}

int sfmm_server_drain(sfmm_server *sv)
{
    if (!sv) return SFMM_ERR_LOGIC;
    std::unique_lock<std::mutex> lock(sv->mu);
    sv->cv_done.wait(lock, [&] { return sv->unfinished == 0; });
    return SFMM_OK;
}

int sfmm_server_request_info(sfmm_server *sv, uint64_t rid, double *times, int *rc)
{
    if (!sv || !times) return SFMM_ERR_LOGIC;
    std::lock_guard<std::mutex> lock(sv->mu);
    if (rid >= sv->all.size()) return SFMM_ERR_LOGIC;
    const Request &r = *sv->all[rid];
    if (!r.finished) return SFMM_ERR_LOGIC;
    times[0] = r.arrival_ms;
    times[1] = r.admitted_ms;
    times[2] = r.done_ms;
    times[3] = r.gpu_ms;
    times[4] = r.noticed_ms;
    if (rc) *rc = r.rc;
    return SFMM_OK;
}

int sfmm_server_counters(sfmm_server *sv, uint64_t *out)
{
    if (!sv || !out) return SFMM_ERR_LOGIC;
    std::lock_guard<std::mutex> lock(sv->mu);
    out[0] = sv->max_inflight;
    out[1] = sv->arena->peak();
    out[2] = sv->arena->size();
    for (int k = 0; k < 4; ++k) out[3 + k] = sv->blocked[k];
    return SFMM_OK;
}

double sfmm_server_clock_ms(sfmm_server *sv) { return sv ? sv->now_ms() : 0.0; }

void sfmm_server_destroy(sfmm_server *sv)
{
    if (!sv) return;
    {
        std::unique_lock<std::mutex> lock(sv->mu);
        sv->cv_done.wait(lock, [&] { return sv->unfinished == 0; });
        sv->stop = true;
        sv->cv_prep.notify_all();
        sv->cv_work.notify_all();
    }
    for (auto &t : sv->preparers) t.join();
    sv->scheduler.join();
    for (hipStream_t st : sv->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (hipEvent_t ev : sv->slot_start) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : sv->slot_end) (void)hipEventDestroy(ev);
    (void)hipFree(sv->arena_mem);
    (void)hipHostFree(sv->host_pool);
    (void)hipEventDestroy(sv->ref);
    delete sv;
}

}  // extern "C"
