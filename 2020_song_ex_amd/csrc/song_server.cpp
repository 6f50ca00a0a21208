// song_server.cpp -- continuous-batching serving loop over a resident index.
//
// No counterpart in the reference (its bench_gpu runs fixed batches,
// kernalGraph.h:343-375).  Requests (one query, its k and queue size) are
// queued by any thread; a dispatcher thread admits them into launch slots as
// slots and device memory free up.  Each slot owns a HIP stream, pinned host
// staging and device buffers; a launch searches its requests with per-request
// budgets (song_search_device_mixed semantics), so requests that arrive while
// others run on the GPU start in the next free slot and overlap on the device.
// A request's result depends only on its query, k, queue size and the search
// options, never on what it was batched with.
//
// Memory budget: every admitted request holds row_bytes + 8 + 8*k_max bytes
// of device state (plus its planning workspace in the learned search) until
// its slot completes; admission keeps the sum at or
// under the budget (requests that do not fit wait, counted as deferred).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/song.h"
#include "song_kernels.h"
#include "song_server.h"

namespace song {

using clk = std::chrono::steady_clock;

struct Request {
    int64_t ticket;
    int64_t seq = -1;  // resident mode: request number in the ring
    int32_t k, ef;
    clk::time_point t_arrival;  // submission (or scheduled arrival) time
    clk::time_point t_done;
    bool done = false;
    std::vector<char> query;
    std::vector<int32_t> ids;
    std::vector<float> dists;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    char* h_q = nullptr;
    int32_t* h_ks = nullptr;
    int32_t* h_efs = nullptr;
    int32_t* h_ids = nullptr;
    float* h_dists = nullptr;
    void* d_q = nullptr;
    int32_t* d_ks = nullptr;
    int32_t* d_efs = nullptr;
    int32_t* d_ids = nullptr;
    float* d_dists = nullptr;
    int32_t* d_counter = nullptr;
    int32_t* d_plan = nullptr;  // planning workspace of the learned search
    std::vector<Request*> reqs;
    int64_t bytes = 0;
    bool busy = false;
};

}  // namespace song

struct song_server {
    song_index* ix = nullptr;
    song_server_opts opts{};
    int device = 0;
    bool float_rows = true;  // query rows are floats (L2, IP), not Hamming codes
    size_t row_bytes = 0;
    int64_t req_bytes = 0;
    int capacity = 0;  // requests per slot
    std::vector<song::Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<song::Request*> pending;
    // Outstanding requests by ticket: tickets are handed out in order under
    // the lock, so request t sits at by_ticket[t - first_ticket] (null once
    // taken); taken entries at the front are dropped.  No rehashing: a hash
    // map that grows to millions of outstanding tickets stalls every
    // submitter and the dispatcher while it rehashes under the lock.
    std::deque<song::Request*> by_ticket;
    int64_t first_ticket = 1;
    int64_t next_ticket = 1;
    song::Request* find(int64_t t) const {
        if (t < first_ticket || t >= first_ticket + (int64_t)by_ticket.size()) return nullptr;
        return by_ticket[(size_t)(t - first_ticket)];
    }
    void forget(int64_t t) {
        by_ticket[(size_t)(t - first_ticket)] = nullptr;
        while (!by_ticket.empty() && by_ticket.front() == nullptr) {
            by_ticket.pop_front();
            first_ticket++;
        }
    }
    int64_t inflight_bytes = 0;
    song_server_stats stats{};
    // plain fallback: with a learned search and k_max above what it serves,
    // requests with k > fallback_k run the plain search (plain_opts)
    int32_t fallback_k = 0;
    song_search_opts plain_opts{};
    bool stop = false;
    std::string error;
    std::thread dispatcher;
    // resident mode: the request ring lives in host memory the device reads
    // and writes (fine-grained, coherent); head and the ring descriptor on
    // the device
    bool resident = false;
    int ring = 0;
    char* rq_q = nullptr;
    int32_t* rq_ks = nullptr;
    int32_t* rq_efs = nullptr;
    int32_t* rq_ids = nullptr;
    float* rq_dists = nullptr;
    int64_t* rq_done = nullptr;
    int64_t* rq_tail = nullptr;
    int32_t* rq_stop = nullptr;
    int64_t* d_head = nullptr;
    song::ResidentQueue* d_queue = nullptr;
    // the ring's device-side addresses (hipHostGetDevicePointer)
    void* dv_q = nullptr;
    int32_t* dv_ks = nullptr;
    int32_t* dv_efs = nullptr;
    int32_t* dv_ids = nullptr;
    float* dv_dists = nullptr;
    std::vector<song::Request*> ring_req;  // slot -> request in it
    int64_t published = 0;                 // requests written into the ring
    int64_t oldest = 0;                    // smallest request number not yet retired
    hipStream_t rstream = nullptr;
    hipEvent_t rdone = nullptr;
    bool running = false;
};

namespace {

int sfail(int code, const std::string& m) {
    song::set_last_error(m);
    return code;
}

double ms_between(song::clk::time_point a, song::clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Launch one slot's batch (called by the dispatcher without the lock held).
hipError_t launch_slot(song_server* sv, song::Slot& s, std::string* why) {
    // with the plain fallback, the requests beyond the learned search's k go
    // last and run as a second launch on the same stream (rows stay in
    // s.reqs order, which the retirement reads)
    if (sv->fallback_k > 0)
        std::stable_partition(s.reqs.begin(), s.reqs.end(),
                              [&](const song::Request* r) { return r->k <= sv->fallback_k; });
    const int nb = (int)s.reqs.size();
    int n1 = nb;  // requests of the configured search; the rest fall back
    int efmax1 = 1, efmax2 = 1;
    for (int i = 0; i < nb; i++) {
        song::Request* r = s.reqs[i];
        std::memcpy(s.h_q + (size_t)i * sv->row_bytes, r->query.data(), sv->row_bytes);
        s.h_ks[i] = r->k;
        s.h_efs[i] = r->ef;
        const bool plain = sv->fallback_k > 0 && r->k > sv->fallback_k;
        if (plain && n1 == nb) n1 = i;
        int& efm = plain ? efmax2 : efmax1;
        efm = std::max(efm, std::max(r->ef, r->k));
    }
    const int kst = sv->opts.k_max;
    hipError_t e = hipMemcpyAsync(s.d_q, s.h_q, (size_t)nb * sv->row_bytes, hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.d_ks, s.h_ks, (size_t)nb * 4, hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.d_efs, s.h_efs, (size_t)nb * 4, hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) return e;
    int rc = SONG_OK;
    if (n1 > 0)
        rc = song::launch_mixed_on_stream(sv->ix, s.d_q, n1, s.d_ks, s.d_efs, kst, efmax1, &sv->opts.search, s.d_ids,
                                          s.d_dists, s.d_counter, s.d_plan, s.stream,
                                          sv->fallback_k > 0 ? sv->fallback_k : kst);
    if (rc == SONG_OK && n1 < nb)
        rc = song::launch_mixed_on_stream(sv->ix, (const char*)s.d_q + (size_t)n1 * sv->row_bytes, nb - n1,
                                          s.d_ks + n1, s.d_efs + n1, kst, efmax2, &sv->plain_opts,
                                          s.d_ids + (size_t)n1 * kst, s.d_dists + (size_t)n1 * kst, s.d_counter,
                                          nullptr, s.stream);
    if (rc != SONG_OK) {
        *why = song_last_error();
        return hipErrorInvalidValue;
    }
    e = hipMemcpyAsync(s.h_ids, s.d_ids, (size_t)nb * kst * 4, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.h_dists, s.d_dists, (size_t)nb * kst * 4, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    return e;
}

void dispatcher_loop(song_server* sv) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lock(sv->mu);
    for (;;) {
        // 1. retire finished slots
        for (auto& s : sv->slots) {
            if (!s.busy) continue;
            lock.unlock();
            hipError_t q = hipEventQuery(s.done);
            lock.lock();
            if (q == hipErrorNotReady) continue;
            const auto now = song::clk::now();
            const int kst = sv->opts.k_max;
            for (size_t i = 0; i < s.reqs.size(); i++) {
                song::Request* r = s.reqs[i];
                r->ids.assign(s.h_ids + i * kst, s.h_ids + i * kst + r->k);
                r->dists.assign(s.h_dists + i * kst, s.h_dists + i * kst + r->k);
                r->t_done = now;
                r->done = true;
                sv->stats.completed++;
            }
            if (q != hipSuccess) sv->error = std::string("search launch failed: ") + hipGetErrorString(q);
            sv->inflight_bytes -= s.bytes;
            s.bytes = 0;
            s.reqs.clear();
            s.busy = false;
            sv->stats.in_flight = 0;
            for (auto& t : sv->slots) sv->stats.in_flight += (int64_t)t.reqs.size();
            sv->cv_done.notify_all();
        }
        // 2. admit pending requests into free slots
        bool launched = false;
        for (auto& s : sv->slots) {
            if (s.busy || sv->pending.empty()) continue;
            while (!sv->pending.empty() && (int)s.reqs.size() < sv->capacity &&
                   sv->inflight_bytes + sv->req_bytes <= sv->opts.memory_budget) {
                s.reqs.push_back(sv->pending.front());
                sv->pending.pop_front();
                s.bytes += sv->req_bytes;
                sv->inflight_bytes += sv->req_bytes;
            }
            if (s.reqs.empty()) {
                sv->stats.deferred++;
                break;  // the budget is full: wait for a slot to retire
            }
            sv->stats.peak_bytes = std::max(sv->stats.peak_bytes, sv->inflight_bytes);
            sv->stats.launches++;
            s.busy = true;
            sv->stats.in_flight += (int64_t)s.reqs.size();
            lock.unlock();
            std::string why;
            hipError_t e = launch_slot(sv, s, &why);
            lock.lock();
            if (e != hipSuccess) sv->error = "launch: " + (why.empty() ? std::string(hipGetErrorString(e)) : why);
            launched = true;
        }
        bool any_busy = false;
        for (auto& s : sv->slots) any_busy = any_busy || s.busy;
        if (sv->stop && !any_busy && sv->pending.empty()) break;
        if (!launched) {
            if (any_busy) {
                lock.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(10));
                lock.lock();
            } else {
                sv->cv_work.wait_for(lock, std::chrono::milliseconds(5),
                                     [&] { return sv->stop || !sv->pending.empty(); });
            }
        }
    }
}

// Resident mode: admit requests into free ring slots, publish them, keep the
// grid running while requests are outstanding, retire finished slots.
void dispatcher_loop_resident(song_server* sv) {
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lock(sv->mu);
    const int kst = sv->opts.k_max;
    song::clk::time_point stop_seen{};
    for (;;) {
        bool progress = false;
        // 1. admit: request number `published` goes to slot published % ring
        //    once that slot's previous request has retired
        int admitted = 0;
        while (!sv->pending.empty()) {
            const int slot = (int)(sv->published % sv->ring);
            if (sv->ring_req[slot] != nullptr) break;
            if (sv->inflight_bytes + sv->req_bytes > sv->opts.memory_budget) {
                sv->stats.deferred++;
                break;
            }
            song::Request* r = sv->pending.front();
            sv->pending.pop_front();
            std::memcpy(sv->rq_q + (size_t)slot * sv->row_bytes, r->query.data(), sv->row_bytes);
            sv->rq_ks[slot] = r->k;
            sv->rq_efs[slot] = r->ef;
            r->seq = sv->published++;
            sv->ring_req[slot] = r;
            sv->inflight_bytes += sv->req_bytes;
            sv->stats.in_flight++;
            admitted++;
        }
        if (admitted) {
            // the slots' contents before the new tail
            __atomic_store_n(sv->rq_tail, sv->published, __ATOMIC_RELEASE);
            sv->stats.peak_bytes = std::max(sv->stats.peak_bytes, sv->inflight_bytes);
            progress = true;
        }
        // 2. the grid: relaunch when it has left and requests are outstanding
        if (sv->running && hipEventQuery(sv->rdone) != hipErrorNotReady) sv->running = false;
        if (!sv->running && sv->oldest < sv->published && sv->error.empty()) {
            int rc = song::launch_resident_on_stream(sv->ix, sv->dv_q, sv->ring, sv->dv_ks, sv->dv_efs, kst,
                                                     sv->opts.ef_max, &sv->opts.search, sv->dv_ids, sv->dv_dists,
                                                     sv->d_queue, sv->rstream);
            hipError_t e = rc == SONG_OK ? hipEventRecord(sv->rdone, sv->rstream) : hipErrorInvalidValue;
            if (rc != SONG_OK || e != hipSuccess) {
                sv->error = std::string("resident launch failed: ") + song_last_error();
                sv->cv_done.notify_all();
            } else {
                sv->running = true;
                sv->stats.launches++;
            }
        }
        // 3. retire finished requests (any order); oldest tracks the first unfinished
        const auto now = song::clk::now();
        for (int64_t q = sv->oldest; q < sv->published; q++) {
            const int slot = (int)(q % sv->ring);
            song::Request* r = sv->ring_req[slot];
            if (r == nullptr || r->seq != q) continue;  // retired already
            if (__atomic_load_n(sv->rq_done + slot, __ATOMIC_ACQUIRE) != q + 1) continue;
            r->ids.assign(sv->rq_ids + (size_t)slot * kst, sv->rq_ids + (size_t)slot * kst + r->k);
            r->dists.assign(sv->rq_dists + (size_t)slot * kst, sv->rq_dists + (size_t)slot * kst + r->k);
            r->t_done = now;
            r->done = true;
            sv->ring_req[slot] = nullptr;
            sv->inflight_bytes -= sv->req_bytes;
            sv->stats.in_flight--;
            sv->stats.completed++;
            progress = true;
        }
        while (sv->oldest < sv->published && sv->ring_req[(int)(sv->oldest % sv->ring)] == nullptr) sv->oldest++;
        if (progress) sv->cv_done.notify_all();
        if (sv->stop && sv->pending.empty() && sv->oldest == sv->published) break;
        if (!sv->error.empty() && sv->stop) break;
        if (sv->stop) {
            // destroy waits at most a few seconds for outstanding requests
            if (stop_seen == song::clk::time_point{}) stop_seen = now;
            if (ms_between(stop_seen, now) > 5000.0) {
                sv->error = "server destroyed with requests outstanding";
                sv->cv_done.notify_all();
                break;
            }
        }
        if (!progress) {
            if (sv->oldest < sv->published) {
                lock.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(5));
                lock.lock();
            } else {
                sv->cv_work.wait_for(lock, std::chrono::milliseconds(5),
                                     [&] { return sv->stop || !sv->pending.empty(); });
            }
        }
    }
}

void free_resident(song_server* sv) {
    if (sv->rq_stop) __atomic_store_n(sv->rq_stop, 1, __ATOMIC_RELEASE);
    if (sv->rstream) (void)hipStreamSynchronize(sv->rstream);
    for (void* p : {(void*)sv->rq_q, (void*)sv->rq_ks, (void*)sv->rq_efs, (void*)sv->rq_ids, (void*)sv->rq_dists,
                    (void*)sv->rq_done, (void*)sv->rq_tail, (void*)sv->rq_stop})
        if (p) (void)hipHostFree(p);
    if (sv->d_head) (void)hipFree(sv->d_head);
    if (sv->d_queue) (void)hipFree(sv->d_queue);
    if (sv->rdone) (void)hipEventDestroy(sv->rdone);
    if (sv->rstream) (void)hipStreamDestroy(sv->rstream);
}

// Ring buffers for resident mode; returns a hipError_t.
hipError_t setup_resident(song_server* sv) {
    const size_t R = (size_t)sv->ring, kst = (size_t)sv->opts.k_max;
    const unsigned fl = hipHostMallocCoherent | hipHostMallocMapped;
    hipError_t e = hipStreamCreateWithFlags(&sv->rstream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->rdone, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->rq_q, R * sv->row_bytes, fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->rq_ks, R * 4, fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->rq_efs, R * 4, fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->rq_ids, R * kst * 4, fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->rq_dists, R * kst * 4, fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->rq_done, R * 8, fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->rq_tail, 64, fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sv->rq_stop, 64, fl);
    if (e == hipSuccess) e = hipMalloc((void**)&sv->d_head, 64);
    if (e == hipSuccess) e = hipMalloc((void**)&sv->d_queue, sizeof(song::ResidentQueue));
    if (e != hipSuccess) return e;
    void *dtail = nullptr, *dstop = nullptr, *ddone = nullptr;
    e = hipHostGetDevicePointer(&sv->dv_q, sv->rq_q, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&sv->dv_ks, sv->rq_ks, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&sv->dv_efs, sv->rq_efs, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&sv->dv_ids, sv->rq_ids, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&sv->dv_dists, sv->rq_dists, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&ddone, sv->rq_done, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&dtail, sv->rq_tail, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&dstop, sv->rq_stop, 0);
    if (e != hipSuccess) return e;
    std::memset(sv->rq_q, 0, R * sv->row_bytes);
    for (size_t i = 0; i < R; i++) {
        sv->rq_ks[i] = 1;
        sv->rq_efs[i] = 1;
        sv->rq_done[i] = 0;
    }
    *sv->rq_tail = 0;
    *sv->rq_stop = 0;
    e = hipMemsetAsync(sv->d_head, 0, 64, sv->rstream);
    song::ResidentQueue q{};
    q.head = sv->d_head;
    q.tail = (const int64_t*)dtail;
    q.stop = (const int32_t*)dstop;
    q.done = (int64_t*)ddone;
    q.ring = sv->ring;
    const int idle_ms = sv->opts.idle_ms > 0 ? sv->opts.idle_ms : 20;
    q.idle_ticks = (int64_t)idle_ms * 100000;  // the real-time clock runs at 100 MHz
    if (e == hipSuccess) e = hipMemcpyAsync(sv->d_queue, &q, sizeof q, hipMemcpyHostToDevice, sv->rstream);
    if (e == hipSuccess) e = hipStreamSynchronize(sv->rstream);
    sv->ring_req.assign(R, nullptr);
    return e;
}

void free_slot(song::Slot& s) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    for (void* p : {(void*)s.h_q, (void*)s.h_ks, (void*)s.h_efs, (void*)s.h_ids, (void*)s.h_dists})
        if (p) (void)hipHostFree(p);
    for (void* p : {(void*)s.d_q, (void*)s.d_ks, (void*)s.d_efs, (void*)s.d_ids, (void*)s.d_dists, (void*)s.d_counter,
                    (void*)s.d_plan})
        if (p) (void)hipFree(p);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
}

}  // namespace

extern "C" {

int song_server_create(song_index* ix, const song_server_opts* o, song_server** out) {
    if (!ix || !out) return sfail(SONG_ERR_ARG, "null argument");
    *out = nullptr;
    song_server_opts opts;
    if (o) opts = *o;
    else std::memset(&opts, 0, sizeof opts);
    if (opts.max_batch <= 0) opts.max_batch = 1024;
    if (opts.n_streams <= 0) opts.n_streams = 4;
    if (opts.memory_budget <= 0) opts.memory_budget = (int64_t)256 << 20;
    if (opts.k_max <= 0) opts.k_max = 100;
    if (opts.ef_max <= 0) opts.ef_max = 512;
    if (opts.n_streams > 32) return sfail(SONG_ERR_ARG, "n_streams must be <= 32");
    if (opts.resident != 0 && opts.resident != 1) return sfail(SONG_ERR_ARG, "resident must be 0 or 1");
    // The resident grid is switched off: in its device test it returned stale
    // request rows for some ring slots (the device cached the host-written
    // ring) and some of its waves did not leave at shutdown.  Until both are
    // understood the server refuses the mode instead of running it.
    if (opts.resident) return sfail(SONG_ERR_ARG, "the resident serving kernel is disabled in this build");
    if (opts.resident && (opts.search.learned & 5))
        return sfail(SONG_ERR_ARG, "the resident server supports the learned stop only (learned bit 1)");
    if (opts.idle_ms < 0) return sfail(SONG_ERR_ARG, "idle_ms must be >= 0");
    if (opts.k_max > 512 || opts.ef_max > 512) return sfail(SONG_ERR_ARG, "k_max and ef_max must be <= 512");
    int32_t stride = 0;
    int rc = song_index_row_stride(ix, &stride);
    if (rc) return rc;
    // The learned search serves k <= LEARNED_K_MAX; with a larger k_max the
    // requests above it fall back to the plain search (the same options
    // without the learned ones).  Options every launch would refuse are
    // refused here.
    constexpr int32_t LEARNED_K_MAX = 64;
    int32_t fallback_k = 0;
    song_search_opts plain_opts = opts.search;
    if (opts.search.learned && opts.k_max > LEARNED_K_MAX && !opts.resident) {
        fallback_k = LEARNED_K_MAX;
        plain_opts.learned = 0;
        plain_opts.stop_threshold = 0.f;
        plain_opts.stop_min_hops = plain_opts.stop_max_hops = plain_opts.stop_confirm = 0;
        plain_opts.stop_interval = 1;
        plain_opts.stop_late_hops = plain_opts.stop_late_interval = 0;
        plain_opts.result_cap = 0;
        plain_opts.entry_late_hops = 0;
        if ((rc = song::check_search_opts(ix, fallback_k, opts.ef_max, &opts.search)) != SONG_OK) return rc;
        if ((rc = song::check_search_opts(ix, opts.k_max, opts.ef_max, &plain_opts)) != SONG_OK) return rc;
    } else if ((rc = song::check_search_opts(ix, opts.k_max, opts.ef_max, &opts.search)) != SONG_OK) {
        return rc;
    }
    int metric = 0, device = 0;
    song::index_info(ix, &metric, &device);
    song_server* sv = new song_server();
    sv->ix = ix;
    sv->opts = opts;
    sv->fallback_k = fallback_k;
    sv->plain_opts = plain_opts;
    sv->device = device;
    sv->float_rows = metric != SONG_METRIC_HAMMING;
    sv->row_bytes = (size_t)stride * (metric == SONG_METRIC_HAMMING ? 8 : 4);
    sv->req_bytes = (int64_t)sv->row_bytes + 8 + 8 * (int64_t)opts.k_max +
                    (opts.search.learned ? 4 * song::plan_ws_words(1) : 0);
    sv->capacity = (int)std::max<int64_t>(1, std::min<int64_t>(opts.max_batch, opts.memory_budget / sv->req_bytes));
    if (opts.memory_budget < sv->req_bytes) {
        delete sv;
        return sfail(SONG_ERR_ARG, "memory_budget is smaller than one request");
    }
    if (hipSetDevice(device) != hipSuccess) {
        delete sv;
        return sfail(SONG_ERR_HIP, "hipSetDevice");
    }
    if (opts.resident) {
        sv->resident = true;
        sv->ring = (int)std::max<int64_t>(
            1, std::min<int64_t>((int64_t)opts.max_batch * opts.n_streams, opts.memory_budget / sv->req_bytes));
        sv->capacity = sv->ring;
        hipError_t e = setup_resident(sv);
        if (e != hipSuccess) {
            free_resident(sv);
            delete sv;
            return sfail(SONG_ERR_ALLOC, std::string("resident ring: ") + hipGetErrorString(e));
        }
        sv->dispatcher = std::thread(dispatcher_loop_resident, sv);
        *out = sv;
        return SONG_OK;
    }
    sv->slots.resize(opts.n_streams);
    const size_t cap = (size_t)sv->capacity, kst = (size_t)opts.k_max;
    for (auto& s : sv->slots) {
        hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_q, cap * sv->row_bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_ks, cap * 4, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_efs, cap * 4, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_ids, cap * kst * 4, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_dists, cap * kst * 4, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&s.d_q, cap * sv->row_bytes);
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_ks, cap * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_efs, cap * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_ids, cap * kst * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_dists, cap * kst * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_counter, 64);
        if (e == hipSuccess && opts.search.learned)
            e = hipMalloc((void**)&s.d_plan, (size_t)song::plan_ws_words((int64_t)cap) * 4);
        if (e != hipSuccess) {
            for (auto& t : sv->slots) free_slot(t);
            delete sv;
            return sfail(SONG_ERR_ALLOC, std::string("server buffers: ") + hipGetErrorString(e));
        }
    }
    sv->dispatcher = std::thread(dispatcher_loop, sv);
    *out = sv;
    return SONG_OK;
}

static int submit_at(song_server* sv, const void* query, int32_t k, int32_t ef, song::clk::time_point arrival,
                     int64_t* ticket) {
    if (!sv || !query || !ticket) return sfail(SONG_ERR_ARG, "null argument");
    if (k < 1 || k > sv->opts.k_max) return sfail(SONG_ERR_ARG, "k out of range for this server");
    if (std::max(ef, k) > sv->opts.ef_max) return sfail(SONG_ERR_ARG, "queue size above the server's ef_max");
    if (sv->float_rows) {
        const float* qf = (const float*)query;
        for (size_t i = 0; i < sv->row_bytes / 4; i++)
            if (!std::isfinite(qf[i])) return sfail(SONG_ERR_ARG, "query must be finite (NaN or Inf found)");
    }
    song::Request* r = new song::Request();
    r->k = k;
    r->ef = std::max(ef, k);
    r->t_arrival = arrival;
    r->query.assign((const char*)query, (const char*)query + sv->row_bytes);
    int64_t t;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        if (sv->stop) {
            delete r;
            return sfail(SONG_ERR_STATE, "server is shutting down");
        }
        t = r->ticket = sv->next_ticket++;
        sv->pending.push_back(r);
        sv->by_ticket.push_back(r);
        sv->stats.submitted++;
    }
    sv->cv_work.notify_one();
    *ticket = t;
    return SONG_OK;
}

int song_server_submit(song_server* sv, const void* query, int32_t k, int32_t ef, int64_t* ticket) {
    return submit_at(sv, query, k, ef, song::clk::now(), ticket);
}

static int take_result(song_server* sv, song::Request* r, int32_t* ids_out, float* dists_out, double* latency_ms) {
    if (ids_out) std::memcpy(ids_out, r->ids.data(), r->ids.size() * 4);
    if (dists_out) std::memcpy(dists_out, r->dists.data(), r->dists.size() * 4);
    if (latency_ms) *latency_ms = ms_between(r->t_arrival, r->t_done);
    sv->forget(r->ticket);
    delete r;
    return 1;
}

int song_server_poll(song_server* sv, int64_t ticket, int32_t* ids_out, float* dists_out, double* latency_ms) {
    if (!sv) return sfail(SONG_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lock(sv->mu);
    song::Request* r = sv->find(ticket);
    if (!r) return sfail(SONG_ERR_ARG, "unknown ticket");
    if (!sv->error.empty()) return sfail(SONG_ERR_HIP, sv->error);
    if (!r->done) return 0;
    return take_result(sv, r, ids_out, dists_out, latency_ms);
}

int song_server_wait(song_server* sv, int64_t ticket, int32_t timeout_ms, int32_t* ids_out, float* dists_out,
                     double* latency_ms) {
    if (!sv) return sfail(SONG_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lock(sv->mu);
    song::Request* r = sv->find(ticket);
    if (!r) return sfail(SONG_ERR_ARG, "unknown ticket");
    const bool ok = sv->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms),
                                         [&] { return r->done || !sv->error.empty(); });
    if (!sv->error.empty()) return sfail(SONG_ERR_HIP, sv->error);
    if (!ok) return 0;
    return take_result(sv, r, ids_out, dists_out, latency_ms);
}

int song_server_stats_get(song_server* sv, song_server_stats* out) {
    if (!sv || !out) return sfail(SONG_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lock(sv->mu);
    *out = sv->stats;
    out->memory_budget = sv->opts.memory_budget;
    out->slot_capacity = sv->capacity;
    return SONG_OK;
}

void song_server_destroy(song_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->dispatcher.joinable()) sv->dispatcher.join();
    // the resident grid: every wave sees the flag at its next claim
    if (sv->rq_stop) __atomic_store_n(sv->rq_stop, 1, __ATOMIC_RELEASE);
    for (auto& s : sv->slots) free_slot(s);
    if (sv->resident) free_resident(sv);
    for (song::Request* r : sv->by_ticket) delete r;  // pending and unretired requests included
    delete sv;
}

int song_server_run_poisson(song_server* sv, const void* queries, int32_t nq, const int32_t* ks, const int32_t* efs,
                            double rate_qps, uint64_t seed, int32_t* ids_out, float* dists_out, double* latency_ms_out,
                            double* wall_ms_out) {
    if (!sv || (nq > 0 && (!queries || !ks || !efs || !latency_ms_out)) || nq < 0 || !(rate_qps > 0))
        return sfail(SONG_ERR_ARG, "bad arguments");
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_qps / 1e6);  // microseconds
    std::vector<int64_t> tickets((size_t)nq);
    std::vector<song::clk::time_point> sched((size_t)nq);
    const size_t hb = sv->row_bytes;  // callers pass queries in the index row layout
    const auto t0 = song::clk::now();
    double t_us = 0;
    for (int i = 0; i < nq; i++) {
        t_us += gap(rng);
        sched[i] = t0 + std::chrono::microseconds((int64_t)t_us);
    }
    for (int i = 0; i < nq; i++) {
        while (song::clk::now() < sched[i]) {
            const auto left = sched[i] - song::clk::now();
            if (left > std::chrono::microseconds(200)) std::this_thread::sleep_for(left / 2);
        }
        // latency counts from the scheduled arrival, not from when the generator got to it
        int rc = submit_at(sv, (const char*)queries + (size_t)i * hb, ks[i], efs[i], sched[i], &tickets[i]);
        if (rc) return rc;
    }
    const int kst = sv->opts.k_max;
    auto t_last = t0;
    for (int i = 0; i < nq; i++) {
        std::vector<int32_t> ids(ks[i]);
        std::vector<float> d(ks[i]);
        int rc = song_server_wait(sv, tickets[i], 60000, ids.data(), d.data(), &latency_ms_out[i]);
        if (rc < 0) return rc;
        if (rc == 0) return sfail(SONG_ERR_STATE, "request timed out");
        if (ids_out) {
            std::fill(ids_out + (size_t)i * kst, ids_out + (size_t)(i + 1) * kst, -1);
            std::copy(ids.begin(), ids.end(), ids_out + (size_t)i * kst);
        }
        if (dists_out) {
            std::fill(dists_out + (size_t)i * kst, dists_out + (size_t)(i + 1) * kst, INFINITY);
            std::copy(d.begin(), d.end(), dists_out + (size_t)i * kst);
        }
        const auto done = sched[i] + std::chrono::microseconds((int64_t)(latency_ms_out[i] * 1000.0));
        if (done > t_last) t_last = done;
    }
    if (wall_ms_out) *wall_ms_out = ms_between(t0, t_last);
    return SONG_OK;
}

}  // extern "C"
