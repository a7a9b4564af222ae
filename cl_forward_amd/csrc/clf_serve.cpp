// clf_serve.cpp — the serving loop of the C ABI (clf_server_* in clf.h).
//
// Submitters copy each event into a pinned input arena and queue it.  One
// dispatcher thread owns every HIP call: whenever one of the slots (a
// stream with resident device buffers sized for a full batch) is free, it
// takes the longest ready prefix of the queue that fits the slot, copies
// it to the device in one or two DMA transfers (arena order is queue
// order, so a batch is one contiguous arena range unless the ring wraps),
// and enqueues the same batched search as clf_batch_run.  Results come back
// in two steps: per-event stats and track offsets, then exactly the packed
// tracks into one of a pool of pinned output buffers (two per slot), so a
// slot takes its next batch as soon as its tracks are on the host.  An
// output buffer returns to the pool when every request whose tracks it
// holds has been released.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "clf_host.h"

using namespace clf;

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr int kMaxSlots = 8;
constexpr size_t kBlobAlign = 256;
constexpr int64_t kEventBlobOverhead = 4096;  // header allowance per event in a slot
constexpr int64_t kEventCapOverhead = 64 * 48 + 64;

enum ReqState : int { R_FREE = 0, R_WRITING, R_PENDING, R_RUNNING, R_DONE };

struct Request {
    uint64_t ticket = 0;
    uint64_t tag = 0;
    std::atomic<int> state{R_FREE};
    size_t arena_off = 0;  // first byte of the blob in the arena
    size_t span = 0;       // aligned blob bytes
    uint64_t arena_seq = 0;
    int32_t S = 0, N = 0;
    int64_t cap = 0;
    double cost = 0.0;
    int out_buf = -1;  // output buffer holding its tracks once finished
    clf_served out{};
};

// An arena allocation in allocation order; space returns to the ring from
// the front once the front's batch has been copied to the device.
struct ArenaEntry {
    size_t taken;  // bytes taken from the ring, wrap gap included
    bool freed;
};

// A slot's batch: RUNNING (copies in, search, stats out), DRAINED (its input
// arena space returned, waiting for an output buffer), COPYING (tracks out).
enum SlotPhase { S_FREE = 0, S_RUNNING, S_DRAINED, S_COPYING };

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev_run = nullptr, ev_copy = nullptr;
    Resident dev;
    PinnedBuf h_desc, h_order, h_results, h_prefix;
    std::vector<int32_t> reqs;  // request indices in batch order
    SlotPhase phase = S_FREE;   // dispatcher thread only
    int out_buf = -1;           // output buffer of the batch being copied
    int64_t total = 0;          // tracks of the batch
};

struct OutBuf {
    PinnedBuf mem;
    int outstanding = 0;  // requests holding tracks here, not yet released
    bool busy = false;
};

}  // namespace

struct clf_server {
    clf_context* ctx = nullptr;
    int device = 0, n_cu = 0;
    clf_launch_config cfg{};
    clf_server_config conf{};
    int n_slots = 0;
    size_t slot_blob_bytes = 0;
    int64_t slot_cap = 0;

    PinnedBuf arena;
    size_t arena_head = 0, arena_used = 0;
    std::deque<ArenaEntry> arena_fifo;
    uint64_t arena_base_seq = 0;  // sequence number of arena_fifo.front()

    std::vector<Request> reqs;
    uint64_t req_mask = 0;
    int req_shift = 0;
    std::vector<int32_t> free_reqs;
    uint64_t next_seq = 1;
    std::deque<int32_t> pending;

    Slot slots[kMaxSlots];
    std::vector<OutBuf> outs;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_completed;
    std::thread dispatcher;
    bool stop = false;
    clf_status error = CLF_OK;
    std::string error_msg;

    bool collect = false;  // replay: finished requests also go to `completed`
    std::deque<int32_t> completed;
    int64_t n_batches = 0, n_batched_events = 0;
};

namespace {

Request* find_request(clf_server* s, uint64_t ticket) {
    if (ticket == 0) return nullptr;
    Request& r = s->reqs[ticket & s->req_mask];
    return r.ticket == ticket ? &r : nullptr;
}

// Gives arena space back from the front of the allocation order.  Lock held.
void arena_reclaim(clf_server* s) {
    while (!s->arena_fifo.empty() && s->arena_fifo.front().freed) {
        s->arena_used -= s->arena_fifo.front().taken;
        s->arena_fifo.pop_front();
        ++s->arena_base_seq;
    }
    if (s->arena_fifo.empty()) s->arena_head = 0, s->arena_used = 0;
}

void arena_free(clf_server* s, const Request& r) {
    s->arena_fifo[r.arena_seq - s->arena_base_seq].freed = true;
}

// Ends every request that has not finished with `st` (device failure or
// shutdown).  Lock held.
void fail_all(clf_server* s, clf_status st, const std::string& msg) {
    if (s->error == CLF_OK) {
        s->error = st;
        s->error_msg = msg;
    }
    const int64_t t = now_ns();
    for (size_t i = 0; i < s->reqs.size(); ++i) {
        Request& r = s->reqs[i];
        const int state = r.state.load();
        if (state == R_FREE || state == R_DONE) continue;
        r.out.ticket = r.ticket;
        r.out.tag = r.tag;
        r.out.status = st;
        r.out.n_tracks = 0;
        r.out.tracks = nullptr;
        r.out.t_done_ns = t;
        r.out_buf = -1;
        r.state.store(R_DONE);
        if (s->collect) s->completed.push_back((int32_t)i);
    }
    s->pending.clear();
    s->cv_done.notify_all();
    s->cv_completed.notify_all();
}

bool slot_busy(const Slot& sl) { return sl.phase != S_FREE; }

// Picks the next batch for slot `k` and enqueues it.  Returns false when
// nothing was launched.
clf_status launch_batch(clf_server* s, int k, bool device_idle, bool& launched) {
    launched = false;
    Slot& sl = s->slots[k];
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (sl.phase != S_FREE || s->pending.empty() || s->error != CLF_OK) return CLF_OK;
        sl.reqs.clear();  // the previous batch has been published
        Request& first = s->reqs[s->pending.front()];
        if (first.state.load(std::memory_order_acquire) != R_PENDING) return CLF_OK;
        if (!device_idle && s->conf.max_wait_us > 0) {
            // another batch is running: wait a little for this one to fill
            int ready = 0;
            for (int32_t idx : s->pending) {
                if (s->reqs[idx].state.load(std::memory_order_acquire) != R_PENDING) break;
                if (++ready >= s->conf.max_batch_events) break;
            }
            if (ready < s->conf.max_batch_events &&
                now_ns() - first.out.t_submit_ns < (int64_t)s->conf.max_wait_us * 1000)
                return CLF_OK;
        }
        size_t blob = 0;
        int64_t hits = 0, cap = 0;
        while (!s->pending.empty() && (int)sl.reqs.size() < s->conf.max_batch_events) {
            const int32_t idx = s->pending.front();
            Request& r = s->reqs[idx];
            if (r.state.load(std::memory_order_acquire) != R_PENDING) break;
            if (blob + r.span > s->slot_blob_bytes || hits + r.N > s->conf.max_batch_hits ||
                cap + r.cap > s->slot_cap)
                break;
            s->pending.pop_front();
            r.state.store(R_RUNNING);
            sl.reqs.push_back(idx);
            blob += r.span;
            hits += r.N;
            cap += r.cap;
        }
    }
    const int n = (int)sl.reqs.size();
    if (n == 0) return CLF_OK;

    // descriptors and the arena runs to copy
    EventDesc* d = sl.h_desc.as<EventDesc>();
    int32_t* order = sl.h_order.as<int32_t>();
    struct Run {
        size_t src, dst, len;
    };
    Run runs[4];
    int n_runs = 0;
    size_t pos = 0, prev_end = (size_t)-1;
    int64_t hit_off = 0, track_off = 0;
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i) {
        const Request& r = s->reqs[sl.reqs[i]];
        if (r.arena_off != prev_end) {
            if (n_runs == 4) return fail(CLF_E_DEVICE, "serving: batch spans too many arena runs");
            runs[n_runs++] = {r.arena_off, pos, 0};
        }
        runs[n_runs - 1].len += r.span;
        d[i].blob_off = (int64_t)pos;
        d[i].hit_off = hit_off;
        d[i].track_off = track_off;
        d[i].n_hits = r.N;
        d[i].n_sensors = r.S;
        pos += r.span;
        prev_end = r.arena_off + r.span;
        hit_off += r.N;
        track_off += r.cap;
        cost[i] = r.cost;
    }
    std::iota(order, order + n, 0);
    std::stable_sort(order, order + n, [&](int x, int y) { return cost[x] > cost[y]; });

    hipStream_t st = sl.stream;
    const uint8_t* arena = s->arena.as<uint8_t>();
    for (int i = 0; i < n_runs; ++i)
        HIP_TRY(hipMemcpyAsync(sl.dev.blobs.as<uint8_t>() + runs[i].dst, arena + runs[i].src,
                               runs[i].len, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(sl.dev.desc.p, d, sizeof(EventDesc) * (size_t)n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(sl.dev.order.p, order, sizeof(int32_t) * (size_t)n,
                           hipMemcpyHostToDevice, st));
    clf_status cs = enqueue_search(sl.dev, n, CLF_MODE_BATCHED, s->cfg, s->n_cu, st, nullptr, nullptr);
    if (cs != CLF_OK) return cs;
    HIP_TRY(hipMemcpyAsync(sl.h_results.p, sl.dev.results.p, sizeof(EventResult) * (size_t)n,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(sl.h_prefix.p, sl.dev.prefix.p, sizeof(int64_t) * ((size_t)n + 1),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(sl.ev_run, st));
    const int64_t t = now_ns();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (int32_t idx : sl.reqs) s->reqs[idx].out.t_launch_ns = t;
        sl.phase = S_RUNNING;
        ++s->n_batches;
        s->n_batched_events += n;
    }
    launched = true;
    return CLF_OK;
}

// Advances a busy slot when its device work has finished.
clf_status poll_slot(clf_server* s, int k, bool& progress) {
    Slot& sl = s->slots[k];
    if (sl.phase == S_RUNNING) {
        hipError_t q = hipEventQuery(sl.ev_run);
        if (q == hipErrorNotReady) return CLF_OK;
        if (q != hipSuccess) return hip_fail(q, "serving: batch run");
        const int n = (int)sl.reqs.size();
        const int64_t total = sl.h_prefix.as<int64_t>()[n];
        int64_t cap = 0;
        for (int32_t idx : sl.reqs) cap += s->reqs[idx].cap;
        if (total < 0 || total > cap) return fail(CLF_E_DEVICE, "serving: inconsistent track count");
        sl.total = total;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (int32_t idx : sl.reqs) arena_free(s, s->reqs[idx]);
            arena_reclaim(s);
        }
        sl.phase = S_DRAINED;
        progress = true;
    }
    if (sl.phase == S_DRAINED) {
        int ob = -1;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (size_t i = 0; i < s->outs.size(); ++i)
                if (!s->outs[i].busy) {
                    ob = (int)i;
                    s->outs[i].busy = true;
                    break;
                }
        }
        if (ob < 0) return CLF_OK;  // every output buffer is held: retry later
        progress = true;
        sl.out_buf = ob;
        const int64_t total = sl.total;
        PinnedBuf& mem = s->outs[ob].mem;
        HIP_TRY(mem.ensure(sizeof(clf_track) * (size_t)std::max<int64_t>(total, 1)));
        if (total > 0)
            HIP_TRY(hipMemcpyAsync(mem.p, sl.dev.packed.p, sizeof(clf_track) * (size_t)total,
                                   hipMemcpyDeviceToHost, sl.stream));
        HIP_TRY(hipEventRecord(sl.ev_copy, sl.stream));
        sl.phase = S_COPYING;
    }
    if (sl.phase == S_COPYING) {
        hipError_t q = hipEventQuery(sl.ev_copy);
        if (q == hipErrorNotReady) return CLF_OK;
        if (q != hipSuccess) return hip_fail(q, "serving: result copy");
        const int n = (int)sl.reqs.size();
        const int64_t* pre = sl.h_prefix.as<int64_t>();
        const clf_event_stats* stats = sl.h_results.as<clf_event_stats>();
        const uint8_t* packed = s->outs[sl.out_buf].mem.as<uint8_t>();
        const int64_t t = now_ns();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (int i = 0; i < n; ++i) {
                const int32_t idx = sl.reqs[i];
                Request& r = s->reqs[idx];
                r.out.ticket = r.ticket;
                r.out.tag = r.tag;
                r.out.status = CLF_OK;
                r.out.batch_events = n;
                r.out.stats = stats[i];
                r.out.n_tracks = pre[i + 1] - pre[i];
                r.out.tracks = packed + sizeof(clf_track) * (size_t)pre[i];
                r.out.t_done_ns = t;
                r.out_buf = sl.out_buf;
                r.state.store(R_DONE);
                if (s->collect) s->completed.push_back(idx);
            }
            s->outs[sl.out_buf].outstanding = n;
            sl.out_buf = -1;
            sl.phase = S_FREE;
        }
        s->cv_done.notify_all();
        s->cv_completed.notify_all();
        progress = true;
    }
    return CLF_OK;
}

void dispatch_loop(clf_server* s) {
    (void)hipSetDevice(s->device);
    for (;;) {
        bool progress = false;
        clf_status st = CLF_OK;
        for (int k = 0; k < s->n_slots && st == CLF_OK; ++k)
            if (slot_busy(s->slots[k])) st = poll_slot(s, k, progress);
        bool device_idle = true;
        for (int k = 0; k < s->n_slots; ++k)
            if (s->slots[k].phase == S_RUNNING) device_idle = false;
        bool stopping;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            stopping = s->stop;
        }
        if (!stopping) {
            for (int k = 0; k < s->n_slots && st == CLF_OK; ++k) {
                bool launched = false;
                st = launch_batch(s, k, device_idle, launched);
                if (launched) {
                    progress = true;
                    device_idle = false;
                }
            }
        }
        if (st != CLF_OK) {
            std::lock_guard<std::mutex> lk(s->mu);
            fail_all(s, st, std::string("serving: ") + clf_last_error());
        }
        bool busy = false;
        for (int k = 0; k < s->n_slots; ++k) busy = busy || slot_busy(s->slots[k]);
        if (st != CLF_OK) {
            // drain what is still on the device before leaving
            for (int k = 0; k < s->n_slots; ++k) (void)hipStreamSynchronize(s->slots[k].stream);
            return;
        }
        if (stopping && !busy) return;
        if (progress) continue;
        if (busy) {
            std::this_thread::yield();
            continue;
        }
        // idle device: sleep until a submit, a release or the wait policy
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_work.wait_for(lk, std::chrono::microseconds(s->pending.empty() ? 2000 : 20));
    }
}

}  // namespace

extern "C" {

void clf_server_default_config(clf_server_config* cfg) {
    if (!cfg) return;
    cfg->slots = 4;
    cfg->max_batch_events = 256;
    cfg->max_batch_hits = 1 << 20;
    cfg->max_wait_us = 100;
    cfg->max_requests = 65536;
    cfg->arena_bytes = (int64_t)512 << 20;
}

void clf_server_destroy(clf_server* s) {
    if (!s) return;
    if (s->dispatcher.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->dispatcher.join();
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        fail_all(s, CLF_E_INVALID_ARGUMENT, "server stopped");
    }
    (void)hipSetDevice(s->device);
    for (int k = 0; k < kMaxSlots; ++k) {
        Slot& sl = s->slots[k];
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        sl.dev.release();
        sl.h_desc.release();
        sl.h_order.release();
        sl.h_results.release();
        sl.h_prefix.release();
        if (sl.ev_run) (void)hipEventDestroy(sl.ev_run);
        if (sl.ev_copy) (void)hipEventDestroy(sl.ev_copy);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    for (OutBuf& o : s->outs) o.mem.release();
    s->arena.release();
    delete s;
}

clf_status clf_server_create(clf_context* ctx, const clf_server_config* cfg, clf_server** out) {
    if (!ctx || !out) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    clf_server_config c;
    clf_server_default_config(&c);
    if (cfg) c = *cfg;
    if (c.slots < 1 || c.slots > kMaxSlots || c.max_batch_events < 1 ||
        c.max_batch_events > (1 << 20) || c.max_batch_hits < 1 || c.max_wait_us < 0 ||
        c.max_requests < 1 || c.max_requests > (1 << 24) || c.arena_bytes < (int64_t)(1 << 20))
        return fail(CLF_E_INVALID_ARGUMENT, "bad server configuration");
    clf_launch_config lc;
    clf_context_get_config(ctx, &lc);
    const int device = context_device(ctx);
    clf_server* s = new clf_server();
    s->ctx = ctx;
    s->cfg = lc;
    s->n_cu = clf_context_num_cus(ctx);
    s->conf = c;
    s->n_slots = c.slots;
    s->slot_blob_bytes = (size_t)(16 * c.max_batch_hits + kEventBlobOverhead * c.max_batch_events);
    s->slot_cap = c.max_batch_hits + kEventCapOverhead * (int64_t)c.max_batch_events;
    int shift = 0;
    while ((1ll << shift) < c.max_requests) ++shift;
    s->req_shift = shift;
    s->req_mask = (1ull << shift) - 1ull;
    s->reqs = std::vector<Request>((size_t)1 << shift);
    s->free_reqs.reserve(s->reqs.size());
    for (int64_t i = (int64_t)c.max_requests - 1; i >= 0; --i) s->free_reqs.push_back((int32_t)i);

    hipError_t e = hipSetDevice(s->device = device);
    auto chk = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    chk(s->arena.ensure((size_t)c.arena_bytes));
    for (int k = 0; k < s->n_slots && e == hipSuccess; ++k) {
        Slot& sl = s->slots[k];
        chk(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        chk(hipEventCreateWithFlags(&sl.ev_run, hipEventDisableTiming));
        chk(hipEventCreateWithFlags(&sl.ev_copy, hipEventDisableTiming));
        chk(sl.dev.ensure(c.max_batch_events, s->slot_blob_bytes, c.max_batch_hits, s->slot_cap));
        chk(sl.h_desc.ensure(sizeof(EventDesc) * (size_t)c.max_batch_events));
        chk(sl.h_order.ensure(sizeof(int32_t) * (size_t)c.max_batch_events));
        chk(sl.h_results.ensure(sizeof(EventResult) * (size_t)c.max_batch_events));
        chk(sl.h_prefix.ensure(sizeof(int64_t) * ((size_t)c.max_batch_events + 1)));
    }
    s->outs = std::vector<OutBuf>(2 * (size_t)s->n_slots);
    for (OutBuf& o : s->outs) chk(o.mem.ensure(sizeof(clf_track) * (size_t)(c.max_batch_hits / 4 + 1024)));
    if (e != hipSuccess) {
        clf_server_destroy(s);
        return hip_fail(e, "server allocation");
    }
    s->dispatcher = std::thread(dispatch_loop, s);
    *out = s;
    return CLF_OK;
}

clf_status clf_server_submit(clf_server* s, const uint8_t* event, size_t size, uint64_t tag,
                             uint64_t* ticket) {
    if (!s || !ticket) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    *ticket = 0;
    int32_t S = 0, N = 0;
    if (!validate_event(event, size, S, N)) return fail(CLF_E_INVALID_ARGUMENT, "malformed event");
    const size_t bytes = event_bytes(S, N);
    const size_t span = align_up(bytes, kBlobAlign);
    const int64_t cap = CLF_EVENT_TRACK_CAPACITY(N, S);
    if (span > s->slot_blob_bytes || N > s->conf.max_batch_hits || cap > s->slot_cap)
        return fail(CLF_E_CAPACITY, "event larger than a serving batch");
    Request* r = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->error != CLF_OK) return fail(s->error, s->error_msg);
        if (s->stop) return fail(CLF_E_INVALID_ARGUMENT, "server stopped");
        if (s->free_reqs.empty()) return fail(CLF_E_CAPACITY, "request table full");
        const size_t acap = s->arena.cap;
        size_t off, taken;
        if (s->arena_head + span <= acap) {
            off = s->arena_head;
            taken = span;
        } else {
            off = 0;
            taken = (acap - s->arena_head) + span;
        }
        if (s->arena_used + taken > acap) return fail(CLF_E_CAPACITY, "input arena full");
        s->arena_head = off + span == acap ? 0 : off + span;
        s->arena_used += taken;
        const int32_t idx = s->free_reqs.back();
        s->free_reqs.pop_back();
        r = &s->reqs[idx];
        r->ticket = (s->next_seq++ << s->req_shift) | (uint64_t)idx;
        r->tag = tag;
        r->arena_off = off;
        r->span = span;
        r->arena_seq = s->arena_base_seq + s->arena_fifo.size();
        s->arena_fifo.push_back({taken, false});
        r->S = S;
        r->N = N;
        r->cap = cap;
        r->out_buf = -1;
        r->out = clf_served{};
        r->state.store(R_WRITING);
        s->pending.push_back(idx);
        *ticket = r->ticket;
    }
    std::memcpy(s->arena.as<uint8_t>() + r->arena_off, event, bytes);
    r->cost = event_cost(event, S, N);
    r->out.t_submit_ns = now_ns();
    // a failure of the server may have ended the request meanwhile
    int expected = R_WRITING;
    r->state.compare_exchange_strong(expected, R_PENDING, std::memory_order_release);
    s->cv_work.notify_one();
    return CLF_OK;
}

clf_status clf_server_wait(clf_server* s, uint64_t ticket, int32_t timeout_ms, clf_served* out) {
    if (!s || !out) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    Request* r = find_request(s, ticket);
    if (!r) return fail(CLF_E_INVALID_ARGUMENT, "unknown ticket");
    auto done = [&] { return r->state.load() == R_DONE; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
        return fail(CLF_E_CAPACITY, "timed out");
    }
    *out = r->out;
    return CLF_OK;
}

clf_status clf_server_release(clf_server* s, uint64_t ticket) {
    if (!s) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        Request* r = find_request(s, ticket);
        if (!r || r->state.load() != R_DONE)
            return fail(CLF_E_INVALID_ARGUMENT, "ticket unknown or not finished");
        if (r->out_buf >= 0) {
            OutBuf& o = s->outs[r->out_buf];
            if (--o.outstanding == 0) o.busy = false;
        }
        r->out_buf = -1;
        r->ticket = 0;
        r->state.store(R_FREE);
        s->free_reqs.push_back((int32_t)(r - s->reqs.data()));
    }
    s->cv_work.notify_one();
    return CLF_OK;
}

uint64_t clf_tracks_digest(const uint8_t* tracks, int64_t n_tracks) {
    const uint64_t P = 0x9E3779B97F4A7C15ull;
    uint64_t h[8];
    for (int l = 0; l < 8; ++l) h[l] = 0x243F6A8885A308D3ull * (uint64_t)(2 * l + 1);
    const int64_t bytes = n_tracks > 0 && tracks ? n_tracks * (int64_t)sizeof(clf_track) : 0;
    int64_t i = 0;
    for (; i + 64 <= bytes; i += 64) {
        uint64_t w[8];
        std::memcpy(w, tracks + i, 64);
        for (int l = 0; l < 8; ++l) {
            h[l] = (h[l] ^ w[l]) * P;
            h[l] ^= h[l] >> 29;
        }
    }
    for (int l = 0; i < bytes; i += 4, l = (l + 1) & 7) {  // tail: 32-bit words
        uint32_t w;
        std::memcpy(&w, tracks + i, 4);
        h[l] = (h[l] ^ w) * P;
        h[l] ^= h[l] >> 29;
    }
    uint64_t r = (uint64_t)n_tracks * P;
    for (int l = 0; l < 8; ++l) {
        r = (r ^ h[l]) * P;
        r ^= r >> 32;
    }
    return r;
}

clf_status clf_server_replay(clf_server* s, int32_t n_pool, const uint8_t* const* pool,
                             const size_t* sizes, int64_t n, const int32_t* pick,
                             const int64_t* arrival_ns, int32_t submit_threads,
                             clf_replay_record* rec, clf_replay_summary* sum) {
    if (!s || n_pool < 1 || !pool || !sizes || n < 0 || (n > 0 && (!pick || !arrival_ns || !rec)) ||
        submit_threads < 1 || submit_threads > 64)
        return fail(CLF_E_INVALID_ARGUMENT, "bad replay argument");
    for (int64_t i = 0; i < n; ++i) {
        if (pick[i] < 0 || pick[i] >= n_pool) return fail(CLF_E_INVALID_ARGUMENT, "pick out of range");
        if (arrival_ns[i] < 0) return fail(CLF_E_INVALID_ARGUMENT, "negative arrival time");
        rec[i] = clf_replay_record{-1, -1, 0, 0, -1, 0, 0};
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->collect) return fail(CLF_E_INVALID_ARGUMENT, "a replay is already running");
        s->collect = true;
        s->completed.clear();
    }
    const int64_t b0 = s->n_batches, e0 = s->n_batched_events;
    const int64_t start = now_ns() + 2000000;  // 2 ms for the threads to start
    std::atomic<int64_t> remaining{n};
    std::atomic<int64_t> max_lag{0};
    std::atomic<int64_t> last_done{start};

    auto submitter = [&](int t) {
        for (int64_t i = t; i < n; i += submit_threads) {
            const int64_t target = start + arrival_ns[i];
            int64_t now = now_ns();
            if (target - now > 120000)  // sleep most of a long gap, spin the rest
                std::this_thread::sleep_for(std::chrono::nanoseconds(target - now - 80000));
            while ((now = now_ns()) < target) {
            }
            const int64_t lag = now - target;
            int64_t m = max_lag.load();
            while (lag > m && !max_lag.compare_exchange_weak(m, lag)) {
            }
            uint64_t ticket = 0;
            const int p = pick[i];
            clf_status st = clf_server_submit(s, pool[p], sizes[p], (uint64_t)i, &ticket);
            if (st != CLF_OK) {
                rec[i].status = st;
                if (remaining.fetch_sub(1) == 1) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->cv_completed.notify_all();
                }
            }
        }
    };
    auto collector = [&]() {
        for (;;) {
            int32_t idx;
            clf_served o;
            {
                std::unique_lock<std::mutex> lk(s->mu);
                s->cv_completed.wait(lk, [&] { return !s->completed.empty() || remaining.load() <= 0; });
                if (s->completed.empty()) return;
                idx = s->completed.front();
                s->completed.pop_front();
                o = s->reqs[idx].out;
            }
            const int64_t i = (int64_t)o.tag;
            const int64_t arr = start + arrival_ns[i];
            clf_replay_record& r = rec[i];
            r.status = o.status;
            r.latency_ns = o.t_done_ns - arr;
            r.wait_ns = o.t_launch_ns > 0 ? o.t_launch_ns - arr : -1;
            r.n_tracks = (int32_t)o.n_tracks;
            r.flags = o.stats.flags;
            r.batch_events = o.batch_events;
            r.digest = o.status == CLF_OK ? clf_tracks_digest(o.tracks, o.n_tracks) : 0;
            int64_t ld = last_done.load();
            while (o.t_done_ns > ld && !last_done.compare_exchange_weak(ld, o.t_done_ns)) {
            }
            (void)clf_server_release(s, o.ticket);
            if (remaining.fetch_sub(1) == 1) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->cv_completed.notify_all();
            }
        }
    };
    std::vector<std::thread> threads;
    const int n_collect = 4;
    for (int c = 0; c < n_collect; ++c) threads.emplace_back(collector);
    for (int t = 0; t < submit_threads; ++t) threads.emplace_back(submitter, t);
    for (auto& th : threads) th.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->collect = false;
        s->completed.clear();
    }
    if (sum) {
        std::vector<int64_t> lat;
        lat.reserve((size_t)n);
        int64_t rejected = 0;
        for (int64_t i = 0; i < n; ++i) {
            if (rec[i].status == CLF_OK) lat.push_back(rec[i].latency_ns);
            else ++rejected;
        }
        std::sort(lat.begin(), lat.end());
        auto q = [&](double f) -> int64_t {
            if (lat.empty()) return 0;
            size_t k = (size_t)std::ceil(f * (double)lat.size()) - 1;
            return lat[std::min(k, lat.size() - 1)];
        };
        const int64_t first = n > 0 ? start + arrival_ns[0] : start;
        sum->wall_s = (double)(last_done.load() - first) * 1e-9;
        sum->completed = (int64_t)lat.size();
        sum->rejected = rejected;
        sum->events_per_s = sum->wall_s > 0 ? (double)lat.size() / sum->wall_s : 0.0;
        sum->p50_ns = q(0.50);
        sum->p90_ns = q(0.90);
        sum->p99_ns = q(0.99);
        sum->max_ns = lat.empty() ? 0 : lat.back();
        std::lock_guard<std::mutex> lk(s->mu);
        sum->batches = s->n_batches - b0;
        sum->mean_batch_events =
            sum->batches > 0 ? (double)(s->n_batched_events - e0) / (double)sum->batches : 0.0;
        sum->max_submit_lag_ns = (double)max_lag.load();
    }
    return CLF_OK;
}

}  // extern "C"
