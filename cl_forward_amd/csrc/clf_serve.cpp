// clf_serve.cpp — the serving loop of the C ABI (clf_server_* in clf.h).
//
// Submitters copy each event into a pinned input arena (a ring, freed in
// allocation order) and queue it; any number of threads may submit.  Two
// engines take it from there:
//
// Resident: a persistent kernel (serve_kernel) follows a descriptor ring in
// pinned memory.  Submit copies the words of the event that the search
// reads (all but the hit IDs) into pinned staging; the server's thread
// moves each run of newly submitted events into the device input
// arena with one copy-engine transfer and, once it has landed, publishes
// their descriptors (release store of each sequence word).  A workgroup
// claims a descriptor, copies the event from the device arena into its
// scratch, searches it, writes the tracks into the device output ring at a
// position it reserves with one atomic, and appends a completion record.
// One host thread collects completions, copies the finished, contiguous part
// of the output ring back into its pinned mirror with the copy engine, hands
// out the results, and relaunches the kernel when it has left for lack of
// work.  Admission is by device memory: the input arena, and the output
// ring's room for every admitted event's capacity bound.
//
// Batches: a dispatcher thread owns every HIP call.  Whenever one of the
// slots (a stream with resident device buffers sized for a full batch) is
// free, it takes the longest ready prefix of the queue that fits the slot,
// copies it to the device in one or two DMA transfers (arena order is queue
// order, so a batch is one contiguous arena range unless the ring wraps),
// and enqueues the same batched search as clf_batch_run.  Per-event stats
// and track offsets come back first, then exactly the packed tracks, into
// one of a pool of pinned output buffers (two per slot), so a slot takes its
// next batch as soon as its tracks are on the host.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <functional>
#include <mutex>
#include <numeric>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include <immintrin.h>

#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include "clf_host.h"

using namespace clf;

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr int kMaxSlots = 8;
constexpr int32_t kMaxSubmitMany = 64;  // events per clf_server_submit_many call
constexpr size_t kBlobAlign = 256;
constexpr int64_t kEventBlobOverhead = 4096;  // header allowance per event in a slot
constexpr int64_t kEventCapOverhead = 64 * 48 + 64;
// Resident engine: copy-engine transfers in flight per direction.  While
// they are, newly submitted (or finished) events accumulate and go in the
// next transfer, so that the per-transfer host cost is spread over more
// events the higher the load.
constexpr size_t kCopiesInFlight = 2;
// Output copies move whole 64 KiB-aligned pieces of the ring: a copy of a
// few bytes, or at an odd offset, may be done with a kernel instead of a
// copy engine, and such a kernel cannot start while the resident kernel
// holds every compute unit.
constexpr uint64_t kOutGrain = uint64_t(64) << 10;

enum ReqState : int { R_FREE = 0, R_WRITING, R_PENDING, R_RUNNING, R_DONE };

// A ring of bytes handed out in allocation order and returned in any order;
// space comes back from the front of the allocation order.
struct Ring {
    struct Entry {
        size_t taken;  // bytes taken, wrap gap included
        bool freed;
    };
    size_t cap = 0, head = 0, used = 0;
    std::deque<Entry> fifo;
    uint64_t base_seq = 0;  // sequence number of fifo.front()

    bool fits(size_t span, size_t& off, size_t& taken) const {
        if (head + span <= cap) {
            off = head;
            taken = span;
        } else {
            off = 0;
            taken = (cap - head) + span;
        }
        return used + taken <= cap;
    }
    uint64_t commit(size_t off, size_t span, size_t taken) {
        head = off + span == cap ? 0 : off + span;
        used += taken;
        fifo.push_back({taken, false});
        return base_seq + fifo.size() - 1;
    }
    void release(uint64_t seq) {
        fifo[seq - base_seq].freed = true;
        while (!fifo.empty() && fifo.front().freed) {
            used -= fifo.front().taken;
            fifo.pop_front();
            ++base_seq;
        }
        if (fifo.empty()) head = 0, used = 0;
    }
};

struct Request {
    uint64_t ticket = 0;
    uint64_t tag = 0;
    std::atomic<int> state{R_FREE};
    size_t in_off = 0;  // first byte of the blob in the input arena
    size_t span = 0;    // aligned blob bytes
    uint64_t in_seq = 0;
    bool in_live = false;
    uint64_t desc = 0;   // resident: its descriptor number
    int32_t S = 0, N = 0;
    int64_t cap = 0;
    double cost = 0.0;
    int out_buf = -1;  // batches: output buffer holding its tracks once finished
    // resident: the capacity bound in bytes, which the output ring reserves
    // until completion; then its tracks' entry in the ring's live list
    uint64_t out_cap_bytes = 0;
    bool out_pending = false;  // admitted, not yet completed (counted in out_pend)
    bool out_live = false;     // has an entry in the live list
    uint64_t live_seq = 0;
    clf_served out{};
};

// A process-wide pool of fine-grained device buffers (coherent with the
// copy engines while a kernel runs).  Buffers are reused by later servers
// and never freed before the process ends.
struct FinePool {
    struct Buf {
        int device;
        void* p;
        size_t bytes;
        bool busy;
    };
    std::mutex mu;
    std::vector<Buf> bufs;
    hipError_t take(int device, size_t bytes, void** out) {
        std::lock_guard<std::mutex> lk(mu);
        for (Buf& b : bufs)
            if (!b.busy && b.device == device && b.bytes >= bytes) {
                b.busy = true;
                *out = b.p;
                return hipSuccess;
            }
        void* p = nullptr;
        hipError_t e = hipExtMallocWithFlags(&p, bytes, hipDeviceMallocFinegrained);
        if (e != hipSuccess) return e;
        bufs.push_back({device, p, bytes, true});
        *out = p;
        return hipSuccess;
    }
    void give_back(void* p) {
        std::lock_guard<std::mutex> lk(mu);
        for (Buf& b : bufs)
            if (b.p == p) b.busy = false;
    }
};
FinePool g_fine;

// Copy-engine transfers through the HSA runtime, with completion signals of
// their own.  HIP streams share a few hardware queues per process, and a
// transfer enqueued on a stream that shares the resident kernel's queue would
// wait for the kernel to leave; these transfers are on no queue.
struct Dma {
    hsa_agent_t gpu{}, cpu{};
    bool ready = false;
    std::vector<hsa_signal_t> sigs;
    std::vector<int> free_sigs;
    std::mutex mu;
};

struct AgentSearch {
    uint32_t domain, bus, dev;
    Dma* dma;
    bool found_gpu, found_cpu;
};

hsa_status_t visit_agent(hsa_agent_t agent, void* data) {
    AgentSearch* a = static_cast<AgentSearch*>(data);
    hsa_device_type_t type;
    if (hsa_agent_get_info(agent, HSA_AGENT_INFO_DEVICE, &type) != HSA_STATUS_SUCCESS) return HSA_STATUS_SUCCESS;
    if (type == HSA_DEVICE_TYPE_CPU && !a->found_cpu) {
        a->dma->cpu = agent;
        a->found_cpu = true;
    } else if (type == HSA_DEVICE_TYPE_GPU && !a->found_gpu) {
        uint32_t bdf = 0, domain = 0;
        if (hsa_agent_get_info(agent, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_BDFID, &bdf) == HSA_STATUS_SUCCESS &&
            hsa_agent_get_info(agent, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_DOMAIN, &domain) == HSA_STATUS_SUCCESS &&
            domain == a->domain && ((bdf >> 8) & 0xffu) == a->bus && ((bdf >> 3) & 0x1fu) == a->dev) {
            a->dma->gpu = agent;
            a->found_gpu = true;
        }
    }
    return HSA_STATUS_SUCCESS;
}

// Finds the HSA agents of HIP device `device` and of the host, and makes
// `n_signals` completion signals.
clf_status dma_open(Dma& d, int device, int n_signals) {
    if (hsa_init() != HSA_STATUS_SUCCESS) return fail(CLF_E_DEVICE, "serving: hsa_init failed");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    AgentSearch a{(uint32_t)prop.pciDomainID, (uint32_t)prop.pciBusID, (uint32_t)prop.pciDeviceID, &d, false,
                  false};
    (void)hsa_iterate_agents(visit_agent, &a);
    if (!a.found_gpu || !a.found_cpu)
        return fail(CLF_E_DEVICE, "serving: the device's copy-engine agent was not found");
    d.sigs.resize((size_t)n_signals);
    for (int k = 0; k < n_signals; ++k) {
        if (hsa_signal_create(0, 0, nullptr, &d.sigs[(size_t)k]) != HSA_STATUS_SUCCESS)
            return fail(CLF_E_OUT_OF_MEMORY, "serving: hsa_signal_create failed");
        d.free_sigs.push_back(k);
    }
    d.ready = true;
    return CLF_OK;
}

void dma_close(Dma& d) {
    for (hsa_signal_t s : d.sigs)
        if (s.handle) (void)hsa_signal_destroy(s);
    d.sigs.clear();
    d.free_sigs.clear();
    if (d.ready) (void)hsa_shut_down();
    d.ready = false;
}

int dma_take(Dma& d) {
    std::lock_guard<std::mutex> lk(d.mu);
    if (d.free_sigs.empty()) return -1;
    const int k = d.free_sigs.back();
    d.free_sigs.pop_back();
    return k;
}
void dma_give(Dma& d, int k) {
    std::lock_guard<std::mutex> lk(d.mu);
    d.free_sigs.push_back(k);
}

clf_status hsa_fail(hsa_status_t e, const char* what) {
    const char* msg = nullptr;
    (void)hsa_status_string(e, &msg);
    return fail(CLF_E_DEVICE, std::string(what) + ": " + (msg ? msg : "unknown HSA status"));
}

// 1: finished, 0: in flight, -1: failed.
int dma_state(Dma& d, int k) {
    const hsa_signal_value_t v = hsa_signal_load_scacquire(d.sigs[(size_t)k]);
    return v == 0 ? 1 : (v > 0 ? 0 : -1);
}

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
    bool resident = true;

    PinnedBuf arena;
    Ring in_ring;
    std::vector<Request> reqs;
    uint64_t req_mask = 0;
    int req_shift = 0;
    std::vector<int32_t> free_reqs;
    uint64_t next_seq = 1;

    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_completed;
    std::thread worker;  // dispatcher (batches) or completion collector (resident)
    bool stop = false;
    clf_status error = CLF_OK;
    std::string error_msg;
    bool collect = false;  // replay: finished requests also go to `completed`
    std::deque<int32_t> completed;
    int64_t n_batches = 0, n_batched_events = 0;

    // batches engine
    int n_slots = 0;
    size_t slot_blob_bytes = 0;
    int64_t slot_cap = 0;
    std::deque<int32_t> pending;
    Slot slots[kMaxSlots];
    std::vector<OutBuf> outs;

    // resident engine
    PinnedBuf ring_mem, comp_mem, stop_mem;  // coherent pinned: descriptors, completions, stop word
    void* d_in = nullptr;                    // fine-grained device input arena (same offsets as `arena`)
    void* d_out = nullptr;                   // fine-grained device output ring
    PinnedBuf h_out;                         // pinned mirror of the output ring, plus out_ext bytes
    uint64_t out_ring = 0, out_ext = 0;
    // output ring accounting, absolute byte positions (lock held): completed
    // bytes, capacity bytes of admitted requests not yet completed, and the
    // released prefix
    uint64_t out_done = 0, out_pend = 0, out_tail = 0;
    struct Live {
        uint64_t pos, size;
        bool released;
    };
    std::deque<Live> live;  // ready allocations in position order, from out_tail on
    uint64_t live_base = 0;
    uint32_t ring_mask = 0, comp_mask = 0;
    uint64_t next_desc = 0;
    std::vector<int32_t> desc_req;  // request of each descriptor slot, -1 free
    std::atomic<uint64_t> published{0};
    uint64_t consumed = 0;
    uint32_t comp_next = 0;
    DevBuf ctl, mailbox, idle_ring, wg_blobs, wg_desc, wg_cand, wg_used, wg_tracks, wg_ttf_a, wg_ttf_b, wg_weak,
        wg_results;
    int grid = 0;
    uint32_t idle_mask = 0;
    int64_t wg_blob_bytes = 0, wg_hits = 0, wg_cap = 0;
    // copies between the pinned and the device arenas, in both directions
    Dma dma;
    hipStream_t rstream = nullptr;
    hipEvent_t ev_kernel = nullptr;
    bool kernel_running = false;
    int64_t launches = 0;  // resident: kernel launches (the resident thread's own count)
    int rate_khz = 0;      // resident: device wall clock
    // device phase sums over collected completions (wall-clock ticks)
    uint64_t dev_in = 0, dev_search = 0, dev_out = 0, dev_n = 0;
    double tick_ns = 10.0;
};

namespace {

Request* find_request(clf_server* s, uint64_t ticket) {
    if (ticket == 0) return nullptr;
    Request& r = s->reqs[ticket & s->req_mask];
    return r.ticket == ticket ? &r : nullptr;
}

void free_input(clf_server* s, Request& r) {  // lock held
    if (r.in_live) {
        s->in_ring.release(r.in_seq);
        r.in_live = false;
    }
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

// ---------------------------------------------------------------------------
// Batches engine
// ---------------------------------------------------------------------------
// Returns a finished request's output and its table entry.  Lock held.
bool release_locked(clf_server* s, uint64_t ticket) {
    Request* r = find_request(s, ticket);
    if (!r || r->state.load() != R_DONE) return false;
    if (r->out_buf >= 0) {
        OutBuf& o = s->outs[r->out_buf];
        if (--o.outstanding == 0) o.busy = false;
    }
    r->out_buf = -1;
    if (r->out_live) {
        // its tracks' room in the output ring comes back from the front
        s->live[r->live_seq - s->live_base].released = true;
        while (!s->live.empty() && s->live.front().released) {
            s->out_tail = s->live.front().pos + s->live.front().size;
            s->live.pop_front();
            ++s->live_base;
        }
        r->out_live = false;
    }
    free_input(s, *r);
    r->ticket = 0;
    r->state.store(R_FREE);
    s->free_reqs.push_back((int32_t)(r - s->reqs.data()));
    return true;
}

bool slot_busy(const Slot& sl) { return sl.phase != S_FREE; }

// Picks the next batch for slot `k` and enqueues it; `launched` tells
// whether anything was.
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
    int64_t hit_off = 0, track_off = 0, max_slice = 0;
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i) {
        const Request& r = s->reqs[sl.reqs[i]];
        if (r.in_off != prev_end) {
            if (n_runs == 4) return fail(CLF_E_DEVICE, "serving: batch spans too many arena runs");
            runs[n_runs++] = {r.in_off, pos, 0};
        }
        runs[n_runs - 1].len += r.span;
        d[i].blob_off = (int64_t)pos;
        d[i].hit_off = hit_off;
        d[i].track_off = track_off;
        d[i].n_hits = r.N;
        d[i].n_sensors = r.S;
        pos += r.span;
        prev_end = r.in_off + r.span;
        hit_off += r.N;
        track_off += r.cap;
        max_slice = std::max(max_slice, r.cap);
        cost[i] = r.cost;
    }
    sl.dev.work_records = s->slot_cap;  // what the slot's buffers hold
    sl.dev.max_slice = max_slice;
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
    int64_t batch_hits = 0;
    for (int i = 0; i < n; ++i) batch_hits += std::max(0, d[i].n_hits);
    clf_status cs = enqueue_search(sl.dev, n, CLF_MODE_BATCHED, resolve_launch(s->cfg, batch_hits, n),
                                   s->n_cu, st, nullptr, nullptr);
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
            const int64_t t = now_ns();
            std::lock_guard<std::mutex> lk(s->mu);
            for (int32_t idx : sl.reqs) {
                free_input(s, s->reqs[idx]);
                s->reqs[idx].out.t_complete_ns = t;
            }
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
                r.out.n_tracks = stats[i].n_tracks;
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
            {
                std::lock_guard<std::mutex> lk(s->mu);
                fail_all(s, st, std::string("serving: ") + clf_last_error());
            }
            // drain what is still on the device before leaving
            for (int k = 0; k < s->n_slots; ++k) (void)hipStreamSynchronize(s->slots[k].stream);
            return;
        }
        bool busy = false;
        for (int k = 0; k < s->n_slots; ++k) busy = busy || slot_busy(s->slots[k]);
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

// ---------------------------------------------------------------------------
// Resident engine
// ---------------------------------------------------------------------------
ServeArgs serve_args(clf_server* s) {
    ServeArgs a{};
    SearchArgs& b = a.base;
    b.blobs = s->wg_blobs.as<uint8_t>();
    b.ev = s->wg_desc.as<EventDesc>();
    b.order = nullptr;
    b.first_item = 0;
    b.n_items = 0;
    b.queue_head = nullptr;
    b.grain = 1;
    b.cand = s->wg_cand.as<int4>();
    b.used = s->wg_used.as<uint8_t>();
    b.tracks = s->wg_tracks.as<clf_track>();
    b.ttf_a = s->wg_ttf_a.as<TtfEntry>();
    b.ttf_b = s->wg_ttf_b.as<TtfEntry>();
    b.weak = s->wg_weak.as<uint4>();
    b.slot_records = 0;  // each worker's slices: its descriptor's track_off
    b.slot_base = 0;
    b.results = s->wg_results.as<EventResult>();
    b.packed = nullptr;  // serve_one copies the tracks to the request itself
    b.track_base = nullptr;
    b.track_alloc = nullptr;
    b.done = nullptr;  // no run end: the serving kernel never leaves its queue
    b.done_total = 0;
    b.reset_queue = nullptr;
    b.reset_n = 0;
    b.track_total = nullptr;
    b.fill_only = 0;
    b.prof = nullptr;
    void* p = nullptr;
    (void)hipHostGetDevicePointer(&p, s->ring_mem.p, 0);
    a.ring = static_cast<const ServeDesc*>(p);
    a.ring_mask = s->ring_mask;
    a.in_arena = static_cast<const uint8_t*>(s->d_in);
    a.out_arena = static_cast<uint8_t*>(s->d_out);
    a.in_bytes = s->arena.cap;
    a.out_bytes = s->out_ring;
    (void)hipHostGetDevicePointer(&p, s->comp_mem.p, 0);
    a.comp = static_cast<CompRec*>(p);
    a.comp_mask = s->comp_mask;
    (void)hipHostGetDevicePointer(&p, s->stop_mem.p, 0);
    a.host_stop = static_cast<const uint32_t*>(p);
    a.ctl = s->ctl.as<uint32_t>();
    a.mailbox = s->mailbox.as<uint32_t>();
    a.idle_ring = s->idle_ring.as<uint64_t>();
    a.idle_mask = s->idle_mask;
    a.wg_blob_bytes = s->wg_blob_bytes;
    a.wg_hits = s->wg_hits;
    a.wg_cap = s->wg_cap;
    int rate_khz = 0;
    (void)hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, s->device);
    if (rate_khz <= 0) rate_khz = 100000;
    s->tick_ns = 1.0e6 / (double)rate_khz;
    a.idle_ticks = (uint64_t)s->conf.idle_us * (uint64_t)rate_khz / 1000ull;
    a.pause_ticks = 0;
    s->rate_khz = rate_khz;
    return a;
}

clf_status launch_resident(clf_server* s) {
    // per-launch state: QUIT, IDLE_TAIL, BEAT, the mailboxes and the idle ring
    HIP_TRY(launch_zero_i32(s->ctl.as<int32_t>() + CTL_QUIT, 1, s->rstream));
    HIP_TRY(launch_zero_i32(s->ctl.as<int32_t>() + CTL_IDLE_TAIL, CTL_OUT_HEAD - CTL_IDLE_TAIL, s->rstream));
    HIP_TRY(launch_zero_i32(s->mailbox.as<int32_t>(), s->grid, s->rstream));
    HIP_TRY(launch_zero_i32(s->idle_ring.as<int32_t>(), 2 * (int)(s->idle_mask + 1u), s->rstream));
    ServeArgs a = serve_args(s);
    if (s->launches++ == 0)  // the testing stall, first launch only
        a.pause_ticks = (uint64_t)s->conf.feeder_pause_us * (uint64_t)s->rate_khz / 1000ull;
    HIP_TRY(launch_serve(a, s->cfg.block_size, s->cfg.lds_hits, s->grid, s->rstream));
    HIP_TRY(hipEventRecord(s->ev_kernel, s->rstream));
    s->kernel_running = true;
    std::lock_guard<std::mutex> lk(s->mu);
    ++s->n_batches;  // for the resident engine: launches of the kernel
    return CLF_OK;
}

// Ends a resident request (lock held): its result is ready, or it failed.
void finish_resident(clf_server* s, int32_t idx, clf_status st, const EventResult* res,
                     const uint8_t* tracks, int64_t t_done) {
    Request& r = s->reqs[idx];
    free_input(s, r);
    s->desc_req[r.desc & s->ring_mask] = -1;
    r.out.ticket = r.ticket;
    r.out.tag = r.tag;
    r.out.status = st;
    r.out.batch_events = 1;
    if (res) std::memcpy(&r.out.stats, res, sizeof(clf_event_stats));
    r.out.n_tracks = st == CLF_OK && res ? res->n_tracks : 0;
    r.out.tracks = st == CLF_OK ? tracks : nullptr;
    r.out.t_done_ns = t_done;
    r.state.store(R_DONE);
    if (s->collect) s->completed.push_back(idx);
}

// The resident engine's view of the output ring on the completion thread:
// completed allocations not yet contiguous with the copied part, the
// contiguous end, what has been copied back, and the copies in flight.
struct OutCopy {
    struct Done {
        uint64_t pos, size;
        int32_t req;
        EventResult res;
        bool operator>(const Done& o) const { return pos > o.pos; }
    };
    std::priority_queue<Done, std::vector<Done>, std::greater<Done>> waiting;
    uint64_t ready = 0;  // every allocation below is complete
    std::vector<Done> batch;  // ready, not yet in a copy
    struct InFlight {
        int sig;
        std::vector<Done> reqs;
    };
    std::deque<InFlight> flights;
    // the signal of the last copy issued (the next one waits for it), and
    // of the last one finished (kept until its successor has finished too,
    // since that one may still be waiting on it)
    int last_issued = -1, last_done = -1;
};

// Takes finished requests off the completion ring, in order, and moves the
// output ring's completed prefix toward the host.
bool collect_completions(clf_server* s, std::vector<CompRec>& buf, OutCopy& oc) {
    CompRec* ring = s->comp_mem.as<CompRec>();
    buf.clear();
    for (;;) {
        CompRec* c = &ring[s->comp_next & s->comp_mask];
        if (__atomic_load_n(&c->seq, __ATOMIC_ACQUIRE) != s->comp_next + 1u) break;
        buf.push_back(*c);
        ++s->comp_next;
        if (buf.size() >= 4096) break;
    }
    if (buf.empty()) return false;
    const int64_t t = now_ns();
    std::lock_guard<std::mutex> lk(s->mu);
    for (const CompRec& rec : buf) {
        ++s->consumed;
        ++s->n_batched_events;
        s->dev_in += rec.t_copied - rec.t_claim;
        s->dev_search += rec.t_searched - rec.t_copied;
        s->dev_out += rec.t_written - rec.t_searched;
        ++s->dev_n;
        if (rec.req < 0 || (size_t)rec.req >= s->reqs.size()) continue;
        Request& r = s->reqs[rec.req];
        if (r.state.load() != R_PENDING) continue;  // already ended by a failure
        r.out.t_complete_ns = t;
        const uint64_t size = (uint64_t)std::max(rec.res.n_tracks, 0) * sizeof(clf_track);
        if (r.out_pending) {
            s->out_pend -= r.out_cap_bytes;
            r.out_pending = false;
        }
        if (rec.status != COMP_OK) {
            finish_resident(s, rec.req, CLF_E_DEVICE, &rec.res, nullptr, t);
            continue;
        }
        s->out_done += size;
        if (size == 0) {
            finish_resident(s, rec.req, CLF_OK, &rec.res, s->h_out.as<uint8_t>(), t);
            continue;
        }
        oc.waiting.push({rec.out_pos, size, rec.req, rec.res});
    }
    // allocations are contiguous: the next ready one starts where the last ended
    while (!oc.waiting.empty() && oc.waiting.top().pos == oc.ready) {
        const OutCopy::Done d = oc.waiting.top();
        oc.waiting.pop();
        oc.ready = d.pos + d.size;
        s->live.push_back({d.pos, d.size, false});
        Request& r = s->reqs[d.req];
        r.out_live = true;
        r.live_seq = s->live_base + s->live.size() - 1;
        oc.batch.push_back(d);
    }
    return true;
}

// Copies the batch's part of the output ring [batch[0].pos, ready) into the
// pinned mirror at the same offsets (an allocation that runs past the
// ring's end is also copied past the mirror's end, so that its tracks are
// contiguous there); results are handed out when the copy has finished.
// Each piece is widened to kOutGrain boundaries.  That is safe: below the
// batch the device still holds the bytes already copied (live results are
// never overwritten, by the admission rule), and above it lie positions
// that no live result occupies in the mirror, which their own copies
// overwrite later (the copies run in order on one stream).
clf_status copy_back(clf_server* s, OutCopy& oc) {
    if (oc.batch.empty() || oc.flights.size() >= kCopiesInFlight) return CLF_OK;
    const int k = dma_take(s->dma);
    if (k < 0) return CLF_OK;  // every signal in use: next round
    const uint64_t R = s->out_ring, G = kOutGrain;
    const uint64_t lo = oc.batch.front().pos, hi = oc.ready;
    uint8_t* dev = static_cast<uint8_t*>(s->d_out);
    uint8_t* host = s->h_out.as<uint8_t>();
    struct Piece {
        uint8_t* dst;
        uint64_t a, b;
    };
    Piece pieces[8];
    int np = 0;
    auto piece = [&](uint8_t* dst, uint64_t from, uint64_t to, uint64_t limit) {
        pieces[np++] = {dst, from / G * G, std::min(limit, (to + G - 1) / G * G)};
    };
    const uint64_t a = lo % R, len = hi - lo;
    if (a + len <= R) {
        piece(host, a, a + len, R);
    } else {
        piece(host, a, R, R);
        piece(host, 0, a + len - R, R);
    }
    for (const OutCopy::Done& d : oc.batch) {
        const uint64_t b0 = d.pos % R;
        if (b0 + d.size > R) {  // this one wraps: its tail again past the mirror's end
            if (b0 + d.size - R > s->out_ext) {
                dma_give(s->dma, k);
                return fail(CLF_E_DEVICE, "serving: output larger than the mirror");
            }
            if (np < 8) piece(host + R, 0, b0 + d.size - R, s->out_ext);
        }
    }
    // in order after the previous copy (the widened pieces of consecutive
    // copies overlap, and the later one must land last)
    hsa_signal_t sig = s->dma.sigs[(size_t)k];
    hsa_signal_store_relaxed(sig, np);
    const int ndep = oc.last_issued >= 0 ? 1 : 0;
    hsa_signal_t dep = ndep ? s->dma.sigs[(size_t)oc.last_issued] : hsa_signal_t{0};
    for (int i = 0; i < np; ++i) {
        const hsa_status_t e = hsa_amd_memory_async_copy(pieces[i].dst + pieces[i].a, s->dma.cpu, dev + pieces[i].a,
                                                         s->dma.gpu, pieces[i].b - pieces[i].a, ndep,
                                                         ndep ? &dep : nullptr, sig);
        if (e != HSA_STATUS_SUCCESS) {
            char what[256];
            std::snprintf(what, sizeof(what), "serving: output copy could not be issued (dst %p src %p bytes %llu, %d deps)",
                          (void*)(pieces[i].dst + pieces[i].a), (void*)(dev + pieces[i].a),
                          (unsigned long long)(pieces[i].b - pieces[i].a), ndep);
            return hsa_fail(e, what);
        }
    }
    oc.last_issued = k;
    oc.flights.push_back({k, std::move(oc.batch)});
    oc.batch.clear();
    return CLF_OK;
}

// Hands out the results of finished copies, oldest first.
clf_status finish_copies(clf_server* s, OutCopy& oc, bool& any) {
    any = false;
    while (!oc.flights.empty()) {
        const int st = dma_state(s->dma, oc.flights.front().sig);
        if (st == 0) break;
        if (st < 0) return fail(CLF_E_DEVICE, "serving: output copy failed");
        OutCopy::InFlight f = std::move(oc.flights.front());
        oc.flights.pop_front();
        if (oc.last_done >= 0) dma_give(s->dma, oc.last_done);
        oc.last_done = f.sig;
        if (oc.last_issued == f.sig && oc.flights.empty()) {
            // nothing waits on it any more, and the next copy need not
            oc.last_issued = -1;
            dma_give(s->dma, f.sig);
            oc.last_done = -1;
        }
        const int64_t t = now_ns();
        std::lock_guard<std::mutex> lk(s->mu);
        for (const OutCopy::Done& d : f.reqs) {
            if (s->reqs[d.req].state.load() != R_PENDING) continue;
            finish_resident(s, d.req, CLF_OK, &d.res, s->h_out.as<uint8_t>() + d.pos % s->out_ring, t);
        }
        any = true;
    }
    return CLF_OK;
}

// The upload step of the resident loop: runs of newly submitted events (in
// descriptor order, each already copied into the staging arena by its
// submitter) go to the device input arena in one transfer per contiguous
// run, at most kCopiesInFlight transfers at a time; when a transfer has
// finished, the descriptors of its events are published.
struct Upload {
    uint64_t next = 0;       // first descriptor not yet sent
    uint64_t published = 0;  // first descriptor not yet published
    struct Sent {
        int sig;
        uint64_t d_end;
    };
    std::deque<Sent> sent;
    struct Run {
        size_t off, len;
    };
    std::vector<Run> runs;
};

clf_status upload_step(clf_server* s, Upload& u, bool& progress) {
    // transfers that have finished: publish their descriptors
    while (!u.sent.empty()) {
        const int st = dma_state(s->dma, u.sent.front().sig);
        if (st == 0) break;
        if (st < 0) return fail(CLF_E_DEVICE, "serving: input copy failed");
        const uint64_t end = u.sent.front().d_end;
        dma_give(s->dma, u.sent.front().sig);
        u.sent.pop_front();
        const int64_t t = now_ns();
        ServeDesc* ring = s->ring_mem.as<ServeDesc>();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (; u.published < end; ++u.published) {
                const int32_t idx = s->desc_req[u.published & s->ring_mask];
                if (idx < 0 || s->reqs[idx].state.load() != R_PENDING) continue;  // ended by a failure
                Request& r = s->reqs[idx];
                ServeDesc* sd = ring + (u.published & s->ring_mask);
                sd->in_off = r.in_off;
                sd->reserved = 0;
                sd->n_hits = r.N;
                sd->n_sensors = r.S;
                sd->req = idx;
                sd->out_cap = (int32_t)CLF_EVENT_TRACK_CAPACITY(r.N, r.S);
                r.out.t_launch_ns = t;
                __atomic_store_n(&sd->seq, (uint32_t)(u.published + 1), __ATOMIC_RELEASE);
            }
        }
        s->published.store(u.published);
        progress = true;
    }
    if (u.sent.size() >= kCopiesInFlight) return CLF_OK;
    // the newly submitted events whose staging copies are complete
    uint64_t d_end = u.next;
    u.runs.clear();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        while (d_end < s->next_desc && u.runs.size() < 64) {
            const int32_t idx = s->desc_req[d_end & s->ring_mask];
            if (idx < 0) break;
            const Request& r = s->reqs[idx];
            if (r.desc != d_end || r.state.load(std::memory_order_acquire) != R_PENDING) break;
            if (!u.runs.empty() && u.runs.back().off + u.runs.back().len == r.in_off)
                u.runs.back().len += r.span;
            else
                u.runs.push_back({r.in_off, r.span});
            ++d_end;
        }
    }
    if (d_end == u.next) return CLF_OK;
    const int k = dma_take(s->dma);
    if (k < 0) return CLF_OK;
    hsa_signal_t sig = s->dma.sigs[(size_t)k];
    hsa_signal_store_relaxed(sig, (hsa_signal_value_t)u.runs.size());
    for (const Upload::Run& run : u.runs) {
        const hsa_status_t e = hsa_amd_memory_async_copy(static_cast<uint8_t*>(s->d_in) + run.off, s->dma.gpu,
                                                         s->arena.as<uint8_t>() + run.off, s->dma.cpu, run.len,
                                                         0, nullptr, sig);
        if (e != HSA_STATUS_SUCCESS) return hsa_fail(e, "serving: input copy could not be issued");
    }
    u.sent.push_back({k, d_end});
    u.next = d_end;
    progress = true;
    return CLF_OK;
}

void resident_loop(clf_server* s) {
    (void)hipSetDevice(s->device);
    std::vector<CompRec> buf;
    buf.reserve(4096);
    OutCopy oc;
    Upload up;
    int64_t last_query = 0;
    for (;;) {
        bool progress = false;
        clf_status st = upload_step(s, up, progress);
        progress = collect_completions(s, buf, oc) || progress;
        bool copied = false;
        if (st == CLF_OK) st = finish_copies(s, oc, copied);
        if (st == CLF_OK) st = copy_back(s, oc);
        progress = progress || copied;
        if (copied || progress) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->cv_done.notify_all();
            if (s->collect) s->cv_completed.notify_all();
        }
        bool stopping, outstanding;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            stopping = s->stop;
            outstanding = s->consumed < s->published.load();
        }
        // has the kernel left?  (asked at most every 100 us while it has work)
        const int64_t now = now_ns();
        if (st == CLF_OK && s->kernel_running && (!outstanding || stopping || now - last_query > 100000)) {
            last_query = now;
            hipError_t q = hipEventQuery(s->ev_kernel);
            if (q == hipSuccess) {
                s->kernel_running = false;
                progress = collect_completions(s, buf, oc) || progress;
                std::lock_guard<std::mutex> lk(s->mu);
                outstanding = s->consumed < s->published.load();
            } else if (q != hipErrorNotReady) {
                st = hip_fail(q, "serving: resident kernel");
            }
        }
        if (st == CLF_OK && !s->kernel_running && !stopping && outstanding) {
            st = launch_resident(s);
            progress = true;
        }
        if (st != CLF_OK) {
            std::lock_guard<std::mutex> lk(s->mu);
            fail_all(s, st, std::string("serving: ") + clf_last_error());
            __atomic_store_n(s->stop_mem.as<uint32_t>(), 1u, __ATOMIC_RELEASE);
        }
        const bool copies_left = !oc.flights.empty() || !oc.batch.empty() || !up.sent.empty();
        if ((stopping || st != CLF_OK) && !s->kernel_running && !copies_left) return;
        if (st != CLF_OK) {
            (void)hipStreamSynchronize(s->rstream);
            // let the copies in flight end before their buffers can go
            for (const Upload::Sent& x : up.sent)
                while (dma_state(s->dma, x.sig) == 0) std::this_thread::yield();
            for (const OutCopy::InFlight& x : oc.flights)
                while (dma_state(s->dma, x.sig) == 0) std::this_thread::yield();
            s->kernel_running = false;
            return;
        }
        if (progress) continue;
        if (s->kernel_running || outstanding || copies_left) {
            std::this_thread::yield();
            continue;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_work.wait_for(lk, std::chrono::microseconds(1000));
    }
}

clf_status create_resident(clf_server* s) {
    const clf_server_config& c = s->conf;
    int per_cu = 0;
    HIP_TRY(serve_occupancy(s->cfg.block_size, s->cfg.lds_hits, &per_cu));
    if (per_cu < 1) return fail(CLF_E_DEVICE, "serving kernel does not fit on a compute unit");
    s->grid = per_cu * s->n_cu;  // every workgroup resident at once; 0 is the feeder
    s->wg_hits = c.max_event_hits;
    s->wg_cap = track_slice(s->wg_hits, CLF_MAX_SENSORS);
    s->wg_blob_bytes = (int64_t)align_up(event_bytes(CLF_MAX_SENSORS, (int32_t)s->wg_hits), kBlobAlign);
    const size_t g = (size_t)s->grid;
    int ring_n = 1, comp_n = 1;
    while (ring_n < 2 * c.max_requests) ring_n <<= 1;
    while (comp_n < c.max_requests) comp_n <<= 1;
    s->ring_mask = (uint32_t)ring_n - 1u;
    s->comp_mask = (uint32_t)comp_n - 1u;
    s->desc_req.assign((size_t)ring_n, -1);
    // the output ring (a multiple of kOutGrain) and its pinned mirror, which
    // has room past its end for the largest event's tracks
    s->out_ring = (uint64_t)c.out_arena_bytes / kOutGrain * kOutGrain;
    s->out_ext = ((uint64_t)CLF_EVENT_TRACK_CAPACITY(s->wg_hits, CLF_MAX_SENSORS) * sizeof(clf_track) +
                  kOutGrain - 1) / kOutGrain * kOutGrain;
    const unsigned coh = hipHostMallocCoherent | hipHostMallocMapped;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    chk(g_fine.take(s->device, s->arena.cap, &s->d_in));
    chk(g_fine.take(s->device, (size_t)s->out_ring, &s->d_out));
    chk(s->h_out.ensure((size_t)(s->out_ring + s->out_ext)));
    chk(s->ring_mem.ensure(sizeof(ServeDesc) * (size_t)ring_n, coh));
    chk(s->comp_mem.ensure(sizeof(CompRec) * (size_t)comp_n, coh));
    chk(s->stop_mem.ensure(256, coh));
    chk(s->ctl.ensure(sizeof(uint32_t) * CTL_WORDS));
    int idle_n = 1;
    while (idle_n < s->grid) idle_n <<= 1;
    s->idle_mask = (uint32_t)idle_n - 1u;
    chk(s->mailbox.ensure(sizeof(uint32_t) * g));
    chk(s->idle_ring.ensure(sizeof(uint64_t) * (size_t)idle_n));
    chk(s->wg_blobs.ensure(g * (size_t)s->wg_blob_bytes));
    chk(s->wg_desc.ensure(sizeof(EventDesc) * g));
    chk(s->wg_cand.ensure(sizeof(int4) * g * (size_t)s->wg_hits));
    chk(s->wg_used.ensure(g * (size_t)s->wg_hits));
    chk(s->wg_tracks.ensure(sizeof(clf_track) * g * (size_t)s->wg_cap));
    chk(s->wg_ttf_a.ensure(sizeof(TtfEntry) * g * (size_t)s->wg_cap));
    chk(s->wg_ttf_b.ensure(sizeof(TtfEntry) * g * (size_t)s->wg_cap));
    chk(s->wg_weak.ensure(sizeof(uint4) * g * (size_t)s->wg_cap));
    chk(s->wg_results.ensure(sizeof(EventResult) * g));
    chk(hipStreamCreateWithFlags(&s->rstream, hipStreamNonBlocking));
    chk(hipEventCreateWithFlags(&s->ev_kernel, hipEventDisableTiming));
    if (e != hipSuccess) return hip_fail(e, "resident server allocation");
    clf_status ds = dma_open(s->dma, s->device, 64);
    if (ds != CLF_OK) return ds;
    // touch every page once so that no request pays for it
    std::memset(s->h_out.p, 0, s->h_out.cap);
    std::memset(s->ring_mem.p, 0, s->ring_mem.cap);
    std::memset(s->comp_mem.p, 0, s->comp_mem.cap);
    std::memset(s->stop_mem.p, 0, s->stop_mem.cap);
    HIP_TRY(launch_zero_i32(s->ctl.as<int32_t>(), CTL_WORDS, s->rstream));
    HIP_TRY(hipStreamSynchronize(s->rstream));
    return CLF_OK;
}

clf_status create_batches(clf_server* s) {
    const clf_server_config& c = s->conf;
    s->n_slots = c.slots;
    s->slot_blob_bytes = (size_t)(16 * c.max_batch_hits + kEventBlobOverhead * c.max_batch_events);
    s->slot_cap = c.max_batch_hits + kEventCapOverhead * (int64_t)c.max_batch_events;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
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
    if (e != hipSuccess) return hip_fail(e, "server allocation");
    return CLF_OK;
}

// Writes the words of an event blob that the search reads, packed
// (SearchedWords), to dst: what the resident engine stages and uploads.
// 16-byte words to a 16-byte aligned staging destination with
// non-temporal stores: the copy engine reads the staging arena soon after,
// and lines left dirty in the host's caches would have to be snooped
void stream_copy16(uint8_t* dst, const uint8_t* src, size_t n16) {
    __m128i* d = reinterpret_cast<__m128i*>(dst);
    const __m128i* q = reinterpret_cast<const __m128i*>(src);
    for (size_t i = 0; i < n16; ++i) _mm_stream_si128(d + i, _mm_loadu_si128(q + i));
}
// The words the search reads (header and sensor table, then the
// coordinates from the word holding X[0], then the last bytes), packed; the
// fence orders the non-temporal stores before the request's publication.
void stage_searched(uint8_t* dst, const uint8_t* blob, const SearchedWords& w) {
    stream_copy16(dst, blob, (size_t)std::min(w.h16, w.m16));
    if (w.m16 > w.h16) stream_copy16(dst + 16 * w.h16, blob + 16 * w.x16, (size_t)(w.m16 - w.h16));
    std::memcpy(dst + 16 * w.m16, blob + 16 * w.n16, (size_t)(w.bytes - 16 * w.n16));
    _mm_sfence();
}

// Submits events[0..n) in order and stops at the first one that is
// malformed or does not fit: `*accepted` of them are queued, and the status
// is that of the first one refused (CLF_OK if none was).  The requests, arena
// spans and descriptors of all of them are taken under one acquisition of
// the server lock; the blobs are copied and the descriptors published after
// it, in order.
clf_status submit_events(clf_server* s, int32_t n, const uint8_t* const* events, const size_t* sizes,
                         const uint64_t* tags, uint64_t* tickets, int32_t* accepted) {
    struct Sized {
        int32_t S, N;
        size_t bytes, staged, span;  // blob, staged form, its arena span
        int64_t cap;
        uint64_t out_cap_bytes;
    };
    Sized z[kMaxSubmitMany];
    Request* rq[kMaxSubmitMany];
    *accepted = 0;
    clf_status refused = CLF_OK;
    const char* why = nullptr;
    int32_t valid = 0;
    for (; valid < n; ++valid) {
        Sized& e = z[valid];
        if (!validate_event(events[valid], sizes[valid], e.S, e.N)) {
            refused = CLF_E_INVALID_ARGUMENT, why = "malformed event";
            break;
        }
        e.bytes = event_bytes(e.S, e.N);
        // the resident engine ships only the words the search reads
        e.staged = s->resident ? (size_t)searched_words(e.S, e.N).packed : e.bytes;
        e.span = align_up(e.staged, kBlobAlign);
        e.cap = track_slice(e.N, e.S);
        e.out_cap_bytes = (uint64_t)CLF_EVENT_TRACK_CAPACITY(e.N, e.S) * sizeof(clf_track);
        if (s->resident) {
            if (e.N > s->wg_hits || e.cap > s->wg_cap ||
                (int64_t)align_up(e.bytes, kBlobAlign) > s->wg_blob_bytes ||
                e.out_cap_bytes > s->out_ring) {
                refused = CLF_E_CAPACITY, why = "event larger than the serving kernel's scratch";
                break;
            }
        } else if (e.span > s->slot_blob_bytes || e.N > s->conf.max_batch_hits || e.cap > s->slot_cap) {
            refused = CLF_E_CAPACITY, why = "event larger than a serving batch";
            break;
        }
    }
    int32_t taken_n = 0;
    uint64_t d0 = 0;
    std::string err_msg;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        d0 = s->next_desc;
        for (; taken_n < valid; ++taken_n) {
            const Sized& e = z[taken_n];
            if (s->error != CLF_OK) {
                refused = s->error, err_msg = s->error_msg;
                break;
            }
            if (s->stop) {
                refused = CLF_E_INVALID_ARGUMENT, why = "server stopped";
                break;
            }
            if (s->free_reqs.empty()) {
                refused = CLF_E_CAPACITY, why = "request table full";
                break;
            }
            size_t off, taken;
            if (!s->in_ring.fits(e.span, off, taken)) {
                refused = CLF_E_CAPACITY, why = "input arena full";
                break;
            }
            uint64_t d = 0;
            if (s->resident) {
                d = s->next_desc;
                if (s->desc_req[d & s->ring_mask] >= 0) {
                    refused = CLF_E_CAPACITY, why = "descriptor ring full";
                    break;
                }
                // the output ring holds every admitted event's capacity bound
                // beyond what completed events have used, from its released
                // prefix on, so a worker's reservation never overruns it
                if (s->out_done + s->out_pend + e.out_cap_bytes > s->out_tail + s->out_ring) {
                    refused = CLF_E_CAPACITY, why = "device output ring full";
                    break;
                }
            }
            const int32_t idx = s->free_reqs.back();
            s->free_reqs.pop_back();
            Request* r = &s->reqs[idx];
            r->ticket = (s->next_seq++ << s->req_shift) | (uint64_t)idx;
            r->tag = tags[taken_n];
            r->in_off = off;
            r->span = e.span;
            r->in_seq = s->in_ring.commit(off, e.span, taken);
            r->in_live = true;
            if (s->resident) {
                r->desc = d;
                s->desc_req[d & s->ring_mask] = idx;
                ++s->next_desc;
                r->out_cap_bytes = e.out_cap_bytes;
                r->out_pending = true;
                r->out_live = false;
                s->out_pend += e.out_cap_bytes;
            }
            r->S = e.S;
            r->N = e.N;
            r->cap = e.cap;
            r->out_buf = -1;
            r->out = clf_served{};
            r->state.store(R_WRITING);
            if (!s->resident) s->pending.push_back(idx);
            tickets[taken_n] = r->ticket;
            rq[taken_n] = r;
        }
    }
    (void)d0;
    for (int32_t k = 0; k < taken_n; ++k) {
        Request* r = rq[k];
        const Sized& e = z[k];
        if (s->resident)
            stage_searched(s->arena.as<uint8_t>() + r->in_off, events[k], searched_words(e.S, e.N));
        else
            std::memcpy(s->arena.as<uint8_t>() + r->in_off, events[k], e.bytes);
        r->cost = event_cost(events[k], e.S, e.N);
        r->out.t_submit_ns = now_ns();
        // a failure of the server may have ended the request meanwhile; it
        // is then never uploaded (the kernel has been told to stop)
        int expected = R_WRITING;
        (void)r->state.compare_exchange_strong(expected, R_PENDING, std::memory_order_acq_rel);
    }
    *accepted = taken_n;
    if (taken_n > 0) s->cv_work.notify_one();
    if (refused != CLF_OK) return why ? fail(refused, why) : fail(refused, err_msg);
    return CLF_OK;
}

}  // namespace

extern "C" {

void clf_server_default_config(clf_server_config* cfg) {
    if (!cfg) return;
    cfg->engine = CLF_SERVE_RESIDENT;
    cfg->max_requests = 65536;
    cfg->arena_bytes = (int64_t)512 << 20;
    cfg->out_arena_bytes = (int64_t)2 << 30;  // device output ring
    cfg->max_event_hits = 32768;
    cfg->idle_us = 5000;
    cfg->slots = 4;
    cfg->max_batch_events = 256;
    cfg->max_batch_hits = 1 << 20;
    cfg->max_wait_us = 100;
    cfg->feeder_pause_us = 0;
}

void clf_server_destroy(clf_server* s) {
    if (!s) return;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        if (s->stop_mem.p) __atomic_store_n(s->stop_mem.as<uint32_t>(), 1u, __ATOMIC_RELEASE);
        s->cv_work.notify_all();
        s->worker.join();
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
    if (s->rstream) (void)hipStreamSynchronize(s->rstream);
    DevBuf* dev[] = {&s->ctl,       &s->mailbox,  &s->idle_ring, &s->wg_blobs,
                     &s->wg_desc,   &s->wg_cand,  &s->wg_used,   &s->wg_tracks,
                     &s->wg_ttf_a,  &s->wg_ttf_b, &s->wg_weak,   &s->wg_results};
    for (DevBuf* d : dev) d->release();
    if (s->d_in) g_fine.give_back(s->d_in);
    if (s->d_out) g_fine.give_back(s->d_out);
    dma_close(s->dma);
    if (s->ev_kernel) (void)hipEventDestroy(s->ev_kernel);
    if (s->rstream) (void)hipStreamDestroy(s->rstream);
    s->h_out.release();
    s->ring_mem.release();
    s->comp_mem.release();
    s->stop_mem.release();
    s->arena.release();
    delete s;
}

clf_status clf_server_create(clf_context* ctx, const clf_server_config* cfg, clf_server** out) {
    if (!ctx || !out) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    clf_server_config c;
    clf_server_default_config(&c);
    if (cfg) c = *cfg;
    if ((c.engine != CLF_SERVE_RESIDENT && c.engine != CLF_SERVE_BATCHES) || c.max_requests < 1 ||
        c.max_requests > (1 << 22) || c.arena_bytes < (int64_t)(1 << 20))
        return fail(CLF_E_INVALID_ARGUMENT, "bad server configuration");
    if (c.engine == CLF_SERVE_RESIDENT &&
        (c.max_event_hits < 1 || c.max_event_hits > CLF_MAX_SENSOR_HITS * 64 || c.idle_us < 1 ||
         c.idle_us > 10000000 || c.out_arena_bytes < (int64_t)(1 << 20) || c.feeder_pause_us < 0 ||
         c.feeder_pause_us > 10000000))
        return fail(CLF_E_INVALID_ARGUMENT, "bad resident server configuration");
    if (c.engine == CLF_SERVE_BATCHES &&
        (c.slots < 1 || c.slots > kMaxSlots || c.max_batch_events < 1 ||
         c.max_batch_events > (1 << 20) || c.max_batch_hits < 1 || c.max_wait_us < 0))
        return fail(CLF_E_INVALID_ARGUMENT, "bad batch server configuration");
    clf_launch_config lc;
    clf_context_get_config(ctx, &lc);
    const int device = context_device(ctx);
    clf_server* s = new clf_server();
    s->ctx = ctx;
    s->n_cu = clf_context_num_cus(ctx);
    s->conf = c;
    s->resident = c.engine == CLF_SERVE_RESIDENT;
    // the resident kernel's launch is fixed at creation: an automatic
    // configuration takes the shape for events of a few thousand hits
    // (512 x 256; 1024-thread workers served about 515k events/s against
    // 705k); the batch engine resolves it per batch
    s->cfg = s->resident ? resolve_launch(lc, 0, 0) : lc;
    s->device = device;
    int shift = 0;
    while ((1ll << shift) < c.max_requests) ++shift;
    s->req_shift = shift;
    s->req_mask = (1ull << shift) - 1ull;
    s->reqs = std::vector<Request>((size_t)1 << shift);
    s->free_reqs.reserve(s->reqs.size());
    for (int64_t i = (int64_t)c.max_requests - 1; i >= 0; --i) s->free_reqs.push_back((int32_t)i);

    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = s->arena.ensure((size_t)c.arena_bytes);  // staging: read by the copy engines only
    if (e != hipSuccess) {
        clf_server_destroy(s);
        return hip_fail(e, "server arena");
    }
    std::memset(s->arena.p, 0, s->arena.cap);  // touch every page once
    s->in_ring.cap = s->arena.cap;
    clf_status st = s->resident ? create_resident(s) : create_batches(s);
    if (st != CLF_OK) {
        clf_server_destroy(s);
        return st;
    }
    s->worker = std::thread(s->resident ? resident_loop : dispatch_loop, s);
    *out = s;
    return CLF_OK;
}

clf_status clf_server_submit(clf_server* s, const uint8_t* event, size_t size, uint64_t tag,
                             uint64_t* ticket) {
    if (!s || !ticket) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    *ticket = 0;
    int32_t accepted = 0;
    return submit_events(s, 1, &event, &size, &tag, ticket, &accepted);
}

clf_status clf_server_submit_many(clf_server* s, int32_t n, const uint8_t* const* events,
                                  const size_t* sizes, const uint64_t* tags, uint64_t* tickets,
                                  int32_t* n_accepted) {
    if (n_accepted) *n_accepted = 0;
    if (!s || !n_accepted || n < 0 || n > kMaxSubmitMany ||
        (n > 0 && (!events || !sizes || !tags || !tickets)))
        return fail(CLF_E_INVALID_ARGUMENT, "bad submit_many argument");
    for (int32_t k = 0; k < n; ++k) tickets[k] = 0;
    return submit_events(s, n, events, sizes, tags, tickets, n_accepted);
}

clf_status clf_server_submit_framed(clf_server* s, const uint8_t* frame, size_t size, uint64_t tag,
                                    uint64_t* ticket) {
    const uint8_t* data = nullptr;
    size_t n = 0;
    clf_status st = clf_unframe(frame, size, nullptr, nullptr, &data, &n);
    if (st != CLF_OK) return st;
    return clf_server_submit(s, data, n, tag, ticket);
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
    // a request ended by a server failure carries the failure's message
    if (out->status != CLF_OK && s->error == out->status && !s->error_msg.empty())
        (void)fail((clf_status)out->status, s->error_msg);
    return CLF_OK;
}

clf_status clf_server_release(clf_server* s, uint64_t ticket) {
    if (!s) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!release_locked(s, ticket))
            return fail(CLF_E_INVALID_ARGUMENT, "ticket unknown or not finished");
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
        rec[i] = clf_replay_record{-1, -1, 0, 0, -1, 0, 0, -1};
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->collect) return fail(CLF_E_INVALID_ARGUMENT, "a replay is already running");
        s->collect = true;
        s->completed.clear();
    }
    int64_t b0, e0;
    uint64_t din0, dsr0, dout0, dn0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        b0 = s->n_batches;
        e0 = s->n_batched_events;
        din0 = s->dev_in;
        dsr0 = s->dev_search;
        dout0 = s->dev_out;
        dn0 = s->dev_n;
    }
    std::atomic<int64_t> submit_ns{0}, digest_ns{0};
    const int64_t start = now_ns() + 2000000;  // 2 ms for the threads to start
    std::atomic<int64_t> remaining{n};
    std::atomic<int64_t> max_lag{0};
    std::atomic<int64_t> last_done{start};

    // Thread t submits arrivals t, t + threads, ...; when it runs late, the
    // arrivals of its share already due go in one clf_server_submit_many.
    auto submitter = [&](int t) {
        const uint8_t* ev[kMaxSubmitMany];
        size_t sz[kMaxSubmitMany];
        uint64_t tg[kMaxSubmitMany], tk[kMaxSubmitMany];
        for (int64_t i = t; i < n;) {
            const int64_t target = start + arrival_ns[i];
            int64_t now = now_ns();
            if (target - now > 120000)  // sleep most of a long gap, spin the rest
                std::this_thread::sleep_for(std::chrono::nanoseconds(target - now - 80000));
            while ((now = now_ns()) < target) {
            }
            int32_t m = 0;
            for (int64_t j = i; j < n && m < kMaxSubmitMany && (m == 0 || start + arrival_ns[j] <= now);
                 j += submit_threads) {
                const int64_t lag = now - (start + arrival_ns[j]);
                int64_t ml = max_lag.load();
                while (lag > ml && !max_lag.compare_exchange_weak(ml, lag)) {
                }
                const int p = pick[j];
                ev[m] = pool[p];
                sz[m] = sizes[p];
                tg[m] = (uint64_t)j;
                ++m;
            }
            i += (int64_t)m * submit_threads;
            for (int32_t k = 0; k < m;) {
                int32_t acc = 0;
                const clf_status st = clf_server_submit_many(s, m - k, ev + k, sz + k, tg + k, tk + k, &acc);
                k += acc;
                if (st == CLF_OK) break;
                rec[tg[k]].status = st;  // arrival tg[k] was refused; the rest are tried again
                ++k;
                if (remaining.fetch_sub(1) == 1) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->cv_completed.notify_all();
                }
            }
            submit_ns.fetch_add(now_ns() - now);
        }
    };
    auto collector = [&]() {
        std::vector<clf_served> got;
        got.reserve(64);
        for (;;) {
            got.clear();
            {
                std::unique_lock<std::mutex> lk(s->mu);
                s->cv_completed.wait(lk, [&] { return !s->completed.empty() || remaining.load() <= 0; });
                if (s->completed.empty()) return;
                while (!s->completed.empty() && got.size() < 64) {
                    got.push_back(s->reqs[s->completed.front()].out);
                    s->completed.pop_front();
                }
            }
            const int64_t c0 = now_ns();
            int64_t ld_max = 0;
            for (const clf_served& o : got) {
                const int64_t i = (int64_t)o.tag;
                const int64_t arr = start + arrival_ns[i];
                clf_replay_record& r = rec[i];
                r.status = o.status;
                r.latency_ns = o.t_done_ns - arr;
                r.wait_ns = o.t_launch_ns > 0 ? o.t_launch_ns - arr : -1;
                r.n_tracks = (int32_t)o.n_tracks;
                r.flags = o.stats.flags;
                r.batch_events = o.batch_events;
                r.device_ns = o.t_complete_ns > 0 && o.t_launch_ns > 0 ? o.t_complete_ns - o.t_launch_ns : -1;
                r.digest = o.status == CLF_OK ? clf_tracks_digest(o.tracks, o.n_tracks) : 0;
                ld_max = std::max(ld_max, o.t_done_ns);
            }
            int64_t ld = last_done.load();
            while (ld_max > ld && !last_done.compare_exchange_weak(ld, ld_max)) {
            }
            {
                std::lock_guard<std::mutex> lk(s->mu);
                for (const clf_served& o : got) (void)release_locked(s, o.ticket);
            }
            s->cv_work.notify_one();
            digest_ns.fetch_add(now_ns() - c0);
            if (remaining.fetch_sub((int64_t)got.size()) == (int64_t)got.size()) {
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
        if (s->resident)
            sum->mean_batch_events = sum->completed > 0 ? 1.0 : 0.0;
        else
            sum->mean_batch_events =
                sum->batches > 0 ? (double)(s->n_batched_events - e0) / (double)sum->batches : 0.0;
        sum->max_submit_lag_ns = (double)max_lag.load();
        const double dn = (double)(s->dev_n - dn0);
        const double tk = s->tick_ns * 1e-3;
        sum->dev_copy_in_us = dn > 0 ? (double)(s->dev_in - din0) * tk / dn : 0.0;
        sum->dev_search_us = dn > 0 ? (double)(s->dev_search - dsr0) * tk / dn : 0.0;
        sum->dev_copy_out_us = dn > 0 ? (double)(s->dev_out - dout0) * tk / dn : 0.0;
        const double busy_us =
            (double)((s->dev_in - din0) + (s->dev_search - dsr0) + (s->dev_out - dout0)) * tk;
        const double workers = s->grid > 1 ? (double)(s->grid - 1) : 1.0;
        sum->dev_busy_frac = s->resident && sum->wall_s > 0 ? busy_us * 1e-6 / (workers * sum->wall_s) : 0.0;
        sum->host_submit_us = n > 0 ? (double)submit_ns.load() * 1e-3 / (double)n : 0.0;
        sum->host_digest_us = lat.empty() ? 0.0 : (double)digest_ns.load() * 1e-3 / (double)lat.size();
    }
    return CLF_OK;
}

}  // extern "C"
