// clf_api.cpp — host side of the C ABI (include/clf.h): contexts, resident
// batches, launches, graphs, read-back and the host presort.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <numeric>
#include <string>
#include <vector>

#include "clf_host.h"

using namespace clf;

static_assert(sizeof(EventResult) == sizeof(clf_event_stats),
              "the kernel's result record is the C ABI's clf_event_stats");

namespace {
thread_local std::string g_last_error;

// The no-allocation scratch array of each device: its address is looked up
// once per device, with that device current, when the first context on it
// is created, and shared by every context on the device.
constexpr int kMaxDevices = 64;
struct ScratchSlot {
    bool looked_up = false;
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t err = hipSuccess;
};
std::mutex g_scratch_mu;
ScratchSlot g_scratch[kMaxDevices];

// Events per call of clf_search_noalloc (host tables sized at creation).
constexpr int kNoAllocMaxEvents = 4096;
}  // namespace

namespace clf {

clf_status fail(clf_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

clf_status hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? CLF_E_OUT_OF_MEMORY : CLF_E_DEVICE;
}

bool validate_event(const uint8_t* b, size_t size, int32_t& S, int32_t& N) {
    if (!b || size < 8) return false;
    S = rd_i32(b);
    N = rd_i32(b + 4);
    if (S < 0 || S > CLF_MAX_SENSORS || N < 0 || N > CLF_MAX_EVENT_HITS) return false;
    if (size < event_bytes(S, N)) return false;
    for (int s = 0; s < S; ++s) {
        const int32_t st = rd_i32(b + 8 + 4 * ((size_t)S + s));
        const int32_t n = rd_i32(b + 8 + 4 * (2 * (size_t)S + s));
        if (n < 0 || n > CLF_MAX_SENSOR_HITS || st < 0 || (int64_t)st + n > N) return false;
    }
    return true;
}

double event_cost(const uint8_t* b, int32_t S, int32_t N) {
    double c = N;
    for (int s = 2; s < S; ++s) {
        const double a = rd_i32(b + 8 + 4 * (2 * (size_t)S + s));
        const double p = rd_i32(b + 8 + 4 * (2 * (size_t)S + s - 2));
        c += a * p;
    }
    return c;
}

hipError_t Resident::ensure(int32_t n, size_t blob_bytes, int64_t hits, int64_t cap) {
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    chk(blobs.ensure(blob_bytes));
    chk(desc.ensure(sizeof(EventDesc) * (size_t)n));
    chk(order.ensure(sizeof(int32_t) * (size_t)n));
    chk(queue.ensure(sizeof(int32_t) * queue_words(n)));
    chk(cand.ensure(sizeof(int4) * (size_t)hits));
    chk(used.ensure((size_t)hits));
    chk(tracks.ensure(sizeof(clf_track) * (size_t)cap));
    chk(ttf_a.ensure(sizeof(TtfEntry) * (size_t)cap));
    chk(ttf_b.ensure(sizeof(TtfEntry) * (size_t)cap));
    chk(weak.ensure(sizeof(uint4) * (size_t)cap));
    chk(results.ensure(sizeof(EventResult) * (size_t)n));
    chk(prefix.ensure(sizeof(int64_t) * ((size_t)n + 1)));
    chk(packed.ensure(sizeof(clf_track) * (size_t)cap));
    return e;
}

void Resident::release() {
    DevBuf* all[] = {&blobs, &desc,  &order, &queue,   &cand,   &used,  &tracks,
                     &ttf_a, &ttf_b, &weak,  &results, &prefix, &packed};
    for (DevBuf* d : all) d->release();
}

void Resident::keep_larger(Resident& o) {
    DevBuf* mine[] = {&blobs, &desc,  &order, &queue,   &cand,   &used,  &tracks,
                      &ttf_a, &ttf_b, &weak,  &results, &prefix, &packed};
    DevBuf* theirs[] = {&o.blobs, &o.desc,  &o.order, &o.queue,   &o.cand,   &o.used,  &o.tracks,
                        &o.ttf_a, &o.ttf_b, &o.weak,  &o.results, &o.prefix, &o.packed};
    for (int i = 0; i < 13; ++i) mine[i]->keep_larger(*theirs[i]);
}

SearchArgs Resident::args(int32_t n) const {
    SearchArgs a;
    a.blobs = blobs.as<uint8_t>();
    a.ev = desc.as<EventDesc>();
    a.order = order.as<int32_t>();
    a.first_item = 0;
    a.n_items = n;
    a.queue_head = queue_counters();
    a.grain = 1;
    a.cand = cand.as<int4>();
    a.used = used.as<uint8_t>();
    a.tracks = tracks.as<clf_track>();
    a.ttf_a = ttf_a.as<TtfEntry>();
    a.ttf_b = ttf_b.as<TtfEntry>();
    a.weak = weak.as<uint4>();
    a.slot_records = 0;
    a.slot_base = 0;
    a.results = results.as<EventResult>();
    // prefix[e]: first packed record of event e; prefix[n]: the number of
    // packed records, stored by the run's last workgroup from the allocator
    // (queue header words 0-1); `done` is header word 2
    a.packed = packed.as<clf_track>();
    a.track_base = prefix.as<int64_t>();
    a.track_alloc = reinterpret_cast<unsigned long long*>(queue.as<int32_t>());
    a.done = reinterpret_cast<unsigned int*>(queue.as<int32_t>() + 2);
    a.done_total = 0;  // the workgroups of the run, set where it is launched
    a.reset_queue = queue_counters();
    a.reset_n = n + 1;
    a.track_total = prefix.as<int64_t>() + n;
    a.fill_only = 0;
    a.prof = nullptr;
    return a;
}

int resolved_grid(const clf_launch_config& cfg, int n_cu, int n_items) {
    int grid = cfg.grid_size;
    if (grid <= 0) {
        int per_cu = 0;
        if (occupancy(cfg.block_size, cfg.lds_hits, &per_cu) != hipSuccess || per_cu <= 0)
            per_cu = 1;
        grid = n_cu * per_cu;
    }
    return std::max(1, std::min(grid, n_items));
}

clf_launch_config resolve_launch(const clf_launch_config& cfg, int64_t hits, int32_t n) {
    clf_launch_config r = cfg;
    if (r.block_size == 0) {
        // the tuned launches of the two workload shapes: 2.8k-hit events
        // (batch1000: 0.86 ms at 512 x 256 against 1.39 ms at 1024 x 512) and
        // 7.1k-hit events (stress1000: 2.67 ms at 1024 x 512 against 5.66 ms)
        const bool large = n > 0 && hits > kAutoLargeEventHits * (int64_t)n;
        r.block_size = large ? 1024 : 512;
        r.lds_hits = large ? 512 : 256;
    } else if (r.lds_hits == 0) {
        r.lds_hits = 256;
    }
    return r;
}

int64_t slot_records(const Resident& r, int64_t slots) {
    return r.max_slice > 0 && slots * r.max_slice <= r.work_records ? r.max_slice : 0;
}

clf_status enqueue_search(const Resident& r, int32_t n, clf_mode mode,
                          const clf_launch_config& cfg, int n_cu, hipStream_t st, hipEvent_t t0,
                          hipEvent_t t1, hipStream_t side, hipEvent_t fork, hipEvent_t join,
                          bool clear) {
    if (n == 0) return CLF_OK;
    SearchArgs a = r.args(n);
    a.grain = std::max(1, cfg.grain);
    // the allocator, the run's count of finished workgroups and the queue
    // counters (one per launch, n + 1 so plain mode has one per event)
    if (clear) HIP_TRY(launch_zero_i32(r.queue.as<int32_t>(), (int)Resident::queue_words(n), st));
    if (t0) HIP_TRY(hipEventRecord(t0, st));
    const int heavy = side && fork && join && cfg.heavy_events > 0 ? std::min(cfg.heavy_events, n) : 0;
    if (mode == CLF_MODE_BATCHED && heavy > 0) {
        // the costliest events (the front of the cost order) in a launch of
        // larger workgroups, one per event, on the side stream; the rest in
        // the main launch; joined before packing
        HIP_TRY(hipEventRecord(fork, st));
        HIP_TRY(hipStreamWaitEvent(side, fork, 0));
        // work slots: the heavy launch's first, then the main launch's
        const int grid = heavy < n ? resolved_grid(cfg, n_cu, n - heavy) : 0;
        a.slot_records = slot_records(r, (int64_t)heavy + grid);
        a.done_total = heavy + grid;  // both launches end the run together
        SearchArgs h = a;
        h.first_item = 0;
        h.n_items = heavy;
        h.queue_head = r.queue_counters() + 1;
        HIP_TRY(launch_search(h, cfg.heavy_block, cfg.lds_hits, heavy, side));
        HIP_TRY(hipEventRecord(join, side));
        if (heavy < n) {
            a.first_item = heavy;
            a.n_items = n - heavy;
            a.slot_base = heavy;
            HIP_TRY(launch_search(a, cfg.block_size, cfg.lds_hits, grid, st));
        }
        HIP_TRY(hipStreamWaitEvent(st, join, 0));
    } else if (mode == CLF_MODE_BATCHED) {
        const int grid = resolved_grid(cfg, n_cu, n);
        a.slot_records = slot_records(r, grid);
        a.done_total = grid;
        HIP_TRY(launch_search(a, cfg.block_size, cfg.lds_hits, grid, st));
    } else {
        // plain: event i alone, in its own launch of one workgroup; the
        // launches run one after another on the stream and share one slot
        // (the last one ends the run)
        a.order = nullptr;
        a.n_items = 1;
        a.slot_records = slot_records(r, 1);
        a.done_total = n;
        for (int i = 0; i < n; ++i) {
            a.first_item = i;
            a.queue_head = r.queue_counters() + 1 + i;
            HIP_TRY(launch_search(a, cfg.block_size, cfg.lds_hits, 1, st));
        }
    }
    if (t1) HIP_TRY(hipEventRecord(t1, st));
    return CLF_OK;
}

}  // namespace clf

struct clf_context {
    int device = 0;
    char name[256] = {0};
    hipStream_t stream = nullptr;
    int n_cu = 0;
    // automatic launch shape, direct launches (use_graph 0)
    clf_launch_config cfg = {0, 0, 1, 0, 0, 0, 512};
    hipStream_t side = nullptr;  // the heavy-event launch of the batched mode
    hipEvent_t fork = nullptr, join = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    // buffers of destroyed batches, reused by the next batch of this context
    // (the one-call clf_search then allocates nothing in steady state)
    std::mutex spare_mu;
    Resident spare;
    PinnedBuf spare_blobs, spare_results, spare_prefix, spare_packed;
    // clf_search_noalloc: the device's scratch array and host tables for
    // its event descriptors, sized at creation
    void* na_scratch = nullptr;
    size_t na_bytes = 0;
    hipError_t na_err = hipSuccess;
    std::vector<EventDesc> na_desc;
    std::vector<int32_t> na_order;
    std::vector<double> na_cost;
    // search_each: one event per chunk of the read-back, created on first
    // use and kept (each_mu holds them for a call)
    std::mutex each_mu;
    std::vector<hipEvent_t> each_ev;
};

struct clf_batch {
    clf_context* ctx = nullptr;
    int32_t n = 0;
    int64_t total_hits = 0, total_cap = 0;
    size_t blob_bytes = 0;
    Resident dev;
    PinnedBuf h_blobs, h_results, h_prefix, h_packed;
    std::vector<EventDesc> h_desc;
    std::vector<int32_t> h_order;
    // graph of one batched run, rebuilt when the launch configuration changes
    hipGraphExec_t graph = nullptr;
    clf_launch_config graph_cfg = {};
    bool results_valid = false;
    // the queue words are at zero for the next run: a run has been enqueued
    // (each leaves them so), or they were cleared; fresh buffers hold anything
    bool words_clear = false;
};

namespace clf {
int context_device(const clf_context* ctx) { return ctx->device; }
}  // namespace clf

namespace {

hipError_t current_device(clf_context* ctx) { return hipSetDevice(ctx->device); }

// The kernels of one run, enqueued on the context stream (clear: with the
// clear of the queue words before them).
clf_status enqueue_run(clf_context* ctx, clf_batch* b, clf_mode mode, bool time_search,
                       bool clear = true) {
    return enqueue_search(b->dev, b->n, mode, resolve_launch(ctx->cfg, b->total_hits, b->n),
                          ctx->n_cu, ctx->stream,
                          time_search ? ctx->ev2 : nullptr, time_search ? ctx->ev3 : nullptr,
                          ctx->side, ctx->fork, ctx->join, clear);
}

bool same_cfg(const clf_launch_config& x, const clf_launch_config& y) {
    return x.block_size == y.block_size && x.grid_size == y.grid_size && x.grain == y.grain &&
           x.lds_hits == y.lds_hits && x.heavy_events == y.heavy_events &&
           x.heavy_block == y.heavy_block;
}

clf_status run(clf_context* ctx, clf_batch* b, clf_mode mode) {
    HIP_TRY(current_device(ctx));
    b->results_valid = false;
    if (mode == CLF_MODE_BATCHED && ctx->cfg.use_graph && b->n > 0) {
        const clf_launch_config lc = resolve_launch(ctx->cfg, b->total_hits, b->n);
        if (b->graph && !same_cfg(b->graph_cfg, lc)) {
            (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
        }
        if (!b->graph) {
            // the graph holds the search launch(es) alone: every run ends
            // with its queue words at zero, so only a batch's first run
            // needs them cleared, before it
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            clf_status s = enqueue_run(ctx, b, mode, false, false);
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (s != CLF_OK) {
                if (g) (void)hipGraphDestroy(g);
                return s;
            }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
            b->graph_cfg = lc;
        }
        if (!b->words_clear)
            HIP_TRY(launch_zero_i32(b->dev.queue.as<int32_t>(), (int)Resident::queue_words(b->n),
                                    ctx->stream));
        HIP_TRY(hipGraphLaunch(b->graph, ctx->stream));
        b->words_clear = true;
        return CLF_OK;
    }
    // a direct launch: the clear only before the batch's first run
    clf_status s = enqueue_run(ctx, b, mode, false, !b->words_clear);
    if (s == CLF_OK) b->words_clear = true;
    return s;
}

}  // namespace

extern "C" {

const char* clf_status_string(clf_status s) {
    switch (s) {
        case CLF_OK: return "ok";
        case CLF_E_INVALID_ARGUMENT: return "invalid argument";
        case CLF_E_DEVICE: return "device error";
        case CLF_E_OUT_OF_MEMORY: return "out of memory";
        case CLF_E_NO_DEVICE: return "no device";
        case CLF_E_CAPACITY: return "capacity exceeded";
    }
    return "unknown status";
}

const char* clf_last_error(void) { return g_last_error.c_str(); }

const char* clf_version(void) { return "clf 0.1.0 (gfx950)"; }

clf_status clf_context_create(int device, clf_context** out) {
    if (!out) return fail(CLF_E_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(CLF_E_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(CLF_E_NO_DEVICE, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device));
    clf_context* ctx = new clf_context();
    ctx->device = device;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete ctx;
        return hip_fail(e, "hipGetDeviceProperties");
    }
    ctx->n_cu = prop.multiProcessorCount;
    if (device < kMaxDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        ScratchSlot& s = g_scratch[device];
        if (!s.looked_up) {
            s.err = noalloc_scratch(&s.ptr, &s.bytes);
            if (s.err != hipSuccess) s.ptr = nullptr;
            s.looked_up = true;
        }
        ctx->na_scratch = s.ptr;
        ctx->na_bytes = s.bytes;
        ctx->na_err = s.err;
    } else {
        ctx->na_err = hipErrorInvalidDevice;
    }
    ctx->na_desc.resize(kNoAllocMaxEvents);
    ctx->na_order.resize(kNoAllocMaxEvents);
    ctx->na_cost.resize(kNoAllocMaxEvents);
    std::snprintf(ctx->name, sizeof(ctx->name), "%s", prop.name[0] ? prop.name : prop.gcnArchName);
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev2);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev3);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        clf_context_destroy(ctx);
        return hip_fail(e, "stream/event creation");
    }
    *out = ctx;
    return CLF_OK;
}

void clf_context_destroy(clf_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
    if (ctx->ev3) (void)hipEventDestroy(ctx->ev3);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    for (hipEvent_t e : ctx->each_ev) (void)hipEventDestroy(e);
    ctx->spare.release();
    ctx->spare_blobs.release();
    ctx->spare_results.release();
    ctx->spare_prefix.release();
    ctx->spare_packed.release();
    delete ctx;
}

clf_status clf_context_set_config(clf_context* ctx, const clf_launch_config* cfg) {
    if (!ctx || !cfg) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    if (cfg->block_size != 0 &&
        !config_supported(cfg->block_size, cfg->lds_hits == 0 ? 256 : cfg->lds_hits))
        return fail(CLF_E_INVALID_ARGUMENT,
                    "unsupported (block_size, lds_hits): block 0 (automatic) or 64/128/256/512/1024, "
                    "lds_hits 0/256/512");
    if (cfg->lds_hits != 0 && cfg->lds_hits != 256 && cfg->lds_hits != 512)
        return fail(CLF_E_INVALID_ARGUMENT, "lds_hits must be 0, 256 or 512");
    if (cfg->grid_size < 0) return fail(CLF_E_INVALID_ARGUMENT, "grid_size must be >= 0");
    if (cfg->grain < 1 || cfg->grain > 64) return fail(CLF_E_INVALID_ARGUMENT, "grain must be 1..64");
    if (cfg->heavy_events < 0 ||
        (cfg->heavy_events > 0 && !(config_supported(cfg->heavy_block, 256) &&
                                    config_supported(cfg->heavy_block, 512))))
        return fail(CLF_E_INVALID_ARGUMENT, "heavy_events must be >= 0 with a supported heavy_block");
    ctx->cfg = *cfg;
    return CLF_OK;
}

clf_status clf_context_get_config(const clf_context* ctx, clf_launch_config* cfg) {
    if (!ctx || !cfg) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    *cfg = ctx->cfg;
    return CLF_OK;
}

int clf_context_num_cus(const clf_context* ctx) { return ctx ? ctx->n_cu : 0; }

const char* clf_context_device_name(const clf_context* ctx) { return ctx ? ctx->name : ""; }

clf_status clf_context_synchronize(clf_context* ctx) {
    if (!ctx) return fail(CLF_E_INVALID_ARGUMENT, "null context");
    HIP_TRY(current_device(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return CLF_OK;
}

int64_t clf_event_track_capacity(const uint8_t* event, size_t size) {
    int32_t S, N;
    if (!validate_event(event, size, S, N)) return -1;
    return CLF_EVENT_TRACK_CAPACITY(N, S);
}

void clf_batch_destroy(clf_batch* b) {
    if (!b) return;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->ctx) {
        clf_context* c = b->ctx;
        std::lock_guard<std::mutex> lk(c->spare_mu);
        c->spare.keep_larger(b->dev);
        c->spare_blobs.keep_larger(b->h_blobs);
        c->spare_results.keep_larger(b->h_results);
        c->spare_prefix.keep_larger(b->h_prefix);
        c->spare_packed.keep_larger(b->h_packed);
    }
    b->dev.release();
    b->h_blobs.release();
    b->h_results.release();
    b->h_prefix.release();
    b->h_packed.release();
    delete b;
}

}  // extern "C"

namespace {

// Copies the events into the batch's pinned staging and from there to the
// device, on the context stream.  Large batches are gathered by several
// host threads in chunks of about 2 MiB, and each chunk's upload is issued
// as soon as the chunk is complete, so that the host copies and the DMA
// transfer overlap.
hipError_t upload_blobs(clf_context* ctx, clf_batch* b, const uint8_t* const* events) {
    const int n = b->n;
    uint8_t* hb = b->h_blobs.as<uint8_t>();
    uint8_t* db = b->dev.blobs.as<uint8_t>();
    hipStream_t st = ctx->stream;
    auto gather = [&](int i) {
        const EventDesc& d = b->h_desc[i];
        if (d.n_hits >= 0) std::memcpy(hb + d.blob_off, events[i], event_bytes(d.n_sensors, d.n_hits));
    };
    const int threads = copy_threads(b->blob_bytes);
    if (threads <= 1) {
        for (int i = 0; i < n; ++i) gather(i);
        return b->blob_bytes ? hipMemcpyAsync(db, hb, b->blob_bytes, hipMemcpyHostToDevice, st)
                             : hipSuccess;
    }
    // chunk k: events [first[k], first[k + 1]), staging bytes [lo[k], lo[k + 1])
    std::vector<int> first;
    std::vector<size_t> lo;
    const size_t kChunk = size_t(2) << 20;
    for (int i = 0; i < n; ++i) {
        const size_t at = (size_t)b->h_desc[i].blob_off;
        if (b->h_desc[i].n_hits < 0) continue;
        if (first.empty() || at - lo.back() >= kChunk) {
            first.push_back(i);
            lo.push_back(at);
        }
    }
    const int nc = (int)first.size();
    first.push_back(n);
    lo.push_back(b->blob_bytes);
    std::unique_ptr<std::atomic<int>[]> done(new std::atomic<int>[nc]);
    for (int k = 0; k < nc; ++k) done[k].store(0);
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k; (k = next.fetch_add(1)) < nc;) {
            for (int i = first[k]; i < first[k + 1]; ++i) gather(i);
            done[k].store(1, std::memory_order_release);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(work);
    hipError_t e = hipSuccess;
    for (int k = 0; k < nc; ++k) {
        while (!done[k].load(std::memory_order_acquire)) std::this_thread::yield();
        if (e == hipSuccess)
            e = hipMemcpyAsync(db + lo[k], hb + lo[k], lo[k + 1] - lo[k], hipMemcpyHostToDevice, st);
    }
    for (std::thread& t : pool) t.join();
    return e;
}

// clf_batch_create; without `sync` the uploads are only enqueued.
clf_status batch_create(clf_context* ctx, int32_t n, const uint8_t* const* events,
                        const size_t* sizes, clf_batch** out, bool sync) {
    if (!ctx || !out || n < 0 || (n > 0 && (!events || !sizes)))
        return fail(CLF_E_INVALID_ARGUMENT, "null argument or negative count");
    *out = nullptr;
    HIP_TRY(current_device(ctx));
    clf_batch* b = new clf_batch();
    b->ctx = ctx;
    b->n = n;
    {
        std::lock_guard<std::mutex> lk(ctx->spare_mu);
        b->dev.keep_larger(ctx->spare);
        b->h_blobs.keep_larger(ctx->spare_blobs);
        b->h_results.keep_larger(ctx->spare_results);
        b->h_prefix.keep_larger(ctx->spare_prefix);
        b->h_packed.keep_larger(ctx->spare_packed);
    }
    b->h_desc.resize(n);
    std::vector<double> cost(n, 0.0);
    size_t off = 0;
    int64_t hits = 0, cap = 0, max_slice = 0;
    for (int i = 0; i < n; ++i) {
        int32_t S = 0, N = 0;
        EventDesc& d = b->h_desc[i];
        if (validate_event(events[i], sizes[i], S, N)) {
            d.blob_off = (int64_t)off;
            d.hit_off = hits;
            d.track_off = cap;
            d.n_hits = N;
            d.n_sensors = S;
            off += align_up(8 + 12 * (size_t)S + 16 * (size_t)N, 256);
            hits += N;
            cap += track_slice(N, S);
            max_slice = std::max(max_slice, track_slice(N, S));
            cost[i] = event_cost(events[i], S, N);
        } else {
            d.blob_off = 0;
            d.hit_off = hits;
            d.track_off = cap;
            d.n_hits = -1;
            d.n_sensors = 0;
        }
    }
    b->blob_bytes = off;
    b->total_hits = hits;
    b->total_cap = cap;
    b->h_order.resize(n);
    std::iota(b->h_order.begin(), b->h_order.end(), 0);
    std::stable_sort(b->h_order.begin(), b->h_order.end(),
                     [&](int x, int y) { return cost[x] > cost[y]; });

    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    chk(b->h_blobs.ensure(off));
    chk(b->dev.ensure(n, off, hits, cap));
    b->dev.work_records = cap;
    b->dev.max_slice = max_slice;
    chk(b->h_results.ensure(sizeof(EventResult) * (size_t)n));
    chk(b->h_prefix.ensure(sizeof(int64_t) * ((size_t)n + 1)));
    if (e != hipSuccess) {
        clf_batch_destroy(b);
        return hip_fail(e, "batch allocation");
    }
    hipStream_t st = ctx->stream;
    chk(upload_blobs(ctx, b, events));
    chk(hipMemcpyAsync(b->dev.desc.p, b->h_desc.data(), sizeof(EventDesc) * (size_t)n,
                       hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(b->dev.order.p, b->h_order.data(), sizeof(int32_t) * (size_t)n,
                       hipMemcpyHostToDevice, st));
    if (sync) chk(hipStreamSynchronize(st));
    if (e != hipSuccess) {
        clf_batch_destroy(b);
        return hip_fail(e, "batch upload");
    }
    *out = b;
    return CLF_OK;
}

}  // namespace

extern "C" {

clf_status clf_batch_create(clf_context* ctx, int32_t n, const uint8_t* const* events,
                            const size_t* sizes, clf_batch** out) {
    return batch_create(ctx, n, events, sizes, out, true);
}

clf_status clf_batch_run(clf_context* ctx, clf_batch* b, clf_mode mode) {
    if (!ctx || !b || b->ctx != ctx) return fail(CLF_E_INVALID_ARGUMENT, "bad context or batch");
    if (mode != CLF_MODE_PLAIN && mode != CLF_MODE_BATCHED)
        return fail(CLF_E_INVALID_ARGUMENT, "unknown mode");
    return run(ctx, b, mode);
}

clf_status clf_batch_results(clf_context* ctx, clf_batch* b, clf_result* out) {
    if (!ctx || !b || !out || b->ctx != ctx) return fail(CLF_E_INVALID_ARGUMENT, "bad argument");
    HIP_TRY(current_device(ctx));
    const int n = b->n;
    hipStream_t st = ctx->stream;
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(b->h_results.p, b->dev.results.p, sizeof(EventResult) * (size_t)n,
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(b->h_prefix.p, b->dev.prefix.p, sizeof(int64_t) * ((size_t)n + 1),
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const int64_t total = b->h_prefix.as<int64_t>()[n];
        if (total < 0 || total > b->total_cap) return fail(CLF_E_DEVICE, "inconsistent track count");
        const clf_event_stats* stv = b->h_results.as<clf_event_stats>();
        for (int i = 0; i < n; ++i) {
            const int64_t a0 = b->h_prefix.as<int64_t>()[i];
            if (a0 < 0 || stv[i].n_tracks < 0 || a0 + stv[i].n_tracks > total)
                return fail(CLF_E_DEVICE, "inconsistent track offsets");
        }
        HIP_TRY(b->h_packed.ensure(sizeof(clf_track) * (size_t)total));
        if (total > 0)
            HIP_TRY(hipMemcpyAsync(b->h_packed.p, b->dev.packed.p, sizeof(clf_track) * (size_t)total,
                                   hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    } else {
        HIP_TRY(b->h_prefix.ensure(sizeof(int64_t)));
        b->h_prefix.as<int64_t>()[0] = 0;
    }
    out->n_events = n;
    out->stats = b->h_results.as<clf_event_stats>();
    out->track_offset = b->h_prefix.as<int64_t>();
    out->tracks = b->h_packed.p ? b->h_packed.as<uint8_t>() : nullptr;
    b->results_valid = true;
    return CLF_OK;
}

clf_status clf_batch_time(clf_context* ctx, clf_batch* b, clf_mode mode, int32_t warmup,
                          int32_t reps, clf_timing* out) {
    if (!ctx || !b || !out || b->ctx != ctx || reps <= 0 || warmup < 0)
        return fail(CLF_E_INVALID_ARGUMENT, "bad argument");
    HIP_TRY(current_device(ctx));
    for (int i = 0; i < warmup; ++i) {
        clf_status s = run(ctx, b, mode);
        if (s != CLF_OK) return s;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < reps; ++i) {
        clf_status s = run(ctx, b, mode);
        if (s != CLF_OK) return s;
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out->ms_run = ms / (float)reps;
    // search kernel(s) alone, launched directly after a clear of the queue
    // words, with events right around them
    float total = 0.f;
    for (int i = 0; i < reps; ++i) {
        clf_status s = enqueue_run(ctx, b, mode, true);
        if (s != CLF_OK) return s;
        HIP_TRY(hipEventSynchronize(ctx->ev3));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, ctx->ev2, ctx->ev3));
        total += t;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    out->ms_search = total / (float)reps;
    out->reps = reps;
    b->results_valid = false;
    return CLF_OK;
}

clf_status clf_batch_profile(clf_context* ctx, clf_batch* b, clf_mode mode, uint64_t* out,
                             double* tick_ns) {
    if (!ctx || !b || b->ctx != ctx || (b->n > 0 && !out))
        return fail(CLF_E_INVALID_ARGUMENT, "bad argument");
    if (mode != CLF_MODE_PLAIN && mode != CLF_MODE_BATCHED)
        return fail(CLF_E_INVALID_ARGUMENT, "unknown mode");
    HIP_TRY(current_device(ctx));
    const int n = b->n;
    int rate_khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
    if (tick_ns) *tick_ns = rate_khz > 0 ? 1.0e6 / (double)rate_khz : 0.0;
    if (n == 0) return CLF_OK;
    DevBuf prof;
    HIP_TRY(prof.ensure(sizeof(unsigned long long) * (size_t)n * kProfSlots));
    struct Guard {
        DevBuf& p;
        ~Guard() { p.release(); }
    } guard{prof};
    hipStream_t st = ctx->stream;
    SearchArgs a = b->dev.args(n);
    a.prof = prof.as<unsigned long long>();
    const clf_launch_config lc = resolve_launch(ctx->cfg, b->total_hits, n);
    HIP_TRY(launch_zero_i32(b->dev.queue.as<int32_t>(), (int)Resident::queue_words(n), st));
    if (mode == CLF_MODE_BATCHED) {
        const int grid = resolved_grid(lc, ctx->n_cu, n);
        a.slot_records = slot_records(b->dev, grid);
        a.done_total = grid;
        HIP_TRY(launch_search(a, lc.block_size, lc.lds_hits, grid, st));
    } else {
        a.order = nullptr;
        a.n_items = 1;
        a.slot_records = slot_records(b->dev, 1);
        a.done_total = n;
        for (int i = 0; i < n; ++i) {
            a.first_item = i;
            a.queue_head = b->dev.queue_counters() + 1 + i;
            HIP_TRY(launch_search(a, lc.block_size, lc.lds_hits, 1, st));
        }
    }
    HIP_TRY(hipMemcpyAsync(out, prof.p, sizeof(unsigned long long) * (size_t)n * kProfSlots,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    b->results_valid = false;
    return CLF_OK;
}

clf_status clf_probe_copy(clf_context* ctx, int64_t bytes, int32_t reps, float* ms_per_copy) {
    if (!ctx || bytes < 16 || reps < 1) return fail(CLF_E_INVALID_ARGUMENT, "bad argument");
    HIP_TRY(current_device(ctx));
    DevBuf src, dst;
    struct Guard {
        DevBuf& s;
        DevBuf& d;
        ~Guard() {
            s.release();
            d.release();
        }
    } guard{src, dst};
    HIP_TRY(src.ensure((size_t)bytes));
    HIP_TRY(dst.ensure((size_t)bytes));
    HIP_TRY(launch_zero_i32(src.as<int32_t>(), (int)std::min<int64_t>(bytes / 4, 1 << 30), ctx->stream));
    HIP_TRY(launch_probe_copy(src.p, dst.p, bytes, ctx->stream));  // warm-up
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < reps; ++i) HIP_TRY(launch_probe_copy(src.p, dst.p, bytes, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms_per_copy) *ms_per_copy = ms / (float)reps;
    return CLF_OK;
}

}  // extern "C"

namespace {

// Windows of every hit through fill mode `mode` of the search kernel (1:
// binary searches over the blob in HBM; 2: the search's LDS ring path).
clf_status fill_windows(clf_context* ctx, int32_t n, const uint8_t* const* events,
                        const size_t* sizes, int32_t* const* cand_h1, int32_t* const* cand_h2,
                        uint32_t* flags, int mode) {
    if (!ctx || n < 0 || (n > 0 && (!cand_h1 || !cand_h2)))
        return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    clf_batch* b = nullptr;
    clf_status s = clf_batch_create(ctx, n, events, sizes, &b);
    if (s != CLF_OK) return s;
    struct Guard {
        clf_batch* b;
        ~Guard() { clf_batch_destroy(b); }
    } guard{b};
    if (n == 0) return CLF_OK;
    hipStream_t st = ctx->stream;
    HIP_TRY(launch_zero_i32(b->dev.queue.as<int32_t>(), (int)Resident::queue_words(n), st));
    SearchArgs a = b->dev.args(n);
    a.fill_only = mode;
    const clf_launch_config lc = resolve_launch(ctx->cfg, b->total_hits, n);
    a.done_total = resolved_grid(lc, ctx->n_cu, n);
    HIP_TRY(launch_search(a, lc.block_size, lc.lds_hits, a.done_total, st));
    std::vector<int4> cand((size_t)std::max<int64_t>(b->total_hits, 1));
    std::vector<EventResult> res(n);
    if (b->total_hits > 0)
        HIP_TRY(hipMemcpyAsync(cand.data(), b->dev.cand.p, sizeof(int4) * (size_t)b->total_hits,
                               hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(res.data(), b->dev.results.p, sizeof(EventResult) * (size_t)n,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) {
        const EventDesc& d = b->h_desc[i];
        if (flags) flags[i] = res[i].flags;
        if (d.n_hits < 0) continue;
        for (int h = 0; h < d.n_hits; ++h) {
            const int4 c = cand[(size_t)d.hit_off + h];
            cand_h1[i][2 * h] = c.x;
            cand_h1[i][2 * h + 1] = c.y;
            cand_h2[i][2 * h] = c.z;
            cand_h2[i][2 * h + 1] = c.w;
        }
    }
    return CLF_OK;
}

}  // namespace

extern "C" {

clf_status clf_probe_link(clf_context* ctx, int64_t bytes, int32_t reps, double* h2d_gbps,
                          double* d2h_gbps) {
    if (!ctx || bytes < 1 || reps < 1) return fail(CLF_E_INVALID_ARGUMENT, "bad argument");
    HIP_TRY(current_device(ctx));
    DevBuf dev;
    PinnedBuf host;
    struct Guard {
        DevBuf& d;
        PinnedBuf& h;
        ~Guard() {
            d.release();
            h.release();
        }
    } guard{dev, host};
    HIP_TRY(dev.ensure((size_t)bytes));
    HIP_TRY(host.ensure((size_t)bytes));
    std::memset(host.p, 1, (size_t)bytes);
    double gbps[2] = {0.0, 0.0};
    for (int dir = 0; dir < 2; ++dir) {
        auto copy = [&]() {
            return dir == 0 ? hipMemcpyAsync(dev.p, host.p, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream)
                            : hipMemcpyAsync(host.p, dev.p, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream);
        };
        HIP_TRY(copy());  // warm-up
        HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        for (int i = 0; i < reps; ++i) HIP_TRY(copy());
        HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        gbps[dir] = ms > 0.f ? (double)bytes * reps / (ms * 1e-3) / 1e9 : 0.0;
    }
    if (h2d_gbps) *h2d_gbps = gbps[0];
    if (d2h_gbps) *d2h_gbps = gbps[1];
    return CLF_OK;
}

clf_status clf_fill_candidates(clf_context* ctx, int32_t n, const uint8_t* const* events,
                               const size_t* sizes, int32_t* const* cand_h1,
                               int32_t* const* cand_h2, uint32_t* flags) {
    return fill_windows(ctx, n, events, sizes, cand_h1, cand_h2, flags, 1);
}

clf_status clf_fill_candidates_search_path(clf_context* ctx, int32_t n,
                                           const uint8_t* const* events, const size_t* sizes,
                                           int32_t* const* cand_h1, int32_t* const* cand_h2,
                                           uint32_t* flags) {
    return fill_windows(ctx, n, events, sizes, cand_h1, cand_h2, flags, 2);
}

clf_status clf_search_noalloc(clf_context* ctx, clf_mode mode, int32_t n,
                              const uint8_t* const* events, const size_t* sizes, uint8_t* tracks,
                              int64_t track_capacity, int64_t* track_offset,
                              clf_event_stats* stats) {
    if (!ctx || n < 0 || (n > 0 && (!events || !sizes || !stats)) || !track_offset ||
        track_capacity < 0 || (track_capacity > 0 && !tracks))
        return fail(CLF_E_INVALID_ARGUMENT, "null argument or negative count");
    if (mode != CLF_MODE_PLAIN && mode != CLF_MODE_BATCHED)
        return fail(CLF_E_INVALID_ARGUMENT, "unknown mode");
    if (!ctx->na_scratch) return hip_fail(ctx->na_err, "no-allocation scratch (hipGetSymbolAddress)");
    if (n > kNoAllocMaxEvents)
        return fail(CLF_E_CAPACITY, "clf_search_noalloc takes at most " +
                                        std::to_string(kNoAllocMaxEvents) + " events per call, got " +
                                        std::to_string(n));
    HIP_TRY(current_device(ctx));
    // event table, in the context's host arrays
    EventDesc* desc = ctx->na_desc.data();
    int32_t* order = ctx->na_order.data();
    double* cost = ctx->na_cost.data();
    size_t off = 0;
    int64_t hits = 0, cap = 0, max_slice = 0;
    for (int i = 0; i < n; ++i) {
        int32_t S = 0, N = 0;
        EventDesc& d = desc[i];
        d.hit_off = hits;
        d.track_off = cap;
        cost[i] = 0.0;
        if (validate_event(events[i], sizes[i], S, N)) {
            d.blob_off = (int64_t)off;
            d.n_hits = N;
            d.n_sensors = S;
            off += align_up(event_bytes(S, N), 256);
            hits += N;
            cap += track_slice(N, S);
            max_slice = std::max(max_slice, track_slice(N, S));
            cost[i] = event_cost(events[i], S, N);
        } else {
            d.blob_off = 0;
            d.n_hits = -1;
            d.n_sensors = 0;
        }
        order[i] = i;
    }
    // largest estimated cost first, ties by index: the order clf_batch_create
    // gives (std::sort allocates nothing)
    std::sort(order, order + n, [&](int x, int y) { return cost[x] != cost[y] ? cost[x] > cost[y] : x < y; });
    // the call's buffers, carved out of the scratch array
    size_t at = 0;
    auto carve = [&](size_t bytes) {
        const size_t p = at;
        at = align_up(at + std::max<size_t>(bytes, 1), 256);
        return p;
    };
    const size_t o_blobs = carve(off), o_desc = carve(sizeof(EventDesc) * (size_t)n),
                 o_order = carve(sizeof(int32_t) * (size_t)n),
                 o_queue = carve(sizeof(int32_t) * Resident::queue_words(n)),
                 o_cand = carve(sizeof(int4) * (size_t)hits), o_used = carve((size_t)hits),
                 o_tracks = carve(sizeof(clf_track) * (size_t)cap),
                 o_ttf_a = carve(sizeof(TtfEntry) * (size_t)cap),
                 o_ttf_b = carve(sizeof(TtfEntry) * (size_t)cap),
                 o_weak = carve(sizeof(uint4) * (size_t)cap),
                 o_results = carve(sizeof(EventResult) * (size_t)n),
                 o_prefix = carve(sizeof(int64_t) * ((size_t)n + 1)),
                 o_packed = carve(sizeof(clf_track) * (size_t)cap);
    if (at > ctx->na_bytes)
        return fail(CLF_E_CAPACITY, "these events need " + std::to_string(at) +
                                        " bytes of device scratch; clf_search_noalloc has " +
                                        std::to_string(ctx->na_bytes));
    uint8_t* base = static_cast<uint8_t*>(ctx->na_scratch);
    Resident r;  // a view: nothing here owns or frees these pointers
    DevBuf* bufs[] = {&r.blobs, &r.desc,  &r.order, &r.queue,   &r.cand,   &r.used,  &r.tracks,
                      &r.ttf_a, &r.ttf_b, &r.weak,  &r.results, &r.prefix, &r.packed};
    const size_t offs[] = {o_blobs, o_desc,  o_order, o_queue,   o_cand,   o_used,  o_tracks,
                           o_ttf_a, o_ttf_b, o_weak,  o_results, o_prefix, o_packed};
    for (int k = 0; k < 13; ++k) bufs[k]->p = base + offs[k];
    r.work_records = cap;
    r.max_slice = max_slice;
    hipStream_t st = ctx->stream;
    for (int i = 0; i < n; ++i) {
        if (desc[i].n_hits < 0) continue;
        HIP_TRY(hipMemcpyAsync(base + o_blobs + desc[i].blob_off, events[i],
                               event_bytes(desc[i].n_sensors, desc[i].n_hits),
                               hipMemcpyHostToDevice, st));
    }
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(r.desc.p, desc, sizeof(EventDesc) * (size_t)n, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(r.order.p, order, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    }
    clf_status s = enqueue_search(r, n, mode, resolve_launch(ctx->cfg, hits, n), ctx->n_cu, st, nullptr,
                                  nullptr, ctx->side, ctx->fork, ctx->join);
    if (s != CLF_OK) return s;
    if (n == 0) {
        track_offset[0] = 0;
        return CLF_OK;
    }
    HIP_TRY(hipMemcpyAsync(stats, r.results.p, sizeof(EventResult) * (size_t)n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(track_offset, r.prefix.p, sizeof(int64_t) * ((size_t)n + 1),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const int64_t total = track_offset[n];
    if (total < 0 || total > cap) return fail(CLF_E_DEVICE, "inconsistent track count");
    for (int i = 0; i < n; ++i)
        if (track_offset[i] < 0 || stats[i].n_tracks < 0 || track_offset[i] + stats[i].n_tracks > total)
            return fail(CLF_E_DEVICE, "inconsistent track offsets");
    if (total > track_capacity)
        return fail(CLF_E_CAPACITY, "the events have " + std::to_string(total) +
                                        " tracks; the output holds " + std::to_string(track_capacity));
    if (total > 0) {
        HIP_TRY(hipMemcpyAsync(tracks, r.packed.p, sizeof(clf_track) * (size_t)total,
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return CLF_OK;
}

clf_status clf_search(clf_context* ctx, clf_mode mode, int32_t n, const uint8_t* const* events,
                      const size_t* sizes, clf_result* out, clf_batch** batch_out) {
    if (!ctx || !out || !batch_out) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    *batch_out = nullptr;
    if (mode != CLF_MODE_PLAIN && mode != CLF_MODE_BATCHED)
        return fail(CLF_E_INVALID_ARGUMENT, "unknown mode");
    // upload, search and read-back in stream order with no host wait in
    // between, and no graph for a batch that runs once
    clf_batch* b = nullptr;
    clf_status s = batch_create(ctx, n, events, sizes, &b, false);
    if (s != CLF_OK) return s;
    s = enqueue_run(ctx, b, mode, false);
    if (s == CLF_OK) s = clf_batch_results(ctx, b, out);
    if (s != CLF_OK) {
        clf_batch_destroy(b);
        return s;
    }
    *batch_out = b;
    return CLF_OK;
}

}  // extern "C"

namespace clf {

clf_status search_each(clf_context* ctx, clf_mode mode, int32_t n, const uint8_t* const* events,
                       const size_t* sizes,
                       const std::function<void(int, const uint8_t*, size_t)>& emit,
                       clf_batch** keep) {
    if (!ctx || n < 0 || (n > 0 && (!events || !sizes)))
        return fail(CLF_E_INVALID_ARGUMENT, "null argument or negative count");
    if (keep) *keep = nullptr;
    if (mode != CLF_MODE_PLAIN && mode != CLF_MODE_BATCHED)
        return fail(CLF_E_INVALID_ARGUMENT, "unknown mode");
    clf_batch* b = nullptr;
    clf_status s = batch_create(ctx, n, events, sizes, &b, false);
    if (s != CLF_OK) return s;
    // the batch goes to `keep` only when the call succeeds
    struct Guard {
        clf_batch*& b;
        clf_batch** keep;
        bool ok = false;
        ~Guard() {
            if (keep && ok) *keep = b;
            else clf_batch_destroy(b);
        }
    } guard{b, keep};
    s = enqueue_run(ctx, b, mode, false);
    if (s != CLF_OK) return s;
    if (n == 0) {
        guard.ok = true;
        return CLF_OK;
    }
    std::lock_guard<std::mutex> lk(ctx->each_mu);
    hipStream_t st = ctx->stream;
    // records of the packed output per read-back chunk (2 MiB)
    constexpr int64_t kChunkRecs = (int64_t(2) << 20) / (int64_t)sizeof(clf_track);
    const size_t est_bytes = sizeof(clf_track) * (size_t)b->total_cap;
    const int threads = est_bytes >= (size_t(4) << 20) ? 8 : 0;

    // The hand-over: events in the order of their records; event j waits
    // for the chunk holding its last record.  `ready` is -1 until the order
    // is known, then the number of chunks in host memory; on an error the
    // workers leave without handing anything over.
    std::vector<int> by_off(n);
    std::vector<int> need(n);
    std::atomic<int> ready{-1}, next{0};
    std::atomic<bool> failed{false};
    const int64_t* pre = nullptr;
    const clf_event_stats* stv = nullptr;
    const uint8_t* host = nullptr;
    auto hand_over = [&]() {
        // the order (by_off, need) is read only once it is published
        while (!failed.load(std::memory_order_acquire) && ready.load(std::memory_order_acquire) < 0)
            std::this_thread::yield();
        if (failed.load(std::memory_order_acquire)) return;
        for (int j; (j = next.fetch_add(1)) < n;) {
            const int i = by_off[j];
            while (!failed.load(std::memory_order_acquire) && ready.load(std::memory_order_acquire) <= need[i])
                std::this_thread::yield();
            if (failed.load(std::memory_order_acquire)) return;
            emit(i, host + sizeof(clf_track) * (size_t)pre[i],
                 sizeof(clf_track) * (size_t)std::max(stv[i].n_tracks, 0));
        }
    };
    // the workers start while the search runs; whatever happens after, they
    // are joined (told to leave first if the order was never published)
    std::vector<std::thread> pool;
    struct Join {
        std::vector<std::thread>& pool;
        std::atomic<bool>& failed;
        std::atomic<int>& ready;
        ~Join() {
            if (ready.load() < 0) failed.store(true, std::memory_order_release);
            for (std::thread& t : pool) t.join();
        }
    } join{pool, failed, ready};
    for (int t = 0; t < threads; ++t) pool.emplace_back(hand_over);
    auto stop = [&](clf_status e) {
        failed.store(true, std::memory_order_release);
        return e;
    };

    hipError_t e = hipMemcpyAsync(b->h_results.p, b->dev.results.p, sizeof(EventResult) * (size_t)n,
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->h_prefix.p, b->dev.prefix.p, sizeof(int64_t) * ((size_t)n + 1),
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return stop(hip_fail(e, "search read-back"));
    pre = b->h_prefix.as<int64_t>();
    stv = b->h_results.as<clf_event_stats>();
    const int64_t total = pre[n];
    if (total < 0 || total > b->total_cap) return stop(fail(CLF_E_DEVICE, "inconsistent track count"));
    for (int i = 0; i < n; ++i)
        if (pre[i] < 0 || stv[i].n_tracks < 0 || pre[i] + stv[i].n_tracks > total)
            return stop(fail(CLF_E_DEVICE, "inconsistent track offsets"));
    e = b->h_packed.ensure(sizeof(clf_track) * (size_t)std::max<int64_t>(total, 1));
    if (e != hipSuccess) return stop(hip_fail(e, "pinned read-back buffer"));
    host = b->h_packed.as<uint8_t>();
    std::iota(by_off.begin(), by_off.end(), 0);
    std::sort(by_off.begin(), by_off.end(), [&](int x, int y) { return pre[x] < pre[y]; });
    for (int i = 0; i < n; ++i)
        need[i] = stv[i].n_tracks > 0 ? (int)((pre[i] + stv[i].n_tracks - 1) / kChunkRecs) : -1;
    const int chunks = (int)((total + kChunkRecs - 1) / kChunkRecs);
    while ((int)ctx->each_ev.size() < chunks) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return stop(hip_fail(e, "hipEventCreateWithFlags"));
        ctx->each_ev.push_back(ev);
    }
    const uint8_t* dev = b->dev.packed.as<uint8_t>();
    for (int k = 0; k < chunks && e == hipSuccess; ++k) {
        const int64_t r0 = k * kChunkRecs, r1 = std::min(total, r0 + kChunkRecs);
        e = hipMemcpyAsync(b->h_packed.as<uint8_t>() + sizeof(clf_track) * (size_t)r0,
                           dev + sizeof(clf_track) * (size_t)r0, sizeof(clf_track) * (size_t)(r1 - r0),
                           hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(ctx->each_ev[k], st);
    }
    if (e != hipSuccess) return stop(hip_fail(e, "chunked read-back"));
    ready.store(0, std::memory_order_release);
    for (int k = 0; k < chunks; ++k) {
        e = hipEventSynchronize(ctx->each_ev[k]);
        if (e != hipSuccess) return stop(hip_fail(e, "chunked read-back"));
        ready.store(k + 1, std::memory_order_release);
    }
    if (threads == 0) hand_over();
    guard.ok = true;
    return CLF_OK;
}

}  // namespace clf
