// clf_host.h — host-side pieces shared by the C ABI (clf_api.cpp) and the
// serving loop (clf_serve.cpp): buffers, header checks, the resident device
// buffers of one batch and the launch sequence of one search.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "clf.h"
#include "clf_device.h"

namespace clf {

// Kernel launchers (clf_kernels.hip).
hipError_t launch_search(const SearchArgs& a, int block, int lds_hits, int grid, hipStream_t stream);
bool config_supported(int block, int lds_hits);
// Zeros p[0..n) and, if given, the 64-bit counter *alloc (one kernel).
hipError_t launch_zero_i32(int32_t* p, int n, hipStream_t stream, unsigned long long* alloc = nullptr);
hipError_t occupancy(int block, int lds_hits, int* blocks_per_cu);
hipError_t launch_probe_copy(const void* src, void* dst, int64_t bytes, hipStream_t stream);
hipError_t launch_serve(const ServeArgs& a, int block, int lds_hits, int grid, hipStream_t stream);
hipError_t serve_occupancy(int block, int lds_hits, int* blocks_per_cu);
// Device address and size of the no-allocation scratch array of the
// current device (hipGetSymbolAddress; call once per device, with it current).
hipError_t noalloc_scratch(void** ptr, size_t* bytes);

// Error reporting: the message is kept per thread for clf_last_error().
clf_status fail(clf_status st, const std::string& msg);
clf_status hip_fail(hipError_t e, const char* what);

#define HIP_TRY(expr)                                               \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return ::clf::hip_fail(_e, #expr);    \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes == 0) bytes = 256;
        if (bytes <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
    // keeps the larger of the two allocations here and frees the other
    void keep_larger(DevBuf& o) {
        if (o.cap > cap) {
            release();
            p = o.p;
            cap = o.cap;
            o.p = nullptr;
            o.cap = 0;
        } else {
            o.release();
        }
    }
};

// Pinned host memory.  `flags` as for hipHostMalloc; hipHostMallocCoherent
// gives memory that a running kernel and the host may both access.
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes, unsigned flags = hipHostMallocDefault) {
        if (bytes == 0) bytes = 256;
        if (bytes <= cap) return hipSuccess;
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
        }
        hipError_t e = hipHostMalloc(&p, bytes, flags);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
    void keep_larger(PinnedBuf& o) {
        if (o.cap > cap) {
            release();
            p = o.p;
            cap = o.cap;
            o.p = nullptr;
            o.cap = 0;
        } else {
            o.release();
        }
    }
};

// Runs fn(k) for every k in [0, n) on `threads` host threads (the caller
// among them), each taking the next k from a shared counter.
template <class F>
void parallel_for(int n, int threads, F&& fn) {
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k; (k = next.fetch_add(1, std::memory_order_relaxed)) < n;) fn(k);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < std::min(threads, n); ++t) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
}

// Host threads for copies of `bytes` bytes: one below 4 MiB, else up to 8
// (no more than the machine has).
inline int copy_threads(size_t bytes) {
    if (bytes < (size_t(4) << 20)) return 1;
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(8u, hw ? hw : 1u));
}

inline int32_t rd_i32(const uint8_t* p) {
    int32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bytes of a well-formed event blob with S sensors and N hits.
inline size_t event_bytes(int32_t S, int32_t N) { return 8 + 12 * (size_t)S + 16 * (size_t)N; }

// Records of one event's slice of the per-track scratch (tracks, ttf lists,
// weak entries): the capacity bound rounded up to 4 records, so that every
// slice of the 100-byte records starts 16-byte aligned.
inline int64_t track_slice(int32_t N, int32_t S) {
    return (CLF_EVENT_TRACK_CAPACITY(N, S) + 3) & ~(int64_t)3;
}

// Header checks of one blob; fills S and N.  Same rules as the oracle.
bool validate_event(const uint8_t* b, size_t size, int32_t& S, int32_t& N);

// Estimated relative cost of an event (hits on sensors two apart), used
// only to order a work queue largest first.
double event_cost(const uint8_t* b, int32_t S, int32_t N);

// Device buffers of one batch of events.  Per-hit and per-track scratch is
// sliced per event by EventDesc::hit_off / track_off.
struct Resident {
    DevBuf blobs, desc, order, queue, cand, used, tracks, ttf_a, ttf_b, weak, results, prefix,
        packed;
    // the batch these buffers hold: the records of its per-event work
    // slices (the capacity `ensure` was given) and its largest slice
    int64_t work_records = 0, max_slice = 0;
    // n events, blob_bytes of packed blobs, hits in total, track capacity
    hipError_t ensure(int32_t n, size_t blob_bytes, int64_t hits, int64_t cap);
    void release();
    void keep_larger(Resident& o);
    SearchArgs args(int32_t n) const;
    // The queue words: a 16-byte header (the packed-output allocator, a
    // uint64, then the run's count of finished workgroups) and n + 1 queue
    // counters (plain mode takes one per event, counter 1 + i).
    static constexpr int kQueueHeaderWords = 4;
    static size_t queue_words(int32_t n) { return (size_t)kQueueHeaderWords + (size_t)n + 1; }
    int32_t* queue_counters() const { return queue.as<int32_t>() + kQueueHeaderWords; }
};

// HIP device ordinal of a context.
int context_device(const clf_context* ctx);

// Persistent grid for n queued events under a launch configuration.
int resolved_grid(const clf_launch_config& cfg, int n_cu, int n_items);

// The launch a configuration gives a batch of n events with `hits` hits:
// the configuration itself when block_size is set (lds_hits 0 meaning 256),
// or, for block_size 0 (automatic), the tuned shape for the batch's mean
// event size: 1024 threads with 512-hit slots above kAutoLargeEventHits
// hits per event, else 512 threads with 256-hit slots.
constexpr int64_t kAutoLargeEventHits = 4096;
clf_launch_config resolve_launch(const clf_launch_config& cfg, int64_t hits, int32_t n);

// Records per work slot when `slots` workgroups each take one slot of the
// batch's largest slice and the slots fit where its per-event slices lie
// (every workgroup then reuses its slot, event after event, so the slices'
// lines stay in L2 between events); 0 when they do not fit (one slice per
// event).
int64_t slot_records(const Resident& r, int64_t slots);

// Enqueues one search of n resident events on `stream`: with `clear`, the
// clear of the queue words (a run leaves them at zero, but fresh buffers
// hold anything), then the search kernel(s), which pack each event's
// tracks (track_base[e], the `prefix` buffer, holds where they start; the
// run's last workgroup writes the total to prefix[n]).  cfg must be
// resolved (resolve_launch).  t0/t1 (nullable) are recorded right around
// the search kernel(s).  side/fork/join (nullable): the stream and events
// of the heavy-event launch (cfg.heavy_events); without them everything
// runs in one launch.
clf_status enqueue_search(const Resident& r, int32_t n, clf_mode mode,
                          const clf_launch_config& cfg, int n_cu, hipStream_t stream,
                          hipEvent_t t0, hipEvent_t t1, hipStream_t side = nullptr,
                          hipEvent_t fork = nullptr, hipEvent_t join = nullptr,
                          bool clear = true);

// The one-call search of clf_search with the tracks handed over event by
// event: emit(i, records, bytes) runs for every event i exactly once (bytes
// 0 for an event without tracks) before this returns, on up to eight host
// threads, as soon as the part of the packed output holding its records is
// in host memory: the output comes back in 2 MiB chunks, and the events
// whose records have arrived are handed over while the next chunks are in
// flight.  keep (nullable): receives the batch, for the caller to destroy.
clf_status search_each(clf_context* ctx, clf_mode mode, int32_t n, const uint8_t* const* events,
                       const size_t* sizes,
                       const std::function<void(int, const uint8_t*, size_t)>& emit,
                       clf_batch** keep = nullptr);

}  // namespace clf
