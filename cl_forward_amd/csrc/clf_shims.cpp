// clf_shims.cpp — the reference's C++ entry points (GpuPixelSearchByTriplet.h,
// KernelInvoker.h) as thin shims over the library (include/clf.h, and the
// event-by-event read-back of clf_host.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <thread>
#include <vector>

#include "GpuPixelSearchByTriplet.h"
#include "KernelInvoker.h"
#include "clf.h"
#include "clf_host.h"

namespace {

std::mutex g_mu;
clf_context* g_ctx = nullptr;
int g_verbosity = 0;  // 3 on the standalone path, 0 on the Gaudi path

// The reference's contract for device errors: message, then exit(-1)
// (Tools.h:27-34, KernelInvoker.cpp:56-68).
void fatal(clf_status s, const char* what) {
    std::cerr << "Error in function " << what << std::endl
              << "Error string: " << clf_status_string(s) << " (" << clf_last_error() << ")"
              << std::endl;
    std::exit(-1);
}

clf_context* context() {
    if (!g_ctx) {
        clf_status s = clf_context_create(0, &g_ctx);  // DEVICE_NUMBER 0
        if (s != CLF_OK) fatal(s, "clf_context_create");
    }
    return g_ctx;
}

// The reference's timing report (KernelInvoker.cpp:136-193, 245-250, with
// calcResults, Tools.cpp:82-107): four experiments of four timed runs of
// the search kernel, then mean and standard deviation per experiment in
// single precision.  Its experiments vary the local size in y (1..4), which
// decides its results; here the results never depend on the launch, so the
// four experiments are the workgroup sizes 64, 128, 256 and 512, printed as
// (64, waves).  The configured launch is restored afterwards.
struct TimingReport {
    int waves[4] = {1, 2, 4, 8};
    std::vector<float> ms[4];
};

void run_timing(clf_context* ctx, clf_batch* batch, TimingReport& rep) {
    clf_launch_config saved;
    clf_context_get_config(ctx, &saved);
    for (int i = 0; i < 4; ++i) {
        clf_launch_config c = saved;
        c.block_size = 64 * rep.waves[i];
        c.grid_size = 0;
        if (clf_context_set_config(ctx, &c) != CLF_OK) continue;
        std::cout << i + 1 << ": " << std::flush;
        for (int j = 0; j < 4; ++j) {
            clf_timing t;
            clf_status s = clf_batch_time(ctx, batch, CLF_MODE_BATCHED, 0, 1, &t);
            if (s != CLF_OK) fatal(s, "clf_batch_time");
            rep.ms[i].push_back(t.ms_search);
            std::cout << "." << std::flush;
        }
        std::cout << std::endl;
    }
    clf_context_set_config(ctx, &saved);
}

void print_timing(const TimingReport& rep) {
    std::cout << std::endl << "Time averages:" << std::endl;
    for (int i = 0; i < 4; ++i) {
        const std::vector<float>& v = rep.ms[i];
        if (v.empty()) continue;
        float mean = 0.0f, variance = 0.0f;
        for (float x : v) {
            mean += x;
            variance += x * x;
        }
        mean /= v.size();
        variance = (variance / v.size()) - (mean * mean);
        const float deviation = std::sqrt(variance);
        std::cout << " nthreads (64, " << rep.waves[i] << "): " << mean << " ms (std dev "
                  << deviation << ")" << std::endl;
    }
}

int invoke_locked(int start, int count, const std::vector<const std::vector<uint8_t>*>& input,
                  std::vector<std::vector<uint8_t> >& output) {
    clf_context* ctx = context();
    std::vector<const uint8_t*> ptrs(count);
    std::vector<size_t> sizes(count);
    for (int i = 0; i < count; ++i) {
        const std::vector<uint8_t>* ev = input[start + i];
        ptrs[i] = ev && !ev->empty() ? ev->data() : nullptr;
        sizes[i] = ev ? ev->size() : 0;
    }
    if (g_verbosity >= 3)
        std::cout << "Invoking kernels on your " << clf_context_device_name(ctx) << std::endl;
    // each event's records into its output vector (KernelInvoker.cpp:197-206)
    // as soon as they are in host memory: the read-back comes in chunks and
    // the copies into the vectors run on several host threads beside it
    clf_batch* batch = nullptr;
    clf_status s = clf::search_each(
        ctx, CLF_MODE_BATCHED, count, ptrs.data(), sizes.data(),
        [&](int i, const uint8_t* src, size_t bytes) { output[start + i].assign(src, src + bytes); },
        g_verbosity >= 3 ? &batch : nullptr);
    if (s != CLF_OK) fatal(s, "clf_search");
    TimingReport timing;
    if (batch) {
        if (count > 0) run_timing(ctx, batch, timing);
        clf_batch_destroy(batch);
    }
    if (g_verbosity >= 3 && count > 0) print_timing(timing);
    return 0;
}

}  // namespace

int invokeParallelSearch(const int startingEvent, const int eventsToProcess,
                         const std::vector<const std::vector<uint8_t>*>& input,
                         std::vector<std::vector<uint8_t> >& output) {
    if (startingEvent < 0 || eventsToProcess < 0 ||
        (size_t)startingEvent + (size_t)eventsToProcess > input.size() ||
        (size_t)startingEvent + (size_t)eventsToProcess > output.size()) {
        std::cerr << "invokeParallelSearch: event range out of bounds" << std::endl;
        std::exit(-1);
    }
    std::lock_guard<std::mutex> lock(g_mu);
    return invoke_locked(startingEvent, eventsToProcess, input, output);
}

int gpuPixelSearchByTripletInvocation(const std::vector<const std::vector<uint8_t>*>& input,
                                      std::vector<std::vector<uint8_t> >& output) {
    if (g_verbosity >= 3)
        std::cout << "Invoking gpuPixelSearchByTriplet with " << input.size() << " events"
                  << std::endl;
    output.resize(input.size());
    const int max_events_per_launch = 16000;  // GpuPixelSearchByTriplet.cpp:54
    for (size_t i = 0; i < input.size(); i += max_events_per_launch) {
        const size_t n = std::min(input.size() - i, (size_t)max_events_per_launch);
        invokeParallelSearch((int)i, (int)n, input, output);
    }
    return 0;
}

int gpuPixelSearchByTriplet(const std::vector<const std::vector<uint8_t>*>& input,
                            std::vector<std::vector<uint8_t> >& output) {
    std::cout << std::fixed << std::setprecision(2);  // GpuPixelSearchByTriplet.cpp:33-34
    g_verbosity = 0;
    return gpuPixelSearchByTripletInvocation(input, output);
}

int independent_execute(const std::vector<std::vector<uint8_t> >& input,
                        std::vector<std::vector<uint8_t> >& output) {
    std::vector<const std::vector<uint8_t>*> converted(input.size());
    for (size_t i = 0; i < input.size(); ++i) converted[i] = &input[i];
    std::cout << std::fixed << std::setprecision(2);  // GpuPixelSearchByTriplet.cpp:15-16
    g_verbosity = 3;
    // in place through the const reference, like preorder_by_x (Tools.cpp:57-63)
    for (size_t i = 0; i < input.size(); ++i) {
        std::vector<uint8_t>& ev = const_cast<std::vector<uint8_t>&>(input[i]);
        if (!ev.empty()) (void)clf_presort_event(ev.data(), ev.size());
    }
    return gpuPixelSearchByTripletInvocation(converted, output);
}

void independent_post_execute(const std::vector<std::vector<uint8_t> >& output) {
    if (g_verbosity >= 3)
        std::cout << std::endl << "Size of output: " << output.size() << " entries" << std::endl;
}
