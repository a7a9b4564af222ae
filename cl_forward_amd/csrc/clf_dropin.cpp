// clf_dropin.cpp — a C++ caller of the reference's drop-in entry points, the
// way the Gaudi offload handler calls them: events in host vectors, one call
// of gpuPixelSearchByTriplet (GpuPixelSearchByTriplet.cpp:29-37, chunks of
// at most 16000 events, :44-65) or of invokeParallelSearch for a range of
// them (KernelInvoker.cpp:14-31), tracks back in host vectors.  Times each
// call with the steady clock (host in, host out) and writes the outputs.
//
//   clf_dropin EVENTS [--entry gaudi|invoke] [--start K] [--count N]
//              [--reps R] [--warmup W] [--out FILE]
//
// EVENTS: a stream of offload frames back to back, each
// int32 name_len; name; int32 data_size; data (IndependentEntrypoint.cpp
// :103-123), data being one event blob.  --out FILE: for every input event
// in order, int64 size then that many bytes of Track records (events
// outside an invoked range come out empty).  Prints one JSON line:
// {"entry", "events", "reps", "ms": [...], "median_ms", "min_ms"}.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "GpuPixelSearchByTriplet.h"
#include "KernelInvoker.h"

namespace {

bool read_frames(const std::string& path, std::vector<std::vector<uint8_t> >& events) {
    std::ifstream f(path.c_str(), std::ifstream::binary);
    if (!f) {
        std::cerr << "clf_dropin: cannot open " << path << std::endl;
        return false;
    }
    std::vector<uint8_t> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t p = 0;
    while (p < raw.size()) {
        int32_t name_len = 0, data_size = 0;
        if (p + 4 > raw.size()) return false;
        std::memcpy(&name_len, raw.data() + p, 4);
        if (name_len < 0 || p + 8 + (size_t)name_len > raw.size()) return false;
        std::memcpy(&data_size, raw.data() + p + 4 + name_len, 4);
        const size_t d = p + 8 + (size_t)name_len;
        if (data_size < 0 || d + (size_t)data_size > raw.size()) return false;
        events.emplace_back(raw.begin() + d, raw.begin() + d + data_size);
        p = d + (size_t)data_size;
    }
    return true;
}

void usage() {
    std::cerr << "Usage: clf_dropin EVENTS [--entry gaudi|invoke] [--start K] [--count N]"
                 " [--reps R] [--warmup W] [--out FILE]"
              << std::endl;
}

}  // namespace

int main(int argc, char* argv[]) {
    std::string in_path, out_path, entry = "gaudi";
    long start = 0, count = -1, reps = 1, warmup = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--entry") entry = next();
        else if (a == "--start") start = std::stol(next());
        else if (a == "--count") count = std::stol(next());
        else if (a == "--reps") reps = std::stol(next());
        else if (a == "--warmup") warmup = std::stol(next());
        else if (a == "--out") out_path = next();
        else if (!a.empty() && a[0] != '-' && in_path.empty()) in_path = a;
        else {
            usage();
            return 2;
        }
    }
    if (in_path.empty() || (entry != "gaudi" && entry != "invoke") || reps < 1 || warmup < 0) {
        usage();
        return 2;
    }
    std::vector<std::vector<uint8_t> > events;
    if (!read_frames(in_path, events)) {
        std::cerr << "clf_dropin: malformed frame stream " << in_path << std::endl;
        return 2;
    }
    std::vector<const std::vector<uint8_t>*> input(events.size());
    for (size_t i = 0; i < events.size(); ++i) input[i] = &events[i];
    if (count < 0) count = (long)events.size() - start;
    if (start < 0 || count < 0 || (size_t)(start + count) > events.size()) {
        std::cerr << "clf_dropin: range out of bounds" << std::endl;
        return 2;
    }

    std::vector<std::vector<uint8_t> > output;
    std::vector<double> ms;
    for (long r = 0; r < warmup + reps; ++r) {
        output.assign(events.size(), std::vector<uint8_t>());
        const auto t0 = std::chrono::steady_clock::now();
        if (entry == "gaudi") gpuPixelSearchByTriplet(input, output);
        else invokeParallelSearch((int)start, (int)count, input, output);
        const auto t1 = std::chrono::steady_clock::now();
        if (r >= warmup) ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }

    if (!out_path.empty()) {
        std::ofstream o(out_path.c_str(), std::ofstream::binary);
        for (const std::vector<uint8_t>& v : output) {
            const int64_t n = (int64_t)v.size();
            o.write(reinterpret_cast<const char*>(&n), 8);
            if (n) o.write(reinterpret_cast<const char*>(v.data()), n);
        }
        if (!o) {
            std::cerr << "clf_dropin: cannot write " << out_path << std::endl;
            return 2;
        }
    }
    std::vector<double> sorted = ms;
    std::sort(sorted.begin(), sorted.end());
    std::printf("{\"entry\": \"%s\", \"events\": %ld, \"reps\": %ld, \"ms\": [", entry.c_str(),
                entry == "gaudi" ? (long)events.size() : count, reps);
    for (size_t i = 0; i < ms.size(); ++i) std::printf("%s%.4f", i ? ", " : "", ms[i]);
    std::printf("], \"median_ms\": %.4f, \"min_ms\": %.4f}\n", sorted[sorted.size() / 2], sorted[0]);
    return 0;
}
