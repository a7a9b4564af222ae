// clf_results.cpp — the reference's text result format and input framing,
// host code of the C ABI (clf_format_tracks, clf_unframe).
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <map>
#include <sstream>
#include <string>

#include "clf_host.h"

using namespace clf;

namespace {

// Nearest key of the z -> module map (KernelInvoker.cpp:284-303): z - 1,
// z + 1, z - 2, z + 2, ...  The map is never empty here.
int closest_module(int z, const std::map<int, int>& m) {
    auto it = m.find(z);
    if (it != m.end()) return it->second;
    for (long long err = 1;; ++err) {
        auto lo = m.find((int)(z - err));
        if (lo != m.end()) return lo->second;
        auto hi = m.find((int)(z + err));
        if (hi != m.end()) return hi->second;
    }
}

}  // namespace

extern "C" {

clf_status clf_format_tracks(const uint8_t* event, size_t size, const uint8_t* tracks,
                             int64_t n_tracks, char* text, size_t cap, size_t* len) {
    int32_t S = 0, N = 0;
    if (!len || n_tracks < 0 || (n_tracks > 0 && !tracks))
        return fail(CLF_E_INVALID_ARGUMENT, "null argument or negative count");
    if (!validate_event(event, size, S, N)) return fail(CLF_E_INVALID_ARGUMENT, "malformed event");
    const uint8_t* zs = event + 8;
    const uint8_t* ids = event + 8 + 12 * (size_t)S;
    const uint8_t* X = ids + 4 * (size_t)N;
    const uint8_t* Y = X + 4 * (size_t)N;
    const uint8_t* Z = Y + 4 * (size_t)N;
    auto f32 = [](const uint8_t* p, size_t i) {
        float v;
        std::memcpy(&v, p + 4 * i, 4);
        return v;
    };
    // z -> module map (KernelInvoker.cpp:214-231): sensor Zs, later sensors
    // winning ties; then every hit's truncated z, in hit order, mapped to the
    // nearest key present at that moment (so the map grows as it goes)
    std::map<int, int> zmod;
    for (int j = 0; j < S; ++j) zmod[rd_i32(zs + 4 * (size_t)j)] = j;
    if (zmod.empty() && N > 0) return fail(CLF_E_INVALID_ARGUMENT, "hits without sensors");
    for (int j = 0; j < N; ++j) {
        const int z = (int)f32(Z, (size_t)j);
        if (zmod.find(z) == zmod.end()) zmod[z] = closest_module(z, zmod);
    }
    // printTrack (KernelInvoker.cpp:254-275) into a stream with default
    // formatting, as the reference's std::ofstream
    std::ostringstream out;
    for (int64_t j = 0; j < n_tracks; ++j) {
        clf_track t;
        std::memcpy(&t, tracks + sizeof(clf_track) * (size_t)j, sizeof(clf_track));
        if (t.hitsNum > CLF_MAX_TRACK_SIZE)
            return fail(CLF_E_INVALID_ARGUMENT, "track with more than 24 hits");
        out << "Track #" << j << ", length " << (int)t.hitsNum << std::endl;
        for (uint32_t i = 0; i < t.hitsNum; ++i) {
            const int h = (int)t.hits[i];
            if (h < 0 || h >= N) return fail(CLF_E_INVALID_ARGUMENT, "hit index out of range");
            uint32_t id;
            std::memcpy(&id, ids + 4 * (size_t)h, 4);
            const float x = f32(X, (size_t)h), y = f32(Y, (size_t)h), z = f32(Z, (size_t)h);
            const int module = zmod.at((int)z);
            out << " " << std::setw(8) << id << " (" << h << ")"
                << " module " << std::setw(2) << module << ", x " << std::setw(6) << x << ", y "
                << std::setw(6) << y << ", z " << std::setw(6) << z << std::endl;
        }
        out << std::endl;
    }
    const std::string s = out.str();
    *len = s.size();
    if (text && cap > 0) {
        const size_t k = s.size() < cap - 1 ? s.size() : cap - 1;
        std::memcpy(text, s.data(), k);
        text[k] = '\0';
        if (s.size() >= cap) return fail(CLF_E_CAPACITY, "text buffer too small");
    }
    return CLF_OK;
}

clf_status clf_unframe(const uint8_t* frame, size_t size, const uint8_t** name, int32_t* name_len,
                       const uint8_t** data, size_t* data_size) {
    if (!frame || !data || !data_size) return fail(CLF_E_INVALID_ARGUMENT, "null argument");
    if (size < 8) return fail(CLF_E_INVALID_ARGUMENT, "frame shorter than its two length words");
    const int32_t nl = rd_i32(frame);
    if (nl < 0 || (size_t)nl > size - 8) return fail(CLF_E_INVALID_ARGUMENT, "bad name length");
    const int32_t ds = rd_i32(frame + 4 + nl);
    if (ds < 0 || (size_t)ds > size - 8 - (size_t)nl)
        return fail(CLF_E_INVALID_ARGUMENT, "bad data size");
    if (name) *name = frame + 4;
    if (name_len) *name_len = nl;
    *data = frame + 8 + nl;
    *data_size = (size_t)ds;
    return CLF_OK;
}

}  // extern "C"
