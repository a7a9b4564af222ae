// clf_presort.cpp — x-presort of an event blob in place (the standalone
// path's preorder_by_x, Tools.cpp:46-66).
//
// The partition sequence is the reference's (Tools.cpp:117-157): pivot =
// first element, the right index skips elements "> pivot", the left index
// skips elements "<= pivot", swap, and finally the pivot moves to the right
// index.  It has to be the same sequence, not just any sort, because the
// order of tied x values decides which hit index the tracker reports.  The
// two sub-ranges of a partition are disjoint, so they are visited from an
// explicit stack instead of by recursion (no stack depth limit on sorted
// input, where the first-element pivot degenerates).
#include <cstdint>
#include <cstring>
#include <vector>

#include "clf.h"

namespace {

struct HitColumns {
    float* x;
    float* y;
    float* z;
    uint32_t* id;
    void swap(int i, int j) {
        std::swap(x[i], x[j]);
        std::swap(y[i], y[j]);
        std::swap(z[i], z[j]);
        std::swap(id[i], id[j]);
    }
};

int partition_first_pivot(HitColumns& c, int lo, int hi) {
    const float pivot = c.x[lo];
    int left = lo, right = hi;
    while (left < right) {
        while (c.x[right] > pivot) --right;
        while (left < right && c.x[left] <= pivot) ++left;
        if (left < right) c.swap(left, right);
    }
    c.swap(right, lo);
    return right;
}

void sort_range(HitColumns& c, int lo, int hi, std::vector<std::pair<int, int>>& stack) {
    stack.clear();
    stack.emplace_back(lo, hi);
    while (!stack.empty()) {
        const auto [l, h] = stack.back();
        stack.pop_back();
        if (!(l < h)) continue;
        const int p = partition_first_pivot(c, l, h);
        stack.emplace_back(l, p - 1);
        stack.emplace_back(p + 1, h);
    }
}

}  // namespace

extern "C" int clf_presort_event(uint8_t* blob, size_t size) {
    if (!blob || size < 8) return -1;
    int32_t S, N;
    std::memcpy(&S, blob, 4);
    std::memcpy(&N, blob + 4, 4);
    if (S < 0 || S > CLF_MAX_SENSORS || N < 0 || N > CLF_MAX_EVENT_HITS) return -1;
    const size_t need = 8 + 12 * (size_t)S + 16 * (size_t)N;
    if (size < need) return -1;
    std::vector<int32_t> nums(S);
    std::memcpy(nums.data(), blob + 8 + 8 * (size_t)S, 4 * (size_t)S);
    // aligned copies of the four hit columns
    std::vector<uint32_t> id(N);
    std::vector<float> x(N), y(N), z(N);
    const size_t off = 8 + 12 * (size_t)S;
    std::memcpy(id.data(), blob + off, 4 * (size_t)N);
    std::memcpy(x.data(), blob + off + 4 * (size_t)N, 4 * (size_t)N);
    std::memcpy(y.data(), blob + off + 8 * (size_t)N, 4 * (size_t)N);
    std::memcpy(z.data(), blob + off + 12 * (size_t)N, 4 * (size_t)N);
    // A NaN x never satisfies either scan, so the reference's partition
    // never ends on it; such an event is rejected instead.
    for (int i = 0; i < N; ++i)
        if (x[i] != x[i]) return -1;
    HitColumns c{x.data(), y.data(), z.data(), id.data()};
    std::vector<std::pair<int, int>> stack;
    // sensors are taken as contiguous from 0, as Tools.cpp:53-63 does
    int acc = 0;
    for (int s = 0; s < S; ++s) {
        const int n = nums[s];
        if (n < 0 || acc + n > N) return -1;
        sort_range(c, acc, acc + n - 1, stack);
        acc += n;
    }
    std::memcpy(blob + off, id.data(), 4 * (size_t)N);
    std::memcpy(blob + off + 4 * (size_t)N, x.data(), 4 * (size_t)N);
    std::memcpy(blob + off + 8 * (size_t)N, y.data(), 4 * (size_t)N);
    std::memcpy(blob + off + 12 * (size_t)N, z.data(), 4 * (size_t)N);
    return 0;
}
