// ref_harness.cpp — C-ABI entry points over the reference's own host code,
// compiled together with the reference's Tools.cpp and Logger.cpp where they
// lie (see oracle/Makefile, target "ref").  TEST INFRASTRUCTURE ONLY: used by
// tests/ to pin the oracle's presort and layout parser against the
// reference itself.  The output library lives in oracle/_ref/ (untracked).
#include <cstdint>
#include <cstring>
#include <vector>

#include "Tools.h"

extern int* h_no_sensors;
extern int* h_no_hits;
extern int* h_sensor_Zs;
extern int* h_sensor_hitStarts;
extern int* h_sensor_hitNums;
extern unsigned int* h_hit_IDs;
extern float* h_hit_Xs;
extern float* h_hit_Ys;
extern float* h_hit_Zs;

extern "C" {

// preorder_by_x (Tools.cpp:46-66) over a batch of one event, in place.
int ref_preorder_by_x(uint8_t* blob, size_t size) {
    std::vector<uint8_t> ev(blob, blob + size);
    std::vector<const std::vector<uint8_t>*> batch(1, &ev);
    preorder_by_x(batch);
    std::memcpy(blob, ev.data(), size);
    return 0;
}

// setHPointersFromInput (Tools.cpp:68-80): byte offsets of the nine views.
int ref_layout_offsets(uint8_t* blob, size_t size, int64_t* offsets9) {
    setHPointersFromInput(blob, size);
    const uint8_t* base = blob;
    const void* ptrs[9] = {h_no_sensors, h_no_hits,  h_sensor_Zs, h_sensor_hitStarts, h_sensor_hitNums,
                           h_hit_IDs,    h_hit_Xs,   h_hit_Ys,    h_hit_Zs};
    for (int i = 0; i < 9; ++i) offsets9[i] = (const uint8_t*)ptrs[i] - base;
    return 0;
}

}  // extern "C"
