// clf_device.h — records shared by the HIP kernels and the host library.
#pragma once

#include <stdint.h>

#include "clf_params.h"

namespace clf {

// Where one event of a batch lives.  Offsets are set by the host from the
// events' headers (Kernel.cl:613-624 reads the same header on the device).
struct EventDesc {
    int64_t blob_off;   // byte offset of the event blob in the batch input
    int64_t hit_off;    // first slot of the event in per-hit scratch
    int64_t track_off;  // first record of the event in per-track scratch
    int32_t n_hits;     // N, or -1 for a blob the host rejected
    int32_t n_sensors;  // S
};

// Per-event outcome.  Layout is clf_event_stats of clf.h.
struct EventResult {
    int32_t n_tracks;
    uint32_t flags;
    int32_t n_tracklets;
    int32_t n_weak;
    int32_t max_ttf_live;
    int32_t n_full;      // tracks that reached 24 hits
    int32_t h0_visits;
    int32_t n_overflow;  // 24-hit tracks that found a 25th hit (CLF_FLAG_TRACK_FULL)
};

// A track or tracklet to follow.  "word" is the reference's ttf word
// (bit 31 tracklet, bits 28-30 skipped modules, bits 0-27 index,
// Kernel.cl:222-225).  The last two hits' coordinates ride along so that
// forwarding never gathers them again, and so do a tracklet's three hit
// indices, so that no tracklet record has to be read back.
struct TtfEntry {
    uint32_t word;
    uint32_t nhits;
    uint32_t t0, t1, t2;  // tracklet hits (nhits == 3)
    float x0, y0, z0;     // second-to-last hit
    float x1, y1, z1;     // last hit
    uint32_t pad;
};

// Candidate windows of one hit, relative to the start of the sensor they
// point into, -1 = none.  W1Rec: the hit as h0 (window on the sensor two
// below); W2Rec: the hit as h1.
struct W1Rec {
    int16_t lo, hi;
};
struct W2Rec {
    int16_t lo, hi;
};

struct SearchArgs {
    const uint8_t* blobs;
    const EventDesc* ev;
    const int32_t* order;  // dequeue order of events (nullptr: identity)
    int32_t first_item;    // items [first_item, first_item + n_items) of order
    int32_t n_items;
    int32_t* queue_head;   // work-queue counter, zero before the run
    int32_t grain;         // consecutive queue items taken per dequeue (>= 1)
    int4* cand;            // per hit, 16 B: stage (a) output of the fill-only
                           // check; otherwise W1Rec[N] then W2Rec[N] of
                           // sensors too large for an LDS slot
    uint8_t* used;         // per hit (events above the LDS bitmask)
    // Work slices: the tracks being built, the two ttf lists and the weak
    // tracklets' hits, each a slice of records at one offset.  With
    // slot_records > 0 a workgroup's slice is slot slot_base + blockIdx.x
    // (slot_records records each), reused by every event the workgroup
    // takes; with 0 each event has its own at its descriptor's track_off.
    clf_track* tracks;
    TtfEntry* ttf_a;
    TtfEntry* ttf_b;
    uint4* weak;
    int64_t slot_records;
    int32_t slot_base;
    EventResult* results;  // per event
    // Packed output (nullptr: none, the caller copies from `tracks`): at the
    // end of event e its workgroup reserves n_tracks records at
    // *track_alloc (zero before the run), records the first one in
    // track_base[e] and copies the event's tracks there.
    clf_track* packed;
    int64_t* track_base;
    unsigned long long* track_alloc;
    // End of a run (done nullptr: none).  `done` counts the workgroups that
    // have left, over every launch of the run; the one that makes it
    // done_total stores the allocator's final value in *track_total, then
    // clears the allocator, the run's queue counters (reset_queue[0,
    // reset_n)) and `done`: the next run finds them at zero with no
    // clearing launch before it.
    unsigned int* done;
    int32_t done_total;
    int32_t* reset_queue;
    int32_t reset_n;
    int64_t* track_total;
    int32_t fill_only;     // stop after the candidate windows (stage (a) check)
    // Optional per-event phase profile (nullptr: off): kProfSlots counters per
    // event, see the PROF_* indices in clf_kernels.hip.
    unsigned long long* prof;
};

constexpr int kProfSlots = 16;

// ---- Resident serving kernel --------------------------------------------
// A request published by the host in pinned memory once its event blob is
// in the device input arena.  `seq` is written last (release): descriptor d
// of the ring is valid once seq == d + 1.
// The part of an event blob the search reads, as the serving input arena
// holds it: the blob's 16-byte words [0, h16) (header and sensor table),
// then its words [x16, n16) (from the one holding X[0]: the hit IDs between
// are never read), then its last bytes - bytes - 16 n16 - past word n16.
// Word j < m16 of the packed form is blob word j < h16 ? j : j + x16 - h16.
struct SearchedWords {
    int64_t bytes, h16, x16, n16, m16, packed;
};
constexpr SearchedWords searched_words(int32_t S, int32_t N) {
    SearchedWords w{};
    w.bytes = 8 + 12 * (int64_t)S + 16 * (int64_t)N;
    w.n16 = w.bytes >> 4;
    w.h16 = (8 + 12 * (int64_t)S + 15) >> 4;
    const int64_t x = (8 + 12 * (int64_t)S + 4 * (int64_t)N) >> 4;
    w.x16 = x > w.h16 ? x : w.h16;
    w.m16 = w.h16 + (w.n16 - w.x16);
    w.packed = 16 * w.m16 + (w.bytes - 16 * w.n16);
    return w;
}

struct alignas(16) ServeDesc {
    uint64_t in_off;   // event, in its searched-words form (SearchedWords), in
                       // the device input arena (256-byte aligned)
    uint64_t reserved;
    int32_t n_hits;
    int32_t n_sensors;
    int32_t req;       // host request index, echoed in the completion
    int32_t out_cap;   // Track records the request may produce (its capacity bound)
    uint32_t seq;
    uint32_t pad[3];
};
static_assert(sizeof(ServeDesc) == 48, "the kernel reads a descriptor as 16-byte words");

// A finished request, appended by the device to a ring in pinned memory;
// valid once seq == position + 1.  Its tracks are the n_tracks records at
// ring byte position out_pos (modulo the output ring's size) of the device
// output ring.
enum CompStatus { COMP_OK = 0, COMP_INVALID = 1 };  // INVALID: a descriptor outside the limits

struct CompRec {
    uint32_t seq;
    int32_t req;
    int32_t wg;        // workgroup that ran it
    int32_t status;    // CompStatus
    EventResult res;
    uint64_t out_pos;  // absolute byte position of its tracks in the output ring
    // device wall clock: claimed, event copied in, searched, tracks written
    uint64_t t_claim, t_copied, t_searched, t_written;
};

// Control words in device memory.  HEAD (descriptors handed to workers),
// TAIL (descriptors seen published), DONE (completions written) and the
// output ring's allocator OUT_HEAD (64-bit, words 8-9) persist across
// launches; QUIT, IDLE_TAIL, LEFT (workers that have left) and the
// feeder's heartbeat BEAT (64-bit, words 6-7) are cleared before each
// launch.
enum ServeCtl {
    CTL_HEAD = 0,
    CTL_TAIL = 1,
    CTL_QUIT = 2,
    CTL_DONE = 3,
    CTL_IDLE_TAIL = 4,
    CTL_LEFT = 5,
    CTL_BEAT = 6,
    CTL_OUT_HEAD = 8,
    CTL_WORDS = 16
};

struct ServeArgs {
    SearchArgs base;           // per-workgroup scratch: ev[w], results[w], slices by ev[w]
    const ServeDesc* ring;     // pinned, ring_mask + 1 entries
    uint32_t ring_mask;
    const uint8_t* in_arena;   // device (fine-grained), in_bytes: the events
    uint8_t* out_arena;        // device (fine-grained), out_bytes: the output ring
    uint64_t in_bytes, out_bytes;
    CompRec* comp;             // pinned, comp_mask + 1 entries
    uint32_t comp_mask;
    const uint32_t* host_stop; // pinned: nonzero asks the kernel to finish
    uint32_t* ctl;             // device, CTL_WORDS
    uint32_t* mailbox;         // device, one per workgroup: last descriptor + 1 handed to it
    uint64_t* idle_ring;       // device: (position + 1) << 32 | workgroup of idle workers
    uint32_t idle_mask;
    int64_t wg_blob_bytes;     // scratch per workgroup
    int64_t wg_hits;
    int64_t wg_cap;
    uint64_t idle_ticks;       // wall-clock ticks without work before leaving
    uint64_t pause_ticks;      // testing: the feeder stalls this long, without a
                               // heartbeat, before its first hand-out (0: never)
};

}  // namespace clf
