/*
 * clf_params.h — algorithm constants, on-device record types and per-event
 * status flags of the search-by-triplet VELO pixel tracker.
 *
 * One header shared by the HIP kernels, the C-ABI host library, the C++
 * entry-point shims and the CPU oracle, so every side compiles the same
 * numbers.  Plain C (no C++ features) so gcc can compile the oracle with it.
 *
 * Values follow KernelDefinitions.h:9-34 of the reference (cuts, capacities)
 * and its Track record KernelDefinitions.h:59-62 (4 + 24*4 = 100 bytes).
 */
#ifndef CLF_PARAMS_H
#define CLF_PARAMS_H

#include <stdint.h>

/* ---- Cuts (KernelDefinitions.h:26-33) -------------------------------- */
#define CLF_PARAM_MAXXSLOPE            0.7f     /* PARAM_MAXXSLOPE            */
#define CLF_PARAM_MAXYSLOPE            0.7f     /* PARAM_MAXYSLOPE            */
#define CLF_PARAM_MAXXSLOPE_CANDIDATES 0.7f     /* PARAM_MAXXSLOPE_CANDIDATES */
#define CLF_PARAM_TOLERANCE            0.6f     /* PARAM_TOLERANCE            */
#define CLF_PARAM_TOLERANCE_CANDIDATES 0.6f     /* PARAM_TOLERANCE_CANDIDATES */
#define CLF_MAX_SCATTER                0.000016f /* MAX_SCATTER               */
#define CLF_MAX_SKIPPED_MODULES        3        /* MAX_SKIPPED_MODULES        */

/* ---- Capacities of the reference (KernelDefinitions.h:15-23) ---------- *
 * The reference has no bounds checks on these.  This implementation has no
 * such capacity of its own except the 24-slot Track record (an output
 * format constraint); it raises a per-event flag whenever the reference's
 * capacity would have been exceeded (see CLF_FLAG_* below).              */
#define CLF_MAX_TRACKS                 3000     /* MAX_TRACKS per event       */
#define CLF_MAX_TRACK_SIZE             24       /* MAX_TRACK_SIZE             */
#define CLF_TTF_MODULO                 2000     /* TTF_MODULO ring length     */

/* ---- Launch geometry whose semantics are reproduced ------------------ *
 * The reference's returned tracks are the ones of its last timing run,
 * local size (64, 4) (KernelInvoker.cpp:149-156,197-206).  The 64-wide h0
 * chunks of the gather (Kernel.cl:720-754) and the 4-way split of h1
 * candidates in seeding (Kernel.cl:435-436) are part of the results.     */
#define CLF_GATHER_CHUNK               64       /* NUMTHREADS_X               */
#define CLF_SEED_ROWS                  4        /* local size in y of the final run */

/* ---- ttf word encoding (Kernel.cl:222-225,351,566) ------------------- */
#define CLF_TTF_TRACKLET_BIT           0x80000000u
#define CLF_TTF_SKIP_MASK              0x70000000u
#define CLF_TTF_SKIP_SHIFT             28
#define CLF_TTF_INDEX_MASK             0x0FFFFFFFu
#define CLF_TTF_KEEP_MASK              0x8FFFFFFFu

/* ---- Limits of this implementation ----------------------------------- */
#define CLF_MAX_SENSORS                256      /* sensors per event          */
#define CLF_MAX_SENSOR_HITS            32767    /* hits on one sensor (16-bit
                                                   window indices)            */
#define CLF_MAX_EVENT_HITS             (1 << 27)

/* Output record, byte-identical to the reference's struct Track. */
typedef struct clf_track {
    uint32_t hitsNum;
    uint32_t hits[CLF_MAX_TRACK_SIZE];
} clf_track;

/* Upper bound on tracklets, tracks and live ttf entries of one event with
 * N hits and S sensors: every gather chunk visits at most 64 h0 and a
 * sensor step runs at most ceil(HN/64) chunks, so one step visits at most
 * HN + 63 h0; each visit makes at most one tracklet; every track comes from
 * one tracklet.                                                          */
#define CLF_EVENT_TRACK_CAPACITY(N, S) ((int64_t)(N) + 64 * (int64_t)(S) + 64)

/* ---- Per-event status flags ------------------------------------------ */
/* A track that already holds 24 hits found a 25th.  The reference marks
 * that hit used (Kernel.cl:316, defined) and then appends it past its
 * private 24-slot record (Kernel.cl:321, undefined behaviour).  Here the
 * mark is made, the 24-hit record stays as stored and the track is not
 * followed further; each such track is counted (n_overflow).  Everything
 * the reference does before that append is reproduced: a 24-hit track is
 * re-queued and followed, skips included, like any other.               */
#define CLF_FLAG_TRACK_FULL      (1u << 0)
/* More than MAX_TRACKS output tracks: the reference writes into the next
 * event's slice (Kernel.cl:609,627).                                       */
#define CLF_FLAG_TRACKS_OVER     (1u << 1)
/* More than TTF_MODULO live ttf entries in one step: the reference's ring
 * overwrites entries it has not read yet (Kernel.cl:222,344).              */
#define CLF_FLAG_TTF_OVER        (1u << 2)
/* More tracklets than hits: the reference writes past its slice
 * (Kernel.cl:554,638).                                                     */
#define CLF_FLAG_TRACKLETS_OVER  (1u << 3)
/* More weak entries than hits (Kernel.cl:360,787).                         */
#define CLF_FLAG_WEAK_OVER       (1u << 4)
/* Two sensors' hit-z ranges touch, a hit z equals the integer Z of the
 * sensor two above, or a coordinate is not finite: a fit denominator can be
 * zero and the reference's multiply-by-bool blends (Kernel.cl:34,304,518)
 * can turn a running best into NaN.  Outside the supported domain.        */
#define CLF_FLAG_DEGENERATE_Z    (1u << 5)
/* Some sensor is not sorted by x (informational: handled exactly).        */
#define CLF_FLAG_UNSORTED        (1u << 6)
/* The input blob is malformed; no tracks were produced.                   */
#define CLF_FLAG_INVALID         (1u << 7)

/* Flags under which the reference's own result is undefined or corrupt. */
#define CLF_FLAGS_REFERENCE_UNDEFINED                                   \
    (CLF_FLAG_TRACK_FULL | CLF_FLAG_TRACKS_OVER | CLF_FLAG_TTF_OVER |   \
     CLF_FLAG_TRACKLETS_OVER | CLF_FLAG_WEAK_OVER | CLF_FLAG_DEGENERATE_Z | \
     CLF_FLAG_INVALID)

#endif /* CLF_PARAMS_H */
