/*
 * clf_oracle.h — CPU restatement of the reference's search-by-triplet path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; the product library never links, loads or calls it.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark use
 * it.
 *
 * Every function restates the reference literally, in its operation order,
 * FP32, without contraction (built with -ffp-contract=off), one event on one
 * thread.  Where the reference is nondeterministic (atomic slot order) the
 * restatement fixes one valid order, the "lane order" described in
 * clf_oracle.c; the HIP path produces the same order.
 */
#ifndef CLF_ORACLE_H
#define CLF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clf_oracle_stats {
    int32_t  n_tracks;        /* output tracks                                */
    uint32_t flags;           /* CLF_FLAG_* of clf_params.h                   */
    int32_t  n_tracklets;     /* tracklets made by seeding                    */
    int32_t  n_weak;          /* weak entries                                 */
    int32_t  max_ttf_live;    /* largest ring occupancy over the steps        */
    int32_t  n_full;          /* tracks that reached 24 hits                  */
    int32_t  n_overflow;      /* 24-hit tracks that found a 25th hit          */
    int32_t  pad;
    int64_t  seed_h0_visits;  /* gathered h0, with the gather's repeats       */
    int64_t  seed_pairs;      /* (h0,h1) pairs evaluated (live, unused h1)    */
    int64_t  seed_triplets;   /* h2 window members evaluated for those pairs  */
    int64_t  fwd_fits;        /* forwarding fits (followed track x hits)      */
} clf_oracle_stats;

/* Candidate windows of every hit (Kernel.cl:50-178).  cand_h1 and cand_h2
 * hold 2*N int32 each; -1 where the reference leaves its -1 fill
 * (KernelInvoker.cpp:162-163).  Returns 0, or -1 for a malformed blob. */
int clf_oracle_fill_candidates(const uint8_t* blob, size_t size,
                               int32_t* cand_h1, int32_t* cand_h2);

/* Whole event (Kernel.cl:597-816 with local size (64,4)).  Writes up to
 * track_capacity 100-byte records to out_tracks; stats->n_tracks is the
 * count.  Returns 0, -1 for a malformed blob, -2 if track_capacity is too
 * small (stats still filled). */
int clf_oracle_search_event(const uint8_t* blob, size_t size,
                            uint8_t* out_tracks, int64_t track_capacity,
                            clf_oracle_stats* stats);

/* Runs clf_oracle_search_event over n events on n_threads host threads
 * (one event per task).  out_tracks[i] must hold track_capacity[i] records.
 * Returns the number of events that failed. */
int clf_oracle_search_batch(int n, const uint8_t* const* blobs,
                            const size_t* sizes, uint8_t* const* out_tracks,
                            const int64_t* track_capacity,
                            clf_oracle_stats* stats, int n_threads);

/* x-presort of one event in place (Tools.cpp:46-66, 117-164): per sensor,
 * the same partition sequence, so tied x keep the reference's order. */
int clf_oracle_presort(uint8_t* blob, size_t size);

/* The h0 gather of one sensor step alone (Kernel.cl:720-754): hn0 hits,
 * used[0..hn0) their flags.  Writes the visited element indices in the
 * restatement's lane order to visits (room for hn0 + 64) and returns their
 * number; with more than 128 hits some unused hits are skipped and some
 * visited twice, as in the reference. */
int clf_oracle_gather(int hn0, const uint8_t* used, int32_t* visits);

/* Per-event status flags computable from the input alone
 * (CLF_FLAG_DEGENERATE_Z, CLF_FLAG_UNSORTED, CLF_FLAG_INVALID). */
uint32_t clf_oracle_input_flags(const uint8_t* blob, size_t size);

#ifdef __cplusplus
}
#endif
#endif
