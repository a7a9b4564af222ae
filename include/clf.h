/*
 * clf.h — C ABI of the MI355X search-by-triplet VELO pixel tracker.
 *
 * Plain pointers and sizes only.  Every entry point returns a clf_status;
 * nothing here calls exit() (the reference-signature C++ shims in
 * GpuPixelSearchByTriplet.h keep the reference's exit(-1) contract on top).
 *
 * Input: per-event byte blobs in the reference layout (Tools.cpp:68-80):
 *   int32 S; int32 N; int32 sensor_Zs[S]; int32 hitStarts[S];
 *   int32 hitNums[S]; uint32 hit_IDs[N]; float32 X[N], Y[N], Z[N]
 * Output: per event, T records of struct Track {uint32 hitsNum;
 * uint32 hits[24];} (KernelDefinitions.h:59-62, 100 bytes), hits being
 * indices into the event's hit arrays, unused slots zero.
 *
 * Which reference interface each entry point replaces is noted on it.
 */
#ifndef CLF_H
#define CLF_H

#include <stddef.h>
#include <stdint.h>

#include "clf_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum clf_status {
    CLF_OK = 0,
    CLF_E_INVALID_ARGUMENT = 1, /* bad pointer, count or configuration         */
    CLF_E_DEVICE = 2,           /* HIP runtime or kernel launch error          */
    CLF_E_OUT_OF_MEMORY = 3,    /* device or pinned host allocation failed     */
    CLF_E_NO_DEVICE = 4,        /* no HIP device, or the ordinal is invalid    */
    CLF_E_CAPACITY = 5          /* request beyond a stated maximum             */
} clf_status;

typedef enum clf_mode {
    /* One kernel launch per event, one workgroup each: the plain path. */
    CLF_MODE_PLAIN = 0,
    /* One persistent launch over the whole batch: workgroups pull events
     * from an atomic queue ordered by estimated cost. */
    CLF_MODE_BATCHED = 1
} clf_mode;

typedef struct clf_context clf_context;
typedef struct clf_batch clf_batch;

/* Per-event outcome (mirrors the kernel's result record). */
typedef struct clf_event_stats {
    int32_t n_tracks;     /* output tracks (the reference's atomics[e])    */
    uint32_t flags;       /* CLF_FLAG_* of clf_params.h                    */
    int32_t n_tracklets;  /* tracklets made by seeding                     */
    int32_t n_weak;       /* weak entries                                  */
    int32_t max_ttf_live; /* largest ttf ring span over the sensor steps   */
    int32_t n_full;       /* tracks that reached 24 hits (a full record)   */
    int32_t h0_visits;    /* h0 visited by the gather, repeats included    */
    int32_t n_overflow;   /* 24-hit tracks that found a 25th hit: the hit
                             is marked used and the track stops there, where
                             the reference writes past its record
                             (CLF_FLAG_TRACK_FULL)                         */
} clf_event_stats;

/* Launch parameters (no effect on results). */
/* A context starts with block_size 0 and lds_hits 0: the launch shape is
 * then chosen per batch from its mean event size (1024 threads and 512-hit
 * slots above 4096 hits per event, else 512 threads and 256-hit slots; the
 * resident serving kernel takes the latter).  Results never depend on the
 * launch. */
typedef struct clf_launch_config {
    int32_t block_size; /* 64, 128, 256, 512 or 1024 threads per event
                           workgroup, or 0: chosen per batch               */
    int32_t grid_size;  /* persistent workgroups; 0 = CUs x occupancy        */
    int32_t grain;      /* consecutive queue items (events, in cost order)
                           a workgroup takes per dequeue, 1..64              */
    int32_t use_graph;  /* batched mode: capture the run in a hipGraph and
                           replay it; 0 (the default): launch it directly.
                           Replays run as fast as direct launches with the
                           HIP runtime's graph packet capture off
                           (DEBUG_CLR_GRAPH_PACKET_CAPTURE=0 in the
                           environment before the first HIP call); with it
                           on they measured 4-5 % slower                   */
    int32_t lds_hits;   /* hits per LDS sensor slot, 256 or 512, or 0: with
                           the launch (256 for a set block_size); a sensor
                           with more hits is read from HBM                   */
    int32_t heavy_events; /* batched mode: the heavy_events costliest events
                           run in a second launch of heavy_block threads per
                           event, beside the main launch (0 = one launch)    */
    int32_t heavy_block;  /* any supported block_size                       */
} clf_launch_config;

/* Results of a batch; memory owned by the batch, valid until the batch
 * runs again or is destroyed. */
typedef struct clf_result {
    int32_t n_events;
    const clf_event_stats* stats; /* [n_events]                               */
    const int64_t* track_offset;  /* [n_events + 1]: event i's tracks are the
                                     stats[i].n_tracks records from record
                                     track_offset[i] of `tracks` on; entry
                                     n_events is the number of records.  In
                                     batched mode events are packed in the
                                     order they finish, which may change from
                                     run to run; each event's records do not */
    const uint8_t* tracks;        /* packed 100-byte Track records            */
} clf_result;

typedef struct clf_timing {
    float ms_run;    /* mean wall time of one run on the device stream (all kernels) */
    float ms_search; /* mean time of the search kernel(s) alone                      */
    int32_t reps;
} clf_timing;

const char* clf_status_string(clf_status status);
/* Message of the last failure on this thread ("" if none). */
const char* clf_last_error(void);
/* Library version string. */
const char* clf_version(void);

/* Creates a context on HIP device `device` with its own stream.
 * Replaces the per-call OpenCL platform/context/queue/program set-up of
 * invokeParallelSearch (KernelInvoker.cpp:30-68): nothing is compiled at
 * run time, the gfx950 code object is part of the library. */
clf_status clf_context_create(int device, clf_context** out);
void clf_context_destroy(clf_context* ctx);
clf_status clf_context_set_config(clf_context* ctx, const clf_launch_config* cfg);
clf_status clf_context_get_config(const clf_context* ctx, clf_launch_config* cfg);
/* Compute units of the context's device. */
int clf_context_num_cus(const clf_context* ctx);
/* Blocks until all work enqueued on the context's stream has finished
 * (the reference's clFinish / clWaitForEvents, KernelInvoker.cpp:169,175). */
clf_status clf_context_synchronize(clf_context* ctx);

/* Upper bound on the number of tracks of one event (CLF_EVENT_TRACK_CAPACITY),
 * or -1 for a malformed blob. */
int64_t clf_event_track_capacity(const uint8_t* event, size_t size);

/* Uploads n events into device buffers that stay resident.  Malformed
 * blobs are accepted and come back with CLF_FLAG_INVALID and no tracks.
 * Replaces the buffer creation and uploads of invokeParallelSearch
 * (KernelInvoker.cpp:77-117). */
clf_status clf_batch_create(clf_context* ctx, int32_t n_events, const uint8_t* const* events,
                            const size_t* sizes, clf_batch** out);
void clf_batch_destroy(clf_batch* batch);
/* Enqueues one search of the whole batch on the context's stream (no host
 * synchronisation).  Replaces clEnqueueNDRangeKernel of
 * KernelInvoker.cpp:173. */
clf_status clf_batch_run(clf_context* ctx, clf_batch* batch, clf_mode mode);
/* Waits for the batch and copies results to the host.  Replaces the
 * read-back of KernelInvoker.cpp:196-206. */
clf_status clf_batch_results(clf_context* ctx, clf_batch* batch, clf_result* out);
/* Runs the batch warmup + reps times and reports device-side timings
 * (hipEvents on the context's stream), like the reference's kernel-only
 * profiling-counter timing (KernelInvoker.cpp:175-188). */
clf_status clf_batch_time(clf_context* ctx, clf_batch* batch, clf_mode mode, int32_t warmup,
                          int32_t reps, clf_timing* out);

/* One call: upload, run, read back.  The C-ABI core of
 * invokeParallelSearch (KernelInvoker.cpp:14-275) and of
 * gpuPixelSearchByTripletInvocation (GpuPixelSearchByTriplet.cpp:44-65). */
clf_status clf_search(clf_context* ctx, clf_mode mode, int32_t n_events,
                      const uint8_t* const* events, const size_t* sizes, clf_result* out,
                      clf_batch** batch_out);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc):
 * clf_search without any allocation.  Every device buffer of the call is
 * carved out of one static device array of 128 MiB that the library holds
 * on each device it runs on (its address is looked up once per device when
 * the first context there is created); nothing is allocated or freed, on
 * the device or in pinned memory.  At most 4096 events per call, and what
 * they need (about 2 MB per event of 2,800 hits, 12 MB at 32,768 hits)
 * must fit the array: otherwise CLF_E_CAPACITY with the sizes in
 * clf_last_error(), before anything is enqueued.  The caller passes host
 * inputs as clf_search takes them and host outputs: stats[n_events],
 * track_offset[n_events + 1] (as clf_result's) and tracks (track_capacity
 * records; too few gives CLF_E_CAPACITY after the search, with stats and
 * track_offset filled).  Results are those of clf_search.  Calls share the
 * device's array and leave their contents in it: calls that overlap on one
 * device (from any context or thread) race and must not be made. */
clf_status clf_search_noalloc(clf_context* ctx, clf_mode mode, int32_t n_events,
                              const uint8_t* const* events, const size_t* sizes, uint8_t* tracks,
                              int64_t track_capacity, int64_t* track_offset,
                              clf_event_stats* stats);

/* Stage (a) alone: the candidate windows of every hit, as fillCandidates
 * leaves dev_hit_candidates and dev_hit_h2_candidates (Kernel.cl:50-178,
 * -1 where it writes nothing).  cand_h1[i] and cand_h2[i] hold 2*N_i int32
 * for event i; flags (nullable) receives the input-only CLF_FLAG_* bits. */
clf_status clf_fill_candidates(clf_context* ctx, int32_t n_events, const uint8_t* const* events,
                               const size_t* sizes, int32_t* const* cand_h1,
                               int32_t* const* cand_h2, uint32_t* flags);

/* Verification probe, no counterpart in the reference: the same windows,
 * computed the way the search computes them under the context's launch
 * configuration (both sensors of a window staged into LDS ring slots with
 * their x-bucket index by the configuration's staging code, bucket-started
 * scans; sensors above lds_hits hits read from HBM).  Same output form as
 * clf_fill_candidates, which it must equal. */
clf_status clf_fill_candidates_search_path(clf_context* ctx, int32_t n_events,
                                           const uint8_t* const* events, const size_t* sizes,
                                           int32_t* const* cand_h1, int32_t* const* cand_h2,
                                           uint32_t* flags);

/* Measurement probe, no counterpart in the reference: runs the batch once
 * with per-event phase timers and work counters on.  out receives 16
 * uint64 per event: wall-clock ticks of prologue, staging, forwarding,
 * gather, seeding set-up, row stops, pair evaluation, emission, tail, total,
 * then forwarded entries, gathered h0, (h0,h1) pairs, seeding batches and
 * forwarding tiles.  tick_ns receives the length of one tick. */
clf_status clf_batch_profile(clf_context* ctx, clf_batch* batch, clf_mode mode, uint64_t* out,
                             double* tick_ns);

/* Measurement probe, no counterpart in the reference: copies `bytes`
 * (a multiple of 16) between two scratch buffers with a plain 16-byte-per-lane
 * kernel, `reps` times, and reports the mean duration.  Its known traffic
 * (bytes read + bytes written) calibrates HBM counters. */
clf_status clf_probe_copy(clf_context* ctx, int64_t bytes, int32_t reps, float* ms_per_copy);

/* Measurement probe, no counterpart in the reference: the host link's
 * throughput for `bytes` between pinned host memory and the device, `reps`
 * copies each way on the context stream; GB/s (1e9 bytes per second) of
 * host-to-device and device-to-host copies.  The floor of a call that
 * takes its events from the host and returns tracks to it. */
clf_status clf_probe_link(clf_context* ctx, int64_t bytes, int32_t reps, double* h2d_gbps,
                          double* d2h_gbps);

/* ---- Serving ----------------------------------------------------------
 * A server accepts events one at a time and returns each event's tracks
 * as soon as they are ready.  The reference has no serving loop: its Gaudi
 * offload handler calls gpuPixelSearchByTriplet (GpuPixelSearchByTriplet.cpp
 * :29-37) once per request with a blocking upload, launch and read-back
 * (KernelInvoker.cpp:114-206).  Results are bit-identical to clf_search on
 * the same event.  Two engines:
 *
 *  CLF_SERVE_RESIDENT  a persistent kernel stays on the device while there
 *    is work.  Submit copies the event into pinned staging; an uploader
 *    thread moves the newly submitted events into a device input arena
 *    with the copy engine and then publishes their descriptors in pinned
 *    memory.  One wave of the kernel follows the descriptor ring;
 *    workgroups claim events one by one, copy the event from the device
 *    arena, search it, write its tracks into a device output ring and
 *    append a completion record.  A host thread copies the finished part of
 *    the output ring back with the copy engine, hands out the results, and
 *    relaunches the kernel after it has left for lack of work (idle_us).
 *    Admission is by device memory: the input arena, and room in the output
 *    ring for every admitted event's capacity bound of tracks.
 *  CLF_SERVE_BATCHES   a dispatcher thread keeps `slots` batches in flight,
 *    each on its own stream with resident device buffers, and forms the
 *    next batch from everything that arrived meanwhile; tracks come back
 *    when the whole batch has finished.
 *
 * Threading: clf_server_submit, _wait and _release may be called from any
 * number of threads.  Admission: a submit that would exceed an arena, the
 * request table or the engine's per-event or per-batch limits returns
 * CLF_E_CAPACITY at once (nothing is queued). */
typedef struct clf_server clf_server;

typedef enum clf_serve_engine { CLF_SERVE_RESIDENT = 0, CLF_SERVE_BATCHES = 1 } clf_serve_engine;

typedef struct clf_server_config {
    int32_t engine;           /* clf_serve_engine                               */
    int32_t max_requests;     /* submitted and not yet released                 */
    int64_t arena_bytes;      /* input arena: pinned staging (and, resident, the
                                 device input arena of the same size)           */
    int64_t out_arena_bytes;  /* resident: device output ring (and its pinned
                                 mirror)                                        */
    int32_t max_event_hits;   /* resident: largest event (workgroup scratch)    */
    int32_t idle_us;          /* resident: the kernel leaves after this long
                                 without work                                   */
    int32_t slots;            /* batches: batches in flight (streams), 1..8     */
    int32_t max_batch_events; /* batches: events per batch                      */
    int64_t max_batch_hits;   /* batches: hits per batch; sizes each slot       */
    int32_t max_wait_us;      /* batches: while another batch runs, a batch
                                 waits up to this long to fill; an idle device
                                 launches at once                               */
    int32_t feeder_pause_us;  /* resident, testing: in the first launch the
                                 kernel's feeder stalls this long, with no
                                 heartbeat, before its first hand-out (0 in
                                 use)                                           */
} clf_server_config;

/* A finished request.  `tracks` points into the server's pinned output and
 * stays valid until clf_server_release(ticket). */
typedef struct clf_served {
    uint64_t ticket;
    uint64_t tag;             /* the submitter's tag                            */
    int32_t status;           /* CLF_OK, or the error that ended the request    */
    int32_t batch_events;     /* batches: size of its batch; resident: 1        */
    clf_event_stats stats;
    int64_t n_tracks;
    const uint8_t* tracks;    /* n_tracks * 100 bytes                           */
    int64_t t_submit_ns;      /* steady clock: submit returned                  */
    int64_t t_launch_ns;      /* batches: its batch was enqueued; resident: its
                                 event was in device memory and its descriptor
                                 published                                      */
    int64_t t_done_ns;        /* its tracks were visible on the host            */
    int64_t t_complete_ns;    /* its search had finished: resident, its
                                 completion record was seen on the host;
                                 batches, its batch was seen finished           */
} clf_served;

/* Fills *cfg with the defaults: resident engine, 65536 requests, 512 MiB
 * input arena and 2 GiB output ring, events up to 32768 hits, 5000 us idle;
 * for the batch engine 4 slots of 256 events / 2^20 hits, 100 us.  The
 * resident engine's device arenas are fine-grained device memory, which
 * the library keeps for reuse by later servers until the process ends. */
void clf_server_default_config(clf_server_config* cfg);
/* Starts a server on the context's device with the context's launch
 * configuration.  Allocates every buffer up front. */
clf_status clf_server_create(clf_context* ctx, const clf_server_config* cfg, clf_server** out);
/* Stops the server after the requests in flight have finished and frees
 * everything; no other call on it may be in progress.  Requests not yet
 * finished end with CLF_E_INVALID_ARGUMENT. */
void clf_server_destroy(clf_server* server);
/* Copies one event into the input arena and queues it.  A malformed blob is
 * rejected with CLF_E_INVALID_ARGUMENT. */
clf_status clf_server_submit(clf_server* server, const uint8_t* event, size_t size, uint64_t tag,
                             uint64_t* ticket);
/* Submits events[0..n) (n <= 64) in order, as clf_server_submit would one by
 * one, but takes the server's lock once for all of them; stops at the first
 * event refused.  *n_accepted events are queued (tickets[k] set for those,
 * 0 for the rest) and the status is that of the first refused (CLF_OK if
 * none was).  For a front end that receives several requests at once, or
 * has fallen behind.  No counterpart in the reference, which has no serving
 * front end. */
clf_status clf_server_submit_many(clf_server* server, int32_t n, const uint8_t* const* events,
                                  const size_t* sizes, const uint64_t* tags, uint64_t* tickets,
                                  int32_t* n_accepted);
/* Waits up to timeout_ms (< 0: no limit) for a request; CLF_E_CAPACITY on
 * time-out (the request stays queued). */
clf_status clf_server_wait(clf_server* server, uint64_t ticket, int32_t timeout_ms,
                           clf_served* out);
/* Gives a finished request's output back to the server. */
clf_status clf_server_release(clf_server* server, uint64_t ticket);
/* clf_server_submit of a request in the Gaudi offload framing, the wire
 * format of the reference's .dat files and offload handler
 * (IndependentEntrypoint.cpp:103-123): int32 name_len; name; int32
 * data_size; data (bytes after the data are ignored).  The name is not
 * checked. */
clf_status clf_server_submit_framed(clf_server* server, const uint8_t* frame, size_t size,
                                    uint64_t tag, uint64_t* ticket);

/* Load replay: arrival i submits pool event pick[i] at arrival_ns[i] after
 * the start (open loop, from submit_threads threads), every result is
 * digested (clf_tracks_digest) and released as it arrives.  Latency is
 * counted from the scheduled arrival to the result visible on the host. */
typedef struct clf_replay_record {
    int64_t latency_ns;       /* -1 if rejected                                 */
    int64_t wait_ns;          /* scheduled arrival to dispatch: batches, its
                                 batch's launch; resident, its descriptor's
                                 publication (its event in device memory)       */
    uint64_t digest;
    int32_t n_tracks;
    int32_t status;           /* CLF_OK, or why the submit or the run failed    */
    int32_t batch_events;
    uint32_t flags;
    int64_t device_ns;        /* dispatch to the search seen finished (the rest
                                 of the latency is waiting and copying)         */
} clf_replay_record;

typedef struct clf_replay_summary {
    double wall_s;            /* first arrival to last result                   */
    double events_per_s;      /* completed / wall_s                             */
    int64_t completed, rejected;
    int64_t p50_ns, p90_ns, p99_ns, max_ns;
    int64_t batches;          /* batches: batches launched; resident: launches
                                 of the persistent kernel                       */
    double mean_batch_events; /* batches: events per batch; resident: 1 (each
                                 event is dispatched alone)                     */
    double max_submit_lag_ns; /* how late the submitters ran at worst           */
    /* resident engine, device wall clock per event (means, microseconds):
     * event copy from the device arena into the workgroup's scratch, search,
     * track copy into the output ring and completion */
    double dev_copy_in_us, dev_search_us, dev_copy_out_us;
    double dev_busy_frac;     /* workgroup time with an event / (workers x wall) */
    double host_submit_us;    /* mean submit time per arrival (a late submitter
                                 queues the arrivals already due together, with
                                 clf_server_submit_many)                        */
    double host_digest_us;    /* mean time to digest and release one result     */
} clf_replay_summary;

clf_status clf_server_replay(clf_server* server, int32_t n_pool, const uint8_t* const* pool,
                             const size_t* sizes, int64_t n_arrivals, const int32_t* pick,
                             const int64_t* arrival_ns, int32_t submit_threads,
                             clf_replay_record* records, clf_replay_summary* summary);

/* 64-bit digest of n_tracks packed Track records: eight multiply-xorshift
 * lanes over the 64-bit words of each 64-byte block, the remaining 32-bit
 * words spread over the lanes in turn, then a final mix with the count.
 * Used to compare outputs of different paths without moving them. */
uint64_t clf_tracks_digest(const uint8_t* tracks, int64_t n_tracks);

/* Splits one Gaudi offload frame (IndependentEntrypoint.cpp:103-123):
 * int32 name_len; char name[name_len]; int32 data_size; uint8 data[...].
 * Pointers into `frame`; name and name_len may be null.  Unlike the
 * reference reader, lengths are checked against `size`. */
clf_status clf_unframe(const uint8_t* frame, size_t size, const uint8_t** name, int32_t* name_len,
                       const uint8_t** data, size_t* data_size);

/* The reference's per-event result text, results/<i>.out
 * (KernelInvoker.cpp:209-243 with printTrack :254-275 and
 * findClosestModule :284-303): for each track "Track #<n>, length <h>" and
 * one line per hit with ID, hit index, module, x, y, z in default C++
 * stream formatting, then an empty line.  The z -> module map is built as
 * the reference builds it (sensor Zs, then every hit's truncated z mapped
 * to the nearest key present, in hit order).  *len receives the text
 * length; text (capacity cap, NUL-terminated) may be null to ask for it. */
clf_status clf_format_tracks(const uint8_t* event, size_t size, const uint8_t* tracks,
                             int64_t n_tracks, char* text, size_t cap, size_t* len);

/* Name of the context's device (e.g. for the reference's
 * "Invoking kernels on your <device>" line, KernelInvoker.cpp:145-147). */
const char* clf_context_device_name(const clf_context* ctx);

/* x-presort of one event in place, per sensor, with the reference's
 * partition sequence so that tied x keep its order (preorder_by_x,
 * Tools.cpp:46-66, 117-164).  Host code; 0 or -1 for a malformed blob. */
int clf_presort_event(uint8_t* event, size_t size);

#ifdef __cplusplus
}
#endif
#endif /* CLF_H */
