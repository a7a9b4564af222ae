// clf_kernels.hip — gfx950 kernels of the search-by-triplet VELO pixel tracker.
//
// One workgroup owns one event at a time and pulls events from an atomic
// work queue (persistent kernel).  Inside an event the sensor steps
// f = S-1 .. 4 run in sequence (the reference's Kernel.cl:671-767 chain).
// Each step is forwarding, then the h0 gather, then seeding, spread over
// the workgroup; the candidate windows of the next step, which do not
// depend on the used flags, are computed beside the forwarding by the
// threads it leaves idle, and in the batch kernel wave 0 gathers the
// step's first seeding batch beside it too (gathered again if the
// forwarding marks a hit of the gather sensor).
//
// Data placement.  Step f reads hits of sensors f (h0), f-2 (h1), f-4 (h2
// and the forwarding target) and prepares sensors f-1, f-3 and f-5 for the
// next step, so the hit coordinates of sensors f .. f-5 sit in a ring of six
// LDS slots; every step stages one new sensor (512 x 256 and 1024-thread
// workgroups keep a seventh slot, filled with the next step's sensor during
// the step; see StageThreads).  A sensor with more hits
// than a slot holds (C, a template parameter the autotuner picks) is read
// from HBM instead.  Candidate windows live in small LDS double buffers
// (HBM for oversized sensors), the used flags in an LDS bitmask, and the
// two ttf lists in LDS up to what the configuration's occupancy leaves
// (HBM beyond).  Weak entries and output tracks go to HBM.
//
// Registers.  Values that are the same in every lane are kept in scalar
// registers (uni(): sensor tables, the event descriptor, the wave index,
// block totals, the sensor loop counter), which holds a workgroup to 128
// vector registers: two 512-thread or four 256-thread workgroups per CU.
//
// Stages:
//   (a) candidate windows  Kernel.cl:50-178   window_h1 / window_h2: binary
//       search on x-sorted slices, exact scans for unsorted sensors
//   (c) forwarding         Kernel.cl:199-368  best h2 on sensor f-4 over the
//       x-range a 0.6 mm tolerance allows
//   (b) gather + seeding   Kernel.cl:392-568, 720-754  the (64,4) semantics:
//       64-wide h0 chunks, 4 rows of h1 with the per-row stop
//   (d) used flags, weak   Kernel.cl:313-363, 769-815
//
// Every append (tracks, ttf entries, weak entries) goes through a wave
// ballot / prefix sum in a fixed order instead of a global atomic, so the
// output, order included, is the same on every run.
//
// FP32, built with -ffp-contract=off: every expression keeps the reference's
// operation order.  Restricting a scan to an x-range is exact: it drops only
// hits whose |dx| >= 0.6 mm follows from the monotonicity of IEEE rounding
// (x_window()).
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>

#include "clf_device.h"
#include "clf_params.h"

namespace clf {

constexpr int kWave = 64;
constexpr int kMaxVisits = 128;   // gathered h0 handled per seeding batch (multiple of 64)
constexpr int kStepSensors = 6;   // sensors a step reads: f .. f-5
constexpr int kUsedBits = 32768;  // events up to this many hits keep used flags in LDS
constexpr float kLim = 1.0e5f;    // |coordinate| bound of the supported domain
constexpr int16_t kNoStop = 0x7fff;
constexpr int kBuckets = 128;     // x-bucket index per LDS sensor slot

// Phase profile slots (wall-clock ticks, then work counters), filled when
// SearchArgs::prof is set.  Thread 0 takes a timestamp after the barriers
// that close each phase.
enum ProfSlot {
    PROF_PROLOGUE = 0,    // header, per-sensor checks, ring set-up
    PROF_STAGE = 1,       // staging sensor f-5 into the ring
    PROF_FORWARD = 2,     // forwarding + windows of the next step
    PROF_GATHER = 3,      // the h0 gather (wave 0)
    PROF_SEED_SETUP = 4,  // per-h0 window lookup, row stops, pair prefix
    PROF_STAGE_WAIT = 5,  // the barrier after the staging
    PROF_SEED_EVAL = 6,   // (h0, h1) pairs: h2 range and fits
    PROF_SEED_EMIT = 7,   // ttf entries of the new tracklets
    PROF_TAIL = 8,        // tail + weak filter
    PROF_TOTAL = 9,
    CNT_FWD_ENTRIES = 10,
    CNT_VISITS = 11,
    CNT_PAIRS = 12,
    CNT_BATCHES = 13,
    CNT_FWD_TILES = 14,
    CNT_H2_TESTS = 15,    // h2 candidates fitted by the seeding
};

struct PhaseClock {
    bool on;
    unsigned long long t;
};

// ---------------------------------------------------------------------------
// Wave / block helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
// (the same in every lane of a wave: a scalar, so that branches on it are
// scalar branches)
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// number of set bits of m below this lane
__device__ __forceinline__ int mbcnt(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// index of the k-th (0-based) set bit of m; m must have more than k bits set
__device__ __forceinline__ int kth_set_bit(uint64_t m, int k) {
    int pos = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const int c = __popcll(m & ((1ull << step) - 1ull));
        if (k >= c) {
            k -= c;
            m >>= step;
            pos += step;
        }
    }
    return pos;
}
// Values every lane of the wave holds alike (LDS words read at a uniform
// address, block totals), moved to scalar registers so that they take no
// vector registers.  Only for values that are the same in all active lanes.
// A workgroup barrier whose fences cover LDS only: it waits for the
// wave's LDS accesses (lgkmcnt) but leaves its global stores in flight,
// which a full barrier would wait for.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ int64_t uni(int64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
// Inclusive prefix sum over the 64 lanes of a wave (all lanes active), in
// DPP moves that stay in the vector ALU: shifts by 1, 2, 4, 8 within each
// row of 16 lanes (lanes shifted in from outside the row add 0), then lane
// 15 of rows 0 and 2 added to rows 1 and 3, then lane 31 added to rows 2
// and 3.  A shuffle-based scan takes six LDS-pipeline round trips instead.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
__device__ __forceinline__ int wave_incl_scan(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, kDppRowShr1, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, kDppRowShr2, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, kDppRowShr4, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, kDppRowShr8, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, kDppRowBcast15, 0xa, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, kDppRowBcast31, 0xc, 0xf, false);
    return x;
}
// Minimum and maximum over the 64 lanes of a wave (all lanes active), the
// same DPP pattern with +inf / -inf shifted in; lane 63 holds the result.
template <int Ctrl, int RowMask>
__device__ __forceinline__ float dpp_move(float ident, float x) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(ident), __float_as_int(x), Ctrl, RowMask, 0xf, false));
}
__device__ __forceinline__ void wave_min_max(float& mn, float& mx) {
    mn = fminf(mn, dpp_move<kDppRowShr1, 0xf>(INFINITY, mn));
    mx = fmaxf(mx, dpp_move<kDppRowShr1, 0xf>(-INFINITY, mx));
    mn = fminf(mn, dpp_move<kDppRowShr2, 0xf>(INFINITY, mn));
    mx = fmaxf(mx, dpp_move<kDppRowShr2, 0xf>(-INFINITY, mx));
    mn = fminf(mn, dpp_move<kDppRowShr4, 0xf>(INFINITY, mn));
    mx = fmaxf(mx, dpp_move<kDppRowShr4, 0xf>(-INFINITY, mx));
    mn = fminf(mn, dpp_move<kDppRowShr8, 0xf>(INFINITY, mn));
    mx = fmaxf(mx, dpp_move<kDppRowShr8, 0xf>(-INFINITY, mx));
    mn = fminf(mn, dpp_move<kDppRowBcast15, 0xa>(INFINITY, mn));
    mx = fmaxf(mx, dpp_move<kDppRowBcast15, 0xa>(-INFINITY, mx));
    mn = fminf(mn, dpp_move<kDppRowBcast31, 0xc>(INFINITY, mn));
    mx = fmaxf(mx, dpp_move<kDppRowBcast31, 0xc>(-INFINITY, mx));
    mn = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mn), kWave - 1));
    mx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mx), kWave - 1));
}

// Threads that stage sensors: all but wave 0 for a 512-thread workgroup
// whose other waves cover a slot with as many hits per thread, and for a
// 1024-thread one.  Those waves store the next step's sensor f-6 during
// step f into a seventh ring slot (the step reads six sensors, f .. f-5,
// and the slot of f+1 = f-6 mod 7 is free then): at 512 threads from
// registers loaded at the top of the step (by the waves the step leaves
// idle, or else by these in the forwarding's last tile), at 1024 threads
// straight from HBM after the forwarding.  Other configurations stage
// sensor f-5 with all threads at the top of step f, into the slot of f+1
// (six slots).
template <int B, int C>
struct StageThreads {
    static constexpr int first =
        ((B == 512 && (C + B - kWave - 1) / (B - kWave) == (C + B - 1) / B) || B == 1024) ? kWave : 0;
    static constexpr int count = B - first;
    static constexpr bool during_gather = first > 0;
};
// Configurations whose sensor f-6 is loaded into registers at the top of
// step f and stored into the ring during the step (in the forwarding's last
// tile, or during the gather when there is no forwarding).
template <int B, int C>
constexpr bool kStageInForward = StageThreads<B, C>::during_gather && B <= 512;

template <int B, int C>
struct Ring {
    static constexpr int slots = StageThreads<B, C>::during_gather ? kStepSensors + 1 : kStepSensors;
};

template <int B, int C>
struct SharedCore {
    static constexpr int W = B / kWave;
    // per-sensor data of the current event
    int32_t s_start[CLF_MAX_SENSORS];
    int32_t s_num[CLF_MAX_SENSORS];
    int32_t s_Z[CLF_MAX_SENSORS];
    float s_zmin[CLF_MAX_SENSORS];
    float s_zmax[CLF_MAX_SENSORS];
    int32_t s_sorted[CLF_MAX_SENSORS];
    // ring of sensor slots: hit coordinates of sensors f .. f-5
    float rx[Ring<B, C>::slots][C];
    float ry[Ring<B, C>::slots][C];
    float rz[Ring<B, C>::slots][C];
    // x-bucket index of each slot (sorted sensors): the first hit of bucket
    // >= j, bucket(x) = clamp(floor((x - x0) * scale))
    float bk_x0[Ring<B, C>::slots], bk_scale[Ring<B, C>::slots];
    int16_t bk_start[Ring<B, C>::slots][kBuckets + 1];
    // windows: sensor g's h1 windows in w1[g & 1], its h2 windows in
    // w2[g & 1] (each buffer lives from its write to its last read)
    W1Rec w1[2][C];
    W2Rec w2[2][C];
    // which of the h2 windows in w2[k] are empty (bit per hit), and how many
    uint32_t w2_empty[2][C / 32];
    int32_t w2_n_empty[2];
    uint32_t used_bits[kUsedBits / 32];
    int32_t scan[2][3][W];  // two buffers, used in turn by block_scan3
    int32_t item;
    uint32_t flags;
    int32_t max_live;
    int32_t n_full;
    int32_t n_overflow;
    // gather state of the current sensor step (Kernel.cl:720-721,753-754)
    uint32_t g_prev, g_shift, g_counter;
    int32_t g_nv, g_more;
    // set when this step's forwarding marks a hit of the gather sensor f
    // (the first batch, gathered during the forwarding, is then redone)
    int32_t g_redo;
    int32_t n_listed;  // entries of the list the next step forwards
    int64_t pack_base;  // first packed record of the current event
    // the current batch of gathered h0 (indices relative to sensor f)
    int32_t v_h0[kMaxVisits];
    int32_t v_first[kMaxVisits];  // h1 window start, relative to sensor f-2
    int32_t v_pre[kMaxVisits + 1];
    float v_x0[kMaxVisits], v_y0[kMaxVisits], v_z0[kMaxVisits], v_dymax[kMaxVisits];
    int16_t v_stop[kMaxVisits * CLF_SEED_ROWS];
    unsigned long long v_best[kMaxVisits];
    unsigned long long prof[kProfSlots];
    // resident serving: the claimed descriptor
    uint64_t sv_in;
    int32_t sv_n, sv_s, sv_req, sv_cap;
    uint32_t sv_d;
};

// LDS a workgroup may take without lowering the workgroups per CU that its
// registers allow (four 256-thread, two 512-thread or one 1024-thread
// workgroup; 512-hit slots hold 256-thread workgroups to two by LDS).
template <int B, int C>
struct LdsBudget {
    static constexpr int value = B >= 1024 ? 163840 : (B >= 512 || C >= 512) ? 81920 : 40960;
};

// Entries of each ttf list kept in LDS (entries beyond go to HBM): what the
// budget leaves after the rest of the workgroup's LDS, less 256 bytes.
template <int B, int C>
struct TtfCap {
    static constexpr int room = LdsBudget<B, C>::value - (int)sizeof(SharedCore<B, C>) - 256;
    static constexpr int value = room > 0 ? room / (2 * 48) : 0;
};

template <int B, int C>
struct Shared : SharedCore<B, C> {
    // the two ttf lists (current and next step), front part: entry i is
    // ttf[k][0][i], ttf[k][1][i], ttf[k][2][i] (16 bytes each), so that
    // consecutive lanes read consecutive 16-byte words
    uint4 ttf[2][3][TtfCap<B, C>::value > 0 ? TtfCap<B, C>::value : 1];
};

template <int B, int C>
__device__ __forceinline__ void prof_mark(Shared<B, C>& sh, PhaseClock& pc, int slot) {
    if (pc.on && threadIdx.x == 0) {
        const unsigned long long now = wall_clock64();
        sh.prof[slot] += now - pc.t;
        pc.t = now;
    }
}
template <int B, int C>
__device__ __forceinline__ void prof_add(Shared<B, C>& sh, const PhaseClock& pc, int slot,
                                         unsigned long long v) {
    if (pc.on && threadIdx.x == 0) sh.prof[slot] += v;
}

// Exclusive prefix counts of up to three predicates over the block, in
// thread order.  Every thread of the block must call it, with the same
// `sel` (flipped here).  The per-wave counts alternate between two buffers:
// a wave can only write a buffer again after the next call's barrier, which
// every wave reaches after its reads of this call, so one barrier suffices.
template <int B, int C>
__device__ __forceinline__ void block_scan3(Shared<B, C>& sh, bool p0, bool p1, bool p2,
                                            int (&pos)[3], int (&tot)[3], int& sel) {
    const int lane = lane_id(), wave = wave_id();
    const uint64_t m[3] = {ballot(p0), ballot(p1), ballot(p2)};
    auto& cnt = sh.scan[sel];
    sel ^= 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pos[k] = mbcnt(m[k]);
        if (lane == 0) cnt[k][wave] = __popcll(m[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int base = 0, t = 0;
        for (int w = 0; w < Shared<B, C>::W; ++w) {
            const int c = cnt[k][w];
            base += (w < wave) ? c : 0;
            t += c;
        }
        pos[k] += base;
        tot[k] = uni(t);
    }
}

// ---------------------------------------------------------------------------
// Searches over an x-sorted slice (generic pointers: LDS or HBM)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lb_ge(const float* x, int lo, int hi, float v) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] >= v) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}
__device__ __forceinline__ int ub_gt(const float* x, int lo, int hi, float v) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] > v) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}
// first i in [lo,hi) with fl(x[i] - c) > t   (monotone in x[i])
__device__ __forceinline__ int first_diff_gt(const float* x, int lo, int hi, float c, float t) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((x[mid] - c) > t) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}
// first i in [lo,hi) with fl(x[i] - c) >= t
__device__ __forceinline__ int first_diff_ge(const float* x, int lo, int hi, float c, float t) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((x[mid] - c) >= t) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// The same searches through a slot's x-bucket index.  bucket() is
// nondecreasing in x (IEEE subtraction and multiplication by scale >= 0 are
// monotone, and so are the truncation and the clamps; NaN goes to 0), so
// every hit before start[bucket(v)] has x < v.  A search starts there (or at
// a bucket whose threshold is below the exact one by a margin, for
// predicates on a rounded difference) and scans with the exact predicate,
// four loads in flight; a bucket holds a few hits, so the scan is short.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bucket_of(float x, float x0, float scale) {
    const float t = (x - x0) * scale;
    return t >= (float)(kBuckets - 1) ? kBuckets - 1 : (t >= 0.f ? (int)t : 0);
}
// first i in [i, hi) with pred(x[i]) (pred false then true along x)
template <class P>
__device__ __forceinline__ int scan_to(const float* x, int i, int hi, P pred) {
    for (; i < hi; i += 4) {
        const float a0 = x[i];
        const float a1 = x[min(i + 1, hi - 1)];
        const float a2 = x[min(i + 2, hi - 1)];
        const float a3 = x[min(i + 3, hi - 1)];
        if (pred(a0)) return i;
        if (i + 1 < hi && pred(a1)) return i + 1;
        if (i + 2 < hi && pred(a2)) return i + 2;
        if (i + 3 < hi && pred(a3)) return i + 3;
    }
    return hi;
}
// start index for a search in [lo, hi) whose predicate fails below v
__device__ __forceinline__ int bucket_start(const int16_t* bs, float x0, float scale, float v, int lo,
                                            int hi) {
    const int st = bs[bucket_of(v, x0, scale)];
    return st < lo ? lo : (st > hi ? hi : st);
}
// threshold below which fl(x - c) > t (and >= t) fail, with a margin that
// covers the rounding of the difference and of the threshold itself
__device__ __forceinline__ float diff_threshold(float c, float t) {
    return (c + t) - (2.0e-3f + 1.0e-5f * (fabsf(c) + fabsf(t)));
}

// Widened bounds [lo, hi] of the x of hits that can pass |x_pred - x| < 0.6.
// x_pred over a sensor is bracketed by pa and pb, its values at the sensor's
// extreme z (the prediction chain is monotone in z under IEEE rounding).  If
// x > pmax + 0.6 exactly, fl(x_pred - x) <= -0.6f and the cut fails; the
// margin covers the rounding of the bound itself.
__device__ __forceinline__ void x_window(float pa, float pb, float& lo, float& hi) {
    lo = fminf(pa, pb) - CLF_PARAM_TOLERANCE;
    lo = lo - (1.0e-3f + fabsf(lo) * 1.0e-6f);
    hi = fmaxf(pa, pb) + CLF_PARAM_TOLERANCE;
    hi = hi + (1.0e-3f + fabsf(hi) * 1.0e-6f);
}

// fitHitToTrack (Kernel.cl:15-35) with its multiply-by-bool blend, split
// at the hit's z: fit_at takes the predictions and the scatter denominator,
// which depend on the candidate only through its z.
__device__ __forceinline__ float fit_at(float x_prediction, float y_prediction, float scatterDenom,
                                        float x2, float y2) {
    const float dx = fabsf(x_prediction - x2);
    const bool tolx = dx < CLF_PARAM_TOLERANCE;
    const float dy = fabsf(y_prediction - y2);
    const bool toly = dy < CLF_PARAM_TOLERANCE;
    const float scatterNum = (dx * dx) + (dy * dy);
    const float scatter = scatterNum * scatterDenom * scatterDenom;
    const bool cond = tolx && toly && (scatter < CLF_MAX_SCATTER);
    return (float)cond * scatter + (float)(!cond) * FLT_MAX;
}
__device__ __forceinline__ float fit_hit_to_track(float tx, float ty, float x0, float y0, float z0,
                                                  float h1_z, float x2, float y2, float z2) {
    const float dz = z2 - z0;
    const float x_prediction = x0 + tx * dz;
    const float y_prediction = y0 + ty * dz;
    const float scatterDenom = 1.f / (z2 - h1_z);
    return fit_at(x_prediction, y_prediction, scatterDenom, x2, y2);
}

// One sensor's hits: coordinates in LDS (a ring slot) or in the event blob.
struct SView {
    const float* x;
    const float* y;
    const float* z;
    int start;
    int n;
    bool sorted;
    float zmin, zmax;
    const int16_t* bstart;  // x-bucket index (LDS slot of a sorted sensor), else nullptr
    float bx0, bscale;
};

__device__ __forceinline__ int sv_lb_ge(const SView& v, int lo, int hi, float val) {
    if (!v.bstart) return lb_ge(v.x, lo, hi, val);
    return scan_to(v.x, bucket_start(v.bstart, v.bx0, v.bscale, val, lo, hi), hi,
                   [=](float a) { return a >= val; });
}
__device__ __forceinline__ int sv_ub_gt(const SView& v, int lo, int hi, float val) {
    if (!v.bstart) return ub_gt(v.x, lo, hi, val);
    return scan_to(v.x, bucket_start(v.bstart, v.bx0, v.bscale, val, lo, hi), hi,
                   [=](float a) { return a > val; });
}
__device__ __forceinline__ int sv_first_diff_gt(const SView& v, int lo, int hi, float c, float t) {
    if (!v.bstart) return first_diff_gt(v.x, lo, hi, c, t);
    return scan_to(v.x, bucket_start(v.bstart, v.bx0, v.bscale, diff_threshold(c, t), lo, hi), hi,
                   [=](float a) { return (a - c) > t; });
}
__device__ __forceinline__ int sv_first_diff_ge(const SView& v, int lo, int hi, float c, float t) {
    if (!v.bstart) return first_diff_ge(v.x, lo, hi, c, t);
    return scan_to(v.x, bucket_start(v.bstart, v.bx0, v.bscale, diff_threshold(c, t), lo, hi), hi,
                   [=](float a) { return (a - c) >= t; });
}

struct EventView {
    int S, N;
    const float* X;
    const float* Y;
    const float* Zh;
    W1Rec* hbm_w1;  // windows of sensors larger than a slot, per hit
    W2Rec* hbm_w2;
};

// L = true: the caller knows sensor s sits in its ring slot, so the pointers
// are LDS pointers the compiler can address with ds_read; otherwise the slot
// or HBM is chosen at run time (generic addressing).
template <bool L, int B, int C>
__device__ __forceinline__ SView sensor_view(Shared<B, C>& sh, const EventView& ev, int s) {
    SView v;
    v.start = uni(sh.s_start[s]);
    v.n = uni(sh.s_num[s]);
    v.sorted = uni((int)sh.s_sorted[s]) != 0;
    v.zmin = uni(sh.s_zmin[s]);
    v.zmax = uni(sh.s_zmax[s]);
    v.bstart = nullptr;
    v.bx0 = 0.f;
    v.bscale = 0.f;
    if (L || v.n <= C) {
        const int k = s % Ring<B, C>::slots;
        v.x = sh.rx[k];
        v.y = sh.ry[k];
        v.z = sh.rz[k];
        // read whether or not the sensor is sorted, so that these reads go
        // out with the ones above instead of after them
        const float bx0 = uni(sh.bk_x0[k]);
        const float bscale = uni(sh.bk_scale[k]);
        if (v.sorted) {
            v.bstart = sh.bk_start[k];
            v.bx0 = bx0;
            v.bscale = bscale;
        }
    } else {
        v.x = ev.X + v.start;
        v.y = ev.Y + v.start;
        v.z = ev.Zh + v.start;
    }
    return v;
}

template <int B, int C>
__device__ __forceinline__ SView hbm_view(Shared<B, C>& sh, const EventView& ev, int s) {
    SView v;
    v.start = uni(sh.s_start[s]);
    v.n = uni(sh.s_num[s]);
    v.sorted = uni((int)sh.s_sorted[s]) != 0;
    v.zmin = uni(sh.s_zmin[s]);
    v.zmax = uni(sh.s_zmax[s]);
    v.x = ev.X + v.start;
    v.y = ev.Y + v.start;
    v.z = ev.Zh + v.start;
    v.bstart = nullptr;
    v.bx0 = 0.f;
    v.bscale = 0.f;
    return v;
}

template <bool L, int B, int C>
__device__ __forceinline__ W1Rec* w1_of(Shared<B, C>& sh, const EventView& ev, int s) {
    return (L || uni(sh.s_num[s]) <= C) ? sh.w1[s & 1] : ev.hbm_w1 + uni(sh.s_start[s]);
}
template <bool L, int B, int C>
__device__ __forceinline__ W2Rec* w2_of(Shared<B, C>& sh, const EventView& ev, int s) {
    return (L || uni(sh.s_num[s]) <= C) ? sh.w2[s & 1] : ev.hbm_w2 + uni(sh.s_start[s]);
}

// ---------------------------------------------------------------------------
// (a) candidate windows (Kernel.cl:73-172), relative to s2.start, -1 = none
// ---------------------------------------------------------------------------
// h1 window of a hit (x0, z0) on sensor f into sensor f-2: the first index
// with |x1-x0| < 0.7|z1-z0|, then the first later index where that fails,
// else the sensor end (:115-135, :161-165).
__device__ int2 window_h1(const SView& s2, float x0, float z0) {
    int2 w = make_int2(-1, -1);
    const int n = s2.n;
    if (n == 0) return w;
    const float* X = s2.x;
    if (s2.sorted && s2.zmin == s2.zmax) {
        // one z on the sensor: the predicate is -D < fl(x1-x0) < D
        const float D = CLF_PARAM_MAXXSLOPE_CANDIDATES * fabsf(s2.zmin - z0);
        const int lo = sv_first_diff_gt(s2, 0, n, x0, -D);
        const int hi = sv_first_diff_ge(s2, lo, n, x0, D);
        if (lo < hi) w = make_int2(lo, hi);
        return w;
    }
    int i = 0;
    if (s2.sorted) {
        // no hit with fl(x1-x0) <= -Bd can pass, Bd >= every 0.7|z1-z0|
        const float a = fabsf(s2.zmin - z0), b = fabsf(s2.zmax - z0);
        const float Bd = CLF_PARAM_MAXXSLOPE_CANDIDATES * fmaxf(a, b);
        i = sv_first_diff_gt(s2, 0, n, x0, -Bd);
    }
    const float* Z = s2.z;
    for (; i < n; ++i) {
        const float dxmax = CLF_PARAM_MAXXSLOPE_CANDIDATES * fabsf(Z[i] - z0);
        if (fabsf(X[i] - x0) < dxmax) break;
    }
    if (i < n) {
        w.x = i;
        for (++i; i < n; ++i) {
            const float dxmax = CLF_PARAM_MAXXSLOPE_CANDIDATES * fabsf(Z[i] - z0);
            if (!(fabsf(X[i] - x0) < dxmax)) break;
        }
        w.y = i;
    }
    return w;
}

// h2 window of a hit (x0, z0) of sensor g acting as h1, into sensor g-2,
// from the sensor Z integers zs0 = Z[g+2], zs2 = Z[g-2] (:85-102, :137-150).
__device__ int2 window_h2(const SView& s2, float x0, float z0, int zs0, int zs2) {
    int2 w = make_int2(-1, -1);
    const float h_dist = fabsf(z0 - (float)zs0);
    const float dxmax = CLF_PARAM_MAXXSLOPE_CANDIDATES * h_dist;
    const float x0_min = x0 - dxmax;
    const float x0_max = x0 + dxmax;
    const float z2_tz = ((float)zs2 - (float)zs0) / (z0 - (float)zs0);
    float xx = x0_max + (x0 - x0_max) * z2_tz;
    const float xmin_h2 = xx - CLF_PARAM_TOLERANCE_CANDIDATES;
    xx = x0_min + (x0 - x0_min) * z2_tz;
    const float xmax_h2 = xx + CLF_PARAM_TOLERANCE_CANDIDATES;
    const int n = s2.n;
    const float* X = s2.x;
    if (n == 0) return w;
    if (s2.sorted) {
        const int first = sv_ub_gt(s2, 0, n, xmin_h2);
        if (first < n) w = make_int2(first, sv_ub_gt(s2, first + 1, n, xmax_h2));
        return w;
    }
    int i = 0;
    for (; i < n; ++i)
        if (X[i] > xmin_h2) break;
    if (i < n) {
        w.x = i;
        for (++i; i < n; ++i)
            if (X[i] > xmax_h2) break;
        w.y = i;
    }
    return w;
}

// Precomputation for a step, one item per hit:
//  items [0, n_w2):        h2 windows of sensor g2's hits into g2-2
//  items [n_w2, n_w2+n_w1): h1 windows of sensor g1's hits into g1-2
template <bool L, int B, int C>
__device__ __forceinline__ void window_item(Shared<B, C>& sh, const EventView& ev, int i, int g2,
                                            int n_w2, int g1) {
    if (i < n_w2) {
        const SView src = sensor_view<L>(sh, ev, g2);
        const SView dst = sensor_view<L>(sh, ev, g2 - 2);
        const int2 w = window_h2(dst, src.x[i], src.z[i], sh.s_Z[g2 + 2], sh.s_Z[g2 - 2]);
        W2Rec r;
        r.lo = (int16_t)w.x;
        r.hi = (int16_t)w.y;
        w2_of<L>(sh, ev, g2)[i] = r;
        if (w.x < 0 && (L || sh.s_num[g2] <= C)) {  // rare: the row stops look here
            atomicOr(&sh.w2_empty[g2 & 1][i >> 5], 1u << (i & 31));
            atomicAdd(&sh.w2_n_empty[g2 & 1], 1);
        }
    } else {
        i -= n_w2;
        const SView src = sensor_view<L>(sh, ev, g1);
        const SView dst = sensor_view<L>(sh, ev, g1 - 2);
        const int2 w = window_h1(dst, src.x[i], src.z[i]);
        W1Rec* r = w1_of<L>(sh, ev, g1) + i;
        r->lo = (int16_t)w.x;
        r->hi = (int16_t)w.y;
    }
}

// ---------------------------------------------------------------------------
// Used flags: LDS bitmask, or bytes in HBM for events above kUsedBits hits
// ---------------------------------------------------------------------------
struct UsedFlags {
    uint32_t* bits;  // LDS
    uint8_t* bytes;  // HBM
    bool lds;
    __device__ __forceinline__ bool get(int h) const {
        return lds ? ((bits[h >> 5] >> (h & 31)) & 1u) != 0 : bytes[h] != 0;
    }
    __device__ __forceinline__ void set(int h) const {
        if (lds) atomicOr(&bits[h >> 5], 1u << (h & 31));
        else bytes[h] = 1;
    }
};

// ---------------------------------------------------------------------------
// The h0 gather of one sensor step, run by wave 0 (Kernel.cl:720-754).
// Chunk c reads elements prev+lane; a lane whose element is used takes
// values from the shared counter (which restarts at 0 each step and keeps
// counting across chunks) and looks at prev + shift + value until it finds
// an unused or out-of-range element.  The r-th such lane in lane order takes
// the r-th of those elements.  Then prev = counter + shift and shift += 64.
// Visits (relative to sensor f) are appended to v_h0 in lane order.
// ---------------------------------------------------------------------------
template <int B, int C>
__device__ int gather_batch(Shared<B, C>& sh, const UsedFlags& used, int start0, uint32_t hn0,
                            bool fresh) {
    const int lane = lane_id();
    // fresh: the step's first batch, from the start of the sensor
    uint32_t prev = 0u, shift = CLF_GATHER_CHUNK, counter = 0u;
    if (!fresh) {
        prev = sh.g_prev;
        shift = sh.g_shift;
        counter = sh.g_counter;
    }
    int nv = 0;
    while (prev < hn0 && nv + kWave <= kMaxVisits) {
        const uint32_t e = prev + (uint32_t)lane;
        const bool inb = e < hn0;
        const bool u = inb && used.get(start0 + (int)e);
        int got = (inb && !u) ? (int)e : -1;
        const uint64_t need = ballot(u);
        const int n_need = __popcll(need);
        const int rank = mbcnt(need);
        int done = 0;
        uint32_t pos = counter;
        while (done < n_need) {
            const uint32_t e2 = prev + shift + pos + (uint32_t)lane;
            const bool inb2 = e2 < hn0;
            const bool un2 = inb2 && !used.get(start0 + (int)e2);
            const uint64_t sm = ballot(!inb2 || un2);
            const int c = __popcll(sm);
            const int val = un2 ? (int)e2 : -1;
            const int k = rank - done;
            const bool mine = u && k >= 0 && k < c;
            const int src = kth_set_bit(sm, mine ? k : 0);
            const int v2 = __shfl(val, src);
            if (mine) got = v2;
            if (done + c >= n_need) {
                counter = pos + (uint32_t)kth_set_bit(sm, n_need - done - 1) + 1u;
                done = n_need;
            } else {
                done += c;
                pos += kWave;
            }
        }
        const uint64_t gm = ballot(got >= 0);
        if (got >= 0) sh.v_h0[nv + mbcnt(gm)] = got;
        nv += __popcll(gm);
        prev = counter + shift;
        shift += kWave;
    }
    if (lane == 0) {
        sh.g_prev = prev;
        sh.g_shift = shift;
        sh.g_counter = counter;
        sh.g_nv = nv;
        sh.g_more = prev < hn0;
    }
    return nv;
}

// A ttf list: its first TtfCap entries in LDS buffer k, the rest in HBM at
// the same index.
struct TtfList {
    TtfEntry* hbm;
    int k;
};
// Entries move as three 16-byte words, so that an entry stays in registers.
// The empty asm statement ends the HBM branch with an instruction the LDS
// branch lacks: without it the compiler sinks the two branches' accesses
// into one flat access from a selected address, which waits on both
// counters and parks the entry in private memory.
static_assert(sizeof(TtfEntry) == 48, "a ttf entry is three 16-byte words");
__device__ __forceinline__ TtfEntry ttf_unpack(uint4 a, uint4 b, uint4 c) {
    TtfEntry e;
    e.word = a.x;
    e.nhits = a.y;
    e.t0 = a.z;
    e.t1 = a.w;
    e.t2 = b.x;
    e.x0 = __uint_as_float(b.y);
    e.y0 = __uint_as_float(b.z);
    e.z0 = __uint_as_float(b.w);
    e.x1 = __uint_as_float(c.x);
    e.y1 = __uint_as_float(c.y);
    e.z1 = __uint_as_float(c.z);
    e.pad = c.w;
    return e;
}
template <int B, int C>
__device__ __forceinline__ TtfEntry ttf_get(Shared<B, C>& sh, const TtfList& l, int i) {
    uint4 a, b, c;
    if (i < TtfCap<B, C>::value) {
        a = sh.ttf[l.k][0][i];
        b = sh.ttf[l.k][1][i];
        c = sh.ttf[l.k][2][i];
    } else {
        const uint4* p = reinterpret_cast<const uint4*>(l.hbm + i);
        a = p[0];
        b = p[1];
        c = p[2];
        __asm__ volatile("");
    }
    return ttf_unpack(a, b, c);
}
template <int B, int C>
__device__ __forceinline__ void ttf_put(Shared<B, C>& sh, const TtfList& l, int i, const TtfEntry& e) {
    const uint4 a = make_uint4(e.word, e.nhits, e.t0, e.t1);
    const uint4 b = make_uint4(e.t2, __float_as_uint(e.x0), __float_as_uint(e.y0), __float_as_uint(e.z0));
    const uint4 c = make_uint4(__float_as_uint(e.x1), __float_as_uint(e.y1), __float_as_uint(e.z1), e.pad);
    if (i < TtfCap<B, C>::value) {
        sh.ttf[l.k][0][i] = a;
        sh.ttf[l.k][1][i] = b;
        sh.ttf[l.k][2][i] = c;
    } else {
        uint4* p = reinterpret_cast<uint4*>(l.hbm + i);
        p[0] = a;
        p[1] = b;
        p[2] = c;
        __asm__ volatile("");
    }
}

struct StepCtx {
    int f;
    SView v0, v2, v4;  // sensors f, f-2, f-4
    const W1Rec* w1;   // h1 windows + row stops of sensor f
    const W2Rec* w2;   // h2 windows of sensor f-2 (relative to f-4)
    const uint32_t* w2_marks;   // their empty-window bitmask (nullptr: read w2)
    const int32_t* w2_n_empty;  // number of marks
    bool fast4;        // sensor f-4 x-sorted and the event in the supported domain
    bool z4;           // fast4, and every hit of sensor f-4 at one nonzero z (v4.zmin)
    bool seeds;        // h1 sensor f-2 >= 4 has h2 windows (Kernel.cl:58,104)
};

// ---------------------------------------------------------------------------
// (b) seeding of one batch of gathered h0 with the (64,4) semantics
// (Kernel.cl:408-567).  Row y of an h0 takes window elements 4j+y and stops
// for good at its first element whose h2 window is empty; a used h1 is
// skipped.  The winner minimises (fit, y, j, h2) lexicographically, which is
// what the rows' strict-less first minima and the lowest-y tie-break pick.
// ---------------------------------------------------------------------------
// Window, row stops and reach of visit v: row y takes window elements 4j+y
// and stops for good at the first whose h1 has no h2 window
// (Kernel.cl:436-457); without h2 windows on sensor f-4 every row stops at
// once.  The h2 windows are read eight at a time so the loads overlap.
// Returns the part of the window a live row can reach.
template <int B, int C>
__device__ __forceinline__ int visit_setup(Shared<B, C>& sh, const StepCtx& sc, int v, float s1_z) {
    const SView& v0 = sc.v0;
    const int e0 = sh.v_h0[v];
    const W1Rec r = sc.w1[e0];
    const int nh = r.lo >= 0 ? r.hi - r.lo : 0;
    int16_t d[4] = {kNoStop, kNoStop, kNoStop, kNoStop};
    if (!sc.seeds) {
        d[0] = d[1] = d[2] = d[3] = 0;
    } else if (sc.w2_marks) {
        // the empty windows of sensor f-2 are marked in a bitmask: the
        // first marked element of each row's residue class
        if (*sc.w2_n_empty > 0 && nh > 0) {
            const int a0 = r.lo, a1 = r.lo + nh;
            for (int wi = a0 >> 5; wi <= (a1 - 1) >> 5; ++wi) {
                uint32_t m = sc.w2_marks[wi];
                if (wi == (a0 >> 5)) m &= ~0u << (a0 & 31);
                if (wi == ((a1 - 1) >> 5) && (a1 & 31)) m &= (1u << (a1 & 31)) - 1u;
                while (m) {
                    const int el = wi * 32 + __builtin_ctz(m) - a0;
                    m &= m - 1u;
                    if (d[el & 3] == kNoStop) d[el & 3] = (int16_t)(el >> 2);
                }
            }
        }
    } else {
        const W2Rec* w2 = sc.w2 + r.lo;
        int found = 0;
        for (int b0 = 0; b0 < nh && found < 4; b0 += 8) {
            // unconditional loads (index clamped into the window) so
            // that all eight are in flight together
            int16_t lo[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) lo[k] = w2[min(b0 + k, nh - 1)].lo;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int el = b0 + k;
                if (el < nh && lo[k] < 0 && d[el & 3] == kNoStop) {
                    d[el & 3] = (int16_t)(el >> 2);
                    ++found;
                }
            }
        }
    }
    int reach = 0;
#pragma unroll
    for (int y = 0; y < CLF_SEED_ROWS; ++y) {
        sh.v_stop[v * CLF_SEED_ROWS + y] = d[y];
        const int lim = d[y] == kNoStop ? nh : min(nh, CLF_SEED_ROWS * (int)d[y] + y);
        reach = max(reach, lim);
    }
    sh.v_first[v] = r.lo;
    const float z0 = v0.z[e0];
    sh.v_x0[v] = v0.x[e0];
    sh.v_y0[v] = v0.y[e0];
    sh.v_z0[v] = z0;
    sh.v_dymax[v] = CLF_PARAM_MAXYSLOPE * fabsf(s1_z - z0);
    sh.v_best[v] = ~0ull;
    return reach;
}

// The setup of a gathered batch, by the wave that gathered it: each visit's
// window, row stops and reach (lane l takes visits l, l+64), and the
// exclusive prefix of the reaches, which numbers the batch's pairs.
template <int B, int C>
__device__ void seed_prepare(Shared<B, C>& sh, const StepCtx& sc, int nv) {
    const int lane = lane_id();
    const SView& v2 = sc.v2;
    // z of the first hit of sensor f-2 (:420-422)
    const float s1_z = v2.n > 0 ? v2.z[0] : 0.f;
    int running = 0;
    for (int c0 = 0; c0 < nv; c0 += kWave) {
        const int v = c0 + lane;
        const int reach = v < nv ? visit_setup<B, C>(sh, sc, v, s1_z) : 0;
        const int incl = wave_incl_scan(reach);
        if (v < nv) sh.v_pre[v] = running + incl - reach;
        running += __builtin_amdgcn_readlane(incl, kWave - 1);
    }
    if (lane == 0) sh.v_pre[nv] = running;
}

// Fits of the batch's (h0, h1) pairs (all threads), then the batch's
// tracklets as ttf entries in visit order (one wave; n_next and n_tracklets
// are advanced in that wave only).
template <int B, int C>
__device__ void seed_batch(Shared<B, C>& sh, const StepCtx& sc, const UsedFlags& used,
                           const TtfList& nxt, int nv, int& n_next, int& n_tracklets, PhaseClock& pc) {
    const int tid = threadIdx.x;
    const SView& v0 = sc.v0;
    const SView& v2 = sc.v2;
    const SView& v4 = sc.v4;
    const int P = uni(sh.v_pre[nv]);
    prof_add(sh, pc, CNT_PAIRS, (unsigned long long)P);

    // pairs (visit v, window element e), p = v_pre[v] + e; each thread walks
    // its pairs in increasing p, so its visit index only moves forward
    int v = 0;
    if (tid < P) {
        int lo = 0, hi = nv;  // last v with v_pre[v] <= tid
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sh.v_pre[mid] <= tid) lo = mid;
            else hi = mid;
        }
        v = lo;
    }
    unsigned n_h2 = 0;
    for (int p = tid; p < P; p += B) {
        while (sh.v_pre[v + 1] <= p) ++v;
        const int e = p - sh.v_pre[v];
        const int y = e & (CLF_SEED_ROWS - 1);
        const int j = e / CLF_SEED_ROWS;
        // the visit's reads, then the h1 reads, each set in flight together
        // before the tests that may skip the pair (e < the visit's reach <=
        // its window size, so h1 is a hit of sensor f-2 in any case)
        const int stop = sh.v_stop[v * CLF_SEED_ROWS + y];
        const int h1 = sh.v_first[v] + e;
        const float x0 = sh.v_x0[v], y0 = sh.v_y0[v], z0 = sh.v_z0[v], dymax = sh.v_dymax[v];
        const bool h1_used = used.get(v2.start + h1);
        const float x1 = v2.x[h1], y1 = v2.y[h1], z1 = v2.z[h1];
        const W2Rec c = sc.w2[h1];
        if (j >= stop || h1_used) continue;
        if (!(fabsf(y1 - y0) < dymax)) continue;  // :509, independent of h2
        const float dz_inverted = 1.f / (z1 - z0);
        int k = c.lo;
        const int k1 = c.hi;
        float hi = INFINITY;
        if (sc.fast4) {
            const float pa = x0 + (x1 - x0) * ((v4.zmin - z0) * dz_inverted);
            const float pb = x0 + (x1 - x0) * ((v4.zmax - z0) * dz_inverted);
            float lo;
            x_window(pa, pb, lo, hi);
            k = sv_lb_ge(v4, k, k1, lo);
        }
        float best = FLT_MAX;
        int bk = -1;
        if (sc.z4) {
            // one z on sensor f-4: the predictions are the same for every
            // candidate, and so is the scatter denominator
            const float zc = v4.zmin;
            const float z2_tz = (zc - z0) * dz_inverted;
            const float x = x0 + (x1 - x0) * z2_tz;
            const float yy = y0 + (y1 - y0) * z2_tz;
            for (; k < k1; ++k) {
                const float x2 = v4.x[k], y2 = v4.y[k];
                if (x2 > hi) break;  // sorted: nothing further can pass
                if (pc.on) ++n_h2;
                const float dx = x - x2;
                const float dy = yy - y2;
                if (fabsf(dx) < CLF_PARAM_TOLERANCE && fabsf(dy) < CLF_PARAM_TOLERANCE) {
                    const float scatterNum = (dx * dx) + (dy * dy);
                    const float scatterDenom = 1.f / (zc - z1);
                    const float scatter = scatterNum * scatterDenom * scatterDenom;
                    if (scatter < CLF_MAX_SCATTER && scatter < best) {
                        best = scatter;
                        bk = k;
                    }
                }
            }
        } else {
            for (; k < k1; ++k) {
                const float x2 = v4.x[k], y2 = v4.y[k], z2 = v4.z[k];
                if (x2 > hi) break;  // sorted: nothing further can pass
                if (pc.on) ++n_h2;
                const float z2_tz = (z2 - z0) * dz_inverted;
                const float x = x0 + (x1 - x0) * z2_tz;
                const float yy = y0 + (y1 - y0) * z2_tz;
                const float dx = x - x2;
                const float dy = yy - y2;
                if (fabsf(dx) < CLF_PARAM_TOLERANCE && fabsf(dy) < CLF_PARAM_TOLERANCE) {
                    const float scatterNum = (dx * dx) + (dy * dy);
                    const float scatterDenom = 1.f / (z2 - z1);
                    const float scatter = scatterNum * scatterDenom * scatterDenom;
                    if (scatter < CLF_MAX_SCATTER && scatter < best) {
                        best = scatter;
                        bk = k;
                    }
                }
            }
        }
        if (bk >= 0) {
            const unsigned long long key = ((unsigned long long)__float_as_uint(best + 0.0f) << 32) |
                                           ((unsigned long long)y << 30) |
                                           ((unsigned long long)j << 16) | (unsigned long long)bk;
            atomicMin(&sh.v_best[v], key);
        }
    }
    if (pc.on && n_h2) atomicAdd(&sh.prof[CNT_H2_TESTS], (unsigned long long)n_h2);
    __syncthreads();
    prof_mark(sh, pc, PROF_SEED_EVAL);

    // tracklets, as ttf entries, in visit order (:550-566)
    if (wave_id() == 0) {
        const int lane = lane_id();
        for (int c0 = 0; c0 < nv; c0 += kWave) {
            const int vv = c0 + lane;
            const bool acc = vv < nv && sh.v_best[vv] != ~0ull;
            const uint64_t m = ballot(acc);
            const int pos = mbcnt(m);
            if (acc) {
                const unsigned long long key = sh.v_best[vv];
                const int y = (int)((key >> 30) & 3ull);
                const int j = (int)((key >> 16) & 0x3fffull);
                const int k = (int)(key & 0xffffull);
                const int h1 = sh.v_first[vv] + CLF_SEED_ROWS * j + y;
                TtfEntry t;
                t.word = CLF_TTF_TRACKLET_BIT | (uint32_t)(n_tracklets + pos);
                t.nhits = 3;
                t.t0 = (uint32_t)(v0.start + sh.v_h0[vv]);
                t.t1 = (uint32_t)(v2.start + h1);
                t.t2 = (uint32_t)(v4.start + k);
                t.x0 = v2.x[h1];
                t.y0 = v2.y[h1];
                t.z0 = v2.z[h1];
                t.x1 = v4.x[k];
                t.y1 = v4.y[k];
                t.z1 = v4.z[k];
                t.pad = 0;
                ttf_put(sh, nxt, n_next + pos, t);
            }
            const int cnt = __popcll(m);
            n_tracklets += cnt;
            n_next += cnt;
        }
    }
    prof_mark(sh, pc, PROF_SEED_EMIT);
}

// Copies sensor s into its ring slot when it fits (threads first .. B-1).
template <int B, int C, int First = 0>
__device__ __forceinline__ void stage_sensor(Shared<B, C>& sh, const EventView& ev, int s) {
    constexpr int kCount = B - First;
    const int t = (int)threadIdx.x - First;
    if (s < 0 || t < 0) return;
    const int n = uni(sh.s_num[s]);
    if (n > C) return;
    const int b = uni(sh.s_start[s]), k = s % Ring<B, C>::slots;
    for (int i = t; i < n; i += kCount) {
        sh.rx[k][i] = ev.X[b + i];
        sh.ry[k][i] = ev.Y[b + i];
        sh.rz[k][i] = ev.Zh[b + i];
    }
    if (!sh.s_sorted[s]) return;
    // x-bucket index: hit i writes start[j] for the buckets j it opens
    const float x0 = n > 0 ? ev.X[b] : 0.f;
    const float span = n > 1 ? ev.X[b + n - 1] - x0 : 0.f;
    const float scale = span > 0.f ? (float)kBuckets / span : 0.f;
    if (t == 0) {
        sh.bk_x0[k] = x0;
        sh.bk_scale[k] = scale;
        if (n == 0)
            for (int j = 0; j <= kBuckets; ++j) sh.bk_start[k][j] = 0;
    }
    for (int i = t; i < n; i += kCount) {
        const int bi = bucket_of(ev.X[b + i], x0, scale);
        const int bp = i > 0 ? bucket_of(ev.X[b + i - 1], x0, scale) : -1;
        for (int j = bp + 1; j <= bi; ++j) sh.bk_start[k][j] = (int16_t)i;
        if (i == n - 1)
            for (int j = bi + 1; j <= kBuckets; ++j) sh.bk_start[k][j] = (int16_t)n;
    }
}

// The staging of one sensor split in two: stage_load reads the sensor's hits
// (and each hit's left neighbour, for the bucket index) into registers a
// step ahead, and stage_store writes them to the ring slot at the next step,
// so that the HBM latency passes while a whole sensor step runs.
template <int B, int C>
struct StagedSensor {
    static constexpr int kThreads = StageThreads<B, C>::count;
    static constexpr int kPer = (C + kThreads - 1) / kThreads;  // hits per staging thread
    float x[kPer], xl[kPer], y[kPer], z[kPer];
    float x_first, x_last;
    int s;  // the sensor, or -1 (nothing staged: absent or read from HBM)
};
// first: the first staging thread (hit r * kThreads + t on thread first + t)
template <int B, int C>
__device__ __forceinline__ void stage_load(const EventView& ev, int s, int n, int b,
                                           StagedSensor<B, C>& g, int first) {
    g.s = -1;
    const int t = (int)threadIdx.x - first;
    if (s < 0 || t < 0 || n > C) return;
    g.s = s;
#pragma unroll
    for (int r = 0; r < StagedSensor<B, C>::kPer; ++r) {
        const int i = t + r * StagedSensor<B, C>::kThreads;
        if (i < n) {
            g.x[r] = ev.X[b + i];
            g.xl[r] = i > 0 ? ev.X[b + i - 1] : 0.f;
            g.y[r] = ev.Y[b + i];
            g.z[r] = ev.Zh[b + i];
        }
    }
    g.x_first = n > 0 ? ev.X[b] : 0.f;
    g.x_last = n > 1 ? ev.X[b + n - 1] : 0.f;
}
// the same, reading sensor s's size and start (s >= 0)
template <int B, int C>
__device__ __forceinline__ void stage_load(const Shared<B, C>& sh, const EventView& ev, int s,
                                           StagedSensor<B, C>& g,
                                           int first = StageThreads<B, C>::first) {
    const int sc = max(s, 0);
    stage_load<B, C>(ev, s, uni(sh.s_num[sc]), uni(sh.s_start[sc]), g, first);
}
template <int B, int C>
__device__ __forceinline__ void stage_store(Shared<B, C>& sh, const StagedSensor<B, C>& g,
                                            int first = StageThreads<B, C>::first) {
    const int s = g.s;
    if (s < 0) return;  // also every thread outside the staging threads
    const int t = (int)threadIdx.x - first;
    const int n = uni(sh.s_num[s]);
    const int k = s % Ring<B, C>::slots;
    const bool sorted = uni((int)sh.s_sorted[s]) != 0;
    const float x0 = g.x_first;
    const float span = n > 1 ? g.x_last - x0 : 0.f;
    const float scale = span > 0.f ? (float)kBuckets / span : 0.f;
    if (sorted && t == 0) {
        sh.bk_x0[k] = x0;
        sh.bk_scale[k] = scale;
        if (n == 0)
            for (int j = 0; j <= kBuckets; ++j) sh.bk_start[k][j] = 0;
    }
#pragma unroll
    for (int r = 0; r < StagedSensor<B, C>::kPer; ++r) {
        const int i = t + r * StagedSensor<B, C>::kThreads;
        if (i < n) {
            sh.rx[k][i] = g.x[r];
            sh.ry[k][i] = g.y[r];
            sh.rz[k][i] = g.z[r];
            if (sorted) {
                // hit i writes start[j] for the buckets j it opens
                const int bi = bucket_of(g.x[r], x0, scale);
                const int bp = i > 0 ? bucket_of(g.xl[r], x0, scale) : -1;
                for (int j = bp + 1; j <= bi; ++j) sh.bk_start[k][j] = (int16_t)i;
                if (i == n - 1)
                    for (int j = bi + 1; j <= kBuckets; ++j) sh.bk_start[k][j] = (int16_t)n;
            }
        }
    }
}

// Clears the empty-window marks of w2 buffer k (all threads; a barrier must
// follow before the windows that fill them).
template <int B, int C>
__device__ __forceinline__ void clear_empty_marks(Shared<B, C>& sh, int k) {
    for (int i = threadIdx.x; i < C / 32; i += B) sh.w2_empty[k][i] = 0u;
    if (threadIdx.x == 0) sh.w2_n_empty[k] = 0;
}

// Number of window items for the step whose h1 sensor is g2 and h0 sensor g1.
template <int B, int C>
__device__ __forceinline__ int n_w2_items(const Shared<B, C>& sh, int g2) {
    return g2 >= 4 ? uni(sh.s_num[g2]) : 0;
}
template <int B, int C>
__device__ __forceinline__ int n_w1_items(const Shared<B, C>& sh, int g1) {
    return g1 >= 4 ? uni(sh.s_num[g1]) : 0;
}

// Track records in an event's slice hold hitsNum and hits[0, hitsNum)
// only; the zero padding of the 100-byte output record (Kernel
// Definitions.h:59-62) is written when the record leaves the slice, so the
// slice is neither written nor read past a track's last hit.  Word k of the
// output records whose slice copy starts at src:
constexpr uint32_t kTrackWords = (uint32_t)(sizeof(clf_track) / 4);
__device__ __forceinline__ uint32_t track_word(const uint32_t* src, uint32_t k) {
    const uint32_t r = k / kTrackWords;
    const uint32_t n = src[r * kTrackWords];  // hitsNum
    return k - r * kTrackWords <= n ? src[k] : 0u;
}

// Sensor-loop state of one event (Kernel.cl:669-767).
struct StepState {
    int n_cur, n_tracks, n_tracklets, n_weak, visits;
    int sel;        // block_scan3 buffer in turn
    TtfList cur;    // entries inserted in the previous step
    TtfList nxt;    // entries inserted in this step
};

// One sensor step f: forwarding, gather, seeding, plus the windows and row
// stops of step f-1.  L: every sensor f .. f-5 sits in the LDS ring.
template <bool L, int B, int C, bool EarlySeed>
__device__ __forceinline__ void run_step(Shared<B, C>& sh, const EventView& ev,
                                         const UsedFlags& used, clf_track* tracks, uint4* weak,
                                         bool degenerate, int f, StepState& st,
                                         const StagedSensor<B, C>& staged, int stage_first,
                                         PhaseClock& pc) {
    const int tid = threadIdx.x, wave = wave_id();
    int& n_cur = st.n_cur;
    int& n_tracks = st.n_tracks;
    int& n_tracklets = st.n_tracklets;
    int& n_weak = st.n_weak;
    int& visits = st.visits;
    TtfList& cur = st.cur;
    TtfList& nxt = st.nxt;
    // what the forwarding needs (sensor f-4) first; the views of f and f-2
    // and the window tables follow the forwarding, so that the window and
    // forwarding chains do not wait for their reads
    StepCtx sc;
    sc.f = f;
    sc.v4 = sensor_view<L>(sh, ev, f - 4);
    sc.fast4 = sc.v4.sorted && !degenerate;
    // zmin == zmax != 0 pins every z of the sensor to the same bits (the
    // range is the exact minimum and maximum; only +0 and -0 compare equal)
    sc.z4 = sc.fast4 && sc.v4.zmin == sc.v4.zmax && sc.v4.zmin != 0.f;
    sc.seeds = f - 2 >= 4;
    const SView& v4 = sc.v4;
    // the seeding's sensors and window tables
    auto seed_views = [&]() {
        sc.v0 = sensor_view<L>(sh, ev, f);
        sc.v2 = sensor_view<L>(sh, ev, f - 2);
        sc.w1 = w1_of<L>(sh, ev, f);
        sc.w2 = w2_of<L>(sh, ev, f - 2);
        const bool marks = f - 2 >= 0 && (L || sh.s_num[f - 2] <= C);
        sc.w2_marks = marks ? sh.w2_empty[(f - 2) & 1] : nullptr;
        sc.w2_n_empty = &sh.w2_n_empty[(f - 2) & 1];
    };
    // When the forwarding fits on waves 1.. (one tile, entry i on thread
    // i + 64), wave 0 gathers and sets up the step's first seeding batch
    // while the other waves forward.  The gather reads the used flags of
    // sensor f; the forwarding marks a hit there only when a tracklet that
    // skipped one or three sensors becomes a track, and then sets g_redo
    // and wave 0 gathers again after the forwarding.  Nothing else the
    // gather or the setup reads is written during the forwarding.
    const bool early = EarlySeed && B > kWave && n_cur <= B - kWave;

    // windows of the next step (h2 windows of f-3, h1 windows of f-1),
    // on the threads the forwarding leaves free (taken from the top); the
    // h1 items start at a wave boundary, so that no wave runs both kinds
    // (one that did would run the two chains one after the other)
    {
        const int g2 = f - 3, g1 = f - 1;
        const int nw2 = g2 >= 4 && f - 1 >= 4 ? n_w2_items(sh, g2) : 0;
        const int nw1 = n_w1_items(sh, g1);
        const int w2_slots = (nw2 + kWave - 1) / kWave * kWave;
        for (int j = B - 1 - tid; j < w2_slots + nw1; j += B) {
            if (j < nw2) window_item<L>(sh, ev, j, g2, nw2, g1);
            else if (j >= w2_slots) window_item<L>(sh, ev, nw2 + (j - w2_slots), g2, nw2, g1);
        }
    }

    if (early && wave == 0) {
        seed_views();
        const int nv0 = gather_batch<B, C>(sh, used, sc.v0.start, (uint32_t)sc.v0.n, true);
        if (nv0 > 0) seed_prepare<B, C>(sh, sc, nv0);
    }
    // stage_first > 0: the staged sensor goes to its slot now, from waves
    // that have no forwarding entries or window items in this step
    if (kStageInForward<B, C> && stage_first > 0) stage_store<B, C>(sh, staged, stage_first);

    // (c) forwarding of the ttf entries of step f+1 (Kernel.cl:210-367)
    int n_next = 0;
    for (int base = 0; base < n_cur; base += B) {
        // a tile that fits on waves 1.. leaves wave 0 out (entry i on thread
        // i + 64, the scan order unchanged), so that wave 0 goes from the
        // scan straight to the gather, with no entries of its own to write
        const int shift = (B > kWave && n_cur - base <= B - kWave) ? kWave : 0;
        const int i = base + tid - shift;
        const bool active = i >= base && i < n_cur;
        TtfEntry t = {};
        bool found = false, f_new = false, f_req = false, f_weak = false;
        int bk = -1;
        if (active) {
            t = ttf_get(sh, cur, i);
            const bool is_tl = (t.word & CLF_TTF_TRACKLET_BIT) != 0;
            const uint32_t skipped = (t.word & CLF_TTF_SKIP_MASK) >> CLF_TTF_SKIP_SHIFT;
            const float td = 1.0f / (t.z1 - t.z0);
            const float txn = (t.x1 - t.x0);
            const float tyn = (t.y1 - t.y0);
            const float tx = txn * td;
            const float ty = tyn * td;
            int k = 0;
            float hi = INFINITY;
            if (sc.fast4 && v4.n > 0) {
                const float pa = t.x0 + tx * (v4.zmin - t.z0);
                const float pb = t.x0 + tx * (v4.zmax - t.z0);
                float lo;
                x_window(pa, pb, lo, hi);
                k = sv_lb_ge(v4, 0, v4.n, lo);
            }
            float best_fit = FLT_MAX;
            if (sc.z4) {
                // one z on sensor f-4: the predictions and the scatter
                // denominator are the same for every candidate
                const float zc = v4.zmin;
                const float dz = zc - t.z0;
                const float xp = t.x0 + tx * dz;
                const float yp = t.y0 + ty * dz;
                const float sd = 1.f / (zc - t.z1);
                for (; k < v4.n; ++k) {
                    const float x2 = v4.x[k], y2 = v4.y[k];
                    if (x2 > hi) break;  // sorted: nothing further can pass
                    const float fit = fit_at(xp, yp, sd, x2, y2);
                    const bool better = fit < best_fit;
                    best_fit = (float)better * fit + (float)(!better) * best_fit;
                    bk = better ? k : bk;
                }
            } else {
                for (; k < v4.n; ++k) {
                    const float x2 = v4.x[k], y2 = v4.y[k], z2 = v4.z[k];
                    if (x2 > hi) break;  // sorted: nothing further can pass
                    const float fit = fit_hit_to_track(tx, ty, t.x0, t.y0, t.z0, t.z1, x2, y2, z2);
                    const bool better = fit < best_fit;
                    best_fit = (float)better * fit + (float)(!better) * best_fit;
                    bk = better ? k : bk;
                }
            }
            if (best_fit != FLT_MAX) {
                if (bk >= 0) {
                    found = true;
                    f_new = is_tl;
                    // re-queued (Kernel.cl:343-345), unless the record is
                    // full: its 25th hit would be appended past the record
                    // (:321); the hit is still marked used (:316) below
                    f_req = t.nhits < CLF_MAX_TRACK_SIZE;
                } else {
                    // NaN best with no hit recorded: only in the
                    // degenerate domain; the track is dropped
                    atomicOr(&sh.flags, CLF_FLAG_DEGENERATE_Z);
                }
            } else if (skipped <= CLF_MAX_SKIPPED_MODULES) {
                f_req = true;
            } else if (t.nhits == 3) {
                f_weak = true;
            }
        }
        // used flags before the scan's barrier: the gather and the seeding
        // of this step read them after it (a tracklet that skipped sensors
        // marks hits of sensors f and f-2 when it becomes a track)
        if (found) {
            used.set(v4.start + bk);
            if (f_new) {
                used.set((int)t.t0);
                used.set((int)t.t1);
                used.set((int)t.t2);
                // sensor f's range read here, not held across the loop
                const uint32_t fs = (uint32_t)sh.s_start[f], fn = (uint32_t)sh.s_num[f];
                if (t.t0 - fs < fn || t.t1 - fs < fn || t.t2 - fs < fn) sh.g_redo = 1;
            }
        }
        int pos[3], tot[3];
        block_scan3<B, C>(sh, f_new, f_req, f_weak, pos, tot, st.sel);
        // the staged sensor f-6 goes to its ring slot (that of f+1, which
        // no part of this step reads) before the last tile's global stores:
        // the vector-memory counter retires loads and stores in issue
        // order, so the wait for the staged loads at the store would
        // otherwise wait for those stores too
        if (kStageInForward<B, C> && stage_first == 0 && base + B >= n_cur) stage_store<B, C>(sh, staged);
        if (active) {
            if (found && !f_req) {
                // a full record found a 25th hit: the record stays as stored
                atomicOr(&sh.flags, CLF_FLAG_TRACK_FULL);
                atomicAdd(&sh.n_overflow, 1);
            } else if (found) {
                const int b = v4.start + bk;
                uint32_t slot;
                if (f_new) {
                    slot = (uint32_t)(n_tracks + pos[0]);
                    uint32_t* r = reinterpret_cast<uint32_t*>(tracks + slot);
                    r[0] = 4;
                    r[1] = t.t0;
                    r[2] = t.t1;
                    r[3] = t.t2;
                    r[4] = (uint32_t)b;
                } else {
                    slot = t.word & CLF_TTF_INDEX_MASK;
                    tracks[slot].hits[t.nhits] = (uint32_t)b;
                    tracks[slot].hitsNum = t.nhits + 1;
                }
                if (t.nhits + 1 == CLF_MAX_TRACK_SIZE) atomicAdd(&sh.n_full, 1);
                TtfEntry n;
                n.word = slot;
                n.nhits = t.nhits + 1;
                n.t0 = n.t1 = n.t2 = 0;
                n.x0 = t.x1;
                n.y0 = t.y1;
                n.z0 = t.z1;
                n.x1 = v4.x[bk];
                n.y1 = v4.y[bk];
                n.z1 = v4.z[bk];
                n.pad = 0;
                ttf_put(sh, nxt, n_next + pos[1], n);
            } else if (f_req) {
                const uint32_t skipped = (t.word & CLF_TTF_SKIP_MASK) >> CLF_TTF_SKIP_SHIFT;
                TtfEntry n = t;
                n.word = ((skipped + 1) << CLF_TTF_SKIP_SHIFT) | (t.word & CLF_TTF_KEEP_MASK);
                ttf_put(sh, nxt, n_next + pos[1], n);
            } else if (f_weak) {
                weak[n_weak + pos[2]] = make_uint4(t.t0, t.t1, t.t2, 0u);
            }
        }
        n_tracks += tot[0];
        n_next += tot[1];
        n_weak += tot[2];
    }
    seed_views();
    // No barrier here: every used flag of the forwarding was set before the
    // last block scan's barrier, and what the forwarding writes after it
    // (tracks, ttf entries of the next list, weak) is not read by the gather
    // or the seeding; the barrier after wave 0's gather orders the rest.
    prof_mark(sh, pc, PROF_FORWARD);

    // (b) gather + seeding (Kernel.cl:720-764)
    const int start0 = sc.v0.start;
    const uint32_t hn0 = (uint32_t)sc.v0.n;
    // Wave 0 gathers a batch and prepares it; after one barrier all waves
    // fit its pairs; after another, wave 0 emits its tracklets and goes on to
    // the next batch while the other waves wait at the next barrier.  Only
    // wave 0 reads what the emit reads, so no barrier follows it: the other
    // waves of the last batch leave for the next step, whose first barrier
    // waits for wave 0.
    bool first_batch = true;
    for (;;) {
        if (wave == 0) {
            // the first batch stands if it was gathered during the
            // forwarding and no flag of sensor f changed meanwhile (g_redo
            // is set before the forwarding scan's barrier)
            if (!(first_batch && early && uni(sh.g_redo) == 0)) {
                const int nv0 = gather_batch<B, C>(sh, used, start0, hn0, first_batch);
                prof_mark(sh, pc, PROF_GATHER);
                if (nv0 > 0) seed_prepare<B, C>(sh, sc, nv0);
            }
        } else if (StageThreads<B, C>::during_gather && first_batch) {
            // sensor f-6, for the next step: from registers loaded at the
            // top of the step (stored in the forwarding when it has entries),
            // or (1024 threads) straight from HBM
            if (!kStageInForward<B, C>) stage_sensor<B, C, StageThreads<B, C>::first>(sh, ev, f - 6);
            else if (n_cur == 0 && stage_first == 0) stage_store<B, C>(sh, staged);
        }
        first_batch = false;
        // what crosses this barrier is in LDS (the batch's visit tables, the
        // staged ring slot, g_nv and g_more); the forwarding's global stores
        // (tracks, weak and ttf entries beyond the lists' LDS parts) are read
        // or rewritten only after the pair fits' full barrier or the next
        // step's, which wait for them, and the used flags in HBM were set
        // before the forwarding scan's barrier
        lds_barrier();
        prof_mark(sh, pc, PROF_SEED_SETUP);
        const int nv = uni(sh.g_nv);
        const bool more = uni(sh.g_more) != 0;
        visits += nv;
        prof_add(sh, pc, CNT_VISITS, (unsigned long long)nv);
        prof_add(sh, pc, CNT_BATCHES, 1ull);
        if (nv > 0) seed_batch<B, C>(sh, sc, used, nxt, nv, n_next, n_tracklets, pc);
        if (!more) break;
    }

    // span of the reference's ttf ring during this step, and the length of
    // the next step's list (exact in wave 0, which emitted the tracklets)
    if (tid == 0) {
        const int live = n_cur + n_next;
        if (live > sh.max_live) sh.max_live = live;
        sh.n_listed = n_next;
    }
    const TtfList tmp = cur;
    cur = nxt;
    nxt = tmp;
    n_cur = n_next;
}

// ---------------------------------------------------------------------------
// One event
// ---------------------------------------------------------------------------
// First record of the work slices of the event the workgroup runs: its
// workgroup's slot, reused event after event, or its own slice.
__device__ __forceinline__ int64_t work_offset(const SearchArgs& a, int64_t event_off) {
    return a.slot_records > 0 ? (int64_t)(a.slot_base + (int)blockIdx.x) * a.slot_records
                              : uni(event_off);
}

// EarlySeed: wave 0 gathers a step's first seeding batch during the
// forwarding, and idle waves store the staged sensor before it (the batch
// kernel).  The serving kernel does neither: it holds its worker state
// across the search, and with the early gather its spill space grew from
// 144 to 672 bytes per lane and each search took 422 us instead of 329 us
// on the serve256 events (the staging by idle waves, 192 bytes and 334 us).
template <int B, int C, bool EarlySeed>
__device__ __forceinline__ void process_event(const SearchArgs& a, int e, Shared<B, C>& sh) {
    constexpr int W = Shared<B, C>::W;
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    // the descriptor, in scalar registers: the sensor loop below then runs
    // as a uniform loop with scalar sensor indices
    const EventDesc dv = a.ev[e];
    EventDesc d;
    d.blob_off = uni(dv.blob_off);
    d.hit_off = uni(dv.hit_off);
    d.track_off = work_offset(a, dv.track_off);
    d.n_hits = uni(dv.n_hits);
    d.n_sensors = uni(dv.n_sensors);
    EventResult* res = a.results + e;
    PhaseClock pc;
    pc.on = a.prof != nullptr;
    pc.t = 0;
    if (pc.on && tid == 0) {
        for (int k = 0; k < kProfSlots; ++k) sh.prof[k] = 0;
        pc.t = wall_clock64();
    }
    const unsigned long long t_start = pc.t;
    if (d.n_hits < 0) {
        if (tid == 0) {
            EventResult r = {};
            r.flags = CLF_FLAG_INVALID;
            *res = r;
            if (a.track_base) a.track_base[e] = 0;
        }
        return;
    }
    const int S = d.n_sensors, N = d.n_hits;
    const int32_t* hdr = (const int32_t*)(a.blobs + d.blob_off);
    const int32_t* gZ = hdr + 2;
    const int32_t* gStart = gZ + S;
    const int32_t* gNum = gStart + S;
    EventView ev;
    ev.S = S;
    ev.N = N;
    ev.X = (const float*)(gNum + S + N);
    ev.Y = ev.X + N;
    ev.Zh = ev.Y + N;
    ev.hbm_w1 = reinterpret_cast<W1Rec*>(a.cand + d.hit_off);
    ev.hbm_w2 = reinterpret_cast<W2Rec*>(ev.hbm_w1 + N);
    clf_track* tracks = a.tracks + d.track_off;
    TtfList cur = {a.ttf_a + d.track_off, 0};
    TtfList nxt = {a.ttf_b + d.track_off, 1};
    uint4* weak = a.weak + d.track_off;
    UsedFlags used;
    used.lds = N <= kUsedBits;
    used.bits = sh.used_bits;
    used.bytes = a.used + d.hit_off;

    for (int s = tid; s < S; s += B) {
        sh.s_start[s] = gStart[s];
        sh.s_num[s] = gNum[s];
        sh.s_Z[s] = gZ[s];
    }
    if (used.lds) {
        for (int i = tid; i < (N + 31) / 32; i += B) sh.used_bits[i] = 0u;
    } else {
        for (int i = tid; i < N; i += B) used.bytes[i] = 0;
    }
    if (tid == 0) {
        sh.flags = 0;
        sh.max_live = 0;
        sh.n_full = 0;
        sh.n_overflow = 0;
        sh.n_listed = 0;
    }
    __syncthreads();

    // per-sensor z range, x order and domain checks, one wave per sensor
    for (int s = wave; s < S; s += W) {
        const int b = sh.s_start[s], n = sh.s_num[s];
        float mn = INFINITY, mx = -INFINITY;
        bool unsorted = false, bad = false;
        for (int i = lane; i < n; i += kWave) {
            const float x = ev.X[b + i], y = ev.Y[b + i], z = ev.Zh[b + i];
            bad |= !(fabsf(x) <= kLim) || !(fabsf(y) <= kLim) || !(fabsf(z) <= kLim);
            mn = fminf(mn, z);
            mx = fmaxf(mx, z);
            if (i + 1 < n && !(x <= ev.X[b + i + 1])) unsorted = true;
        }
        wave_min_max(mn, mx);
        const bool any_unsorted = ballot(unsorted) != 0;
        const bool any_bad = ballot(bad) != 0 || !(fabsf((float)sh.s_Z[s]) <= kLim);
        if (lane == 0) {
            sh.s_zmin[s] = mn;
            sh.s_zmax[s] = mx;
            sh.s_sorted[s] = !any_unsorted;
            uint32_t fl = 0;
            if (any_unsorted) fl |= CLF_FLAG_UNSORTED;
            if (any_bad) fl |= CLF_FLAG_DEGENERATE_Z;
            if (fl) atomicOr(&sh.flags, fl);
        }
    }
    __syncthreads();
    for (int p = tid; p < S * S; p += B) {
        const int s = p / S, t = p - s * S;
        if (s < t && sh.s_num[s] > 0 && sh.s_num[t] > 0 &&
            !(sh.s_zmax[s] < sh.s_zmin[t] || sh.s_zmax[t] < sh.s_zmin[s]))
            atomicOr(&sh.flags, CLF_FLAG_DEGENERATE_Z);
    }
    for (int f = 4 + tid; f <= S - 3; f += B) {
        const float zs0 = (float)sh.s_Z[f + 2];
        if (sh.s_num[f] > 0 && zs0 >= sh.s_zmin[f] && zs0 <= sh.s_zmax[f])
            atomicOr(&sh.flags, CLF_FLAG_DEGENERATE_Z);
    }
    // the ring starts with sensors S-1 .. S-5 (S-6 too when the staging runs
    // during the step, one step earlier)
    constexpr int kEarly = StageThreads<B, C>::during_gather ? 1 : 0;
    for (int s = S - 1; s >= S - 5 - kEarly && s >= 0; --s) stage_sensor<B, C>(sh, ev, s);
    clear_empty_marks<B, C>(sh, 0);
    clear_empty_marks<B, C>(sh, 1);
    __syncthreads();
    const bool degenerate = (sh.flags & CLF_FLAG_DEGENERATE_Z) != 0;

    if (a.fill_only) {
        // stage (a) alone: every hit's windows in the reference's
        // absolute-index form (dev_hit_candidates / dev_hit_h2_candidates).
        // fill_only 1: binary searches over the event blob in HBM;
        // fill_only 2: the search's own path, both sensors of a window
        // staged into their ring slots with their x-bucket index by the
        // staging code of the configuration (sensors above C hits stay in
        // HBM), then the same window searches.
        int4* cand = a.cand + d.hit_off;
        for (int i = tid; i < N; i += B) cand[i] = make_int4(-1, -1, -1, -1);
        __syncthreads();
        for (int f = 4; f < S; ++f) {
            SView v2, v0;
            if (a.fill_only == 2) {
                __syncthreads();  // the ring slots' previous readers are done
                if constexpr (B <= 512) {
                    StagedSensor<B, C> g;
                    stage_load<B, C>(sh, ev, f, g);
                    stage_store<B, C>(sh, g);
                    stage_load<B, C>(sh, ev, f - 2, g);
                    stage_store<B, C>(sh, g);
                } else {
                    stage_sensor<B, C>(sh, ev, f);
                    stage_sensor<B, C>(sh, ev, f - 2);
                }
                __syncthreads();
                v2 = sensor_view<false>(sh, ev, f - 2);
                v0 = sensor_view<false>(sh, ev, f);
            } else {
                v2 = hbm_view(sh, ev, f - 2);
                v0 = hbm_view(sh, ev, f);
            }
            const bool p2 = f <= S - 3;
            for (int i = tid; i < v0.n; i += B) {
                const float x0 = v0.x[i], z0 = v0.z[i];
                const int2 w1 = window_h1(v2, x0, z0);
                int2 w2 = make_int2(-1, -1);
                if (p2) w2 = window_h2(v2, x0, z0, sh.s_Z[f + 2], sh.s_Z[f - 2]);
                cand[v0.start + i] =
                    make_int4(w1.x < 0 ? -1 : v2.start + w1.x, w1.x < 0 ? -1 : v2.start + w1.y,
                              w2.x < 0 ? -1 : v2.start + w2.x, w2.x < 0 ? -1 : v2.start + w2.y);
            }
        }
        if (tid == 0) {
            EventResult r = {};
            r.flags = sh.flags;
            *res = r;
        }
        return;
    }

    // windows and row stops for the first step (f = S-1): h2 windows of
    // sensor S-3, h1 windows of sensor S-1, then the row stops of S-1
    if (S - 1 >= 4) {
        const int nw2 = n_w2_items(sh, S - 3), nw1 = n_w1_items(sh, S - 1);
        bool lds = true;
        for (int s2 = S - 1; s2 >= S - 5 && s2 >= 0; --s2) lds = lds && uni(sh.s_num[s2]) <= C;
        const int w2_slots = (nw2 + kWave - 1) / kWave * kWave;  // h1 items from a wave boundary
        for (int j = tid; j < w2_slots + nw1; j += B) {
            if (j >= nw2 && j < w2_slots) continue;
            const int i = j < nw2 ? j : nw2 + (j - w2_slots);
            if (lds) window_item<true>(sh, ev, i, S - 3, nw2, S - 1);
            else window_item<false>(sh, ev, i, S - 3, nw2, S - 1);
        }
        __syncthreads();
    }
    prof_mark(sh, pc, PROF_PROLOGUE);

    int n_cur = 0, n_tracks = 0, n_tracklets = 0, n_weak = 0, visits = 0;
    int sel = 0;
    // staging one step ahead holds a sensor's hits in registers across the
    // step, which a 1024-thread workgroup (128 registers) pays in spills
    constexpr bool kAhead = B <= 512;
    StagedSensor<B, C> staged;
    staged.s = -1;
    if (kAhead && !kEarly) stage_load<B, C>(sh, ev, S - 6, staged);
    // f - 1 through readfirstlane: the compiler may copy the decrement into
    // the arms of a lane-divergent branch in the step, and their merge would
    // otherwise make the sensor loop, and every sensor index, per-lane
    for (int f = S - 1; f >= 4; f = uni(f - 1)) {
        if (!kEarly) {  // sensor f-5
            if (kAhead) stage_store<B, C>(sh, staged);
            else stage_sensor<B, C>(sh, ev, f - 5);
        }
        clear_empty_marks<B, C>(sh, (f - 3) & 1);  // filled by this step's windows of f-3
        if (tid == 0) sh.g_redo = 0;
        prof_mark(sh, pc, PROF_STAGE);
        __syncthreads();
        // what the step's set-up reads, issued together (each uni() waits
        // for its read): the list length, the sizes of sensors f .. f-5,
        // and sensor f-6's size and start
        n_cur = uni(sh.n_listed);
        int s_n[kStepSensors];
#pragma unroll
        for (int d = 0; d < kStepSensors; ++d) s_n[d] = uni(sh.s_num[max(f - d, 0)]);
        const int n6 = uni(sh.s_num[max(f - 6, 0)]), b6 = uni(sh.s_start[max(f - 6, 0)]);
        // sensor f-6 for the next step, stored at the top of the next step
        // into the slot of sensor f, or during this step into that of f+1
        // (seven slots): by waves that neither forward nor compute windows
        // in this step, before the forwarding, when enough of them are
        // free (one hit each); else by waves 1.. in the forwarding's last
        // tile
        int stage_first = 0;  // > 0: the staging threads start there and store before the forwarding
        if (kStageInForward<B, C>) {
            if (EarlySeed && f - 6 >= 0 && n_cur <= B - kWave) {
                // the step's window items: h2 windows of f-3, h1 windows of f-1
                const int nw2 = f - 3 >= 4 && f - 1 >= 4 ? s_n[3] : 0;
                const int items = (nw2 + kWave - 1) / kWave * kWave + (f - 1 >= 4 ? s_n[1] : 0);
                const int lo = kWave + (n_cur + kWave - 1) / kWave * kWave;
                const int hi = B - (items + kWave - 1) / kWave * kWave;
                // at least one thread: thread `first` writes the slot's
                // bucket scales (and, for an empty sensor, its starts)
                if (hi - lo >= max(n6, 1)) stage_first = lo;
            }
            stage_load<B, C>(ev, f - 6, n6, b6, staged,
                             stage_first > 0 ? stage_first : StageThreads<B, C>::first);
        } else if (kAhead) {
            stage_load<B, C>(ev, f - 6, n6, b6, staged, StageThreads<B, C>::first);
        }
        prof_mark(sh, pc, PROF_STAGE_WAIT);
        prof_add(sh, pc, CNT_FWD_ENTRIES, (unsigned long long)n_cur);
        prof_add(sh, pc, CNT_FWD_TILES, (unsigned long long)((n_cur + B - 1) / B));
        StepState st = {n_cur, n_tracks, n_tracklets, n_weak, visits, sel, cur, nxt};
        // every sensor f .. f-5 (down to 0) in its ring slot: six independent
        // reads (a short-circuit chain here made six LDS round trips a step)
        bool lds_step = true;
#pragma unroll
        for (int d = 0; d < kStepSensors; ++d) lds_step &= s_n[d] <= C;
        if (lds_step)
            run_step<true, B, C, EarlySeed>(sh, ev, used, tracks, weak, degenerate, f, st, staged, stage_first, pc);
        else
            run_step<false, B, C, EarlySeed>(sh, ev, used, tracks, weak, degenerate, f, st, staged, stage_first, pc);
        n_cur = st.n_cur;
        n_tracks = st.n_tracks;
        n_tracklets = st.n_tracklets;
        n_weak = st.n_weak;
        visits = st.visits;
        sel = st.sel;
        cur = st.cur;
        nxt = st.nxt;
    }
    __syncthreads();
    n_cur = uni(sh.n_listed);  // wave 0 may have emitted after the others left

    // tail: tracklet entries of the last step become weak (Kernel.cl:771-792)
    for (int base = 0; base < n_cur; base += B) {
        const int i = base + tid;
        TtfEntry t = {};
        if (i < n_cur) t = ttf_get(sh, cur, i);
        const bool tl = i < n_cur && (t.word & CLF_TTF_TRACKLET_BIT) != 0;
        int pos[3], tot[3];
        block_scan3<B, C>(sh, tl, false, false, pos, tot, sel);
        if (tl) weak[n_weak + pos[0]] = make_uint4(t.t0, t.t1, t.t2, 0u);
        n_weak += tot[0];
    }
    __syncthreads();

    // (d) weak tracklets made of three unused hits become tracks (:796-815)
    for (int base = 0; base < n_weak; base += B) {
        const int i = base + tid;
        uint4 tl = make_uint4(0, 0, 0, 0);
        bool ok = false;
        if (i < n_weak) {
            tl = weak[i];
            ok = !used.get((int)tl.x) && !used.get((int)tl.y) && !used.get((int)tl.z);
        }
        int pos[3], tot[3];
        block_scan3<B, C>(sh, ok, false, false, pos, tot, sel);
        if (ok) {
            uint32_t* r = reinterpret_cast<uint32_t*>(tracks + n_tracks + pos[0]);
            r[0] = 3;
            r[1] = tl.x;
            r[2] = tl.y;
            r[3] = tl.z;
        }
        n_tracks += tot[0];
    }

    if (tid == 0) {
        uint32_t fl = sh.flags;
        if (n_tracks > CLF_MAX_TRACKS) fl |= CLF_FLAG_TRACKS_OVER;
        if (n_tracklets > N) fl |= CLF_FLAG_TRACKLETS_OVER;
        if (n_weak > N) fl |= CLF_FLAG_WEAK_OVER;
        if (sh.max_live > CLF_TTF_MODULO) fl |= CLF_FLAG_TTF_OVER;
        EventResult r;
        r.n_tracks = n_tracks;
        r.flags = fl;
        r.n_tracklets = n_tracklets;
        r.n_weak = n_weak;
        r.max_ttf_live = sh.max_live;
        r.n_full = sh.n_full;
        r.h0_visits = visits;
        r.n_overflow = sh.n_overflow;
        *res = r;
    }
    if (a.packed) {
        // the event's tracks into the packed output, at records reserved
        // with one atomic: written where they are read back, with no pass
        // of their own over the batch.  What the copy needs is re-read here
        // (LDS, the descriptor) rather than held across the sensor loop.
        if (tid == 0) {
            const unsigned long long b = atomicAdd(a.track_alloc, (unsigned long long)n_tracks);
            a.track_base[e] = (int64_t)b;
            sh.pack_base = (int64_t)b;
            sh.item = n_tracks;
        }
        __syncthreads();  // also orders the weak filter's track stores before the reads
        const int64_t base = uni(sh.pack_base);
        const int nt = uni(sh.item);
        const int64_t toff = work_offset(a, a.ev[e].track_off);
        // consecutive lanes write consecutive words; the zero padding is
        // made here (track_word)
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.tracks + toff);
        uint32_t* dst = reinterpret_cast<uint32_t*>(a.packed + base);
        const uint32_t words = (uint32_t)nt * kTrackWords;
        for (uint32_t k = tid; k < words; k += B) dst[k] = track_word(src, k);
    }
    prof_mark(sh, pc, PROF_TAIL);
    if (pc.on && tid == 0) {
        sh.prof[PROF_TOTAL] = pc.t - t_start;
        for (int k = 0; k < kProfSlots; ++k) a.prof[(size_t)e * kProfSlots + k] = sh.prof[k];
    }
}

// The end of a run, by every workgroup after its last dequeue: the one
// that finishes last (the count of finished workgroups reaches done_total
// over all of the run's launches) stores the number of packed records and
// clears the allocator, the run's queue counters and the count, so that
// the next run, or the next replay of a graph holding this launch alone,
// starts from zero.  Each workgroup's allocations and dequeues precede its
// count (fenced), so the last one sees the allocator's final value.
template <int B, int C>
__device__ __forceinline__ void run_end(const SearchArgs& a, Shared<B, C>& sh) {
    if (!a.done || a.done_total <= 0) return;  // no end step: the caller clears
    if (threadIdx.x == 0) {
        __threadfence();
        sh.item = atomicAdd(a.done, 1u) + 1u == (unsigned int)a.done_total ? 1 : 0;
    }
    __syncthreads();
    if (uni(sh.item) == 0) return;
    __threadfence();
    for (int i = threadIdx.x; i < a.reset_n; i += B) a.reset_queue[i] = 0;
    if (threadIdx.x == 0) {
        if (a.track_alloc) {
            *a.track_total = (int64_t)atomicAdd(a.track_alloc, 0ull);
            *a.track_alloc = 0ull;
        }
        *a.done = 0u;
    }
}

// Waves per SIMD the register allocation is held to (512 registers per
// lane / waves).  An event's search time is set by its waves' latency
// chains, so a CU runs more events at once, and a heavy event finishes
// sooner in a wider workgroup, when more waves fit: four per SIMD (128
// registers) lets a CU hold four 256-thread or two 512-thread workgroups,
// which their LDS also allows with 256-hit slots, and two 512-thread
// workgroups with 512-hit slots.
template <int B, int C>
struct MinWaves {
    static constexpr int value = ((B >= 256 && C == 256) || (B == 512 && C == 512)) ? 4 : 1;
};

template <int B, int C>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(MinWaves<B, C>::value)))
void search_kernel(SearchArgs a) {
    __shared__ Shared<B, C> sh;
    const int grain = uni(a.grain);
    for (;;) {
        // `grain` consecutive items of the cost-ordered queue per dequeue
        if (threadIdx.x == 0) sh.item = atomicAdd(a.queue_head, grain);
        __syncthreads();
        const int item0 = uni(sh.item);
        __syncthreads();
        if (item0 >= a.n_items) break;  // every wave of the group leaves together
        const int end = min(item0 + grain, a.n_items);
        for (int item = item0; item < end; item = uni(item + 1)) {
            const int idx = a.first_item + item;
            const int e = uni(a.order ? a.order[idx] : idx);
            process_event<B, C, true>(a, e, sh);
            __syncthreads();
        }
    }
    run_end(a, sh);
}

// ---------------------------------------------------------------------------
// Resident serving kernel
// ---------------------------------------------------------------------------
// Workgroup 0 runs one wave, the feeder; every other workgroup is a worker.
// A worker announces that it is idle by appending an entry to the idle ring
// (one fetch-add, never retried) and then watches its own mailbox.  The
// feeder follows the host's descriptor ring in pinned memory (a descriptor
// is published once its event is in the device input arena) and pairs
// published descriptors with idle workers in order, up to 64 per round trip:
// it claims each worker's idle entry with a compare-and-swap and writes the
// descriptor number into the claimed workers' mailboxes.  A worker reads its
// descriptor, copies the event from the device input arena into its own
// scratch, runs process_event, reserves room for its tracks in the device
// output ring (one atomic on the ring's allocator), writes them there and
// appends a completion record to a ring in pinned memory; the host copies
// the finished part of the output ring back with the copy engine.
//
// Termination: the feeder leaves when the host sets its stop word, when
// every published descriptor has been handed out and nothing has moved for
// idle_ticks, or when every worker has left (CTL_LEFT); it then stores
// CTL_HEAD, from which the next launch resumes, and sets CTL_QUIT.  A
// waiting worker leaves on CTL_QUIT (after one more look at its mailbox: the
// feeder writes every hand-out before CTL_QUIT), or when the feeder's
// heartbeat has been silent for 4 x idle_ticks (a feeder wave held off the
// chip): then it withdraws its idle entry with a compare-and-swap first, and
// if the feeder has claimed the entry already (the swap fails), it stays for
// the hand-out, which the feeder writes right after the claim.  A feeder
// that comes back to withdrawn entries passes them over; once no worker is
// left it leaves with descriptors still waiting, and the host, which sees
// the kernel end with work outstanding, launches it again.  Every wait is
// bounded once the host stops publishing.

// Polls are relaxed (no cache invalidation per poll); the one acquire per
// event is the worker's check of its descriptor's sequence word.
__device__ __forceinline__ uint32_t poll_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t poll_sys(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// index of the first clear bit of m (64 if none)
__device__ __forceinline__ int first_clear(uint64_t m) {
    return ~m == 0ull ? kWave : __builtin_ctzll(~m);
}
// idle-ring entry of position pos: (pos + 1) << 32 | worker, or, once the
// feeder has claimed it or the worker has withdrawn it, one of these
constexpr uint32_t kIdleClaimed = 0xfffffffeu, kIdleGone = 0xffffffffu;
__device__ __forceinline__ uint64_t idle_entry(uint32_t pos, uint32_t who) {
    return ((uint64_t)(pos + 1u) << 32) | (uint64_t)who;
}

// The feeder wave.  While descriptors keep coming it checks the next 64
// sequence words at once (one round trip over the host link); otherwise
// only the next one.
__device__ void serve_feeder(const ServeArgs& a) {
    const int lane = lane_id();
    uint32_t* ctl = a.ctl;
    uint64_t* beat = reinterpret_cast<uint64_t*>(ctl + CTL_BEAT);
    uint32_t t = poll_agent(&ctl[CTL_TAIL]);
    uint32_t h = poll_agent(&ctl[CTL_HEAD]);
    uint32_t ih = 0;  // idle ring read position (the ring is cleared per launch)
    uint64_t last = wall_clock64();
    uint64_t pause = a.pause_ticks;
    const uint32_t n_workers = gridDim.x - 1u;
    bool wide = false;
    for (;;) {
        const uint64_t now = wall_clock64();
        if (lane == 0) __hip_atomic_store(beat, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool moved = false;
        {
            const uint32_t k = wide ? (uint32_t)lane : 0u;
            const bool ok = poll_sys(&a.ring[(t + k) & a.ring_mask].seq) == t + k + 1u;
            const uint64_t m = ballot(ok);
            const int n = wide ? first_clear(m) : (m != 0ull ? 1 : 0);
            wide = n > 0;
            if (n > 0) {
                t += (uint32_t)n;
                moved = true;
            }
        }
        if (t != h) {
            if (pause != 0) {
                // testing: one bounded stall with no heartbeat before the
                // first hand-out, as if this wave had been held off the chip
                const uint64_t p0 = wall_clock64();
                while (wall_clock64() - p0 < pause) __builtin_amdgcn_s_sleep(16);
                pause = 0;
            }
            // the written prefix of the idle ring from ih, at most one entry
            // per waiting descriptor; each lane claims its entry, and the
            // claimed ones take the descriptors in lane order (an entry its
            // worker has withdrawn is passed over)
            uint64_t* slot = &a.idle_ring[(ih + (uint32_t)lane) & a.idle_mask];
            const uint64_t e = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t pos = ih + (uint32_t)lane;
            const bool written = (uint32_t)(e >> 32) == pos + 1u;
            const int nw = min(first_clear(ballot(written)), (int)(t - h));
            bool claimed = false;
            if (lane < nw && (uint32_t)e < kIdleClaimed) {
                uint64_t expect = e;
                claimed = __hip_atomic_compare_exchange_strong(
                    slot, &expect, idle_entry(pos, kIdleClaimed), __ATOMIC_RELAXED,
                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const uint64_t cm = ballot(claimed);
            if (claimed)
                __hip_atomic_store(&a.mailbox[(uint32_t)e], h + (uint32_t)mbcnt(cm) + 1u,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (nw > 0) {
                h += (uint32_t)__popcll(cm);
                ih += (uint32_t)nw;
                moved = true;
            }
        }
        if (moved) {
            last = now;
            continue;
        }
        if (poll_sys(a.host_stop) != 0u) break;
        if (h == t && now - last > a.idle_ticks) break;
        if (poll_agent(&ctl[CTL_LEFT]) >= n_workers) break;
        __builtin_amdgcn_s_sleep(2);
    }
    if (lane == 0) {
        __hip_atomic_store(&ctl[CTL_HEAD], h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl[CTL_TAIL], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl[CTL_QUIT], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int B, int C>
__device__ void serve_one(const ServeArgs& a, int w, Shared<B, C>& sh, uint64_t t_claim) {
    const int tid = threadIdx.x;
    const uint64_t in_off = sh.sv_in;
    const int N = sh.sv_n, S = sh.sv_s, req = sh.sv_req, out_cap = sh.sv_cap;
    const SearchedWords sw = searched_words(S, N);
    const int64_t bytes = sw.bytes;
    // the host checked these at submit; a descriptor outside them is
    // answered with CLF_FLAG_INVALID instead of being read
    const int64_t cap = CLF_EVENT_TRACK_CAPACITY(N, S);
    const bool fits = N >= 0 && S >= 0 && S <= CLF_MAX_SENSORS && N <= a.wg_hits &&
                      bytes <= a.wg_blob_bytes && cap <= a.wg_cap && out_cap >= 0 &&
                      (int64_t)out_cap * (int64_t)sizeof(clf_track) <= (int64_t)a.out_bytes &&
                      in_off + (uint64_t)sw.packed <= a.in_bytes && (in_off & 15u) == 0;
    // the workgroup's own scratch (SearchArgs holds it as read-only input)
    uint8_t* blob = const_cast<uint8_t*>(a.base.blobs) + (size_t)w * (size_t)a.wg_blob_bytes;
    if (fits) {
        // the copy engine wrote the event before the host published its
        // descriptor, which thread 0 read with a system-scope acquire (that
        // invalidates this compute unit's and this XCD's caches) before the
        // barrier every wave passed since
        //
        // copy-in from the device arena, four 16-byte loads in flight per
        // lane: the arena holds the event's searched words packed (the host
        // left out the hit IDs, which the search never reads), and they go
        // to their places in the blob layout the search reads
        const uint4* src = reinterpret_cast<const uint4*>(a.in_arena + in_off);
        uint4* dst = reinterpret_cast<uint4*>(blob);
        const int64_t h16 = sw.h16, m16 = sw.m16, gap = sw.x16 - sw.h16;
        auto at = [&](int64_t j) { return j < h16 ? j : j + gap; };
        int64_t i = tid;
        for (; i + 3 * B < m16; i += 4 * B) {
            const uint4 v0 = src[i], v1 = src[i + B], v2 = src[i + 2 * B], v3 = src[i + 3 * B];
            dst[at(i)] = v0;
            dst[at(i + B)] = v1;
            dst[at(i + 2 * B)] = v2;
            dst[at(i + 3 * B)] = v3;
        }
        for (; i < m16; i += B) dst[at(i)] = src[i];
        // the blob's last bytes past word n16, after the packed words
        const uint32_t* s4 = reinterpret_cast<const uint32_t*>(a.in_arena + in_off) + 4 * m16;
        uint32_t* d4 = reinterpret_cast<uint32_t*>(blob) + 4 * sw.n16;
        for (int64_t k = tid; k < (bytes - 16 * sw.n16) / 4; k += B) d4[k] = s4[k];
    }
    if (tid == 0) {
        EventDesc ed;
        ed.blob_off = (int64_t)w * a.wg_blob_bytes;
        ed.hit_off = (int64_t)w * a.wg_hits;
        ed.track_off = (int64_t)w * a.wg_cap;
        ed.n_hits = fits ? N : -1;
        ed.n_sensors = fits ? S : 0;
        const_cast<EventDesc*>(a.base.ev)[w] = ed;
    }
    __threadfence_block();
    __syncthreads();
    const uint64_t t_copied = wall_clock64();
    process_event<B, C, false>(a.base, w, sh);
    __threadfence_block();
    __syncthreads();
    const uint64_t t_searched = wall_clock64();
    const EventResult r = a.base.results[w];
    const bool ok = fits && r.n_tracks >= 0 && r.n_tracks <= out_cap;
    // copy-out into the device output ring: n_tracks records at a position
    // reserved with one atomic on its allocator (the host's admission keeps
    // the ring from overrunning what it has not copied back and released)
    const int64_t words = ok ? (int64_t)r.n_tracks * (int64_t)(sizeof(clf_track) / 4) : 0;
    if (tid == 0) {
        unsigned long long* head = reinterpret_cast<unsigned long long*>(a.ctl + CTL_OUT_HEAD);
        sh.pack_base = words > 0 ? (int64_t)__hip_atomic_fetch_add(head, (unsigned long long)(4 * words),
                                                                   __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_AGENT)
                                 : 0;
    }
    __syncthreads();
    const uint64_t pos = (uint64_t)uni(sh.pack_base);
    if (words > 0) {
        const uint32_t* ts = reinterpret_cast<const uint32_t*>(a.base.tracks + (size_t)w * a.wg_cap);
        uint32_t* ring = reinterpret_cast<uint32_t*>(a.out_arena);
        const int64_t ring_words = (int64_t)(a.out_bytes / 4);
        const int64_t w0 = (int64_t)((pos / 4) % (uint64_t)ring_words);
        const int64_t first = min(words, ring_words - w0);  // words before the ring's end
        for (int64_t k = tid; k < first; k += B) ring[w0 + k] = track_word(ts, (uint32_t)k);
        for (int64_t k = first + tid; k < words; k += B) ring[k - first] = track_word(ts, (uint32_t)k);
    }
    // every wave's ring stores are complete; thread 0's system-scope release
    // of the completion below then makes them visible to the copy engine
    // (one L2 writeback per event)
    __syncthreads();
    if (tid == 0) {
        const uint32_t p =
            __hip_atomic_fetch_add(&a.ctl[CTL_DONE], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        CompRec* c = a.comp + (p & a.comp_mask);
        c->req = req;
        c->wg = w;
        c->status = ok ? COMP_OK : COMP_INVALID;
        EventResult rr = r;
        if (!ok) rr.n_tracks = 0;
        c->res = rr;
        c->out_pos = pos;
        c->t_claim = t_claim;
        c->t_copied = t_copied;
        c->t_searched = t_searched;
        c->t_written = wall_clock64();
        __hip_atomic_store(&c->seq, p + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

template <int B, int C>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(MinWaves<B, C>::value)))
void serve_kernel(ServeArgs a) {
    __shared__ Shared<B, C> sh;
    if (blockIdx.x == 0) {
        if (threadIdx.x < kWave) serve_feeder(a);
        return;
    }
    const int w = blockIdx.x;
    uint32_t* ctl = a.ctl;
    const uint64_t* beat = reinterpret_cast<const uint64_t*>(ctl + CTL_BEAT);
    uint32_t last_mb = 0;  // thread 0: the mailbox value already taken
    for (;;) {
        if (threadIdx.x == 0) {
            int32_t got = -1;
            uint32_t d = 0;
            const uint32_t pos = __hip_atomic_fetch_add(&ctl[CTL_IDLE_TAIL], 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            uint64_t* slot = &a.idle_ring[pos & a.idle_mask];
            __hip_atomic_store(slot, idle_entry(pos, (uint32_t)w), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t t0 = wall_clock64();
            int nap = 1;  // idle polls back off (64-cycle units, up to ~4k cycles)
            bool claimed = false;  // the feeder took the entry: a hand-out is coming
            for (;;) {
                uint32_t v = poll_agent(&a.mailbox[w]);
                if (v == last_mb && poll_agent(&ctl[CTL_QUIT]) != 0u) {
                    // the feeder has left; it wrote every hand-out before CTL_QUIT
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    v = poll_agent(&a.mailbox[w]);
                    if (v == last_mb) break;
                }
                if (v != last_mb) {
                    last_mb = v;
                    d = v - 1u;
                    got = 1;
                    break;
                }
                if (!claimed) {
                    const uint64_t hb = __hip_atomic_load(beat, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const uint64_t since = hb > t0 ? hb : t0;
                    if ((int64_t)(wall_clock64() - since) > (int64_t)(4 * a.idle_ticks)) {
                        // the feeder is silent: withdraw the entry and leave,
                        // unless it has been claimed, in which case the
                        // feeder writes the mailbox next
                        uint64_t expect = idle_entry(pos, (uint32_t)w);
                        if (__hip_atomic_compare_exchange_strong(
                                slot, &expect, idle_entry(pos, kIdleGone), __ATOMIC_RELAXED,
                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                            break;
                        claimed = true;
                    }
                }
                for (int k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(4);
                nap = nap < 16 ? 2 * nap : 16;
            }
            sh.prof[0] = wall_clock64();
            if (got > 0) {
                const ServeDesc* sd = a.ring + (d & a.ring_mask);
                // acquire: the fields below were written before the sequence word
                const bool seen =
                    __hip_atomic_load(&sd->seq, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) == d + 1u;
                // the descriptor fields in one round trip
                const uint4 q0 = reinterpret_cast<const uint4*>(sd)[0];
                const uint4 q1 = reinterpret_cast<const uint4*>(sd)[1];
                sh.sv_in = (uint64_t)q0.x | ((uint64_t)q0.y << 32);
                sh.sv_n = seen ? (int32_t)q1.x : -1;  // not seen cannot happen: the feeder saw it
                sh.sv_s = (int32_t)q1.y;
                sh.sv_req = (int32_t)q1.z;
                sh.sv_cap = (int32_t)q1.w;
            }
            sh.item = got;
            sh.sv_d = d;
            if (got < 0)  // leaving: counted so that a feeder left alone leaves too
                __hip_atomic_fetch_add(&ctl[CTL_LEFT], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const int got = uni(sh.item);
        const uint64_t t_claim = sh.prof[0];
        __syncthreads();
        if (got < 0) break;
        serve_one<B, C>(a, w, sh, t_claim);
    }
}

// Writes zeros: a batch's queue words (the packed-output allocator, the
// count of finished workgroups and the queue counters) before its first
// run, and other control words (a kernel, not a memset, so that a sequence
// holding it can be captured into a graph).
__global__ void zero_i32_kernel(int32_t* p, int n, unsigned long long* alloc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
    if (alloc && i == 0) *alloc = 0ull;
}

// Plain 16-byte-per-lane copy: the known-size reference kernel against
// which the HBM counters (FETCH_SIZE, WRITE_SIZE) of the search are scaled.
__global__ __launch_bounds__(256) void probe_copy_kernel(const uint4* __restrict__ src,
                                                         uint4* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Device scratch of the entry point that allocates nothing
// (clf_search_noalloc): every device buffer of one of its calls is carved
// out of this array.  Each device that loads the library holds its own copy
// until the process ends.  Host code never uses the array's name: the
// address comes from noalloc_scratch(), once per device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(void** ptr, size_t* bytes) {
    *bytes = kNoAllocScratchBytes;
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

// ---------------------------------------------------------------------------
// Launchers (called from the host library)
// ---------------------------------------------------------------------------
#define CLF_FOR_EACH_CONFIG(X) \
    X(64, 256)                 \
    X(128, 256)                \
    X(256, 256)                \
    X(512, 256)                \
    X(1024, 256)               \
    X(64, 512)                 \
    X(128, 512)                \
    X(256, 512)                \
    X(512, 512)                \
    X(1024, 512)

bool config_supported(int block, int lds_hits) {
#define CLF_CASE(B_, C_) \
    if (block == B_ && lds_hits == C_) return true;
    CLF_FOR_EACH_CONFIG(CLF_CASE)
#undef CLF_CASE
    return false;
}

hipError_t launch_search(const SearchArgs& a, int block, int lds_hits, int grid, hipStream_t stream) {
#define CLF_CASE(B_, C_)                                                                   \
    if (block == B_ && lds_hits == C_) {                                                   \
        hipLaunchKernelGGL((search_kernel<B_, C_>), dim3(grid), dim3(B_), 0, stream, a); \
        return hipGetLastError();                                                          \
    }
    CLF_FOR_EACH_CONFIG(CLF_CASE)
#undef CLF_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_serve(const ServeArgs& a, int block, int lds_hits, int grid, hipStream_t stream) {
#define CLF_CASE(B_, C_)                                                                  \
    if (block == B_ && lds_hits == C_) {                                                  \
        hipLaunchKernelGGL((serve_kernel<B_, C_>), dim3(grid), dim3(B_), 0, stream, a); \
        return hipGetLastError();                                                         \
    }
    CLF_FOR_EACH_CONFIG(CLF_CASE)
#undef CLF_CASE
    return hipErrorInvalidValue;
}

hipError_t serve_occupancy(int block, int lds_hits, int* blocks_per_cu) {
#define CLF_CASE(B_, C_)                                                                    \
    if (block == B_ && lds_hits == C_)                                                      \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,                  \
                                                            serve_kernel<B_, C_>, B_, 0);
    CLF_FOR_EACH_CONFIG(CLF_CASE)
#undef CLF_CASE
    return hipErrorInvalidValue;
}

hipError_t occupancy(int block, int lds_hits, int* blocks_per_cu) {
#define CLF_CASE(B_, C_)                                                                    \
    if (block == B_ && lds_hits == C_)                                                      \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,                  \
                                                            search_kernel<B_, C_>, B_, 0);
    CLF_FOR_EACH_CONFIG(CLF_CASE)
#undef CLF_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_zero_i32(int32_t* p, int n, hipStream_t stream, unsigned long long* alloc) {
    if (n <= 0 && !alloc) return hipSuccess;
    const int m = n > 1 ? n : 1;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, p, n, alloc);
    return hipGetLastError();
}

hipError_t launch_probe_copy(const void* src, void* dst, int64_t bytes, hipStream_t stream) {
    const int64_t n = bytes / 16;
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + 255) / 256 < 256 * 32 ? (n + 255) / 256 : 256 * 32;
    hipLaunchKernelGGL(probe_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                       (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}

}  // namespace clf
