// clf_kernels.hip — gfx950 kernels of the search-by-triplet VELO pixel tracker.
//
// One workgroup owns one event at a time and pulls events from an atomic
// work queue (persistent kernel).  Inside an event the sensor steps
// f = S-1 .. 4 run in sequence (the reference's Kernel.cl:671-767 chain);
// every step is forwarding, then the h0 gather, then seeding, each spread
// over the workgroup.  The stages:
//
//   (a) fill_candidates   Kernel.cl:50-178   h1 / h2 windows per hit, by
//                                             binary search on x-sorted
//                                             sensors, exact scans otherwise
//   (c) forwarding        Kernel.cl:199-368  best h2 on sensor f-4 over the
//                                             x-range a 0.6 mm tolerance
//                                             allows (binary search)
//   (b) gather + seeding  Kernel.cl:392-568, 720-754
//                                             the reference's local size
//                                             (64,4) semantics: 64-wide h0
//                                             chunks, 4 rows of h1 with the
//                                             per-row early stop
//   (d) used flags, weak  Kernel.cl:313-363, 769-815
//
// Every append (tracks, tracklets, ttf entries, weak entries) goes through
// a wave ballot / prefix sum in a fixed order instead of a global atomic,
// so the output — order included — is the same on every run.
//
// FP32, compiled with -ffp-contract=off: each expression keeps the
// reference's operation order.  The restriction of a scan to an x-range
// is exact: it drops only hits whose |dx| >= 0.6 mm is proven from the
// monotonicity of IEEE rounding (see x_window()).
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>

#include "clf_device.h"
#include "clf_params.h"

namespace clf {

constexpr int kWave = 64;
constexpr int kMaxVisits = 256;  // gathered h0 handled per seeding batch
constexpr float kLim = 1.0e5f;   // |coordinate| bound of the supported domain

// ---------------------------------------------------------------------------
// Wave / block helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / kWave; }
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
__device__ __forceinline__ int wave_incl_scan(int x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

template <int B>
struct Shared {
    static constexpr int W = B / kWave;
    int32_t s_start[CLF_MAX_SENSORS];
    int32_t s_num[CLF_MAX_SENSORS];
    int32_t s_Z[CLF_MAX_SENSORS];
    float s_zmin[CLF_MAX_SENSORS];
    float s_zmax[CLF_MAX_SENSORS];
    int32_t s_sorted[CLF_MAX_SENSORS];
    int32_t s_pre[CLF_MAX_SENSORS + 1];
    int32_t scan[3][W];
    int32_t item;
    uint32_t flags;
    int32_t max_live;
    int32_t n_frozen;
    // gather state of the current sensor step (Kernel.cl:720-721,753-754)
    uint32_t g_prev, g_shift, g_counter;
    int32_t g_nv, g_more;
    // the current batch of gathered h0
    int32_t v_h0[kMaxVisits];
    int32_t v_first[kMaxVisits];
    int32_t v_nh[kMaxVisits];
    int32_t v_pre[kMaxVisits + 1];
    float v_x0[kMaxVisits], v_y0[kMaxVisits], v_z0[kMaxVisits], v_dymax[kMaxVisits];
    int32_t v_death[kMaxVisits * CLF_SEED_ROWS];
    unsigned long long v_best[kMaxVisits];
};

// Exclusive prefix counts of up to three predicates over the block, in
// thread order.  Every thread of the block must call it.
template <int B>
__device__ __forceinline__ void block_scan3(Shared<B>& sh, bool p0, bool p1, bool p2, int (&pos)[3],
                                            int (&tot)[3]) {
    const int lane = lane_id(), wave = wave_id();
    const uint64_t m[3] = {ballot(p0), ballot(p1), ballot(p2)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pos[k] = mbcnt(m[k]);
        if (lane == 0) sh.scan[k][wave] = __popcll(m[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int base = 0, t = 0;
        for (int w = 0; w < Shared<B>::W; ++w) {
            const int c = sh.scan[k][w];
            base += (w < wave) ? c : 0;
            t += c;
        }
        pos[k] += base;
        tot[k] = t;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Searches over an x-sorted slice
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

// Widened bounds [lo, hi] of the x of hits that can pass |x_pred - x| < 0.6.
// x_pred over the sensor is bracketed by pa and pb, its values at the
// sensor's extreme z (the prediction chain is monotone in z under IEEE
// rounding).  If x > pmax + 0.6 exactly, fl(x_pred - x) <= -0.6f, so the
// cut fails; the margin covers the rounding of the bound itself.
__device__ __forceinline__ void x_window(float pa, float pb, float& lo, float& hi) {
    lo = fminf(pa, pb) - CLF_PARAM_TOLERANCE;
    lo = lo - (1.0e-3f + fabsf(lo) * 1.0e-6f);
    hi = fmaxf(pa, pb) + CLF_PARAM_TOLERANCE;
    hi = hi + (1.0e-3f + fabsf(hi) * 1.0e-6f);
}

// fitHitToTrack (Kernel.cl:15-35) with its multiply-by-bool blend.
__device__ __forceinline__ float fit_hit_to_track(float tx, float ty, float x0, float y0, float z0,
                                                  float h1_z, float x2, float y2, float z2) {
    const float dz = z2 - z0;
    const float x_prediction = x0 + tx * dz;
    const float dx = fabsf(x_prediction - x2);
    const bool tolx = dx < CLF_PARAM_TOLERANCE;
    const float y_prediction = y0 + ty * dz;
    const float dy = fabsf(y_prediction - y2);
    const bool toly = dy < CLF_PARAM_TOLERANCE;
    const float scatterNum = (dx * dx) + (dy * dy);
    const float scatterDenom = 1.f / (z2 - h1_z);
    const float scatter = scatterNum * scatterDenom * scatterDenom;
    const bool cond = tolx && toly && (scatter < CLF_MAX_SCATTER);
    return (float)cond * scatter + (float)(!cond) * FLT_MAX;
}

struct EventView {
    int S, N;
    const float* X;
    const float* Y;
    const float* Zh;
};

// ---------------------------------------------------------------------------
// (a) candidate windows of one hit h on sensor f >= 4 (Kernel.cl:73-172)
// ---------------------------------------------------------------------------
template <int B>
__device__ int4 fill_one(const Shared<B>& sh, const EventView& ev, int h, int f) {
    const int s2 = f - 2;
    const bool p2 = f <= ev.S - 3;
    const int st2 = sh.s_start[s2], n2 = sh.s_num[s2], end2 = st2 + n2;
    const float x0 = ev.X[h], z0 = ev.Zh[h];
    const float* X = ev.X;
    const float* Zh = ev.Zh;
    const bool sorted2 = sh.s_sorted[s2] != 0;
    int4 c = make_int4(-1, -1, -1, -1);
    if (n2 == 0) return c;

    // h1 window: first index with |x1-x0| < 0.7|z1-z0|, then the first later
    // index where that fails, else the sensor end (:115-135, :161-165).
    if (sorted2 && sh.s_zmin[s2] == sh.s_zmax[s2]) {
        // one z on the sensor: the predicate is -D < fl(x1-x0) < D
        const float D = CLF_PARAM_MAXXSLOPE_CANDIDATES * fabsf(sh.s_zmin[s2] - z0);
        const int lo = first_diff_gt(X, st2, end2, x0, -D);
        const int hi = first_diff_ge(X, lo, end2, x0, D);
        if (lo < hi) {
            c.x = lo;
            c.y = hi;
        }
    } else {
        int i = st2;
        if (sorted2) {
            // no hit with fl(x1-x0) <= -B can pass, B >= every 0.7|z1-z0|
            const float a = fabsf(sh.s_zmin[s2] - z0), b = fabsf(sh.s_zmax[s2] - z0);
            const float Bd = CLF_PARAM_MAXXSLOPE_CANDIDATES * fmaxf(a, b);
            i = first_diff_gt(X, st2, end2, x0, -Bd);
        }
        for (; i < end2; ++i) {
            const float dxmax = CLF_PARAM_MAXXSLOPE_CANDIDATES * fabsf(Zh[i] - z0);
            if (fabsf(X[i] - x0) < dxmax) break;
        }
        if (i < end2) {
            c.x = i;
            for (++i; i < end2; ++i) {
                const float dxmax = CLF_PARAM_MAXXSLOPE_CANDIDATES * fabsf(Zh[i] - z0);
                if (!(fabsf(X[i] - x0) < dxmax)) break;
            }
            c.y = i;
        }
    }

    // h2 window on the same sensor, for h acting as h1 (:85-102, :137-150)
    if (p2) {
        const int zs0 = sh.s_Z[f + 2];
        const int zs2 = sh.s_Z[s2];
        const float h_dist = fabsf(z0 - (float)zs0);
        const float dxmax = CLF_PARAM_MAXXSLOPE_CANDIDATES * h_dist;
        const float x0_min = x0 - dxmax;
        const float x0_max = x0 + dxmax;
        const float z2_tz = ((float)zs2 - (float)zs0) / (z0 - (float)zs0);
        float xx = x0_max + (x0 - x0_max) * z2_tz;
        const float xmin_h2 = xx - CLF_PARAM_TOLERANCE_CANDIDATES;
        xx = x0_min + (x0 - x0_min) * z2_tz;
        const float xmax_h2 = xx + CLF_PARAM_TOLERANCE_CANDIDATES;
        if (sorted2) {
            const int first = ub_gt(X, st2, end2, xmin_h2);
            if (first < end2) {
                c.z = first;
                c.w = ub_gt(X, first + 1, end2, xmax_h2);
            }
        } else {
            int i = st2;
            for (; i < end2; ++i)
                if (X[i] > xmin_h2) break;
            if (i < end2) {
                c.z = i;
                for (++i; i < end2; ++i)
                    if (X[i] > xmax_h2) break;
                c.w = i;
            }
        }
    }
    return c;
}

// ---------------------------------------------------------------------------
// The h0 gather of one sensor step, run by wave 0 (Kernel.cl:720-754).
// Chunk c reads elements prev+lane; a lane whose element is used takes
// values from the shared counter (which restarts at 0 each step and keeps
// counting across chunks) and looks at prev + shift + value until it finds
// an unused or out-of-range element.  The r-th such lane in lane order
// takes the r-th of those elements.  Then prev = counter + shift and
// shift += 64.  Visits are appended to v_h0 in lane order.
// ---------------------------------------------------------------------------
template <int B>
__device__ void gather_batch(Shared<B>& sh, const uint8_t* used, int start0, uint32_t hn0) {
    const int lane = lane_id();
    uint32_t prev = sh.g_prev, shift = sh.g_shift, counter = sh.g_counter;
    int nv = 0;
    while (prev < hn0 && nv + kWave <= kMaxVisits) {
        const uint32_t e = prev + (uint32_t)lane;
        const bool inb = e < hn0;
        const bool u = inb && used[start0 + (int)e];
        int got = (inb && !u) ? start0 + (int)e : -1;
        const uint64_t need = ballot(u);
        const int n_need = __popcll(need);
        const int rank = mbcnt(need);
        int done = 0;
        uint32_t pos = counter;
        while (done < n_need) {
            const uint32_t e2 = prev + shift + pos + (uint32_t)lane;
            const bool inb2 = e2 < hn0;
            const bool un2 = inb2 && !used[start0 + (int)e2];
            const uint64_t sm = ballot(!inb2 || un2);
            const int c = __popcll(sm);
            const int val = un2 ? start0 + (int)e2 : -1;
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
}

struct StepCtx {
    int f;
    int st2, n2;  // sensor f-2
    int st4, n4;  // sensor f-4
    bool fast4;   // sensor f-4 x-sorted and the event in the supported domain
    float zmin4, zmax4;
};

// ---------------------------------------------------------------------------
// (b) seeding of one batch of gathered h0 with the (64,4) semantics
// (Kernel.cl:408-567).  Row y of an h0 takes window elements 4j+y and stops
// for good at its first element whose h2 window is empty; a used h1 is
// skipped.  The winner minimises (fit, y, j, h2) lexicographically, which is
// what the rows' strict-less first minima and the lowest-y tie-break pick.
// ---------------------------------------------------------------------------
template <int B>
__device__ void seed_batch(Shared<B>& sh, const EventView& ev, const StepCtx& sc, const int4* cand,
                           const uint8_t* used, uint4* tracklets, TtfEntry* nxt, int nv, int& n_next,
                           int& n_tracklets) {
    const int tid = threadIdx.x;
    const float* X = ev.X;
    const float* Y = ev.Y;
    const float* Zh = ev.Zh;
    // z of the first hit of sensor f-2 (:420-422)
    const float s1_z = sc.n2 > 0 ? Zh[sc.st2] : 0.f;

    for (int v = tid; v < nv; v += B) {
        const int h0 = sh.v_h0[v];
        const int4 c = cand[h0];
        sh.v_first[v] = c.x;
        sh.v_nh[v] = (int)(uint32_t)(c.y - c.x);
        const float z0 = Zh[h0];
        sh.v_x0[v] = X[h0];
        sh.v_y0[v] = Y[h0];
        sh.v_z0[v] = z0;
        sh.v_dymax[v] = CLF_PARAM_MAXYSLOPE * fabsf(s1_z - z0);
#pragma unroll
        for (int y = 0; y < CLF_SEED_ROWS; ++y) sh.v_death[v * CLF_SEED_ROWS + y] = 0x7fffffff;
        sh.v_best[v] = ~0ull;
    }
    __syncthreads();
    if (wave_id() == 0) {
        const int lane = lane_id();
        int running = 0;
        for (int c0 = 0; c0 < nv; c0 += kWave) {
            const int v = c0 + lane;
            const int x = v < nv ? sh.v_nh[v] : 0;
            const int incl = wave_incl_scan(x);
            if (v < nv) sh.v_pre[v] = running + incl - x;
            running += __shfl(incl, kWave - 1);
        }
        if (lane == 0) sh.v_pre[nv] = running;
    }
    __syncthreads();
    const int P = sh.v_pre[nv];

    // pairs (visit v, window element e) in a flat index
    auto visit_of = [&](int p) {
        int lo = 0, hi = nv;  // last v with v_pre[v] <= p
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sh.v_pre[mid] <= p) lo = mid;
            else hi = mid;
        }
        return lo;
    };

    // row stop points: first j of row y whose h1 has no h2 window (:453-457)
    for (int p = tid; p < P; p += B) {
        const int v = visit_of(p);
        const int e = p - sh.v_pre[v];
        const int h1 = sh.v_first[v] + e;
        if (cand[h1].z == -1)
            atomicMin(&sh.v_death[v * CLF_SEED_ROWS + (e & (CLF_SEED_ROWS - 1))], e / CLF_SEED_ROWS);
    }
    __syncthreads();

    const float* X4 = X + sc.st4;
    const float* Y4 = Y + sc.st4;
    const float* Z4 = Zh + sc.st4;
    for (int p = tid; p < P; p += B) {
        const int v = visit_of(p);
        const int e = p - sh.v_pre[v];
        const int y = e & (CLF_SEED_ROWS - 1);
        const int j = e / CLF_SEED_ROWS;
        if (j >= sh.v_death[v * CLF_SEED_ROWS + y]) continue;
        const int h1 = sh.v_first[v] + e;
        if (used[h1]) continue;
        const float x0 = sh.v_x0[v], y0 = sh.v_y0[v], z0 = sh.v_z0[v];
        const float x1 = X[h1], y1 = Y[h1], z1 = Zh[h1];
        if (!(fabsf(y1 - y0) < sh.v_dymax[v])) continue;  // :509, independent of h2
        const int4 c = cand[h1];
        const float dz_inverted = 1.f / (z1 - z0);
        int k0 = c.z - sc.st4, k1 = c.w - sc.st4;
        if (sc.fast4) {
            const float pa = x0 + (x1 - x0) * ((sc.zmin4 - z0) * dz_inverted);
            const float pb = x0 + (x1 - x0) * ((sc.zmax4 - z0) * dz_inverted);
            float lo, hi;
            x_window(pa, pb, lo, hi);
            k0 = lb_ge(X4, k0, k1, lo);
            k1 = ub_gt(X4, k0, k1, hi);
        }
        float best = FLT_MAX;
        int bk = -1;
        for (int k = k0; k < k1; ++k) {
            const float x2 = X4[k], y2 = Y4[k], z2 = Z4[k];
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
        if (bk >= 0) {
            const unsigned long long key = ((unsigned long long)__float_as_uint(best + 0.0f) << 32) |
                                           ((unsigned long long)y << 30) |
                                           ((unsigned long long)j << 16) | (unsigned long long)bk;
            atomicMin(&sh.v_best[v], key);
        }
    }
    __syncthreads();

    // tracklets and their ttf entries, in visit order (:550-566)
    for (int base = 0; base < nv; base += B) {
        const int v = base + tid;
        const bool acc = v < nv && sh.v_best[v] != ~0ull;
        int pos[3], tot[3];
        block_scan3<B>(sh, acc, false, false, pos, tot);
        if (acc) {
            const unsigned long long key = sh.v_best[v];
            const int y = (int)((key >> 30) & 3ull);
            const int j = (int)((key >> 16) & 0x3fffull);
            const int k = (int)(key & 0xffffull);
            const int h0 = sh.v_h0[v];
            const int h1 = sh.v_first[v] + CLF_SEED_ROWS * j + y;
            const int h2 = sc.st4 + k;
            const int idx = n_tracklets + pos[0];
            tracklets[idx] = make_uint4((uint32_t)h0, (uint32_t)h1, (uint32_t)h2, 0u);
            TtfEntry t;
            t.word = CLF_TTF_TRACKLET_BIT | (uint32_t)idx;
            t.nhits = 3;
            t.x0 = X[h1];
            t.y0 = Y[h1];
            t.z0 = Zh[h1];
            t.x1 = X[h2];
            t.y1 = Y[h2];
            t.z1 = Zh[h2];
            nxt[n_next + pos[0]] = t;
        }
        n_tracklets += tot[0];
        n_next += tot[0];
    }
}

// ---------------------------------------------------------------------------
// One event
// ---------------------------------------------------------------------------
template <int B>
__device__ void process_event(const SearchArgs& a, int e, Shared<B>& sh) {
    constexpr int W = Shared<B>::W;
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    const EventDesc d = a.ev[e];
    EventResult* res = a.results + e;
    if (d.n_hits < 0) {
        if (tid == 0) {
            EventResult r = {};
            r.flags = CLF_FLAG_INVALID;
            *res = r;
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
    int4* cand = a.cand + d.hit_off;
    uint8_t* used = a.used + d.hit_off;
    clf_track* tracks = a.tracks + d.track_off;
    uint4* tracklets = a.tracklets + d.track_off;
    TtfEntry* cur = a.ttf_a + d.track_off;
    TtfEntry* nxt = a.ttf_b + d.track_off;
    uint32_t* weak = a.weak + d.track_off;

    for (int s = tid; s < S; s += B) {
        sh.s_start[s] = gStart[s];
        sh.s_num[s] = gNum[s];
        sh.s_Z[s] = gZ[s];
    }
    for (int i = tid; i < N; i += B) {
        used[i] = 0;
        cand[i] = make_int4(-1, -1, -1, -1);  // the host's -1 fill (KernelInvoker.cpp:162-163)
    }
    if (tid == 0) {
        sh.flags = 0;
        sh.max_live = 0;
        sh.n_frozen = 0;
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
#pragma unroll
        for (int d2 = 1; d2 < kWave; d2 <<= 1) {
            mn = fminf(mn, __shfl_xor(mn, d2));
            mx = fmaxf(mx, __shfl_xor(mx, d2));
        }
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
    if (tid == 0) {
        int acc = 0;
        for (int f = 0; f < S; ++f) {
            sh.s_pre[f] = acc;
            if (f >= 4) acc += sh.s_num[f];
        }
        sh.s_pre[S] = acc;
    }
    __syncthreads();
    const bool degenerate = (sh.flags & CLF_FLAG_DEGENERATE_Z) != 0;

    // (a) candidate windows of every hit on sensors 4..S-1
    {
        const int total = sh.s_pre[S];
        for (int p = tid; p < total; p += B) {
            int lo = 4, hi = S;  // last f with s_pre[f] <= p
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (sh.s_pre[mid] <= p) lo = mid;
                else hi = mid;
            }
            const int f = lo;
            const int h = sh.s_start[f] + (p - sh.s_pre[f]);
            cand[h] = fill_one<B>(sh, ev, h, f);
        }
    }
    __syncthreads();
    if (a.fill_only) {
        if (tid == 0) {
            EventResult r = {};
            r.flags = sh.flags;
            *res = r;
        }
        return;
    }

    int n_cur = 0, n_tracks = 0, n_tracklets = 0, n_weak = 0, visits = 0;
    for (int f = S - 1; f >= 4; --f) {
        StepCtx sc;
        sc.f = f;
        sc.st2 = sh.s_start[f - 2];
        sc.n2 = sh.s_num[f - 2];
        sc.st4 = sh.s_start[f - 4];
        sc.n4 = sh.s_num[f - 4];
        sc.fast4 = sh.s_sorted[f - 4] && !degenerate;
        sc.zmin4 = sh.s_zmin[f - 4];
        sc.zmax4 = sh.s_zmax[f - 4];
        const float* X4 = ev.X + sc.st4;
        const float* Y4 = ev.Y + sc.st4;
        const float* Z4 = ev.Zh + sc.st4;

        // (c) forwarding of the ttf entries of step f+1 (Kernel.cl:210-367)
        int n_next = 0;
        for (int base = 0; base < n_cur; base += B) {
            const int i = base + tid;
            const bool active = i < n_cur;
            TtfEntry t = {};
            bool found = false, f_new = false, f_req = false, f_weak = false;
            int bk = -1;
            if (active) {
                t = cur[i];
                const bool is_tl = (t.word & CLF_TTF_TRACKLET_BIT) != 0;
                const uint32_t skipped = (t.word & CLF_TTF_SKIP_MASK) >> CLF_TTF_SKIP_SHIFT;
                const float td = 1.0f / (t.z1 - t.z0);
                const float txn = (t.x1 - t.x0);
                const float tyn = (t.y1 - t.y0);
                const float tx = txn * td;
                const float ty = tyn * td;
                int k0 = 0, k1 = sc.n4;
                if (sc.fast4 && sc.n4 > 0) {
                    const float pa = t.x0 + tx * (sc.zmin4 - t.z0);
                    const float pb = t.x0 + tx * (sc.zmax4 - t.z0);
                    float lo, hi;
                    x_window(pa, pb, lo, hi);
                    k0 = lb_ge(X4, 0, sc.n4, lo);
                    k1 = ub_gt(X4, k0, sc.n4, hi);
                }
                float best_fit = FLT_MAX;
                for (int k = k0; k < k1; ++k) {
                    const float fit =
                        fit_hit_to_track(tx, ty, t.x0, t.y0, t.z0, t.z1, X4[k], Y4[k], Z4[k]);
                    const bool better = fit < best_fit;
                    best_fit = (float)better * fit + (float)(!better) * best_fit;
                    bk = better ? k : bk;
                }
                if (best_fit != FLT_MAX) {
                    if (bk >= 0) {
                        found = true;
                        f_new = is_tl;
                        f_req = t.nhits + 1 < CLF_MAX_TRACK_SIZE;
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
            int pos[3], tot[3];
            block_scan3<B>(sh, f_new, f_req, f_weak, pos, tot);
            if (active) {
                const uint32_t idx = t.word & CLF_TTF_INDEX_MASK;
                if (found) {
                    const int b = sc.st4 + bk;
                    used[b] = 1;
                    uint32_t slot;
                    if (f_new) {
                        const uint4 tl = tracklets[idx];
                        used[tl.x] = 1;
                        used[tl.y] = 1;
                        used[tl.z] = 1;
                        slot = (uint32_t)(n_tracks + pos[0]);
                        uint32_t* r = reinterpret_cast<uint32_t*>(tracks + slot);
                        r[0] = 4;
                        r[1] = tl.x;
                        r[2] = tl.y;
                        r[3] = tl.z;
                        r[4] = (uint32_t)b;
#pragma unroll
                        for (int q = 5; q < 1 + CLF_MAX_TRACK_SIZE; ++q) r[q] = 0;
                    } else {
                        slot = idx;
                        tracks[slot].hits[t.nhits] = (uint32_t)b;
                        tracks[slot].hitsNum = t.nhits + 1;
                    }
                    if (f_req) {
                        TtfEntry n;
                        n.word = slot;
                        n.nhits = t.nhits + 1;
                        n.x0 = t.x1;
                        n.y0 = t.y1;
                        n.z0 = t.z1;
                        n.x1 = X4[bk];
                        n.y1 = Y4[bk];
                        n.z1 = Z4[bk];
                        nxt[n_next + pos[1]] = n;
                    } else {
                        atomicOr(&sh.flags, CLF_FLAG_TRACK_FULL);
                        atomicAdd(&sh.n_frozen, 1);
                    }
                } else if (f_req) {
                    const uint32_t skipped = (t.word & CLF_TTF_SKIP_MASK) >> CLF_TTF_SKIP_SHIFT;
                    TtfEntry n = t;
                    n.word = ((skipped + 1) << CLF_TTF_SKIP_SHIFT) | (t.word & CLF_TTF_KEEP_MASK);
                    nxt[n_next + pos[1]] = n;
                } else if (f_weak) {
                    weak[n_weak + pos[2]] = idx;
                }
            }
            n_tracks += tot[0];
            n_next += tot[1];
            n_weak += tot[2];
        }
        __syncthreads();  // used[] of this step visible to the gather and seeding

        // (b) gather + seeding (Kernel.cl:720-764)
        if (tid == 0) {
            sh.g_prev = 0;
            sh.g_shift = CLF_GATHER_CHUNK;
            sh.g_counter = 0;
        }
        __syncthreads();
        const int start0 = sh.s_start[f];
        const uint32_t hn0 = (uint32_t)sh.s_num[f];
        for (;;) {
            if (wave == 0) gather_batch<B>(sh, used, start0, hn0);
            __syncthreads();
            const int nv = sh.g_nv;
            const bool more = sh.g_more != 0;
            visits += nv;
            if (nv > 0)
                seed_batch<B>(sh, ev, sc, cand, used, tracklets, nxt, nv, n_next, n_tracklets);
            __syncthreads();
            if (!more) break;
        }

        // span of the reference's ttf ring during this step
        if (tid == 0) {
            const int live = n_cur + n_next;
            if (live > sh.max_live) sh.max_live = live;
        }
        TtfEntry* tmp = cur;
        cur = nxt;
        nxt = tmp;
        n_cur = n_next;
        __syncthreads();
    }

    // tail: tracklet entries of the last step become weak (Kernel.cl:771-792)
    for (int base = 0; base < n_cur; base += B) {
        const int i = base + tid;
        uint32_t w = 0;
        const bool tl = i < n_cur && ((w = cur[i].word) & CLF_TTF_TRACKLET_BIT) != 0;
        int pos[3], tot[3];
        block_scan3<B>(sh, tl, false, false, pos, tot);
        if (tl) weak[n_weak + pos[0]] = w & CLF_TTF_INDEX_MASK;
        n_weak += tot[0];
    }
    __syncthreads();

    // (d) weak tracklets made of three unused hits become tracks (:796-815)
    for (int base = 0; base < n_weak; base += B) {
        const int i = base + tid;
        uint4 tl = make_uint4(0, 0, 0, 0);
        bool ok = false;
        if (i < n_weak) {
            tl = tracklets[weak[i]];
            ok = !used[tl.x] && !used[tl.y] && !used[tl.z];
        }
        int pos[3], tot[3];
        block_scan3<B>(sh, ok, false, false, pos, tot);
        if (ok) {
            uint32_t* r = reinterpret_cast<uint32_t*>(tracks + n_tracks + pos[0]);
            r[0] = 3;
            r[1] = tl.x;
            r[2] = tl.y;
            r[3] = tl.z;
#pragma unroll
            for (int q = 4; q < 1 + CLF_MAX_TRACK_SIZE; ++q) r[q] = 0;
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
        r.n_frozen = sh.n_frozen;
        r.h0_visits = visits;
        r.reserved = 0;
        *res = r;
    }
}

template <int B>
__global__ __launch_bounds__(B) void search_kernel(SearchArgs a) {
    __shared__ Shared<B> sh;
    for (;;) {
        if (threadIdx.x == 0) sh.item = atomicAdd(a.queue_head, 1);
        __syncthreads();
        const int item = sh.item;
        __syncthreads();
        if (item >= a.n_items) break;  // every wave of the group leaves together
        const int idx = a.first_item + item;
        const int e = a.order ? a.order[idx] : idx;
        process_event<B>(a, e, sh);
        __syncthreads();
    }
}

// Writes zeros: the queue counters of a launch (a kernel, not a memset, so
// that the sequence can be captured into a graph and replayed).
__global__ void zero_i32_kernel(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Exclusive scan of n_tracks over events into track_prefix[0..n] (one block).
__global__ __launch_bounds__(1024) void track_prefix_kernel(const EventResult* res, int n,
                                                            int64_t* prefix) {
    __shared__ int64_t wsum[1024 / kWave];
    __shared__ int64_t carry;
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + tid;
        int64_t x = i < n ? (int64_t)res[i].n_tracks : 0;
        int64_t incl = x;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const int64_t y = __shfl_up(incl, d);
            if (lane >= d) incl += y;
        }
        if (lane == kWave - 1) wsum[wave] = incl;
        __syncthreads();
        int64_t before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (i < n) prefix[i] = before + incl - x;
        __syncthreads();
        if (tid == 1023) carry = before + incl;
        __syncthreads();
    }
    if (tid == 0) prefix[n] = carry;
}

// Copies each event's tracks into one packed array (one block per event).
__global__ __launch_bounds__(256) void pack_tracks_kernel(const EventDesc* ev,
                                                          const EventResult* res,
                                                          const int64_t* prefix,
                                                          const clf_track* tracks,
                                                          uint32_t* packed) {
    const int e = blockIdx.x;
    const int64_t n_words = (int64_t)res[e].n_tracks * (sizeof(clf_track) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(tracks + ev[e].track_off);
    uint32_t* dst = packed + prefix[e] * (int64_t)(sizeof(clf_track) / 4);
    for (int64_t w = threadIdx.x; w < n_words; w += blockDim.x) dst[w] = src[w];
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
// Launchers (called from the host library)
// ---------------------------------------------------------------------------
hipError_t launch_probe_copy(const void* src, void* dst, int64_t bytes, hipStream_t stream) {
    const int64_t n = bytes / 16;
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(probe_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                       (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}

hipError_t launch_search(const SearchArgs& a, int block, int grid, hipStream_t stream) {
    switch (block) {
        case 64:
            hipLaunchKernelGGL(search_kernel<64>, dim3(grid), dim3(64), 0, stream, a);
            break;
        case 128:
            hipLaunchKernelGGL(search_kernel<128>, dim3(grid), dim3(128), 0, stream, a);
            break;
        case 256:
            hipLaunchKernelGGL(search_kernel<256>, dim3(grid), dim3(256), 0, stream, a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_zero_i32(int32_t* p, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launch_pack(const EventDesc* ev, const EventResult* res, int n, int64_t* prefix,
                       const clf_track* tracks, uint32_t* packed, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(track_prefix_kernel, dim3(1), dim3(1024), 0, stream, res, n, prefix);
    hipLaunchKernelGGL(pack_tracks_kernel, dim3(n), dim3(256), 0, stream, ev, res, prefix, tracks,
                       packed);
    return hipGetLastError();
}

hipError_t occupancy(int block, int* blocks_per_cu) {
    switch (block) {
        case 64:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, search_kernel<64>, 64, 0);
        case 128:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, search_kernel<128>, 128, 0);
        case 256:
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, search_kernel<256>, 256, 0);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace clf
