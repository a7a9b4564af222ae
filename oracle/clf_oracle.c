/*
 * clf_oracle.c — CPU restatement of the reference search-by-triplet path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Nothing in the
 * product library (cl_forward_amd/) links, loads or calls this file.  Only
 * tests/, the smoke check and the benchmark's CPU-baseline leg use it.
 *
 * Built with gcc -O2 -ffp-contract=off (oracle/Makefile): FP32 arithmetic in
 * the reference's operation order, no FMA contraction, IEEE division.
 *
 * It restates, line by line in behaviour (not in structure):
 *   fillCandidates      Kernel.cl:50-178
 *   fitHitToTrack       Kernel.cl:15-35
 *   trackForwarding     Kernel.cl:199-368
 *   h0 gather           Kernel.cl:720-754
 *   trackCreation       Kernel.cl:392-568 (local size (64,4), the final run
 *                       whose results the reference returns,
 *                       KernelInvoker.cpp:149-156,197-206)
 *   sensor loop + tail  Kernel.cl:666-815
 *   preorder_by_x       Tools.cpp:46-66,117-164
 *
 * The reference is one work-group per event whose outputs are appended
 * through atomics, so the *order* of tracks, tracklets and ttf entries is
 * nondeterministic while their content is not (reads of hit_used only ever
 * happen on the far side of a barrier from its writes).  This restatement
 * runs the work-items of each barrier phase one after another in "lane
 * order": ttf entries in window order; gathered h0 in lane order within a
 * chunk, where the r-th lane (in lane order) that needs a replacement takes
 * the r-th replacement the shared counter hands out.  That is one of the
 * orders the reference can produce; the HIP path produces this same order.
 *
 * Policies where the reference is undefined (clf_params.h CLF_FLAG_*):
 *   - a track that holds 24 hits is re-queued and followed like any other;
 *     when it finds a hit, that hit is marked used (Kernel.cl:316) and the
 *     track stops there, where the reference appends a 25th hit past its
 *     private Track record (Kernel.cl:321); each such stop is counted;
 *   - no capacity of its own: tracks, tracklets, weak and ttf lists are
 *     sized to the event's upper bound and the reference's capacity
 *     overruns are only flagged.
 */
#include "clf_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../include/clf_params.h"

/* ------------------------------------------------------------------ */
/* Event layout (Tools.cpp:68-80, Kernel.cl:613-624)                   */
/* ------------------------------------------------------------------ */
typedef struct {
    int S, N;
    int32_t *Z, *start, *num;
    uint32_t* id;
    float *x, *y, *z;
    void* mem;
} event_t;

static int32_t rd_i32(const uint8_t* p) {
    int32_t v;
    memcpy(&v, p, 4);
    return v;
}

/* Copies the blob into aligned arrays and validates it.  Returns 0 or -1. */
static int event_parse(const uint8_t* blob, size_t size, event_t* ev) {
    memset(ev, 0, sizeof(*ev));
    if (!blob || size < 8) return -1;
    int S = rd_i32(blob), N = rd_i32(blob + 4);
    if (S < 0 || S > CLF_MAX_SENSORS || N < 0 || N > CLF_MAX_EVENT_HITS) return -1;
    size_t need = 8 + 12 * (size_t)S + 16 * (size_t)N;
    if (size < need) return -1;
    ev->S = S;
    ev->N = N;
    size_t bytes = 12 * (size_t)S + 16 * (size_t)N + 16;
    ev->mem = malloc(bytes);
    if (!ev->mem) return -1;
    memcpy(ev->mem, blob + 8, need - 8);
    ev->Z = (int32_t*)ev->mem;
    ev->start = ev->Z + S;
    ev->num = ev->start + S;
    ev->id = (uint32_t*)(ev->num + S);
    ev->x = (float*)(ev->id + N);
    ev->y = ev->x + N;
    ev->z = ev->y + N;
    for (int s = 0; s < S; ++s) {
        if (ev->num[s] < 0 || ev->num[s] > CLF_MAX_SENSOR_HITS || ev->start[s] < 0 ||
            (int64_t)ev->start[s] + ev->num[s] > N) {
            free(ev->mem);
            ev->mem = NULL;
            return -1;
        }
    }
    return 0;
}

static void event_free(event_t* ev) {
    free(ev->mem);
    ev->mem = NULL;
}

/* Input-only flags: CLF_FLAG_UNSORTED and CLF_FLAG_DEGENERATE_Z. */
static uint32_t input_flags(const event_t* ev) {
    uint32_t flags = 0;
    const float LIM = 1.0e5f;
    float zmin[CLF_MAX_SENSORS], zmax[CLF_MAX_SENSORS];
    for (int s = 0; s < ev->S; ++s) {
        const int b = ev->start[s], n = ev->num[s];
        if (fabsf((float)ev->Z[s]) > LIM) flags |= CLF_FLAG_DEGENERATE_Z;
        zmin[s] = INFINITY;
        zmax[s] = -INFINITY;
        for (int i = 0; i < n; ++i) {
            const float xi = ev->x[b + i], yi = ev->y[b + i], zi = ev->z[b + i];
            /* !(|v| <= LIM) also catches NaN */
            if (!(fabsf(xi) <= LIM) || !(fabsf(yi) <= LIM) || !(fabsf(zi) <= LIM))
                flags |= CLF_FLAG_DEGENERATE_Z;
            if (zi < zmin[s]) zmin[s] = zi;
            if (zi > zmax[s]) zmax[s] = zi;
            if (i + 1 < n && !(xi <= ev->x[b + i + 1])) flags |= CLF_FLAG_UNSORTED;
        }
    }
    /* pairwise disjoint hit-z ranges of non-empty sensors */
    for (int s = 0; s < ev->S; ++s) {
        if (ev->num[s] == 0) continue;
        for (int t = s + 1; t < ev->S; ++t) {
            if (ev->num[t] == 0) continue;
            if (!(zmax[s] < zmin[t] || zmax[t] < zmin[s])) flags |= CLF_FLAG_DEGENERATE_Z;
        }
    }
    /* fillCandidates divides by h0.z - Z[f+2] for 4 <= f <= S-3 */
    for (int f = 4; f <= ev->S - 3; ++f) {
        if (ev->num[f] == 0) continue;
        const float zs0 = (float)ev->Z[f + 2];
        if (zs0 >= zmin[f] && zs0 <= zmax[f]) flags |= CLF_FLAG_DEGENERATE_Z;
    }
    return flags;
}

uint32_t clf_oracle_input_flags(const uint8_t* blob, size_t size) {
    event_t ev;
    if (event_parse(blob, size, &ev) != 0) return CLF_FLAG_INVALID;
    uint32_t f = input_flags(&ev);
    event_free(&ev);
    return f;
}

/* ------------------------------------------------------------------ */
/* fillCandidates (Kernel.cl:50-178)                                    */
/* ------------------------------------------------------------------ */
static void fill_candidates(const event_t* ev, int32_t* c1, int32_t* c2) {
    const int S = ev->S;
    for (int i = 0; i < 2 * ev->N; ++i) c1[i] = c2[i] = -1;
    /* sensors 2 and 3 run the loop but never enter the window code (:104) */
    for (int f = S - 1; f >= 4; --f) {
        const int s2 = f - 2;
        const int p2 = f <= S - 3; /* process_h2_candidates (:62) */
        const int zs0 = p2 ? ev->Z[f + 2] : 0;
        const int zs2 = p2 ? ev->Z[s2] : 0;
        const int st2 = ev->start[s2], n2 = ev->num[s2];
        for (int e = 0; e < ev->num[f]; ++e) {
            const int h0 = ev->start[f] + e;
            const float x0 = ev->x[h0], z0 = ev->z[h0];
            float xmin_h2 = 0.f, xmax_h2 = 0.f;
            if (p2) { /* :85-102 */
                const float h_dist = fabsf(z0 - (float)zs0);
                const float dxmax = CLF_PARAM_MAXXSLOPE_CANDIDATES * h_dist;
                const float x0_min = x0 - dxmax;
                const float x0_max = x0 + dxmax;
                const float z2_tz = ((float)zs2 - (float)zs0) / (z0 - (float)zs0);
                float xx = x0_max + (x0 - x0_max) * z2_tz;
                xmin_h2 = xx - CLF_PARAM_TOLERANCE_CANDIDATES;
                xx = x0_min + (x0 - x0_min) * z2_tz;
                xmax_h2 = xx + CLF_PARAM_TOLERANCE_CANDIDATES;
            }
            int f1 = 0, l1 = 0, f2 = 0, l2 = 0;
            for (int k = 0; k < n2; ++k) { /* :106-157 */
                const int h1 = st2 + k;
                const float x1 = ev->x[h1], z1 = ev->z[h1];
                if (!l1) {
                    const float h_dist = fabsf(z1 - z0);
                    const float dxmax = CLF_PARAM_MAXXSLOPE_CANDIDATES * h_dist;
                    const int tol = fabsf(x1 - x0) < dxmax;
                    if (!f1 && tol) {
                        c1[2 * h0] = h1;
                        f1 = 1;
                    } else if (f1 && !tol) {
                        c1[2 * h0 + 1] = h1;
                        l1 = 1;
                    }
                }
                if (p2 && !l2) {
                    if (!f2 && x1 > xmin_h2) {
                        c2[2 * h0] = h1;
                        f2 = 1;
                    } else if (f2 && x1 > xmax_h2) {
                        c2[2 * h0 + 1] = h1;
                        l2 = 1;
                    }
                }
                if (l1 && (!p2 || l2)) break;
            }
            if (f1 && !l1) c1[2 * h0 + 1] = st2 + n2; /* :161-165 */
            if (p2 && f2 && !l2) c2[2 * h0 + 1] = st2 + n2;
        }
    }
}

int clf_oracle_fill_candidates(const uint8_t* blob, size_t size, int32_t* cand_h1,
                               int32_t* cand_h2) {
    event_t ev;
    if (event_parse(blob, size, &ev) != 0) return -1;
    fill_candidates(&ev, cand_h1, cand_h2);
    event_free(&ev);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Growable int list                                                    */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t* v;
    int64_t n, cap;
} ulist_t;

static int ulist_push(ulist_t* l, uint32_t x) {
    if (l->n == l->cap) {
        int64_t nc = l->cap ? 2 * l->cap : 1024;
        uint32_t* nv = (uint32_t*)realloc(l->v, (size_t)nc * sizeof(uint32_t));
        if (!nv) return -1;
        l->v = nv;
        l->cap = nc;
    }
    l->v[l->n++] = x;
    return 0;
}

/* ------------------------------------------------------------------ */
/* fitHitToTrack (Kernel.cl:15-35), blends kept literal                 */
/* ------------------------------------------------------------------ */
static float fit_hit_to_track(float tx, float ty, float x0, float y0, float z0, float h1_z,
                              float x2, float y2, float z2) {
    const float dz = z2 - z0;
    const float x_prediction = x0 + tx * dz;
    const float dx = fabsf(x_prediction - x2);
    const int tolx = dx < CLF_PARAM_TOLERANCE;
    const float y_prediction = y0 + ty * dz;
    const float dy = fabsf(y_prediction - y2);
    const int toly = dy < CLF_PARAM_TOLERANCE;
    const float scatterNum = (dx * dx) + (dy * dy);
    const float scatterDenom = 1.f / (z2 - h1_z);
    const float scatter = scatterNum * scatterDenom * scatterDenom;
    const int scat = scatter < CLF_MAX_SCATTER;
    const int cond = tolx && toly && scat;
    return (float)cond * scatter + (float)(!cond) * FLT_MAX;
}

/* ------------------------------------------------------------------ */
/* The whole event                                                      */
/* ------------------------------------------------------------------ */
#define UNSET 0xFFFFFFFFu

typedef struct {
    const event_t* ev;
    int32_t *c1, *c2;
    uint8_t* used;
    clf_track* tracks;    /* output records                       */
    int64_t n_tracks, cap_tracks;
    clf_track* tracklets; /* zero-filled, like the host init (:167) */
    int64_t n_tracklets, cap_tracklets;
    ulist_t ttf;  /* every ttf word ever inserted, in order */
    ulist_t weak; /* weak tracklet indices, in order        */
    clf_oracle_stats* st;
    int oom;
} state_t;

static int64_t new_track(state_t* s) {
    if (s->n_tracks >= s->cap_tracks) {
        s->oom = 1;
        return -1;
    }
    return s->n_tracks++;
}

/* trackForwarding for one ttf word (Kernel.cl:221-366). */
static void forward_one(state_t* s, uint32_t fulltrackno, int start4, int num4) {
    const event_t* ev = s->ev;
    const int track_flag = (fulltrackno & CLF_TTF_TRACKLET_BIT) != 0;
    const uint32_t skipped = (fulltrackno & CLF_TTF_SKIP_MASK) >> CLF_TTF_SKIP_SHIFT;
    uint32_t trackno = fulltrackno & CLF_TTF_INDEX_MASK;
    clf_track t = track_flag ? s->tracklets[trackno] : s->tracks[trackno];

    const int n = (int)t.hitsNum;
    const uint32_t h0n = t.hits[n - 2], h1n = t.hits[n - 1];
    const float x0 = ev->x[h0n], y0 = ev->y[h0n], z0 = ev->z[h0n];
    const float h1_x = ev->x[h1n], h1_y = ev->y[h1n], h1_z = ev->z[h1n];
    const float td = 1.0f / (h1_z - z0);
    const float txn = (h1_x - x0);
    const float tyn = (h1_y - y0);
    const float tx = txn * td;
    const float ty = tyn * td;

    float best_fit = FLT_MAX;
    uint32_t best_hit_h2 = UNSET;
    for (int kk = 0; kk < num4; ++kk) {
        const uint32_t h2 = (uint32_t)(start4 + kk);
        const float fit = fit_hit_to_track(tx, ty, x0, y0, z0, h1_z, ev->x[h2], ev->y[h2], ev->z[h2]);
        const int better = fit < best_fit;
        best_fit = (float)better * fit + (float)(!better) * best_fit;
        best_hit_h2 = better ? h2 : best_hit_h2;
    }
    s->st->fwd_fits += num4;

    if (best_fit != FLT_MAX) {
        if (best_hit_h2 == UNSET) {
            /* NaN best without a recorded hit: the reference reads an
             * uninitialised register here.  Only reachable under
             * CLF_FLAG_DEGENERATE_Z; the track is dropped. */
            s->st->flags |= CLF_FLAG_DEGENERATE_Z;
            return;
        }
        s->used[best_hit_h2] = 1; /* :316 */
        if (t.hitsNum >= CLF_MAX_TRACK_SIZE) {
            /* :321 would append a 25th hit past the 24-slot record
             * (undefined).  The mark above is the last defined effect: the
             * stored 24-hit record stays, the track is not re-queued. */
            s->st->flags |= CLF_FLAG_TRACK_FULL;
            s->st->n_overflow++;
            return;
        }
        t.hits[t.hitsNum++] = best_hit_h2;
        if (t.hitsNum <= 4) {
            s->used[t.hits[0]] = 1;
            s->used[t.hits[1]] = 1;
            s->used[t.hits[2]] = 1;
            const int64_t idx = new_track(s);
            if (idx < 0) return;
            trackno = (uint32_t)idx;
        }
        s->tracks[trackno] = t;
        if (t.hitsNum == CLF_MAX_TRACK_SIZE) s->st->n_full++;
        /* :343-345, a full record included */
        if (ulist_push(&s->ttf, trackno)) s->oom = 1;
    } else if (skipped <= CLF_MAX_SKIPPED_MODULES) {
        const uint32_t w = ((skipped + 1) << CLF_TTF_SKIP_SHIFT) | (fulltrackno & CLF_TTF_KEEP_MASK);
        if (ulist_push(&s->ttf, w)) s->oom = 1;
    } else if (t.hitsNum == 3) {
        if (ulist_push(&s->weak, trackno)) s->oom = 1;
    }
}

/* trackCreation for one gathered h0, local size (64, 4)
 * (Kernel.cl:408-567).  Rows y = 0..3 run their own sequential loops; the
 * results meet through best_fits (:530-546). */
static void seed_one(state_t* s, int h0, int f) {
    const event_t* ev = s->ev;
    const int start2 = ev->start[f - 2];
    const int start4 = ev->start[f - 4], num4 = ev->num[f - 4];
    const float x0 = ev->x[h0], y0 = ev->y[h0], z0 = ev->z[h0];
    /* :420-422 — z of the first hit of sensor f-2; only used when the
     * h1 window is non-empty, which needs that sensor to have hits */
    const float s1_z = ev->num[f - 2] > 0 ? ev->z[start2] : 0.f;
    const float h_dist = fabsf(s1_z - z0);
    const float dymax = CLF_PARAM_MAXYSLOPE * h_dist;
    const int first_h1 = s->c1[2 * h0];
    const int last_h1 = s->c1[2 * h0 + 1];
    const uint32_t num_h1 = (uint32_t)(last_h1 - first_h1);

    float row_best[CLF_SEED_ROWS];
    uint32_t row_h1[CLF_SEED_ROWS], row_h2[CLF_SEED_ROWS];
    for (int y = 0; y < CLF_SEED_ROWS; ++y) {
        float best_fit = FLT_MAX;
        uint32_t best_hit_h1 = UNSET, best_hit_h2 = UNSET;
        for (uint32_t j = 0;; ++j) {
            const uint32_t h1_element = CLF_SEED_ROWS * j + (uint32_t)y;
            if (!(h1_element < num_h1)) break; /* inside_bounds &= ... (:437) */
            const int h1 = first_h1 + (int)h1_element;
            const int is_h1_used = s->used[h1];
            float x1 = 0.f, y1 = 0.f, z1 = 0.f, dz_inverted = 0.f;
            if (!is_h1_used) {
                x1 = ev->x[h1];
                y1 = ev->y[h1];
                z1 = ev->z[h1];
                dz_inverted = 1.f / (z1 - z0);
            }
            const int first_h2 = s->c2[2 * h1];
            const int last_h2 = s->c2[2 * h1 + 1];
            if (first_h2 == -1) break; /* the row stops for good (:457) */
            if (is_h1_used) continue;
            s->st->seed_pairs++;
            for (int kk = 0; kk < num4; ++kk) { /* all of f-4, filtered (:485-488) */
                const int h2 = start4 + kk;
                if (!(h2 >= first_h2 && h2 < last_h2)) continue;
                s->st->seed_triplets++;
                const float x2 = ev->x[h2], y2 = ev->y[h2], z2 = ev->z[h2];
                const float z2_tz = (z2 - z0) * dz_inverted;
                const float x = x0 + (x1 - x0) * z2_tz;
                const float yy = y0 + (y1 - y0) * z2_tz;
                const float dx = x - x2;
                const float dy = yy - y2;
                if (fabsf(y1 - y0) < dymax && fabsf(dx) < CLF_PARAM_TOLERANCE &&
                    fabsf(dy) < CLF_PARAM_TOLERANCE) {
                    const float scatterNum = (dx * dx) + (dy * dy);
                    const float scatterDenom = 1.f / (z2 - z1);
                    const float scatter = scatterNum * scatterDenom * scatterDenom;
                    const int condition = scatter < CLF_MAX_SCATTER;
                    const float fit = (float)condition * scatter + (float)(!condition) * FLT_MAX;
                    const int better = fit < best_fit;
                    best_fit = (float)better * fit + (float)(!better) * best_fit;
                    best_hit_h1 = better ? (uint32_t)h1 : best_hit_h1;
                    best_hit_h2 = better ? (uint32_t)h2 : best_hit_h2;
                }
            }
        }
        row_best[y] = best_fit;
        row_h1[y] = best_hit_h1;
        row_h2[y] = best_hit_h2;
    }

    /* :534-546 — every row that holds a result scans the shared best_fits
     * with a strict "<" from FLT_MAX; the first minimal row accepts. */
    int any = 0;
    for (int y = 0; y < CLF_SEED_ROWS; ++y) any |= row_best[y] != FLT_MAX;
    if (!any) return;
    float best = FLT_MAX;
    int winner = -1;
    for (int i = 0; i < CLF_SEED_ROWS; ++i) {
        if (row_best[i] < best) {
            best = row_best[i];
            winner = i;
        }
    }
    if (winner < 0) return;
    if (row_h1[winner] == UNSET || row_h2[winner] == UNSET) {
        s->st->flags |= CLF_FLAG_DEGENERATE_Z;
        return;
    }
    /* :554-566 */
    if (s->n_tracklets >= s->cap_tracklets) {
        s->oom = 1;
        return;
    }
    const int64_t p = s->n_tracklets++;
    clf_track* tl = &s->tracklets[p];
    tl->hitsNum = 3;
    tl->hits[0] = (uint32_t)h0;
    tl->hits[1] = row_h1[winner];
    tl->hits[2] = row_h2[winner];
    if (ulist_push(&s->ttf, CLF_TTF_TRACKLET_BIT | (uint32_t)p)) s->oom = 1;
}

/* The h0 gather (Kernel.cl:720-754).  sh_hit_lastPointer restarts from 0
 * each sensor step (:684-686) and keeps counting across the chunks of the
 * step.  Chunk c reads elements prev+lane; a lane whose element is used
 * takes values v from the shared counter and looks at prev + shift + v until
 * it finds an unused or out-of-range element.  The r-th such lane in lane
 * order takes the r-th of those elements.  After the chunk prev = counter +
 * shift and shift grows by 64 (:753-754). */
typedef struct {
    uint32_t prev, shift, counter;
} gather_t;

static void gather_init(gather_t* g) {
    g->prev = 0;
    g->shift = CLF_GATHER_CHUNK;
    g->counter = 0;
}

/* Fills lane_e[0..63] with the element (0..hn0-1) each lane of the next chunk
 * processes, or -1.  Returns 0 when the step has no chunk left. */
static int gather_next_chunk(gather_t* g, uint32_t hn0, const uint8_t* used,
                             int lane_e[CLF_GATHER_CHUNK]) {
    if (!(g->prev < hn0)) return 0;
    int need[CLF_GATHER_CHUNK], n_need = 0;
    for (int x = 0; x < CLF_GATHER_CHUNK; ++x) {
        const uint32_t e = g->prev + (uint32_t)x;
        lane_e[x] = -1;
        if (e < hn0) {
            if (!used[e]) lane_e[x] = (int)e;
            else need[n_need++] = x;
        }
    }
    for (int r = 0; r < n_need; ++r) {
        for (;;) {
            const uint32_t e = g->prev + g->shift + g->counter++;
            if (!(e < hn0)) break;
            if (!used[e]) {
                lane_e[need[r]] = (int)e;
                break;
            }
        }
    }
    g->prev = g->counter + g->shift;
    g->shift += CLF_GATHER_CHUNK;
    return 1;
}

int clf_oracle_gather(int hn0, const uint8_t* used, int32_t* visits) {
    gather_t g;
    gather_init(&g);
    int lane_e[CLF_GATHER_CHUNK];
    int nv = 0;
    while (gather_next_chunk(&g, (uint32_t)(hn0 < 0 ? 0 : hn0), used, lane_e))
        for (int x = 0; x < CLF_GATHER_CHUNK; ++x)
            if (lane_e[x] >= 0) visits[nv++] = lane_e[x];
    return nv;
}

/* One sensor step of seeding: the gather, then trackCreation for each
 * chunk's h0 in lane order. */
static void seed_step(state_t* s, int f) {
    const event_t* ev = s->ev;
    const int start0 = ev->start[f];
    gather_t g;
    gather_init(&g);
    int lane_e[CLF_GATHER_CHUNK];
    while (gather_next_chunk(&g, (uint32_t)ev->num[f], s->used + start0, lane_e)) {
        for (int x = 0; x < CLF_GATHER_CHUNK; ++x) {
            if (lane_e[x] < 0) continue;
            s->st->seed_h0_visits++;
            seed_one(s, start0 + lane_e[x], f);
        }
    }
}

static int search_event(const event_t* ev, clf_track* out, int64_t out_cap, clf_oracle_stats* st) {
    memset(st, 0, sizeof(*st));
    st->flags = input_flags(ev);
    const int S = ev->S, N = ev->N;
    const int64_t cap = CLF_EVENT_TRACK_CAPACITY(N, S);
    state_t s;
    memset(&s, 0, sizeof(s));
    s.ev = ev;
    s.st = st;
    s.c1 = (int32_t*)malloc(sizeof(int32_t) * (2 * (size_t)N + 2));
    s.c2 = (int32_t*)malloc(sizeof(int32_t) * (2 * (size_t)N + 2));
    s.used = (uint8_t*)calloc((size_t)N + 1, 1);
    s.tracks = (clf_track*)calloc((size_t)cap, sizeof(clf_track));
    s.cap_tracks = cap;
    s.tracklets = (clf_track*)calloc((size_t)cap, sizeof(clf_track));
    s.cap_tracklets = cap;
    int rc = 0;
    if (!s.c1 || !s.c2 || !s.used || !s.tracks || !s.tracklets) {
        rc = -3;
        goto done;
    }
    fill_candidates(ev, s.c1, s.c2);

    int64_t prev_ttf = 0, last_ttf = 0;
    for (int f = S - 1; f >= 4; --f) { /* :671-767 */
        prev_ttf = last_ttf;
        last_ttf = s.ttf.n;
        const int start4 = ev->start[f - 4], num4 = ev->num[f - 4];
        for (int64_t i = prev_ttf; i < last_ttf; ++i) forward_one(&s, s.ttf.v[i], start4, num4);
        /* at f = 5 and 4 the gather still runs, but the h1 lie on sensors
         * 3 and 2 whose h2 windows stay -1, so every row stops at once */
        seed_step(&s, f);
        const int64_t live = s.ttf.n - prev_ttf; /* ring span this step */
        if (live > st->max_ttf_live) st->max_ttf_live = (int32_t)live;
        if (s.oom) break;
    }
    /* :769-792 — tracklet words of the last step become weak */
    prev_ttf = last_ttf;
    last_ttf = s.ttf.n;
    for (int64_t i = prev_ttf; i < last_ttf; ++i) {
        const uint32_t w = s.ttf.v[i];
        if (w & CLF_TTF_TRACKLET_BIT) {
            if (ulist_push(&s.weak, w & CLF_TTF_INDEX_MASK)) s.oom = 1;
        }
    }
    /* :796-815 — weak tracklets whose three hits are all unused */
    for (int64_t i = 0; i < s.weak.n; ++i) {
        const clf_track t = s.tracklets[s.weak.v[i]];
        if (!s.used[t.hits[0]] && !s.used[t.hits[1]] && !s.used[t.hits[2]]) {
            const int64_t idx = new_track(&s);
            if (idx < 0) break;
            s.tracks[idx] = t;
        }
    }
    if (s.oom) {
        rc = -3;
        goto done;
    }
    st->n_tracks = (int32_t)s.n_tracks;
    st->n_tracklets = (int32_t)s.n_tracklets;
    st->n_weak = (int32_t)s.weak.n;
    if (s.n_tracks > CLF_MAX_TRACKS) st->flags |= CLF_FLAG_TRACKS_OVER;
    if (s.n_tracklets > N) st->flags |= CLF_FLAG_TRACKLETS_OVER;
    if (s.weak.n > N) st->flags |= CLF_FLAG_WEAK_OVER;
    if (st->max_ttf_live > CLF_TTF_MODULO) st->flags |= CLF_FLAG_TTF_OVER;
    if (s.n_tracks > out_cap) {
        rc = -2;
        goto done;
    }
    if (out && s.n_tracks) memcpy(out, s.tracks, (size_t)s.n_tracks * sizeof(clf_track));
done:
    free(s.c1);
    free(s.c2);
    free(s.used);
    free(s.tracks);
    free(s.tracklets);
    free(s.ttf.v);
    free(s.weak.v);
    return rc;
}

int clf_oracle_search_event(const uint8_t* blob, size_t size, uint8_t* out_tracks,
                            int64_t track_capacity, clf_oracle_stats* stats) {
    event_t ev;
    clf_oracle_stats local;
    if (!stats) stats = &local;
    if (event_parse(blob, size, &ev) != 0) {
        memset(stats, 0, sizeof(*stats));
        stats->flags = CLF_FLAG_INVALID;
        return -1;
    }
    int rc = search_event(&ev, (clf_track*)out_tracks, track_capacity, stats);
    event_free(&ev);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Thread pool over events (CPU baseline)                               */
/* ------------------------------------------------------------------ */
typedef struct {
    int n;
    const uint8_t* const* blobs;
    const size_t* sizes;
    uint8_t* const* outs;
    const int64_t* caps;
    clf_oracle_stats* stats;
    int next;
    int failed;
    pthread_mutex_t mu;
} batch_t;

static void* batch_worker(void* arg) {
    batch_t* b = (batch_t*)arg;
    for (;;) {
        pthread_mutex_lock(&b->mu);
        const int i = b->next++;
        pthread_mutex_unlock(&b->mu);
        if (i >= b->n) break;
        const int rc = clf_oracle_search_event(b->blobs[i], b->sizes[i], b->outs ? b->outs[i] : NULL,
                                               b->caps ? b->caps[i] : 0, &b->stats[i]);
        if (rc != 0 && !(rc == -2 && !b->outs)) {
            pthread_mutex_lock(&b->mu);
            b->failed++;
            pthread_mutex_unlock(&b->mu);
        }
    }
    return NULL;
}

int clf_oracle_search_batch(int n, const uint8_t* const* blobs, const size_t* sizes,
                            uint8_t* const* out_tracks, const int64_t* track_capacity,
                            clf_oracle_stats* stats, int n_threads) {
    batch_t b;
    b.n = n;
    b.blobs = blobs;
    b.sizes = sizes;
    b.outs = out_tracks;
    b.caps = track_capacity;
    b.stats = stats;
    b.next = 0;
    b.failed = 0;
    pthread_mutex_init(&b.mu, NULL);
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    int started = 0;
    for (int t = 0; t < n_threads; ++t) {
        if (pthread_create(&th[t], NULL, batch_worker, &b) == 0) started++;
        else break;
    }
    if (started == 0) batch_worker(&b);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&b.mu);
    return b.failed;
}

/* ------------------------------------------------------------------ */
/* preorder_by_x (Tools.cpp:46-66) and its partition (Tools.cpp:117-157) */
/* ------------------------------------------------------------------ */
static void swap_hits(float* a, float* b, float* c, uint32_t* d, int i, int j) {
    float t = a[i]; a[i] = a[j]; a[j] = t;
    t = b[i]; b[i] = b[j]; b[j] = t;
    t = c[i]; c[i] = c[j]; c[j] = t;
    uint32_t u = d[i]; d[i] = d[j]; d[j] = u;
}

/* Pivot = first element; right scan skips "> pivot", left scan skips
 * "<= pivot"; the pivot goes to the right index at the end. */
static int partition_first_pivot(float* a, float* b, float* c, uint32_t* d, int lo, int hi) {
    const float pivot = a[lo];
    int left = lo, right = hi;
    while (left < right) {
        while (a[right] > pivot) right--;
        while (left < right && a[left] <= pivot) left++;
        if (left < right) swap_hits(a, b, c, d, left, right);
    }
    swap_hits(a, b, c, d, right, lo);
    return right;
}

/* The two sub-ranges are disjoint, so visiting them from an explicit stack
 * gives the same permutation as the recursion. */
static void sort_range(float* a, float* b, float* c, uint32_t* d, int lo, int hi) {
    int stack_lo[64 + CLF_MAX_SENSOR_HITS / 32], stack_hi[64 + CLF_MAX_SENSOR_HITS / 32];
    int* slo = stack_lo;
    int* shi = stack_hi;
    int* big_lo = NULL;
    int* big_hi = NULL;
    int cap = (int)(sizeof(stack_lo) / sizeof(stack_lo[0]));
    int sp = 0;
    slo[sp] = lo;
    shi[sp] = hi;
    sp++;
    while (sp > 0) {
        sp--;
        const int l = slo[sp], h = shi[sp];
        if (!(l < h)) continue;
        const int p = partition_first_pivot(a, b, c, d, l, h);
        if (sp + 2 > cap) {
            int nc = cap * 2;
            int* nl = (int*)malloc(sizeof(int) * (size_t)nc);
            int* nh = (int*)malloc(sizeof(int) * (size_t)nc);
            memcpy(nl, slo, sizeof(int) * (size_t)sp);
            memcpy(nh, shi, sizeof(int) * (size_t)sp);
            free(big_lo);
            free(big_hi);
            big_lo = slo = nl;
            big_hi = shi = nh;
            cap = nc;
        }
        slo[sp] = l;
        shi[sp] = p - 1;
        sp++;
        slo[sp] = p + 1;
        shi[sp] = h;
        sp++;
    }
    free(big_lo);
    free(big_hi);
}

int clf_oracle_presort(uint8_t* blob, size_t size) {
    event_t ev;
    if (event_parse(blob, size, &ev) != 0) return -1;
    /* hitStarts are taken as contiguous from 0, as Tools.cpp:53-63 does */
    int acc = 0;
    for (int s = 0; s < ev.S; ++s) {
        const int n = ev.num[s];
        if (acc + n > ev.N) break;
        sort_range(ev.x, ev.y, ev.z, ev.id, acc, acc + n - 1);
        acc += n;
    }
    const size_t off = 8 + 12 * (size_t)ev.S;
    memcpy(blob + off, ev.id, 4 * (size_t)ev.N);
    memcpy(blob + off + 4 * (size_t)ev.N, ev.x, 4 * (size_t)ev.N);
    memcpy(blob + off + 8 * (size_t)ev.N, ev.y, 4 * (size_t)ev.N);
    memcpy(blob + off + 12 * (size_t)ev.N, ev.z, 4 * (size_t)ev.N);
    event_free(&ev);
    return 0;
}
