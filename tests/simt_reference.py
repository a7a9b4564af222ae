"""Work-item-level simulation of the reference kernel (tests only).

A second restatement of clSearchByTriplets (Kernel.cl:597-816), written at
the level the reference is written at: a work-group of 64 x Y work-items
with private variables, local memory, global buffers, the atomic counters
(KernelInvoker.cpp:92-108 layout), the 2000-slot ttf ring, and the barriers
(Kernel.cl:432,532,673,700,724,750,769,794) as phase boundaries.  Inside a
phase the work-items run one at a time in a schedule chosen by the caller;
the h0 gather's lanes interleave at each atomic_add (Kernel.cl:739).  Any
schedule is an execution the hardware may produce, so the canonical track
multiset must not depend on it.

Pure Python with numpy float32 scalars (FP32, no contraction, IEEE
division), so only tiny events are practical.  The C oracle is checked
against it; it shares no code with the oracle.
"""
from __future__ import annotations

import numpy as np

F = np.float32
FLT_MAX = F(np.finfo(np.float32).max)
NUMTHREADS_X = 64
MAX_TRACKS = 3000
MAX_TRACK_SIZE = 24
TTF_MODULO = 2000
MAX_SKIPPED_MODULES = 3
PARAM_MAXXSLOPE_CANDIDATES = F(0.7)
PARAM_MAXYSLOPE = F(0.7)
PARAM_TOLERANCE = F(0.6)
PARAM_TOLERANCE_CANDIDATES = F(0.6)
MAX_SCATTER = F(0.000016)


class RingOverflow(Exception):
    pass


class SimtKernel:
    """One work-group running one event."""

    def __init__(self, sensor_Zs, hitStarts, hitNums, X, Y, Z, Y_threads=4, rng=None):
        self.S = len(sensor_Zs)
        self.N = len(X)
        self.Zs = [int(v) for v in sensor_Zs]
        self.starts = [int(v) for v in hitStarts]
        self.nums = [int(v) for v in hitNums]
        self.X = [F(v) for v in X]
        self.Yh = [F(v) for v in Y]
        self.Z = [F(v) for v in Z]
        self.ly = Y_threads
        self.rng = rng
        # global buffers, initialised as KernelInvoker.cpp:160-168 does
        self.hit_used = [False] * self.N
        self.hit_candidates = [-1] * (2 * self.N)
        self.hit_h2_candidates = [-1] * (2 * self.N)
        self.tracks = [None] * MAX_TRACKS
        self.tracklets = [None] * max(self.N, 1)
        self.weak_tracks = [0] * max(self.N, 1)
        self.ttf = [0] * TTF_MODULO
        self.tracks_insertPointer = 0
        self.weaktracks_insertPointer = 0
        self.tracklets_insertPointer = 0
        self.ttf_insertPointer = 0
        self.sh_hit_lastPointer = 0
        self.max_numhits_to_process = 0
        self.best_fits = [FLT_MAX] * (NUMTHREADS_X * self.ly)
        # local memory
        self.sensor_data = [0] * 6
        self.sh_hit_process = [-1] * NUMTHREADS_X
        # tracks whose 25th hit the reference would write past its record
        self.overflows = 0

    # -- schedules ---------------------------------------------------------
    def order(self, items):
        items = list(items)
        if self.rng is not None:
            self.rng.shuffle(items)
        return items

    def items(self):
        return [(x, y) for y in range(self.ly) for x in range(NUMTHREADS_X)]

    # -- fillCandidates (Kernel.cl:50-178) ---------------------------------
    def fill_candidates(self):
        S, bdp = self.S, NUMTHREADS_X * self.ly
        first_sensor = S - 1
        while first_sensor >= 2:
            second_sensor = first_sensor - 2
            process_h1 = first_sensor >= 4
            process_h2 = first_sensor <= S - 3
            z_s0 = self.Zs[first_sensor + 2] if process_h2 else 0
            z_s2 = self.Zs[second_sensor] if process_h2 else 0
            hn = self.nums[first_sensor]
            for i in range((hn + bdp - 1) // bdp):
                for (x, y) in self.order(self.items()):
                    h0_element = bdp * i + y * NUMTHREADS_X + x
                    if not h0_element < hn:
                        continue
                    first_h1 = last_h1 = first_h2 = last_h2 = False
                    h0_index = self.starts[first_sensor] + h0_element
                    h0x, h0z = self.X[h0_index], self.Z[h0_index]
                    hs2, hn2 = self.starts[second_sensor], self.nums[second_sensor]
                    if process_h2:
                        h_dist = abs(h0z - F(z_s0))
                        dxmax = PARAM_MAXXSLOPE_CANDIDATES * h_dist
                        x0_min = h0x - dxmax
                        x0_max = h0x + dxmax
                        z2_tz = (F(z_s2) - F(z_s0)) / (h0z - F(z_s0))
                        xx = x0_max + (h0x - x0_max) * z2_tz
                        xmin_h2 = xx - PARAM_TOLERANCE_CANDIDATES
                        xx = x0_min + (h0x - x0_min) * z2_tz
                        xmax_h2 = xx + PARAM_TOLERANCE_CANDIDATES
                    if first_sensor >= 4:
                        for h1_element in range(hn2):
                            h1_index = hs2 + h1_element
                            h1x, h1z = self.X[h1_index], self.Z[h1_index]
                            if process_h1 and not last_h1:
                                dxmax = PARAM_MAXXSLOPE_CANDIDATES * abs(h1z - h0z)
                                tol = abs(h1x - h0x) < dxmax
                                if not first_h1 and tol:
                                    self.hit_candidates[2 * h0_index] = h1_index
                                    first_h1 = True
                                elif first_h1 and not tol:
                                    self.hit_candidates[2 * h0_index + 1] = h1_index
                                    last_h1 = True
                            if process_h2 and not last_h2:
                                if not first_h2 and h1x > xmin_h2:
                                    self.hit_h2_candidates[2 * h0_index] = h1_index
                                    first_h2 = True
                                elif first_h2 and h1x > xmax_h2:
                                    self.hit_h2_candidates[2 * h0_index + 1] = h1_index
                                    last_h2 = True
                            if (not process_h1 or last_h1) and (not process_h2 or last_h2):
                                break
                        if process_h1 and first_h1 and not last_h1:
                            self.hit_candidates[2 * h0_index + 1] = hs2 + hn2
                        if process_h2 and first_h2 and not last_h2:
                            self.hit_h2_candidates[2 * h0_index + 1] = hs2 + hn2
            first_sensor -= 1

    # -- fitHitToTrack (Kernel.cl:15-35) ------------------------------------
    @staticmethod
    def fit(tx, ty, h0, h1_z, h2):
        dz = h2[2] - h0[2]
        x_prediction = h0[0] + tx * dz
        dx = abs(x_prediction - h2[0])
        tolx = dx < PARAM_TOLERANCE
        y_prediction = h0[1] + ty * dz
        dy = abs(y_prediction - h2[1])
        toly = dy < PARAM_TOLERANCE
        scatterNum = (dx * dx) + (dy * dy)
        scatterDenom = F(1.0) / (h2[2] - h1_z)
        scatter = scatterNum * scatterDenom * scatterDenom
        cond = tolx and toly and scatter < MAX_SCATTER
        with np.errstate(all="ignore"):
            return F(cond) * scatter + F(not cond) * FLT_MAX

    def hit(self, i):
        return (self.X[i], self.Yh[i], self.Z[i])

    def ttf_push(self, value):
        p = self.ttf_insertPointer
        self.ttf_insertPointer += 1
        self.ttf[p % TTF_MODULO] = value

    # -- trackForwarding (Kernel.cl:199-368) for one work-item ---------------
    def forward_item(self, ttf_element, diff_ttf, prev_ttf):
        if not ttf_element < diff_ttf:
            return
        sd = self.sensor_data
        fulltrackno = self.ttf[(prev_ttf + ttf_element) % TTF_MODULO]
        track_flag = (fulltrackno & 0x80000000) == 0x80000000
        skipped_modules = (fulltrackno & 0x70000000) >> 28
        trackno = fulltrackno & 0x0FFFFFFF
        src = self.tracklets if track_flag else self.tracks
        t = {"hitsNum": src[trackno]["hitsNum"], "hits": list(src[trackno]["hits"])}
        n = t["hitsNum"]
        h0 = self.hit(t["hits"][n - 2])
        h1 = self.hit(t["hits"][n - 1])
        td = F(1.0) / (h1[2] - h0[2])
        tx = (h1[0] - h0[0]) * td
        ty = (h1[1] - h0[1]) * td
        best_fit = FLT_MAX
        best_hit_h2 = None
        for kk in range(sd[5]):
            h2_index = sd[2] + kk
            fit = self.fit(tx, ty, h0, h1[2], self.hit(h2_index))
            better = fit < best_fit
            with np.errstate(all="ignore"):
                best_fit = F(better) * fit + F(not better) * best_fit
            best_hit_h2 = h2_index if better else best_hit_h2
        if best_fit != FLT_MAX:
            self.hit_used[best_hit_h2] = True  # :316
            if t["hitsNum"] >= MAX_TRACK_SIZE:
                # :321 writes hits[24], past the private record: the
                # simulation follows the reference up to here and stops
                self.overflows += 1
                return
            t["hits"][t["hitsNum"]] = best_hit_h2
            t["hitsNum"] += 1
            if t["hitsNum"] <= 4:
                for k in range(3):
                    self.hit_used[t["hits"][k]] = True
                trackno = self.tracks_insertPointer
                self.tracks_insertPointer += 1
            self.tracks[trackno] = t
            self.ttf_push(trackno)
        elif skipped_modules <= MAX_SKIPPED_MODULES:
            self.ttf_push(((skipped_modules + 1) << 28) | (fulltrackno & 0x8FFFFFFF))
        elif t["hitsNum"] == 3:
            p = self.weaktracks_insertPointer
            self.weaktracks_insertPointer += 1
            self.weak_tracks[p] = trackno

    # -- the h0 gather (Kernel.cl:724-749), lanes interleaving at each atomic
    def gather(self, sh_hit_prevPointer, shift_lastPointer):
        hn0 = self.sensor_data[3]
        lanes = {}
        for x in range(NUMTHREADS_X):
            sh_element = sh_hit_prevPointer + x
            inside = sh_element < hn0
            h0_index = self.sensor_data[0] + sh_element
            used = self.hit_used[h0_index] if inside else True
            lanes[x] = [inside, used, h0_index]
        active = [x for x in range(NUMTHREADS_X) if lanes[x][0] and lanes[x][1]]
        while active:
            x = active[0] if self.rng is None else active[int(self.rng.integers(len(active)))]
            v = self.sh_hit_lastPointer
            self.sh_hit_lastPointer += 1  # atomic_add
            sh_element = sh_hit_prevPointer + shift_lastPointer + v
            inside = sh_element < hn0
            h0_index = self.sensor_data[0] + sh_element
            used = self.hit_used[h0_index] if inside else True
            lanes[x] = [inside, used, h0_index]
            if not (inside and used):
                active.remove(x)
        for x in range(NUMTHREADS_X):
            inside, used, h0_index = lanes[x]
            self.sh_hit_process[x] = h0_index if (inside and not used) else -1

    # -- trackCreation (Kernel.cl:392-568) ------------------------------------
    def track_creation(self):
        sd = self.sensor_data
        ly = self.ly
        priv = {}
        # phase up to the barrier at :432
        for (x, y) in self.order(self.items()):
            p = {"inside": False, "num": 0, "best_fit": FLT_MAX, "bh1": None, "bh2": None}
            h0_index = self.sh_hit_process[x]
            p["h0_index"] = h0_index
            p["inside"] = h0_index != -1
            if p["inside"]:
                p["h0"] = self.hit(h0_index)
                s1_z = self.Z[sd[1]] if sd[1] < self.N else F(0)
                p["dymax"] = PARAM_MAXYSLOPE * abs(s1_z - p["h0"][2])
                p["first_h1"] = self.hit_candidates[2 * h0_index]
                last_h1 = self.hit_candidates[2 * h0_index + 1]
                p["num"] = (last_h1 - p["first_h1"]) & 0xFFFFFFFF
                self.max_numhits_to_process = max(self.max_numhits_to_process, p["num"])
            priv[(x, y)] = p
        # phase up to :532
        for (x, y) in self.order(self.items()):
            p = priv[(x, y)]
            j = 0
            while j < (self.max_numhits_to_process + ly - 1) // ly:
                h1_element = ly * j + y
                p["inside"] = p["inside"] and h1_element < p["num"]
                is_h1_used = True
                if p["inside"]:
                    h1_index = p["first_h1"] + h1_element
                    is_h1_used = self.hit_used[h1_index]
                    if not is_h1_used:
                        h1 = self.hit(h1_index)
                        dz_inverted = F(1.0) / (h1[2] - p["h0"][2])
                    first_h2 = self.hit_h2_candidates[2 * h1_index]
                    last_h2 = self.hit_h2_candidates[2 * h1_index + 1]
                    p["inside"] = p["inside"] and first_h2 != -1
                if p["inside"] and not is_h1_used:
                    h0 = p["h0"]
                    for kk in range(sd[5]):
                        h2_index = sd[2] + kk
                        if h2_index >= first_h2 and h2_index < last_h2:
                            h2 = self.hit(h2_index)
                            z2_tz = (h2[2] - h0[2]) * dz_inverted
                            xx = h0[0] + (h1[0] - h0[0]) * z2_tz
                            yy = h0[1] + (h1[1] - h0[1]) * z2_tz
                            dx = xx - h2[0]
                            dy = yy - h2[1]
                            if abs(h1[1] - h0[1]) < p["dymax"] and abs(dx) < PARAM_TOLERANCE and \
                                    abs(dy) < PARAM_TOLERANCE:
                                scatterNum = (dx * dx) + (dy * dy)
                                scatterDenom = F(1.0) / (h2[2] - h1[2])
                                scatter = scatterNum * scatterDenom * scatterDenom
                                cond = scatter < MAX_SCATTER
                                with np.errstate(all="ignore"):
                                    fit = F(cond) * scatter + F(not cond) * FLT_MAX
                                    better = fit < p["best_fit"]
                                    p["best_fit"] = F(better) * fit + F(not better) * p["best_fit"]
                                p["bh1"] = h1_index if better else p["bh1"]
                                p["bh2"] = h2_index if better else p["bh2"]
                j += 1
            self.best_fits[x * ly + y] = p["best_fit"]
        # phase after :532
        accept = []
        for (x, y) in self.items():
            p = priv[(x, y)]
            if p["h0_index"] != -1 and p["best_fit"] != FLT_MAX:
                best = FLT_MAX
                winner = -1
                for i in range(ly):
                    fit = self.best_fits[x * ly + i]
                    if fit < best:
                        best = fit
                        winner = i
                if y == winner:
                    accept.append((x, y))
        # lane order unless a random schedule is asked for
        accept.sort()
        for (x, y) in self.order(accept):
            p = priv[(x, y)]
            trackP = self.tracklets_insertPointer
            self.tracklets_insertPointer += 1
            self.tracklets[trackP] = {"hitsNum": 3,
                                      "hits": [self.sh_hit_process[x], p["bh1"], p["bh2"]] + [0] * 21}
            self.ttf_push(0x80000000 | trackP)

    # -- clSearchByTriplets (Kernel.cl:597-816) ------------------------------
    def run(self):
        bdp = NUMTHREADS_X * self.ly
        self.fill_candidates()
        first_sensor = self.S - 1
        last_ttf = 0
        while first_sensor >= 4:
            # phase :673-700
            for i in range(6):
                sensor_number = first_sensor - (i % 3) * 2
                self.sensor_data[i] = (self.starts if i < 3 else self.nums)[sensor_number]
            self.sh_hit_lastPointer = 0
            self.max_numhits_to_process = 0
            prev_ttf = last_ttf
            last_ttf = self.ttf_insertPointer
            diff_ttf = last_ttf - prev_ttf
            # forwarding
            n_iter = (diff_ttf + bdp - 1) // bdp
            for i in range(n_iter):
                for (x, y) in self.order(self.items()):
                    self.forward_item(bdp * i + y * NUMTHREADS_X + x, diff_ttf, prev_ttf)
            if self.ttf_insertPointer - prev_ttf > TTF_MODULO:
                raise RingOverflow()
            # seeding
            sh_hit_prevPointer = 0
            shift_lastPointer = NUMTHREADS_X
            while sh_hit_prevPointer < self.sensor_data[3]:
                self.gather(sh_hit_prevPointer, shift_lastPointer)
                sh_hit_prevPointer = self.sh_hit_lastPointer + shift_lastPointer
                shift_lastPointer += NUMTHREADS_X
                self.track_creation()
            if self.ttf_insertPointer - prev_ttf > TTF_MODULO:
                raise RingOverflow()
            first_sensor -= 1
        # tail :769-792
        prev_ttf = last_ttf
        last_ttf = self.ttf_insertPointer
        diff_ttf = last_ttf - prev_ttf
        for i in range((diff_ttf + bdp - 1) // bdp):
            for (x, y) in self.order(self.items()):
                e = bdp * i + y * NUMTHREADS_X + x
                if e < diff_ttf:
                    full = self.ttf[(prev_ttf + e) % TTF_MODULO]
                    if full & 0x80000000:
                        p = self.weaktracks_insertPointer
                        self.weaktracks_insertPointer += 1
                        self.weak_tracks[p] = full & 0x0FFFFFFF
        # weak filter :796-815
        total = self.weaktracks_insertPointer
        for i in range((total + bdp - 1) // bdp):
            for (x, y) in self.order(self.items()):
                w = bdp * i + y * NUMTHREADS_X + x
                if w < total:
                    t = self.tracklets[self.weak_tracks[w]]
                    if not any(self.hit_used[h] for h in t["hits"][:3]):
                        k = self.tracks_insertPointer
                        self.tracks_insertPointer += 1
                        self.tracks[k] = {"hitsNum": 3, "hits": list(t["hits"])}
        return [tuple(int(h) for h in self.tracks[k]["hits"][: self.tracks[k]["hitsNum"]])
                for k in range(self.tracks_insertPointer)]


def run_event(ev, rng=None, Y_threads=4):
    """Track tuples (hit indices, in output order) of one Event."""
    k = SimtKernel(ev.sensor_Zs, ev.hitStarts, ev.hitNums, ev.X, ev.Y, ev.Z, Y_threads, rng)
    return k.run(), k


def gather_lanes(used, rng=None):
    """The gather of one sensor step alone: visited element indices in lane
    order, chunk after chunk (Kernel.cl:720-754)."""
    n = len(used)
    k = SimtKernel([0], [0], [n], np.zeros(n), np.zeros(n), np.zeros(n), 4, rng)
    k.hit_used = [bool(u) for u in used]
    k.sensor_data = [0, 0, 0, n, 0, 0]
    out = []
    prev, shift = 0, NUMTHREADS_X
    while prev < n:
        k.gather(prev, shift)
        out.extend(h for h in k.sh_hit_process if h != -1)
        prev = k.sh_hit_lastPointer + shift
        shift += NUMTHREADS_X
    return out
