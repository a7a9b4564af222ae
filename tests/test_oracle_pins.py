"""Pins of the CPU oracle (CPU only).

What fixes each part of the oracle, since the reference kernel cannot run
here (OpenCL C needing an OpenCL device runtime, which this image lacks):

* layout + presort: the reference's OWN host code (Tools.cpp compiled from
  its source into oracle/_ref/) and the golden vectors it wrote;
* fillCandidates, forwarding, the gather, (64,4) seeding, ring, tail: a
  second, independent restatement at work-item level with barriers and
  atomics (tests/simt_reference.py), under the lane-order schedule (same
  bytes) and random schedules (same multiset);
* arithmetic: closed forms on noiseless straight lines (clone structure,
  scatter and slope cut boundaries), recomputed cuts, and MC-truth
  efficiency on the synthetic generator.
"""
from __future__ import annotations

import collections
import hashlib
import itertools
import json
import os

import numpy as np
import pytest

from cl_forward_amd import event as ev_mod
from cl_forward_amd import generator
from tests import simt_reference as simt
from tests.helpers import (add_noise, line_event, raw_blob, sparse_line_event, sub_event,
                           window_pin_events)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F = np.float32


# ---------------------------------------------------------------------------
# layout (A1) and presort (A4)
# ---------------------------------------------------------------------------
def test_layout_offsets_match_reference_parser(mcdata_raw):
    with open(os.path.join(GOLDEN, "mcdata_layout_ref.json")) as f:
        ref = json.load(f)
    for k, blob in mcdata_raw.items():
        ev = ev_mod.parse_event(blob)
        S, N = ev.n_sensors, ev.n_hits
        mine = [0, 4, 8, 8 + 4 * S, 8 + 8 * S, 8 + 12 * S, 8 + 12 * S + 4 * N,
                8 + 12 * S + 8 * N, 8 + 12 * S + 12 * N]
        assert mine == ref[k], k
        assert len(blob) == 8 + 12 * S + 16 * N
        assert ev.pack() == blob


def test_dat_framing_roundtrip(mcdata_raw):
    blob = mcdata_raw["0"]
    name, data = ev_mod.unframe(ev_mod.frame(blob) + b"trailing block ignored")
    assert name == ev_mod.HANDLER_NAME and data == blob


def test_oracle_presort_matches_reference_golden(oracle_mod, mcdata_raw, mcdata_sorted):
    for k, blob in mcdata_raw.items():
        assert oracle_mod.presort(blob) == mcdata_sorted[k], k


def _one_sensor_blob(xs):
    n = len(xs)
    return raw_blob(1, n, [0], [0], [n], xs, np.arange(n) * 0.5, np.zeros(n), np.arange(100, 100 + n))


@pytest.mark.parametrize("values", [[1, 1], [2, 1, 1], [3, 3, 1, 2, 2], [1, 2, 3, 4, 5, 6, 7],
                                    [7, 6, 5, 4, 3, 2, 1], [4, 1, 4, 1, 4, 1], [5, 5, 5, 5, 5]])
def test_presort_all_permutations_vs_reference(oracle_mod, values):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    for perm in set(itertools.permutations(values)):
        b = _one_sensor_blob(np.array(perm, np.float32))
        assert oracle_mod.presort(b) == oracle_mod.ref_presort(b), perm


def test_presort_tie_order_hand_traced(oracle_mod):
    # [1, 1]: left scans past the second 1 ("<= pivot"), then the final pivot
    # swap exchanges the two equal keys, so their payloads trade places.
    b = _one_sensor_blob(np.array([1, 1], np.float32))
    ev = ev_mod.parse_event(oracle_mod.presort(b))
    assert ev.hit_IDs.tolist() == [101, 100]


def test_presort_properties(oracle_mod, mcdata_raw):
    for k in ("0", "77"):
        raw = ev_mod.parse_event(mcdata_raw[k])
        srt = ev_mod.parse_event(oracle_mod.presort(mcdata_raw[k]))
        for s in range(raw.n_sensors):
            b, n = int(raw.hitStarts[s]), int(raw.hitNums[s])
            assert np.all(np.diff(srt.X[b:b + n]) >= 0)
            rows = lambda e: sorted(zip(e.X[b:b + n].tolist(), e.Y[b:b + n].tolist(),
                                        e.Z[b:b + n].tolist(), e.hit_IDs[b:b + n].tolist()))
            assert rows(raw) == rows(srt)


# ---------------------------------------------------------------------------
# the gather (A11, Kernel.cl:720-754)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("hn", [0, 1, 63, 64, 65, 127, 128, 129, 130, 200, 257, 400])
def test_gather_vs_work_item_simulation(oracle_mod, hn):
    rng = np.random.default_rng(hn)
    for frac in (0.0, 0.1, 0.5, 0.9, 1.0):
        used = (rng.random(hn) < frac).astype(np.uint8)
        o = oracle_mod.gather(used)
        assert o == simt.gather_lanes(used), (hn, frac)
        for seed in range(3):
            r = simt.gather_lanes(used, rng=np.random.default_rng(seed))
            assert collections.Counter(r) == collections.Counter(o), (hn, frac, seed)


def test_gather_small_sensors_visit_each_unused_hit_once(oracle_mod):
    rng = np.random.default_rng(1)
    for hn in range(0, 129):
        used = (rng.random(hn) < 0.4).astype(np.uint8)
        assert sorted(oracle_mod.gather(used)) == [i for i in range(hn) if not used[i]]


def test_gather_skip_and_repeat_hand_traced(oracle_mod):
    # 300 hits, only element 70 used.  Chunk 0 reads [0,64).  Chunk 1 reads
    # [64,128); lane 6 (element 70) takes counter value 0 and reads
    # 64 + 128 + 0 = 192.  Then prev = 1 + 128 = 129: element 128 is never
    # read and chunk 2 reads 192 again.
    used = np.zeros(300, np.uint8)
    used[70] = 1
    v = oracle_mod.gather(used)
    c = collections.Counter(v)
    assert c[192] == 2 and 128 not in c and 70 not in c
    assert sum(c.values()) == 299


# ---------------------------------------------------------------------------
# whole events against the work-item simulation
# ---------------------------------------------------------------------------
def _compare_with_simt(oracle_mod, ev, schedules=0):
    out, st = oracle_mod.search_event(ev.pack())
    ora = ev_mod.track_lists(out)
    sim, kern = simt.run_event(ev)
    assert sim == ora, f"{len(sim)} vs {len(ora)} tracks"
    for seed in range(schedules):
        s2, _ = simt.run_event(ev, rng=np.random.default_rng(seed))
        assert sorted(s2) == sorted(ora), f"schedule {seed}"
    return st, kern


@pytest.mark.parametrize("key", ["1", "11", "15", "23", "43", "0"])
def test_oracle_vs_simt_mcdata(oracle_mod, mcdata_sorted, key):
    ev = ev_mod.parse_event(mcdata_sorted[key])
    st, kern = _compare_with_simt(oracle_mod, ev)
    assert st["flags"] == 0
    # the candidate windows too
    o1, o2 = oracle_mod.fill_candidates(mcdata_sorted[key])
    assert o1.tolist() == kern.hit_candidates and o2.tolist() == kern.hit_h2_candidates


def test_oracle_vs_simt_mcdata_large(oracle_mod, mcdata_sorted):
    """49.dat: 5,297 hits, sensors up to 146 hits (more than two gather chunks)."""
    _compare_with_simt(oracle_mod, ev_mod.parse_event(mcdata_sorted["49"]))


@pytest.mark.parametrize("key", ["1", "15"])
def test_oracle_vs_simt_mcdata_unsorted(oracle_mod, mcdata_raw, key):
    """Hits as stored (not x-sorted): windows and row stops become irregular."""
    ev = ev_mod.parse_event(mcdata_raw[key])
    st, _ = _compare_with_simt(oracle_mod, ev, schedules=1)
    assert st["flags"] & (1 << 6)


def test_oracle_vs_simt_random_schedules(oracle_mod, mcdata_sorted):
    ev = sub_event(ev_mod.parse_event(mcdata_sorted["0"]), 20, 40)
    _compare_with_simt(oracle_mod, ev, schedules=4)


def test_oracle_vs_simt_gather_anomaly(oracle_mod):
    """A 12-plane event whose sensor 8 gets 250 noise hits: at its step the
    track hits there are already used, the sensor spans five gather chunks,
    and the reference's gather repeats and skips hits."""
    base = sub_event(generator.generate_event(11, n_tracks=40), 36, 48)
    ev = add_noise(base, 8, 250, seed=3)
    # count the repeats the work-item gather produces at that step
    kern = simt.SimtKernel(ev.sensor_Zs, ev.hitStarts, ev.hitNums, ev.X, ev.Y, ev.Z)
    seen = []
    orig = kern.gather

    def spy(prev, shift):
        orig(prev, shift)
        if kern.sensor_data[0] == int(ev.hitStarts[8]):
            seen.extend(h for h in kern.sh_hit_process if h != -1)

    kern.gather = spy
    sim = kern.run()
    repeats = sum(c - 1 for c in collections.Counter(seen).values())
    assert repeats > 0, "the event must exercise the gather's repeats"
    out, st = oracle_mod.search_event(ev.pack())
    assert ev_mod.track_lists(out) == sim
    assert st["seed_h0_visits"] > 0


def _row_stop_event():
    """7 planes 22 mm apart.  h0 at x=0 on sensor 6; sensor 4 holds, in this
    storage order, x = -25, 30, -5, 5, -10 (all inside the 0.7-slope h1
    window); sensor 2 holds one hit at x=-20.  The element at x=30 has an
    empty h2 window (its xmin_h2 = 30 - 0.7*44 - 0.6 = -1.4 lies above every
    x of sensor 2), the element at x=-10 makes the straight triplet."""
    z = [0, 22, 44, 66, 88, 110, 132]
    xs = [[], [], [-20.0], [], [-25.0, 30.0, -5.0, 5.0, -10.0], [], [0.0]]
    starts, nums, X = [], [], []
    for s in range(7):
        starts.append(len(X))
        nums.append(len(xs[s]))
        X += xs[s]
    N = len(X)
    Z = [float(z[s]) for s in range(7) for _ in xs[s]]
    return raw_blob(7, N, z, starts, nums, X, [0.0] * N, Z), N


def test_y_split_matters_and_the_oracle_is_the_y4_run(oracle_mod):
    """Seeding depends on the work-group height Y: row y takes window elements
    4j+y and stops for good at an element with an empty h2 window
    (Kernel.cl:436-457).  With Y=4 the blocker at element 1 stops row 1 only
    and row 0 reaches element 4; with Y=1 the single row stops at element 1.
    The reference returns its Y=4 run (KernelInvoker.cpp:149-156,197-206)."""
    blob, N = _row_stop_event()
    ev = ev_mod.parse_event(blob)
    out, st = oracle_mod.search_event(blob)
    y4, _ = simt.run_event(ev, Y_threads=4)
    y1, _ = simt.run_event(ev, Y_threads=1)
    y2, _ = simt.run_event(ev, Y_threads=2)
    assert ev_mod.track_lists(out) == y4 == [(N - 1, 5, 0)]
    assert y1 == []  # one row: stopped by element 1
    assert y2 == y4  # Y=2: element 4 is in row 0 as well
    c1, c2 = oracle_mod.fill_candidates(blob)
    assert c2[2 * 2] == -1  # the blocker (hit index 2 = x 30) has no h2 window


# ---------------------------------------------------------------------------
# closed forms on noiseless straight lines (uniform 22 mm planes)
# ---------------------------------------------------------------------------
PLANES12 = [-300 + 22 * k for k in range(12)]


def test_straight_lines_give_the_predicted_two_tracks(oracle_mod):
    """Forwarding at step f searches sensor f-4 for entries of step f+1, so a
    line crossing planes K-1..0 is seeded at K-1 and grows from K-5 down one
    plane per step; one step later a second seed (K-2, K-4, K-6) forms (its
    h0 and h1 are still unused, its h2 may be used) and follows the same
    hits: (K-1, K-3, K-5, K-6, ..., 0) and (K-2, K-4, K-6, K-7, ..., 0)."""
    lines = [(-30.0 + 12.0 * i, 8.0 - 3.0 * i, 0.01 * (i - 2), 0.012 * (2 - i)) for i in range(5)]
    ev = line_event(lines, PLANES12)
    out, st = oracle_mod.search_event(ev.pack())
    sensor = ev.sensor_of_hit()
    K = len(PLANES12)
    t1 = [K - 1, K - 3, K - 5] + list(range(K - 6, -1, -1))
    t2 = [K - 2, K - 4, K - 6] + list(range(K - 7, -1, -1))
    got = collections.Counter(tuple(int(sensor[h]) for h in t) for t in ev_mod.track_lists(out))
    assert got == collections.Counter({tuple(t1): len(lines), tuple(t2): len(lines)})
    assert st["flags"] == 0
    sim, _ = simt.run_event(ev)
    assert sim == ev_mod.track_lists(out)


def _displaced(plane, delta, axis):
    """One noiseless line; its hit on `plane` moved by delta along axis."""
    ev = line_event([(1.5, -2.0, 0.02, 0.015)], PLANES12)
    h = int(ev.hitStarts[plane])
    arr = ev.X if axis == "x" else ev.Y
    arr[h] = F(arr[h] + F(delta))
    return ev, h


@pytest.mark.parametrize("axis", ["x", "y"])
def test_forwarding_scatter_cut_boundary(oracle_mod, axis):
    """The line's first track is seeded (K-1, K-3, K-5) and its first
    forwarded hit is on K-6, predicted from the hits on K-3 and K-5 and
    22 mm past the last one: scatter = delta^2 / 22^2 and the cut
    scatter < 1.6e-5 keeps |delta| < 0.0880 mm.  A denominator over the wrong
    hit pair (44 or 66 mm) would move the boundary to 0.176 or 0.264 mm; a
    prediction from the wrong origin would move the hit off for any delta."""
    K = len(PLANES12)
    for delta, inside in ((0.080, True), (0.087, True), (0.089, False), (0.12, False)):
        ev, h = _displaced(K - 6, delta, axis)
        out, _ = oracle_mod.search_event(ev.pack())
        sensor = ev.sensor_of_hit()
        t1 = [t for t in ev_mod.track_lists(out) if int(sensor[t[0]]) == K - 1][0]
        assert (h in t1) == inside, (delta, [int(sensor[x]) for x in t1])
        if not inside:  # skipped one plane, then found the next one
            assert [int(sensor[x]) for x in t1][:5] == [K - 1, K - 3, K - 5, K - 7, K - 8]


@pytest.mark.parametrize("axis", ["x", "y"])
def test_seeding_scatter_cut_boundary(oracle_mod, axis):
    """The seed (K-1, K-3, K-5) predicts its third hit from the first two; the
    third sits 44 mm past the second, so the cut keeps |delta| < 0.176 mm."""
    K = len(PLANES12)
    for delta, inside in ((0.17, True), (0.175, True), (0.177, False), (0.25, False)):
        ev, h = _displaced(K - 5, delta, axis)
        out, _ = oracle_mod.search_event(ev.pack())
        sensor = ev.sensor_of_hit()
        starts = {tuple(int(sensor[x]) for x in t[:3]): t for t in ev_mod.track_lists(out)}
        seeded = (K - 1, K - 3, K - 5) in starts and h in starts[(K - 1, K - 3, K - 5)]
        assert seeded == inside, (delta, sorted(starts))


@pytest.mark.parametrize("slope,seeded", [(0.55, True), (0.69, True), (0.71, False)])
def test_slope_cuts(oracle_mod, slope, seeded):
    """|x1-x0| < 0.7 |z1-z0| bounds the h1 window and |y1-y0| < 0.7 |z1-z0|
    (dymax) the seeds: a line steeper than 0.7 in y is never seeded.  The
    line starts near the origin and the planes are narrow in z so it stays
    within reach."""
    planes = [10 + 4 * k for k in range(10)]
    ev = line_event([(0.5, 0.3, 0.05, slope)], planes)
    out, _ = oracle_mod.search_event(ev.pack())
    assert (len(ev_mod.track_lists(out)) > 0) == seeded


# ---------------------------------------------------------------------------
# candidate windows worked out by hand (A8, Kernel.cl:50-178)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", window_pin_events(), ids=lambda c: c[0])
def test_windows_hand_worked(oracle_mod, case):
    """The h2 window from the sensor Zs and a hit z off its sensor's Z, the
    else-if "last" that keeps one element past xmax, -1 and the sensor end,
    the h1 slope cut one ulp either side, and the h1 window that ends at the
    first failing element although later ones pass (see the derivations in
    tests/helpers.py)."""
    name, blob, want = case
    c1, c2 = oracle_mod.fill_candidates(blob)
    N = len(c1) // 2
    for h in range(N):
        w1, w2 = want.get(h, ((-1, -1), (-1, -1)))
        assert (int(c1[2 * h]), int(c1[2 * h + 1])) == w1, (name, h, "h1")
        assert (int(c2[2 * h]), int(c2[2 * h + 1])) == w2, (name, h, "h2")
    assert oracle_mod.input_flags(blob) == 0


def test_banded_noise_leaves_the_tracks(oracle_mod):
    """Noise in x bands far outside the detector, 1000 mm apart for sensors
    two apart, enters no window and no tolerance: the tracks (as hit-ID
    sets) of the event with 720 such hits per sensor (37,054 hits, above
    the 32,768 the kernel keeps LDS used flags for) are those of the event
    without them."""
    from tests.helpers import banded_noise_event
    from tests.parity import id_sets

    base = generator.plain_event(3)
    big = banded_noise_event(base, 720, 5)
    assert big.n_hits > 32768
    b0, b1 = base.pack(), big.pack()
    assert id_sets(oracle_mod.search_event(b1)[0], b1) == id_sets(oracle_mod.search_event(b0)[0], b0)


# ---------------------------------------------------------------------------
# forwarding skips, the tail and the 24-slot record, worked out by hand
# ---------------------------------------------------------------------------
PLANES17 = [22 * k for k in range(17)]
PLANES48 = [-300 + 22 * k for k in range(48)]


def _plane_tracks(blob, out):
    sensor = ev_mod.parse_event(blob).sensor_of_hit()
    return [[int(sensor[h]) for h in t] for t in ev_mod.track_lists(out)]


@pytest.mark.parametrize("last,want", [(7, [[16, 14, 12, 7]]), (6, [[16, 14, 12]])])
def test_skip_limit(oracle_mod, last, want):
    """A tracklet (16, 14, 12) misses sensors 11, 10, 9 and 8 (steps 15-12,
    skip counts 1-4: a count <= 3 is re-queued, Kernel.cl:349-356).  At step
    11 it searches sensor 7: a hit there extends it to a track; without one
    its count 4 exceeds 3 and, holding three hits, it goes to weak
    (:357-361) and comes out as a 3-hit track (its hits are unused).  A hit
    on sensor 6 is never searched."""
    blob = sparse_line_event([(1.0, 2.0, 0.01, 0.02, [16, 14, 12, last])], PLANES17)
    out, st = oracle_mod.search_event(blob)
    assert _plane_tracks(blob, out) == want
    assert st["n_weak"] == (1 if last == 6 else 0)


def test_tail_takes_only_the_last_step(oracle_mod):
    """Line A (13, 11, 9) misses four sensors and goes to weak in the
    forwarding of step 8.  Line B (8, 6, 4) misses sensors 3 .. 0 and is
    re-queued at step 4, the last step: the tail turns the tracklet entries
    inserted in that step, and only those, into weak entries
    (Kernel.cl:769-792).  Each comes out once, A first (weak order).  A
    tail over all tracklet entries would repeat A five times and B four."""
    blob = sparse_line_event([(0.5, 1.0, 0.01, 0.005, [8, 6, 4]),
                              (60.0, -3.0, 0.02, 0.01, [13, 11, 9])], PLANES17)
    out, st = oracle_mod.search_event(blob)
    assert _plane_tracks(blob, out) == [[13, 11, 9], [8, 6, 4]]
    assert st["n_weak"] == 2 and st["n_tracklets"] == 2


def _one_line_48(drop=()):
    """One noiseless line through 48 planes 22 mm apart (hit index = plane
    index unless planes are dropped)."""
    keep = [k for k in range(48) if k not in drop]
    return sparse_line_event([(1.0, 2.0, 0.01, 0.02, keep)], PLANES48)


def test_full_record_overflow_hand_traced(oracle_mod):
    """Seeds (47, 45, 43) at step 47 and (46, 44, 42) at step 46 (its h2 may
    be used); both then take one plane per step (A.1).  A holds 24 hits at
    step 26 (last hit 22), B at step 25 (last hit 21, taken at the step
    where A, full, finds 21: A marks 21 used, Kernel.cl:316, and stops where
    :321 would write past the record).  At step 24 B finds 20, marks it and
    stops.  Hits 47 .. 20 are then used, so the next seed is (19, 17, 15)
    at step 19 and (18, 16, 14) at step 18, which run to sensor 0.  Without
    the mark of A's and B's 25th hits, hit 20 would stay free and seed
    (20, 18, 16)."""
    blob = _one_line_48()
    out, st = oracle_mod.search_event(blob)
    want = [[47, 45, 43] + list(range(42, 21, -1)),
            [46, 44, 42] + list(range(41, 20, -1)),
            [19, 17, 15] + list(range(14, -1, -1)),
            [18, 16, 14] + list(range(13, -1, -1))]
    assert _plane_tracks(blob, out) == want
    assert st["n_full"] == 2 and st["n_overflow"] == 2 and st["flags"] == 1
    sim, kern = simt.run_event(ev_mod.parse_event(blob))
    assert sim == ev_mod.track_lists(out) and kern.overflows == 2


def test_full_record_keeps_skipping(oracle_mod):
    """As above with plane 21 empty: A, full since step 26, misses it at step
    25 and is re-queued with one skip like any track (:349-356); at step 24
    it finds 20, marks it and stops.  B (23 hits) misses 21 too, takes 20 at
    step 24 (24 hits) and at step 23 finds 19, marks it and stops.  The
    next seeds are (18, 16, 14) and (17, 15, 13)."""
    blob = _one_line_48(drop=(21,))
    out, st = oracle_mod.search_event(blob)
    want = [[47, 45, 43] + list(range(42, 21, -1)),
            [46, 44, 42] + list(range(41, 21, -1)) + [20],
            [18, 16, 14] + list(range(13, -1, -1)),
            [17, 15, 13] + list(range(12, -1, -1))]
    assert _plane_tracks(blob, out) == want
    assert st["n_full"] == 2 and st["n_overflow"] == 2
    sim, kern = simt.run_event(ev_mod.parse_event(blob))
    assert sim == ev_mod.track_lists(out) and kern.overflows == 2


# ---------------------------------------------------------------------------
# recomputed cuts and MC truth on the generator's events
# ---------------------------------------------------------------------------
def _check_track_cuts(ev, tracks):
    sensor = ev.sensor_of_hit()
    X, Y, Z = ev.X, ev.Y, ev.Z
    for t in tracks:
        assert 3 <= len(t) <= 24 and len(set(t)) == len(t)
        s = [int(sensor[h]) for h in t]
        assert s[1] == s[0] - 2 and s[2] == s[0] - 4, s
        assert all(a > b for a, b in zip(s, s[1:])), s
        h0, h1, h2 = t[:3]
        dzi = F(1) / (Z[h1] - Z[h0])
        tz = (Z[h2] - Z[h0]) * dzi
        dx = (X[h0] + (X[h1] - X[h0]) * tz) - X[h2]
        dy = (Y[h0] + (Y[h1] - Y[h0]) * tz) - Y[h2]
        sd = F(1) / (Z[h2] - Z[h1])
        assert abs(dx) < F(0.6) and abs(dy) < F(0.6)
        assert (dx * dx + dy * dy) * sd * sd < F(0.000016)
        for k in range(3, len(t)):
            a, b, c = t[k - 2], t[k - 1], t[k]
            td = F(1) / (Z[b] - Z[a])
            tx, ty = (X[b] - X[a]) * td, (Y[b] - Y[a]) * td
            dz = Z[c] - Z[a]
            ex = abs(X[a] + tx * dz - X[c])
            ey = abs(Y[a] + ty * dz - Y[c])
            sd = F(1) / (Z[c] - Z[b])
            assert ex < F(0.6) and ey < F(0.6)
            assert (ex * ex + ey * ey) * sd * sd < F(0.000016)
            assert s[k - 1] - s[k] <= 1 + 4, s  # at most MAX_SKIPPED_MODULES + 1 planes


def test_tracks_pass_the_cuts(oracle_mod, mcdata_sorted):
    evs = [generator.single_event(), generator.plain_event(3)]
    evs += [ev_mod.parse_event(mcdata_sorted[k]) for k in ("0", "77")]
    for ev in evs:
        out, st = oracle_mod.search_event(ev.pack())
        tracks = ev_mod.track_lists(out)
        assert len(tracks) == st["n_tracks"]
        _check_track_cuts(ev, tracks)
        # weak (3-hit) tracks only ever hold hits no longer track holds
        used = {h for t in tracks if len(t) >= 4 for h in t}
        assert all(not (set(t) & used) for t in tracks if len(t) == 3)


@pytest.mark.parametrize("maker", [lambda: generator.single_event(True),
                                   lambda: generator.batch_event(5, True),
                                   lambda: generator.stress_event(2, True)])
def test_mc_truth_efficiency(oracle_mod, maker):
    ev, truth = maker()
    out, _ = oracle_mod.search_event(ev.pack())
    tracks = ev_mod.track_lists(out)
    part = truth.particle
    found, pure = set(), 0
    for t in tracks:
        p, c = collections.Counter(part[list(t)].tolist()).most_common(1)[0]
        if p >= 0 and c >= 0.7 * len(t):
            found.add(p)
            pure += 1
    nh = np.bincount(part[part >= 0], minlength=truth.n_particles)
    long_ = [i for i in range(truth.n_particles) if nh[i] >= 5]
    eff = np.mean([i in found for i in long_])
    assert eff >= 0.95, eff
    assert pure / len(tracks) >= 0.95


# ---------------------------------------------------------------------------
# flags, malformed input, regression digests
# ---------------------------------------------------------------------------
def test_input_flags(oracle_mod, mcdata_raw, mcdata_sorted):
    assert oracle_mod.input_flags(mcdata_sorted["0"]) == 0
    assert oracle_mod.input_flags(mcdata_raw["0"]) & (1 << 6)
    ev = line_event([(1.0, 1.0, 0.01, 0.01)], PLANES12)
    ev.Z[int(ev.hitStarts[3])] = F(PLANES12[5])  # a hit of sensor 3 at sensor 5's z
    assert oracle_mod.input_flags(ev.pack()) & (1 << 5)
    assert oracle_mod.input_flags(mcdata_raw["0"][:-1]) == 1 << 7
    assert oracle_mod.input_flags(b"\x01\x00") == 1 << 7


def test_malformed_blobs(oracle_mod):
    bad = [b"", b"\x00" * 7, raw_blob(2, 1, [0, 1], [0, 0], [1, 5], [0], [0], [0]),
           raw_blob(1, 1, [0], [-1], [1], [0], [0], [0])]
    for b in bad:
        out, st = oracle_mod.search_event(b)
        assert out == b"" and st["flags"] == 1 << 7


def test_tiny_events(oracle_mod):
    assert oracle_mod.search_event(raw_blob(0, 0, [], [], [], [], [], []))[0] == b""
    four = line_event([(0.0, 1.0, 0.01, 0.01)], PLANES12[:4])
    out, st = oracle_mod.search_event(four.pack())
    assert out == b"" and st["n_tracklets"] == 0


def test_oracle_regression_digests(oracle_mod, mcdata_sorted, oracle_digests):
    for k, blob in mcdata_sorted.items():
        out, st = oracle_mod.search_event(blob)
        assert hashlib.sha256(out).hexdigest() == oracle_digests["mcdata"][k]["sha256"], k
    out, _ = oracle_mod.search_event(generator.single_event().pack())
    assert hashlib.sha256(out).hexdigest() == oracle_digests["single"]["0"]["sha256"]
    for i in range(0, 51, 10):
        out, _ = oracle_mod.search_event(generator.plain_event(i).pack())
        assert hashlib.sha256(out).hexdigest() == oracle_digests["plain51"][str(i)]["sha256"]
