"""Small event builders for tests."""
from __future__ import annotations

import struct

import numpy as np

from cl_forward_amd.event import Event


def sub_event(ev: Event, lo: int, hi: int, stride: int = 1) -> Event:
    """Sensors [lo, hi) of ev, every stride-th hit of each, re-packed."""
    Zs, st, nu, X, Y, Z, ids = [], [], [], [], [], [], []
    for s in range(lo, hi):
        b, n = int(ev.hitStarts[s]), int(ev.hitNums[s])
        idx = np.arange(b, b + n)[::stride]
        st.append(len(X))
        nu.append(len(idx))
        Zs.append(int(ev.sensor_Zs[s]))
        X += ev.X[idx].tolist()
        Y += ev.Y[idx].tolist()
        Z += ev.Z[idx].tolist()
        ids += ev.hit_IDs[idx].tolist()
    return Event(np.array(Zs, np.int32), np.array(st, np.int32), np.array(nu, np.int32),
                 np.array(ids, np.uint32), np.array(X, np.float32), np.array(Y, np.float32),
                 np.array(Z, np.float32))


def add_noise(ev: Event, sensor: int, n: int, seed: int) -> Event:
    """ev with n uniform noise hits added to one sensor (x-sorted, distinct x)."""
    rng = np.random.default_rng(seed)
    Zs, st, nu, X, Y, Z, ids = [], [], [], [], [], [], []
    next_id = int(ev.hit_IDs.max()) + 1 if ev.n_hits else 1
    for s in range(ev.n_sensors):
        b, c = int(ev.hitStarts[s]), int(ev.hitNums[s])
        x, y = ev.X[b:b + c].tolist(), ev.Y[b:b + c].tolist()
        i = ev.hit_IDs[b:b + c].tolist()
        if s == sensor:
            x += rng.uniform(-40, 40, n).astype(np.float32).tolist()
            y += rng.uniform(-40, 40, n).astype(np.float32).tolist()
            i += list(range(next_id, next_id + n))
        order = np.argsort(np.array(x, np.float32), kind="stable")
        st.append(len(X))
        nu.append(len(x))
        Zs.append(int(ev.sensor_Zs[s]))
        X += [x[k] for k in order]
        Y += [y[k] for k in order]
        Z += [float(ev.Z[b]) if c else float(ev.sensor_Zs[s])] * len(x)
        ids += [i[k] for k in order]
    return Event(np.array(Zs, np.int32), np.array(st, np.int32), np.array(nu, np.int32),
                 np.array(ids, np.uint32), np.array(X, np.float32), np.array(Y, np.float32),
                 np.array(Z, np.float32))


def raw_blob(S, N, Z, starts, nums, x, y, z, ids=None) -> bytes:
    ids = np.arange(N, dtype=np.uint32) if ids is None else ids
    return (struct.pack("<ii", S, N) + np.asarray(Z, "<i4").tobytes() +
            np.asarray(starts, "<i4").tobytes() + np.asarray(nums, "<i4").tobytes() +
            np.asarray(ids, "<u4").tobytes() + np.asarray(x, "<f4").tobytes() +
            np.asarray(y, "<f4").tobytes() + np.asarray(z, "<f4").tobytes())


def line_event(points_per_track, z_planes):
    """Event of straight noiseless tracks: points_per_track is a list of
    (x0, y0, tx, ty) lines through z = 0; every plane of z_planes gets one
    hit per line.  Hits are x-sorted per sensor."""
    S = len(z_planes)
    Zs, st, nu, X, Y, Z, ids = [], [], [], [], [], [], []
    hid = 1
    for k, zk in enumerate(z_planes):
        xs = [np.float32(x0 + tx * zk) for (x0, y0, tx, ty) in points_per_track]
        ys = [np.float32(y0 + ty * zk) for (x0, y0, tx, ty) in points_per_track]
        order = np.argsort(np.array(xs, np.float32), kind="stable")
        st.append(len(X))
        nu.append(len(xs))
        Zs.append(int(zk))
        X += [xs[i] for i in order]
        Y += [ys[i] for i in order]
        Z += [float(zk)] * len(xs)
        ids += list(range(hid, hid + len(xs)))
        hid += len(xs)
    return Event(np.array(Zs, np.int32), np.array(st, np.int32), np.array(nu, np.int32),
                 np.array(ids, np.uint32), np.array(X, np.float32), np.array(Y, np.float32),
                 np.array(Z, np.float32))


def sensors_blob(Z, hits) -> bytes:
    """Blob of an event with sensor Zs Z and hits[s] = [(x, y, z), ...] on
    sensor s, stored in the order given; hit IDs 1..N."""
    starts, nums, X, Y, ZH = [], [], [], [], []
    for s in range(len(Z)):
        starts.append(len(X))
        nums.append(len(hits[s]))
        for (x, y, z) in hits[s]:
            X.append(x)
            Y.append(y)
            ZH.append(z)
    N = len(X)
    return raw_blob(len(Z), N, Z, starts, nums, np.array(X, np.float32), np.array(Y, np.float32),
                    np.array(ZH, np.float32), np.arange(1, N + 1, dtype=np.uint32))


def sparse_line_event(lines, z_planes):
    """Blob of straight noiseless lines (x0, y0, tx, ty, planes): each line
    has one hit on each plane index listed in `planes`, at z = z_planes[k]."""
    hits = [[] for _ in z_planes]
    for (x0, y0, tx, ty, planes) in lines:
        for k in planes:
            zk = z_planes[k]
            hits[k].append((float(np.float32(x0 + tx * zk)), float(np.float32(y0 + ty * zk)), float(zk)))
    for h in hits:
        h.sort()
    return sensors_blob([int(z) for z in z_planes], hits)


# ---------------------------------------------------------------------------
# Candidate-window events whose windows are worked out by hand
# (Kernel.cl:50-178).  Each entry: (name, blob, {hit: (h1 window, h2 window)})
# with absolute hit indices, -1 for none; hits not listed have no windows.
# ---------------------------------------------------------------------------
def window_pin_events():
    F = np.float32
    out = []
    Z7 = [-44, -22, 0, 11, 22, 33, 44]
    # (1) h2 window from the sensor Zs Z[6] = 44, Z[2] = 0 and the h1's own
    # z0 = 23.0625, 1.0625 mm off its sensor's Z = 22 (mcdata hits sit up to
    # 2 mm off).  With x0 = 0: h_dist = 20.9375, dxmax = 0.7 h_dist,
    # z2_tz = 44 / 20.9375, xx = dxmax (1 - z2_tz) = -0.7 (44 - 20.9375)
    # = -16.14375, so (xmin, xmax) = -+16.74375.  First x above xmin:
    # -16.7 (index 2); first later x above xmax: 16.8 (index 5).  Reading
    # the hit z as its sensor's 22 would give -+16.0, swapping the two
    # sensor Zs -+15.26.  The h1 window (hit's own z, 0.7 * 23.0625 =
    # 16.14375): only -0.5 passes, so [3, 4).
    xs = [-20.0, -16.8, -16.7, -0.5, 16.7, 16.8, 20.0]
    hits = [[] for _ in Z7]
    hits[2] = [(x, 0.0, 0.0) for x in xs]
    hits[4] = [(0.0, 0.0, 23.0625)]
    out.append(("h2_from_sensor_Z", sensors_blob(Z7, hits), {7: ((3, 4), (2, 5))}))
    # (2) the "last" of an h2 window is set only on an element after the one
    # that set "first" (else-if, Kernel.cl:137-150): sensor 2 holds -30,
    # 17.5, 18.5.  P (x0 = 0, bounds -+16.74375): 17.5 is the first above
    # xmin and also above xmax, so last = the next element, 18.5: window
    # [1, 2) holds 17.5, past xmax.  R (x0 = 17, bounds 0.25625, 33.74375):
    # first 17.5, nothing above xmax, so last = the sensor end, 3; its h1
    # window (|x1 - 17| < 16.14375) is [1, 3) up to the end as well.
    # Q (x0 = 40, xmin 23.25625): no element above xmin, -1.
    hits = [[] for _ in Z7]
    hits[2] = [(-30.0, 0.0, 0.0), (17.5, 0.0, 0.0), (18.5, 0.0, 0.0)]
    hits[4] = [(0.0, 0.0, 23.0625), (17.0, 0.0, 23.0625), (40.0, 0.0, 23.0625)]
    out.append(("h2_else_if_last_and_end", sensors_blob(Z7, hits),
                {3: ((-1, -1), (1, 2)), 4: ((1, 3), (1, 3)), 5: ((-1, -1), (-1, -1))}))
    # (3) h1 window at the slope cut, one ulp either side: h0 (0, z 22),
    # sensor 2 at z 0, D = fl(0.7f * 22).  |x1| < D keeps -prev(D), 0.25
    # and prev(D) and drops -D, D and next(D): window [1, 4).  "<=" would
    # give [0, 5).  The h2 window of the h0 (bounds -+(0.7 * 22 + 0.6) from
    # sensor Zs 44 and 0 with z0 = 22): everything is inside, so
    # [0, end = 6).
    D = F(F(0.7) * F(22.0))
    pD, nD = np.nextafter(D, F(0)), np.nextafter(D, F(100))
    xs = [float(-D), float(-pD), 0.25, float(pD), float(D), float(nD)]
    hits = [[] for _ in Z7]
    hits[2] = [(x, 0.0, 0.0) for x in xs]
    hits[4] = [(0.0, 0.0, 22.0)]
    out.append(("h1_slope_ulp", sensors_blob(Z7, hits), {6: ((1, 4), (0, 6))}))
    # (4) h1 window with hits at two z on one sensor (as in mcdata): h0
    # (0, z 22); sensor 2 holds x -10 at z -10 (0.7 * 32 = 22.4: passes),
    # -9 at z 14 (0.7 * 8 = 5.6: fails), 0 at z 0 and 5 at z 0 (pass).  The
    # window is the first passing element up to the first later failing
    # one, [0, 1), although elements 2 and 3 pass again.  Its h2 window
    # (z0 = 22, bounds -+16.0): -10, -9, 0, 5 all inside, [0, end = 4).
    hits = [[] for _ in Z7]
    hits[2] = [(-10.0, 0.0, -10.0), (-9.0, 0.0, 14.0), (0.0, 0.0, 0.0), (5.0, 0.0, 0.0)]
    hits[4] = [(0.0, 0.0, 22.0)]
    out.append(("h1_two_z_first_failure", sensors_blob(Z7, hits), {4: ((0, 1), (0, 4))}))
    return out


def banded_noise_event(base: Event, per_sensor: int, seed: int) -> Event:
    """base plus per_sensor noise hits on every sensor, in an x band of its
    own far outside the detector: sensor s gets x in [1000 + 500 (s mod 4),
    + 400), so the bands of sensors two apart are 1000 mm apart and no
    noise hit enters an h1 window (|dx| < 0.7 |dz|) or a forwarding
    tolerance.  The noise is still gathered, and its used flags are read,
    at every step.  Hits stay x-sorted per sensor."""
    rng = np.random.default_rng(seed)
    Zs, st, nu, X, Y, Z, ids = [], [], [], [], [], [], []
    next_id = int(base.hit_IDs.max()) + 1 if base.n_hits else 1
    for s in range(base.n_sensors):
        b, c = int(base.hitStarts[s]), int(base.hitNums[s])
        lo = 1000.0 + 500.0 * (s % 4)
        nx = np.sort(rng.uniform(lo, lo + 400.0, per_sensor).astype(np.float32))
        ny = rng.uniform(-40, 40, per_sensor).astype(np.float32)
        z = float(base.Z[b]) if c else float(base.sensor_Zs[s])
        st.append(len(X))
        nu.append(c + per_sensor)
        Zs.append(int(base.sensor_Zs[s]))
        X += base.X[b:b + c].tolist() + nx.tolist()
        Y += base.Y[b:b + c].tolist() + ny.tolist()
        Z += base.Z[b:b + c].tolist() + [z] * per_sensor
        ids += base.hit_IDs[b:b + c].tolist() + list(range(next_id, next_id + per_sensor))
        next_id += per_sensor
    return Event(np.array(Zs, np.int32), np.array(st, np.int32), np.array(nu, np.int32),
                 np.array(ids, np.uint32), np.array(X, np.float32), np.array(Y, np.float32),
                 np.array(Z, np.float32))
