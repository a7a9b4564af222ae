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
