"""Test infrastructure, never imported by the product: a restatement of the
reference's per-event result text, the checker of clf_format_tracks.

Follows KernelInvoker.cpp:209-243 (the results/<i>.out writer),
printTrack (:254-275) and findClosestModule (:284-303):

* the z -> module map starts from the sensor Z integers (a later sensor
  with the same Z wins), then every hit's z, truncated to int, is added in
  hit order, mapped to the key nearest to it among those present at that
  moment, trying z-1 before z+1 at each distance (so keys added earlier
  can be the nearest of later hits);
* each track prints ``Track #<n>, length <hitsNum>`` and one line per hit,
  `` <id:8> (<index>) module <m:2>, x <x:6>, y <y:6>, z <z:6>``, the floats
  in the default C++ stream format (printf ``%g``, six significant digits),
  then an empty line.
"""
from __future__ import annotations

import struct

import numpy as np


def cpp_float(x) -> str:
    """A float as a default-formatted C++ ostream writes it (``%g``)."""
    return "%g" % float(np.float32(x))


def _closest(z: int, m: dict) -> int:
    err = 1
    while True:
        if z - err in m:
            return m[z - err]
        if z + err in m:
            return m[z + err]
        err += 1


def z_to_module(sensor_z, hit_z) -> dict:
    m = {}
    for j, z in enumerate(sensor_z):
        m[int(z)] = j
    if not m and len(hit_z):
        raise ValueError("hits without sensors")
    for z in hit_z:
        zi = int(float(np.float32(z)))  # C++ (int) of a float: toward zero
        if zi not in m:
            m[zi] = _closest(zi, m)
    return m


def format_tracks(event: bytes, tracks: bytes) -> str:
    S, N = struct.unpack_from("<ii", event, 0)
    sensor_z = np.frombuffer(event, "<i4", S, 8)
    off = 8 + 12 * S
    ids = np.frombuffer(event, "<u4", N, off)
    X = np.frombuffer(event, "<f4", N, off + 4 * N)
    Y = np.frombuffer(event, "<f4", N, off + 8 * N)
    Z = np.frombuffer(event, "<f4", N, off + 12 * N)
    m = z_to_module(sensor_z, Z)
    rec = np.frombuffer(tracks, "<u4").reshape(-1, 25) if tracks else np.zeros((0, 25), np.uint32)
    out = []
    for j, r in enumerate(rec):
        hn = int(r[0])
        out.append(f"Track #{j}, length {hn}\n")
        for h in r[1:1 + hn]:
            h = int(h)
            out.append(" %8d (%d) module %2d, x %6s, y %6s, z %6s\n" % (
                int(ids[h]), h, m[int(float(Z[h]))], cpp_float(X[h]), cpp_float(Y[h]),
                cpp_float(Z[h])))
        out.append("\n")
    return "".join(out)
