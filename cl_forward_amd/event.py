"""Per-event input layout, output records and the Gaudi ``.dat`` framing.

Input blob of one event (parsed by the reference in Tools.cpp:68-80 and on
the device in Kernel.cl:613-624)::

    int32  S                 number of sensors
    int32  N                 number of hits
    int32  sensor_Zs[S]
    int32  hitStarts[S]
    int32  hitNums[S]
    uint32 hit_IDs[N]
    float32 X[N], Y[N], Z[N]

8 + 12*S + 16*N bytes, little endian, every field 4 bytes.

Output blob of one event: T records of ``struct Track {uint32 hitsNum;
uint32 hits[24];}`` (KernelDefinitions.h:59-62), 100 bytes each, ``hits``
holding indices into the event's hit arrays, unused slots zero.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

MAX_TRACK_SIZE = 24
TRACK_BYTES = 4 + 4 * MAX_TRACK_SIZE
TRACK_DTYPE = np.dtype([("hitsNum", "<u4"), ("hits", "<u4", (MAX_TRACK_SIZE,))])
assert TRACK_DTYPE.itemsize == TRACK_BYTES

#: Name stored in every reference ``.dat`` file (IndependentEntrypoint.cpp:84-88).
HANDLER_NAME = b"PrPixelCudaHandler"


@dataclass
class Event:
    """One event as SoA arrays (the fields of the input blob)."""

    sensor_Zs: np.ndarray  # int32[S]
    hitStarts: np.ndarray  # int32[S]
    hitNums: np.ndarray  # int32[S]
    hit_IDs: np.ndarray  # uint32[N]
    X: np.ndarray  # float32[N]
    Y: np.ndarray  # float32[N]
    Z: np.ndarray  # float32[N]

    @property
    def n_sensors(self) -> int:
        return int(self.sensor_Zs.shape[0])

    @property
    def n_hits(self) -> int:
        return int(self.hit_IDs.shape[0])

    def pack(self) -> bytes:
        return pack_event(self)

    def sensor_of_hit(self) -> np.ndarray:
        """Sensor index of every hit (hitStarts/hitNums ranges)."""
        out = np.full(self.n_hits, -1, np.int32)
        for s in range(self.n_sensors):
            b, n = int(self.hitStarts[s]), int(self.hitNums[s])
            out[b : b + n] = s
        return out


def event_nbytes(n_sensors: int, n_hits: int) -> int:
    return 8 + 12 * n_sensors + 16 * n_hits


def pack_event(ev: Event) -> bytes:
    S, N = ev.n_sensors, ev.n_hits
    parts = [
        struct.pack("<ii", S, N),
        np.ascontiguousarray(ev.sensor_Zs, "<i4").tobytes(),
        np.ascontiguousarray(ev.hitStarts, "<i4").tobytes(),
        np.ascontiguousarray(ev.hitNums, "<i4").tobytes(),
        np.ascontiguousarray(ev.hit_IDs, "<u4").tobytes(),
        np.ascontiguousarray(ev.X, "<f4").tobytes(),
        np.ascontiguousarray(ev.Y, "<f4").tobytes(),
        np.ascontiguousarray(ev.Z, "<f4").tobytes(),
    ]
    blob = b"".join(parts)
    assert len(blob) == event_nbytes(S, N)
    return blob


def parse_event(blob) -> Event:
    """Views of a blob as arrays (copies, so the blob may be mutated later)."""
    buf = memoryview(bytes(blob))
    if len(buf) < 8:
        raise ValueError("event blob shorter than its 8-byte header")
    S, N = struct.unpack_from("<ii", buf, 0)
    if S < 0 or N < 0 or len(buf) < event_nbytes(S, N):
        raise ValueError(f"event blob of {len(buf)} bytes too short for S={S}, N={N}")
    off = 8

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(buf, dtype, n, off).copy()
        off += 4 * n
        return a

    Zs = take("<i4", S)
    st = take("<i4", S)
    nu = take("<i4", S)
    ids = take("<u4", N)
    X = take("<f4", N)
    Y = take("<f4", N)
    Z = take("<f4", N)
    return Event(Zs, st, nu, ids, X, Y, Z)


def tracks_from_bytes(out) -> np.ndarray:
    """Structured array of Track records from an output blob."""
    b = bytes(out)
    if len(b) % TRACK_BYTES:
        raise ValueError(f"output blob of {len(b)} bytes is not a whole number of tracks")
    return np.frombuffer(b, TRACK_DTYPE).copy()


def track_lists(out) -> list[tuple[int, ...]]:
    """Each track as the tuple of its hit indices, in output order."""
    recs = tracks_from_bytes(out)
    return [tuple(int(h) for h in r["hits"][: int(r["hitsNum"])]) for r in recs]


# --------------------------------------------------------------------------
# Gaudi offload framing (IndependentEntrypoint.cpp:78-123)
# --------------------------------------------------------------------------
def read_dat(path) -> bytes:
    """Event blob of a ``.dat`` file: ``int nameLen; name; int dataSize; data``.

    Bytes after ``dataSize`` (a stored track list of an older tracker in the
    reference's mcdata files) are ignored, as the reference ignores them.
    """
    with open(path, "rb") as f:
        raw = f.read()
    return unframe(raw)[1]


def unframe(raw: bytes) -> tuple[bytes, bytes]:
    if len(raw) < 4:
        raise ValueError("truncated .dat frame")
    (name_len,) = struct.unpack_from("<i", raw, 0)
    if name_len < 0 or len(raw) < 8 + name_len:
        raise ValueError("truncated .dat frame")
    name = raw[4 : 4 + name_len]
    (data_size,) = struct.unpack_from("<i", raw, 4 + name_len)
    start = 8 + name_len
    if data_size < 0 or len(raw) < start + data_size:
        raise ValueError("truncated .dat payload")
    return name, raw[start : start + data_size]


def frame(blob: bytes, name: bytes = HANDLER_NAME) -> bytes:
    return struct.pack("<i", len(name)) + name + struct.pack("<i", len(blob)) + blob


def write_frame_stream(path, blobs) -> None:
    """Events as offload frames back to back (the input of clf_dropin)."""
    with open(path, "wb") as f:
        for b in blobs:
            f.write(frame(bytes(b)))


def read_sized_stream(path) -> list[bytes]:
    """int64 size + bytes records back to back (the output of clf_dropin)."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        (n,) = struct.unpack_from("<q", data, p)
        out.append(data[p + 8 : p + 8 + n])
        p += 8 + n
    return out
