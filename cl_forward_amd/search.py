"""Host-side API over the HIP library, mirroring the reference's entry points.

Reference names (GpuPixelSearchByTriplet.cpp:4-65, KernelInvoker.cpp:14-31):

* :func:`gpu_pixel_search_by_triplet` — Gaudi path, no sort, silent
* :func:`independent_execute` — standalone path, x-presort IN PLACE first
  (the events must be ``bytearray`` so they can be sorted in place)
* :func:`gpu_pixel_search_by_triplet_invocation` — common path, chunks of
  at most 16000 events
* :func:`invoke_parallel_search` — one chunk

Each takes a list of per-event input blobs and returns/fills a list of
per-event output blobs (``T * 100`` bytes of Track records).  Errors raise
:class:`ClfError`.

:class:`Context` and :class:`Batch` expose the resident-buffer API used by
the benchmark and the serving loop.
"""
from __future__ import annotations

import ctypes
import threading

from . import _lib
from ._lib import (MODE_BATCHED, MODE_PLAIN, ClfError, EventStats, LaunchConfig, ReplayRecord,
                   ReplaySummary, Result, Served, ServerConfig, Timing)

TRACK_BYTES = 100
MAX_EVENTS_PER_LAUNCH = 16000  # GpuPixelSearchByTriplet.cpp:54

__all__ = [
    "Context", "Batch", "Server", "ClfError", "tracks_digest", "format_tracks", "unframe", "MODE_PLAIN", "MODE_BATCHED", "gpu_pixel_search_by_triplet",
    "gpu_pixel_search_by_triplet_invocation", "independent_execute", "independent_post_execute",
    "invoke_parallel_search", "presort_event", "track_capacity",
]


def _u8(buf):
    return ctypes.cast(buf, ctypes.POINTER(ctypes.c_uint8))


def _event_arrays(events):
    """Returns (holders, pointer array, size array) for read-only C calls.
    A ``bytes`` event is passed by the address of its own buffer (no copy;
    the library only reads events); any other buffer is copied into a
    ``bytes`` first.  ``holders`` keeps those objects alive for the call."""
    n = len(events)
    holders = []
    ptrs = (ctypes.POINTER(ctypes.c_uint8) * max(n, 1))()
    sizes = (ctypes.c_size_t * max(n, 1))()
    for i, ev in enumerate(events):
        b = ev if isinstance(ev, bytes) else (bytes(ev) if ev is not None else b"")
        if b:
            holders.append(b)
            ptrs[i] = ctypes.cast(ctypes.c_char_p(b), ctypes.POINTER(ctypes.c_uint8))
        sizes[i] = len(b)
    return holders, ptrs, sizes


def _unpack(r):
    """(list of per-event output blobs, list of stats dicts) of a clf_result,
    whose arrays live as long as the batch that filled it."""
    n = r.n_events
    offs = [int(r.track_offset[i]) for i in range(n + 1)]
    total = offs[-1] if n else 0
    stats = []
    for i in range(n):
        s = r.stats[i]
        stats.append({k: int(getattr(s, k)) for k, _ in EventStats._fields_})
    if not total:
        return [b"" for _ in range(n)], stats
    # a view of the batch's host copy of the packed records; each event's
    # records are copied out of it once
    addr = ctypes.cast(r.tracks, ctypes.c_void_p).value
    raw = memoryview((ctypes.c_uint8 * (total * TRACK_BYTES)).from_address(addr)).cast("B")
    outs = [bytes(raw[offs[i] * TRACK_BYTES: (offs[i] + stats[i]["n_tracks"]) * TRACK_BYTES])
            for i in range(n)]
    raw.release()
    return outs, stats


class Context:
    """A HIP device, its stream and launch configuration (``clf_context``)."""

    def __init__(self, device: int = 0, config: dict | None = None):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._L.clf_context_create(int(device), ctypes.byref(h)), "clf_context_create")
        self._h = h
        if config:
            self.set_config(**config)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._L.clf_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        _lib.check(self._L.clf_context_synchronize(self._h), "clf_context_synchronize")

    @property
    def device_name(self) -> str:
        return (self._L.clf_context_device_name(self._h) or b"").decode(errors="replace")

    @property
    def num_cus(self) -> int:
        return int(self._L.clf_context_num_cus(self._h))

    def config(self) -> dict:
        c = LaunchConfig()
        _lib.check(self._L.clf_context_get_config(self._h, ctypes.byref(c)), "get_config")
        return {k: int(getattr(c, k)) for k, _ in LaunchConfig._fields_}

    def set_config(self, **kw):
        cur = self.config()
        cur.update({k: int(v) for k, v in kw.items()})
        c = LaunchConfig(**cur)
        _lib.check(self._L.clf_context_set_config(self._h, ctypes.byref(c)), "set_config")

    def probe_copy(self, nbytes: int, reps: int = 10) -> float:
        """Mean ms of the calibration copy kernel over nbytes (read + write)."""
        ms = ctypes.c_float()
        _lib.check(self._L.clf_probe_copy(self._h, int(nbytes), int(reps), ctypes.byref(ms)),
                   "clf_probe_copy")
        return float(ms.value)

    def probe_link(self, nbytes: int, reps: int = 5):
        """(host-to-device GB/s, device-to-host GB/s) of pinned copies of nbytes."""
        h, d = ctypes.c_double(), ctypes.c_double()
        _lib.check(self._L.clf_probe_link(self._h, int(nbytes), int(reps), ctypes.byref(h),
                                          ctypes.byref(d)), "clf_probe_link")
        return float(h.value), float(d.value)

    def batch(self, events) -> "Batch":
        return Batch(self, events)

    def search(self, events, mode: int = MODE_BATCHED):
        """clf_search: upload, run and read back in stream order (no graph;
        a batch that runs once does not pay for a capture).  Returns (list of
        output blobs, list of stats)."""
        holders, ptrs, sizes = _event_arrays(events)
        r = Result()
        h = ctypes.c_void_p()
        _lib.check(self._L.clf_search(self._h, int(mode), len(events), ptrs, sizes, ctypes.byref(r),
                                      ctypes.byref(h)), "clf_search")
        del holders
        try:
            return _unpack(r)
        finally:
            self._L.clf_batch_destroy(h)

    def search_noalloc(self, events, mode: int = MODE_BATCHED, capacity: int | None = None):
        """clf_search_noalloc: the search without any device allocation (the
        library's static scratch array).  Returns (list of output blobs,
        list of stats); capacity (output records) defaults to the events'
        bound."""
        n = len(events)
        holders, ptrs, sizes = _event_arrays(events)
        if capacity is None:
            capacity = sum(max(track_capacity(e), 0) for e in events)
        tracks = (ctypes.c_uint8 * max(capacity * TRACK_BYTES, 1))()
        offs = (ctypes.c_int64 * (n + 1))()
        stats = (EventStats * max(n, 1))()
        _lib.check(self._L.clf_search_noalloc(self._h, int(mode), n, ptrs, sizes, _u8(tracks),
                                              int(capacity), offs, stats),
                   "clf_search_noalloc")
        del holders
        raw = memoryview(tracks).cast("B")
        st = [{k: int(getattr(stats[i], k)) for k, _ in EventStats._fields_} for i in range(n)]
        outs = [bytes(raw[offs[i] * TRACK_BYTES: (offs[i] + st[i]["n_tracks"]) * TRACK_BYTES])
                for i in range(n)]
        raw.release()
        return outs, st

    def fill_candidates(self, events, search_path: bool = False):
        """Stage (a) alone.  Returns a list of (cand_h1, cand_h2, flags) per event,
        the windows as int32 numpy arrays of 2*N.  search_path: computed as
        the search computes them (LDS ring slots, x-bucket index)."""
        import numpy as np

        n = len(events)
        holders, ptrs, sizes = _event_arrays(events)
        arrs = []
        for ev in events:
            b = bytes(ev)
            N = int(np.frombuffer(b[4:8], "<i4")[0]) if len(b) >= 8 else 0
            N = max(N, 0)
            arrs.append((np.full(2 * N + 2, -7, np.int32), np.full(2 * N + 2, -7, np.int32), N))
        i32p = ctypes.POINTER(ctypes.c_int32)
        p1 = (i32p * max(n, 1))(*[a[0].ctypes.data_as(i32p) for a in arrs])
        p2 = (i32p * max(n, 1))(*[a[1].ctypes.data_as(i32p) for a in arrs])
        flags = (ctypes.c_uint32 * max(n, 1))()
        fn = self._L.clf_fill_candidates_search_path if search_path else self._L.clf_fill_candidates
        _lib.check(fn(self._h, n, ptrs, sizes, p1, p2, flags), "clf_fill_candidates")
        del holders
        return [(a[0][: 2 * a[2]], a[1][: 2 * a[2]], int(flags[i])) for i, a in enumerate(arrs)]


class Batch:
    """Events resident in device memory (``clf_batch``)."""

    def __init__(self, ctx: Context, events):
        self.ctx = ctx
        self._L = ctx._L
        self.n = len(events)
        holders, ptrs, sizes = _event_arrays(events)
        h = ctypes.c_void_p()
        _lib.check(self._L.clf_batch_create(ctx.handle, self.n, ptrs, sizes, ctypes.byref(h)),
                   "clf_batch_create")
        del holders
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.clf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, mode: int = MODE_BATCHED):
        _lib.check(self._L.clf_batch_run(self.ctx.handle, self._h, int(mode)), "clf_batch_run")

    def results(self):
        r = Result()
        _lib.check(self._L.clf_batch_results(self.ctx.handle, self._h, ctypes.byref(r)),
                   "clf_batch_results")
        return _unpack(r)

    PROFILE_FIELDS = ("prologue", "stage", "forward", "gather", "seed_setup", "stage_wait",
                      "seed_eval", "seed_emit", "tail", "total", "fwd_entries", "visits", "pairs",
                      "batches", "fwd_tiles", "h2_tests")

    def profile(self, mode: int = MODE_BATCHED):
        """Per-event phase times (ns) and work counters of one run, as a dict of
        numpy arrays keyed by PROFILE_FIELDS."""
        import numpy as np

        out = np.zeros((max(self.n, 1), 16), np.uint64)
        tick = ctypes.c_double()
        _lib.check(self._L.clf_batch_profile(self.ctx.handle, self._h, int(mode),
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                             ctypes.byref(tick)), "clf_batch_profile")
        out = out[: self.n].astype(np.float64)
        res = {}
        for k, name in enumerate(self.PROFILE_FIELDS):
            res[name] = out[:, k] * (tick.value if k < 10 else 1.0)
        return res

    def time(self, mode: int = MODE_BATCHED, warmup: int = 2, reps: int = 5) -> dict:
        t = Timing()
        _lib.check(self._L.clf_batch_time(self.ctx.handle, self._h, int(mode), int(warmup),
                                          int(reps), ctypes.byref(t)), "clf_batch_time")
        return {"ms_run": float(t.ms_run), "ms_search": float(t.ms_search), "reps": int(t.reps)}


def format_tracks(event: bytes, tracks: bytes) -> str:
    """The reference's results/<i>.out text of one event (clf_format_tracks):
    printTrack lines for each 100-byte Track record of ``tracks``."""
    L = _lib.load()
    ev = (ctypes.c_uint8 * max(len(event), 1)).from_buffer_copy(bytes(event) or b"\0")
    n = len(tracks) // TRACK_BYTES
    tr = (ctypes.c_uint8 * max(len(tracks), 1)).from_buffer_copy(bytes(tracks) or b"\0")
    need = ctypes.c_size_t()
    _lib.check(L.clf_format_tracks(ev, len(event), tr, n, None, 0, ctypes.byref(need)),
               "clf_format_tracks")
    buf = ctypes.create_string_buffer(need.value + 1)
    _lib.check(L.clf_format_tracks(ev, len(event), tr, n, buf, need.value + 1, ctypes.byref(need)),
               "clf_format_tracks")
    return buf.raw[: need.value].decode("ascii")


def unframe(frame: bytes):
    """(name, data) of one Gaudi offload frame (clf_unframe)."""
    L = _lib.load()
    buf = (ctypes.c_uint8 * max(len(frame), 1)).from_buffer_copy(bytes(frame) or b"\0")
    name_p = ctypes.POINTER(ctypes.c_uint8)()
    data_p = ctypes.POINTER(ctypes.c_uint8)()
    nl = ctypes.c_int32()
    ds = ctypes.c_size_t()
    _lib.check(L.clf_unframe(buf, len(frame), ctypes.byref(name_p), ctypes.byref(nl),
                             ctypes.byref(data_p), ctypes.byref(ds)), "clf_unframe")
    base = ctypes.addressof(buf)
    n_off = ctypes.cast(name_p, ctypes.c_void_p).value - base
    d_off = ctypes.cast(data_p, ctypes.c_void_p).value - base
    return bytes(frame[n_off:n_off + nl.value]), bytes(frame[d_off:d_off + ds.value])


def tracks_digest(tracks: bytes) -> int:
    """clf_tracks_digest of packed Track records (len a multiple of 100)."""
    L = _lib.load()
    n = len(tracks) // TRACK_BYTES
    if n == 0:
        return int(L.clf_tracks_digest(None, 0))
    buf = (ctypes.c_uint8 * len(tracks)).from_buffer_copy(tracks)
    return int(L.clf_tracks_digest(buf, n))


class Server:
    """Serving loop (``clf_server``): submit events one at a time, get each
    event's tracks when its batch finishes.  ``config`` keys are the fields
    of clf_server_config; missing ones take the library defaults."""

    def __init__(self, ctx: Context, **config):
        self.ctx = ctx
        self._L = ctx._L
        c = ServerConfig()
        self._L.clf_server_default_config(ctypes.byref(c))
        for k, v in config.items():
            if k not in dict(ServerConfig._fields_):
                raise TypeError(f"unknown server option {k!r}")
            setattr(c, k, int(v))
        h = ctypes.c_void_p()
        _lib.check(self._L.clf_server_create(ctx.handle, ctypes.byref(c), ctypes.byref(h)),
                   "clf_server_create")
        self._h = h
        self.config = {k: int(getattr(c, k)) for k, _ in ServerConfig._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._L.clf_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, event: bytes, tag: int = 0) -> int:
        b = bytes(event)
        buf = (ctypes.c_uint8 * max(len(b), 1)).from_buffer_copy(b or b"\0")
        t = ctypes.c_uint64()
        _lib.check(self._L.clf_server_submit(self._h, buf, len(b), int(tag), ctypes.byref(t)),
                   "clf_server_submit")
        return int(t.value)

    def submit_many(self, events, tags=None):
        """Queues events in order under one acquisition of the server lock
        (clf_server_submit_many, at most 64 a call).  Returns the tickets
        of the events accepted; raises the first refusal's error if any
        event was refused (the ones before it stay queued: their tickets
        are in the exception's `tickets` attribute)."""
        bs = [bytes(e) for e in events]
        n = len(bs)
        tags = list(range(n)) if tags is None else [int(x) for x in tags]
        if len(tags) != n:
            raise ValueError("submit_many: one tag per event")
        bufs = [(ctypes.c_uint8 * max(len(b), 1)).from_buffer_copy(b or b"\0") for b in bs]
        u8p = ctypes.POINTER(ctypes.c_uint8)
        ptrs = (u8p * max(n, 1))(*[ctypes.cast(b, u8p) for b in bufs])
        sizes = (ctypes.c_size_t * max(n, 1))(*[len(b) for b in bs])
        tg = (ctypes.c_uint64 * max(n, 1))(*tags)
        tk = (ctypes.c_uint64 * max(n, 1))()
        acc = ctypes.c_int32()
        st = self._L.clf_server_submit_many(self._h, n, ptrs, sizes, tg, tk, ctypes.byref(acc))
        tickets = [int(tk[k]) for k in range(acc.value)]
        try:
            _lib.check(st, "clf_server_submit_many")
        except Exception as e:
            e.tickets = tickets
            raise
        return tickets

    def submit_framed(self, frame: bytes, tag: int = 0) -> int:
        """submit() of a request in the Gaudi offload framing."""
        b = bytes(frame)
        buf = (ctypes.c_uint8 * max(len(b), 1)).from_buffer_copy(b or b"\0")
        t = ctypes.c_uint64()
        _lib.check(self._L.clf_server_submit_framed(self._h, buf, len(b), int(tag), ctypes.byref(t)),
                   "clf_server_submit_framed")
        return int(t.value)

    def wait(self, ticket: int, timeout_ms: int = -1, release: bool = True):
        """(tracks bytes, stats dict, timing dict) of a request; releases it
        unless release=False."""
        o = Served()
        _lib.check(self._L.clf_server_wait(self._h, int(ticket), int(timeout_ms), ctypes.byref(o)),
                   "clf_server_wait")
        try:
            _lib.check(o.status, "served request")
            raw = ctypes.string_at(o.tracks, o.n_tracks * TRACK_BYTES) if o.n_tracks else b""
            stats = {k: int(getattr(o.stats, k)) for k, _ in EventStats._fields_}
            timing = {"submit_ns": int(o.t_submit_ns), "launch_ns": int(o.t_launch_ns),
                      "done_ns": int(o.t_done_ns), "batch_events": int(o.batch_events)}
        finally:
            if release:
                self.release(ticket)
        return raw, stats, timing

    def release(self, ticket: int):
        _lib.check(self._L.clf_server_release(self._h, int(ticket)), "clf_server_release")

    def replay(self, pool, pick, arrival_ns, submit_threads: int = 4):
        """Open-loop load replay (clf_server_replay).  Returns (records as a
        numpy structured array, summary dict)."""
        import numpy as np

        holders, ptrs, sizes = _event_arrays(pool)
        pick = np.ascontiguousarray(pick, np.int32)
        arr = np.ascontiguousarray(arrival_ns, np.int64)
        n = len(pick)
        recs = (ReplayRecord * max(n, 1))()
        summ = ReplaySummary()
        _lib.check(self._L.clf_server_replay(
            self._h, len(pool), ptrs, sizes, n, pick.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(submit_threads), recs,
            ctypes.byref(summ)), "clf_server_replay")
        del holders
        dt = np.dtype([(k, {ctypes.c_int64: "<i8", ctypes.c_uint64: "<u8", ctypes.c_int32: "<i4",
                             ctypes.c_uint32: "<u4"}[t]) for k, t in ReplayRecord._fields_])
        assert dt.itemsize == ctypes.sizeof(ReplayRecord)
        rec = np.frombuffer(bytes(recs), dt, count=n).copy() if n else np.zeros(0, dt)
        summary = {k: (float(getattr(summ, k)) if t is ctypes.c_double else int(getattr(summ, k)))
                   for k, t in ReplaySummary._fields_}
        return rec, summary


# ---------------------------------------------------------------------------
# Reference entry points
# ---------------------------------------------------------------------------
_default_ctx = None
_default_lock = threading.Lock()
_verbosity = 0


def _context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)  # DEVICE_NUMBER 0 (KernelDefinitions.h:7)
    return _default_ctx


def track_capacity(event: bytes) -> int:
    """Upper bound on the tracks of one event; -1 for a malformed blob."""
    L = _lib.load()
    b = bytes(event)
    h = (ctypes.c_uint8 * max(len(b), 1)).from_buffer_copy(b or b"\0")
    return int(L.clf_event_track_capacity(_u8(h), len(b)))


def presort_event(event: bytearray) -> None:
    """Sorts one event's hits by x per sensor in place (Tools.cpp:46-66)."""
    L = _lib.load()
    if not isinstance(event, bytearray):
        raise TypeError("presort_event sorts in place: pass a bytearray")
    h = (ctypes.c_uint8 * len(event)).from_buffer(event)
    if L.clf_presort_event(_u8(h), len(event)) != 0:
        raise ValueError("malformed event blob (or NaN x, on which the reference's sort never ends)")


def invoke_parallel_search(starting_event: int, events_to_process: int, input, output) -> int:
    """KernelInvoker.cpp:14-275: fills output[starting_event : +events_to_process]."""
    if starting_event < 0 or events_to_process < 0 or \
            starting_event + events_to_process > min(len(input), len(output)):
        raise ValueError("event range out of bounds")
    with _default_lock:
        chunk = input[starting_event: starting_event + events_to_process]
        outs, _ = _context().search(chunk, MODE_BATCHED)
    output[starting_event: starting_event + events_to_process] = outs
    return 0


def gpu_pixel_search_by_triplet_invocation(input, output: list) -> int:
    """GpuPixelSearchByTriplet.cpp:44-65: resizes output, runs chunks of <= 16000 events."""
    if _verbosity >= 3:
        print(f"Invoking gpuPixelSearchByTriplet with {len(input)} events")
    output[:] = [b""] * len(input)
    for i in range(0, len(input), MAX_EVENTS_PER_LAUNCH):
        invoke_parallel_search(i, min(MAX_EVENTS_PER_LAUNCH, len(input) - i), input, output)
    return 0


def gpu_pixel_search_by_triplet(input, output: list) -> int:
    """GpuPixelSearchByTriplet.cpp:29-37: Gaudi path, no presort, silent."""
    global _verbosity
    _verbosity = 0
    return gpu_pixel_search_by_triplet_invocation(input, output)


def independent_execute(input, output: list) -> int:
    """GpuPixelSearchByTriplet.cpp:4-22: presorts every event in place, then searches."""
    global _verbosity
    _verbosity = 3
    for ev in input:
        presort_event(ev)
    return gpu_pixel_search_by_triplet_invocation(input, output)


def independent_post_execute(output) -> None:
    """GpuPixelSearchByTriplet.cpp:24-27."""
    if _verbosity >= 3:
        print(f"\nSize of output: {len(output)} entries")
