"""ctypes binding of cl_forward_amd/libclf.so (the C ABI of include/clf.h).

The library is required: importing the compute entry points without it
raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclf.so")

CLF_OK = 0
MODE_PLAIN = 0
MODE_BATCHED = 1


class EventStats(ctypes.Structure):
    _fields_ = [
        ("n_tracks", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("n_tracklets", ctypes.c_int32),
        ("n_weak", ctypes.c_int32),
        ("max_ttf_live", ctypes.c_int32),
        ("n_full", ctypes.c_int32),
        ("h0_visits", ctypes.c_int32),
        ("n_overflow", ctypes.c_int32),
    ]


class LaunchConfig(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_int32),
        ("grid_size", ctypes.c_int32),
        ("grain", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("lds_hits", ctypes.c_int32),
        ("heavy_events", ctypes.c_int32),
        ("heavy_block", ctypes.c_int32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("n_events", ctypes.c_int32),
        ("stats", ctypes.POINTER(EventStats)),
        ("track_offset", ctypes.POINTER(ctypes.c_int64)),
        ("tracks", ctypes.POINTER(ctypes.c_uint8)),
    ]


class Timing(ctypes.Structure):
    _fields_ = [("ms_run", ctypes.c_float), ("ms_search", ctypes.c_float), ("reps", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("engine", ctypes.c_int32),
        ("max_requests", ctypes.c_int32),
        ("arena_bytes", ctypes.c_int64),
        ("out_arena_bytes", ctypes.c_int64),
        ("max_event_hits", ctypes.c_int32),
        ("idle_us", ctypes.c_int32),
        ("slots", ctypes.c_int32),
        ("max_batch_events", ctypes.c_int32),
        ("max_batch_hits", ctypes.c_int64),
        ("max_wait_us", ctypes.c_int32),
        ("feeder_pause_us", ctypes.c_int32),
    ]


SERVE_RESIDENT = 0
SERVE_BATCHES = 1


class Served(ctypes.Structure):
    _fields_ = [
        ("ticket", ctypes.c_uint64),
        ("tag", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("batch_events", ctypes.c_int32),
        ("stats", EventStats),
        ("n_tracks", ctypes.c_int64),
        ("tracks", ctypes.POINTER(ctypes.c_uint8)),
        ("t_submit_ns", ctypes.c_int64),
        ("t_launch_ns", ctypes.c_int64),
        ("t_done_ns", ctypes.c_int64),
        ("t_complete_ns", ctypes.c_int64),
    ]


class ReplayRecord(ctypes.Structure):
    _fields_ = [
        ("latency_ns", ctypes.c_int64),
        ("wait_ns", ctypes.c_int64),
        ("digest", ctypes.c_uint64),
        ("n_tracks", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("batch_events", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("device_ns", ctypes.c_int64),
    ]


class ReplaySummary(ctypes.Structure):
    _fields_ = [
        ("wall_s", ctypes.c_double),
        ("events_per_s", ctypes.c_double),
        ("completed", ctypes.c_int64),
        ("rejected", ctypes.c_int64),
        ("p50_ns", ctypes.c_int64),
        ("p90_ns", ctypes.c_int64),
        ("p99_ns", ctypes.c_int64),
        ("max_ns", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("mean_batch_events", ctypes.c_double),
        ("max_submit_lag_ns", ctypes.c_double),
        ("dev_copy_in_us", ctypes.c_double),
        ("dev_search_us", ctypes.c_double),
        ("dev_copy_out_us", ctypes.c_double),
        ("dev_busy_frac", ctypes.c_double),
        ("host_submit_us", ctypes.c_double),
        ("host_digest_us", ctypes.c_double),
    ]


class ClfError(RuntimeError):
    """A C-ABI call returned a status other than CLF_OK."""

    def __init__(self, status: int, what: str, message: str):
        super().__init__(f"{what}: status {status} ({message})")
        self.status = status


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m cl_forward_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    u8p = ctypes.POINTER(ctypes.c_uint8)
    u8pp = ctypes.POINTER(u8p)
    sz_p = ctypes.POINTER(ctypes.c_size_t)
    sig = {
        "clf_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "clf_last_error": (ctypes.c_char_p, []),
        "clf_version": (ctypes.c_char_p, []),
        "clf_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "clf_context_destroy": (None, [vp]),
        "clf_context_set_config": (ctypes.c_int, [vp, ctypes.POINTER(LaunchConfig)]),
        "clf_context_get_config": (ctypes.c_int, [vp, ctypes.POINTER(LaunchConfig)]),
        "clf_context_num_cus": (ctypes.c_int, [vp]),
        "clf_context_synchronize": (ctypes.c_int, [vp]),
        "clf_event_track_capacity": (ctypes.c_int64, [u8p, ctypes.c_size_t]),
        "clf_batch_create": (ctypes.c_int, [vp, ctypes.c_int32, u8pp, sz_p, ctypes.POINTER(vp)]),
        "clf_batch_destroy": (None, [vp]),
        "clf_batch_run": (ctypes.c_int, [vp, vp, ctypes.c_int]),
        "clf_batch_results": (ctypes.c_int, [vp, vp, ctypes.POINTER(Result)]),
        "clf_batch_time": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.POINTER(Timing)]),
        "clf_search": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int32, u8pp, sz_p,
                                      ctypes.POINTER(Result), ctypes.POINTER(vp)]),
        "clf_presort_event": (ctypes.c_int, [u8p, ctypes.c_size_t]),
        "clf_probe_copy": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_float)]),
        "clf_batch_profile": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_double)]),
        "clf_fill_candidates": (ctypes.c_int, [vp, ctypes.c_int32, u8pp, sz_p,
                                               ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)),
                                               ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)),
                                               ctypes.POINTER(ctypes.c_uint32)]),
        "clf_fill_candidates_search_path": (ctypes.c_int, [vp, ctypes.c_int32, u8pp, sz_p,
                                                           ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)),
                                                           ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)),
                                                           ctypes.POINTER(ctypes.c_uint32)]),
        "clf_search_noalloc": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int32, u8pp, sz_p, u8p,
                                              ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                              ctypes.POINTER(EventStats)]),
        "clf_probe_link": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "clf_server_default_config": (None, [ctypes.POINTER(ServerConfig)]),
        "clf_server_create": (ctypes.c_int, [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
        "clf_server_destroy": (None, [vp]),
        "clf_server_submit": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, ctypes.c_uint64,
                                             ctypes.POINTER(ctypes.c_uint64)]),
        "clf_server_submit_many": (ctypes.c_int, [vp, ctypes.c_int32, u8pp, sz_p,
                                                  ctypes.POINTER(ctypes.c_uint64),
                                                  ctypes.POINTER(ctypes.c_uint64),
                                                  ctypes.POINTER(ctypes.c_int32)]),
        "clf_server_wait": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(Served)]),
        "clf_server_release": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "clf_server_replay": (ctypes.c_int, [vp, ctypes.c_int32, u8pp, sz_p, ctypes.c_int64,
                                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                             ctypes.c_int32, ctypes.POINTER(ReplayRecord),
                                             ctypes.POINTER(ReplaySummary)]),
        "clf_tracks_digest": (ctypes.c_uint64, [u8p, ctypes.c_int64]),
        "clf_server_submit_framed": (ctypes.c_int, [vp, u8p, ctypes.c_size_t, ctypes.c_uint64,
                                                    ctypes.POINTER(ctypes.c_uint64)]),
        "clf_unframe": (ctypes.c_int, [u8p, ctypes.c_size_t, ctypes.POINTER(u8p),
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(u8p),
                                       ctypes.POINTER(ctypes.c_size_t)]),
        "clf_format_tracks": (ctypes.c_int, [u8p, ctypes.c_size_t, u8p, ctypes.c_int64,
                                             ctypes.c_char_p, ctypes.c_size_t,
                                             ctypes.POINTER(ctypes.c_size_t)]),
        "clf_context_device_name": (ctypes.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


#: every C symbol include/clf.h declares
EXPORTED = (
    "clf_status_string", "clf_last_error", "clf_version", "clf_context_create",
    "clf_context_destroy", "clf_context_set_config", "clf_context_get_config",
    "clf_context_num_cus", "clf_context_synchronize", "clf_event_track_capacity", "clf_batch_create", "clf_batch_destroy",
    "clf_batch_run", "clf_batch_results", "clf_batch_time", "clf_search", "clf_search_noalloc", "clf_presort_event",
    "clf_fill_candidates", "clf_fill_candidates_search_path", "clf_probe_copy", "clf_probe_link", "clf_batch_profile", "clf_server_default_config",
    "clf_server_create", "clf_server_destroy", "clf_server_submit", "clf_server_submit_many", "clf_server_wait",
    "clf_server_release", "clf_server_replay", "clf_tracks_digest", "clf_server_submit_framed",
    "clf_unframe", "clf_format_tracks", "clf_context_device_name",
)


def check(status: int, what: str):
    if status != CLF_OK:
        L = load()
        raise ClfError(status, what, (L.clf_last_error() or b"").decode(errors="replace"))
