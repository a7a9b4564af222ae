"""ctypes binding of the CPU oracle (oracle/libclf_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/, the
smoke check and the benchmark's CPU-baseline leg import this module; the
product package (cl_forward_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclf_oracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_tools.so")
TRACK_BYTES = 100
MAX_SENSORS = 256


class Stats(ctypes.Structure):
    _fields_ = [
        ("n_tracks", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("n_tracklets", ctypes.c_int32),
        ("n_weak", ctypes.c_int32),
        ("max_ttf_live", ctypes.c_int32),
        ("n_full", ctypes.c_int32),
        ("n_overflow", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("seed_h0_visits", ctypes.c_int64),
        ("seed_pairs", ctypes.c_int64),
        ("seed_triplets", ctypes.c_int64),
        ("fwd_fits", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "pad"}


def build(force=False):
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "-s", "libclf_oracle.so"],
                       check=True)
    if os.path.isdir("/root/reference/mcdata") and (force or not os.path.exists(REF_LIB_PATH)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.clf_oracle_fill_candidates.argtypes = [u8p, ctypes.c_size_t,
                                                 ctypes.POINTER(ctypes.c_int32),
                                                 ctypes.POINTER(ctypes.c_int32)]
        L.clf_oracle_fill_candidates.restype = ctypes.c_int
        L.clf_oracle_search_event.argtypes = [u8p, ctypes.c_size_t, u8p, ctypes.c_int64,
                                              ctypes.POINTER(Stats)]
        L.clf_oracle_search_event.restype = ctypes.c_int
        L.clf_oracle_search_batch.argtypes = [ctypes.c_int, ctypes.POINTER(u8p),
                                              ctypes.POINTER(ctypes.c_size_t),
                                              ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_int64),
                                              ctypes.POINTER(Stats), ctypes.c_int]
        L.clf_oracle_search_batch.restype = ctypes.c_int
        L.clf_oracle_presort.argtypes = [u8p, ctypes.c_size_t]
        L.clf_oracle_presort.restype = ctypes.c_int
        L.clf_oracle_input_flags.argtypes = [u8p, ctypes.c_size_t]
        L.clf_oracle_input_flags.restype = ctypes.c_uint32
        L.clf_oracle_gather.argtypes = [ctypes.c_int, u8p, ctypes.POINTER(ctypes.c_int32)]
        L.clf_oracle_gather.restype = ctypes.c_int
        _lib = L
    return _lib


def _u8(buf):
    return ctypes.cast(buf, ctypes.POINTER(ctypes.c_uint8))


def _header(blob):
    S, N = np.frombuffer(bytes(blob[:8]), "<i4", 2)
    return int(S), int(N)


def track_capacity(blob) -> int:
    S, N = _header(blob)
    return N + 64 * S + 64


def fill_candidates(blob: bytes):
    """(cand_h1, cand_h2) int32[2N] each, as Kernel.cl:50-178 leaves them."""
    S, N = _header(blob)
    c1 = np.empty(2 * max(N, 1), np.int32)
    c2 = np.empty(2 * max(N, 1), np.int32)
    b = (ctypes.c_uint8 * len(blob)).from_buffer_copy(blob)
    rc = lib().clf_oracle_fill_candidates(_u8(b), len(blob),
                                          c1.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          c2.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if rc != 0:
        raise ValueError("malformed event blob")
    return c1[: 2 * N], c2[: 2 * N]


def search_event(blob: bytes):
    """(output blob of 100-byte Track records, stats dict) for one event."""
    b = (ctypes.c_uint8 * max(len(blob), 1)).from_buffer_copy(blob) if len(blob) else (ctypes.c_uint8 * 1)()
    try:
        cap = track_capacity(blob)
    except Exception:
        cap = 0
    out = (ctypes.c_uint8 * max(cap * TRACK_BYTES, 1))()
    st = Stats()
    rc = lib().clf_oracle_search_event(_u8(b), len(blob), _u8(out), cap, ctypes.byref(st))
    if rc == -1:
        return b"", st.as_dict()
    if rc != 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return bytes(out)[: st.n_tracks * TRACK_BYTES], st.as_dict()


def search_batch(blobs, n_threads=1, keep_output=True):
    """Oracle over many events on n_threads host threads.  Returns
    (list of output blobs or None, list of stats dicts)."""
    n = len(blobs)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    bufs = [(ctypes.c_uint8 * max(len(b), 1)).from_buffer_copy(b) if len(b) else (ctypes.c_uint8 * 1)()
            for b in blobs]
    caps = (ctypes.c_int64 * n)(*[track_capacity(b) for b in blobs])
    outs = [(ctypes.c_uint8 * (caps[i] * TRACK_BYTES))() for i in range(n)] if keep_output else None
    pb = (u8p * n)(*[_u8(b) for b in bufs])
    ps = (ctypes.c_size_t * n)(*[len(b) for b in blobs])
    po = (u8p * n)(*[_u8(o) for o in outs]) if keep_output else None
    stats = (Stats * n)()
    failed = lib().clf_oracle_search_batch(n, pb, ps, po, caps if keep_output else None,
                                           stats, int(n_threads))
    if failed:
        raise RuntimeError(f"oracle failed on {failed} events")
    sts = [stats[i].as_dict() for i in range(n)]
    if not keep_output:
        return None, sts
    return [bytes(outs[i])[: sts[i]["n_tracks"] * TRACK_BYTES] for i in range(n)], sts


def gather(used) -> list[int]:
    """Visited element indices of one sensor step's h0 gather, in lane order."""
    u = np.ascontiguousarray(np.asarray(used, dtype=np.uint8))
    n = int(u.size)
    out = np.empty(n + 64, np.int32)
    nv = lib().clf_oracle_gather(n, u.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) if n else
                                 ctypes.cast((ctypes.c_uint8 * 1)(), ctypes.POINTER(ctypes.c_uint8)),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out[:nv].tolist()


def presort(blob: bytes) -> bytes:
    b = (ctypes.c_uint8 * len(blob)).from_buffer_copy(blob)
    if lib().clf_oracle_presort(_u8(b), len(blob)) != 0:
        raise ValueError("malformed event blob")
    return bytes(b)


def input_flags(blob: bytes) -> int:
    b = (ctypes.c_uint8 * max(len(blob), 1)).from_buffer_copy(blob) if len(blob) else (ctypes.c_uint8 * 1)()
    return int(lib().clf_oracle_input_flags(_u8(b), len(blob)))


# ---------------------------------------------------------------------------
# The reference's own presort and parser (oracle/_ref), when built
# ---------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_LIB_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.ref_preorder_by_x.argtypes = [u8p, ctypes.c_size_t]
        L.ref_preorder_by_x.restype = ctypes.c_int
        L.ref_layout_offsets.argtypes = [u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int64)]
        L.ref_layout_offsets.restype = ctypes.c_int
        _ref = L
    return _ref


def ref_presort(blob: bytes) -> bytes:
    b = (ctypes.c_uint8 * len(blob)).from_buffer_copy(blob)
    ref_lib().ref_preorder_by_x(_u8(b), len(blob))
    return bytes(b)


def ref_layout_offsets(blob: bytes):
    b = (ctypes.c_uint8 * len(blob)).from_buffer_copy(blob)
    offs = (ctypes.c_int64 * 9)()
    ref_lib().ref_layout_offsets(_u8(b), len(blob), offs)
    return list(offs)
