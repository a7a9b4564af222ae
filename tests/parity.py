"""Comparison helpers for the parity tests.

Two outputs of one event are compared three ways, strongest first:

* bytes   — the 100-byte Track records in output order (the HIP path and the
            oracle both produce the oracle's lane order);
* canonical — the multiset of tracks, each as its tuple of hit indices in
            hit order (the reference's own output order is atomic order, so
            this is the form its results can be compared in);
* id-sets — the multiset of tracks as sets of hit IDs (the weakest form the
            reference's comparison tooling used).
"""
from __future__ import annotations

from collections import Counter

import numpy as np

from cl_forward_amd import event as ev_mod

# counters the GPU path reports next to the oracle's
STAT_PAIRS = (
    ("n_tracks", "n_tracks"),
    ("flags", "flags"),
    ("n_tracklets", "n_tracklets"),
    ("n_weak", "n_weak"),
    ("max_ttf_live", "max_ttf_live"),
    ("n_full", "n_full"),
    ("h0_visits", "seed_h0_visits"),
    ("n_overflow", "n_overflow"),
)


def canonical(out: bytes):
    return sorted(ev_mod.track_lists(out))


def id_sets(out: bytes, blob: bytes):
    ids = ev_mod.parse_event(blob).hit_IDs
    return Counter(frozenset(int(ids[h]) for h in t) for t in ev_mod.track_lists(out))


def describe_diff(a: bytes, b: bytes, limit=5) -> str:
    ca, cb = Counter(ev_mod.track_lists(a)), Counter(ev_mod.track_lists(b))
    only_a = list((ca - cb).elements())[:limit]
    only_b = list((cb - ca).elements())[:limit]
    return (f"{len(a)//100} vs {len(b)//100} tracks; canonical equal={ca == cb}; "
            f"only in first: {only_a}; only in second: {only_b}")


def assert_event_equal(gpu_out: bytes, gpu_stats: dict, ora_out: bytes, ora_stats: dict, tag=""):
    if gpu_out != ora_out:
        raise AssertionError(f"{tag}: output bytes differ: {describe_diff(gpu_out, ora_out)}")
    for g, o in STAT_PAIRS:
        if gpu_stats[g] != ora_stats[o]:
            raise AssertionError(f"{tag}: {g} {gpu_stats[g]} != oracle {o} {ora_stats[o]}")


def windows_equal(g1, g2, o1, o2) -> bool:
    return np.array_equal(np.asarray(g1), np.asarray(o1)) and np.array_equal(np.asarray(g2), np.asarray(o2))
