"""MI355X-native search-by-triplet VELO pixel tracker.

The compute path is cl_forward_amd/libclf.so (hand-written gfx950 HIP
kernels behind the C ABI of include/clf.h).  This package is the host-side
mirror of the reference's entry points plus event I/O and the synthetic
workloads.
"""
from .event import (Event, TRACK_BYTES, TRACK_DTYPE, frame, pack_event, parse_event, read_dat,
                    track_lists, tracks_from_bytes, unframe)

__version__ = "0.1.0"


_COMPUTE_API = (
    "Context", "Batch", "ClfError", "MODE_PLAIN", "MODE_BATCHED", "gpu_pixel_search_by_triplet",
    "gpu_pixel_search_by_triplet_invocation", "independent_execute", "independent_post_execute",
    "invoke_parallel_search", "presort_event", "track_capacity",
)


def __getattr__(name):
    # The compute API is resolved lazily so that event I/O and the generator
    # import on machines without the HIP library; calling it without the
    # library raises ImportError (there is no CPU fallback).
    if name in _COMPUTE_API:
        import importlib

        return getattr(importlib.import_module(".search", __name__), name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
