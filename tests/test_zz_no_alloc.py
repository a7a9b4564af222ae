"""clf_search_noalloc (the entry point that allocates nothing) against
clf_search through the C ABI: byte-identical tracks and identical stats on
two calls in a row, the second smaller, in both modes; no device memory
taken by the call; its capacity errors."""
from __future__ import annotations

import ctypes
import sys

import pytest

from cl_forward_amd import autotune, generator
from cl_forward_amd._lib import ClfError
from cl_forward_amd.search import MODE_BATCHED, MODE_PLAIN, Context
from tests.helpers import banded_noise_event, raw_blob

pytestmark = pytest.mark.gpu

CLF_E_CAPACITY = 5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _device_free_bytes() -> int:
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return int(free.value)


def _inputs():
    batch = generator.workload_blobs("batch1000")
    first = [batch[i] for i in range(0, 1000, 25)]  # 40 events
    first += [generator.plain_event(i).pack() for i in range(3)]
    first.append(b"\x01\x00\x00\x00")  # malformed: CLF_FLAG_INVALID, no tracks
    second = [generator.plain_event(i).pack() for i in (7, 8)] + [batch[999], batch[3]]
    return first, second


@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_PLAIN])
def test_noalloc_equals_clf_search(oracle_mod, mode):
    first, second = _inputs()
    with Context(0) as ctx:
        cfg = autotune.load_tuned(ctx.num_cus, "batch1000") or {}
        ctx.set_config(**cfg)
        for tag, blobs in (("first", first), ("second, smaller", second)):
            _say(f"no-alloc, mode {mode}: {tag} call, {len(blobs)} events")
            want_out, want_st = ctx.search(blobs, mode)
            free0 = _device_free_bytes()
            got_out, got_st = ctx.search_noalloc(blobs, mode)
            free1 = _device_free_bytes()
            assert free1 == free0, f"the call took {free0 - free1} bytes of device memory"
            assert got_out == want_out, tag
            assert got_st == want_st, tag
            for b, o in zip(blobs[:8], got_out[:8]):
                if len(b) > 8:
                    assert o == oracle_mod.search_event(b)[0]


def test_noalloc_capacity_errors():
    with Context(0) as ctx:
        tiny = raw_blob(0, 0, [], [], [], [], [], [])
        _say("no-alloc: 4097 events (over the per-call maximum)")
        with pytest.raises(ClfError) as e:
            ctx.search_noalloc([tiny] * 4097)
        assert e.value.status == CLF_E_CAPACITY
        big = banded_noise_event(generator.plain_event(3), 720, 5).pack()
        _say("no-alloc: 12 events of 37,054 hits (over the scratch array)")
        with pytest.raises(ClfError) as e:
            ctx.search_noalloc([big] * 12)
        assert e.value.status == CLF_E_CAPACITY
        _say("no-alloc: an output of one record for 288 tracks")
        with pytest.raises(ClfError) as e:
            ctx.search_noalloc([generator.single_event().pack()], capacity=1)
        assert e.value.status == CLF_E_CAPACITY
        # the array is usable after the refusals, with the big event alone
        out, st = ctx.search_noalloc([big, generator.single_event().pack()])
        want, wst = ctx.search([big, generator.single_event().pack()])
        assert out == want and st == wst
