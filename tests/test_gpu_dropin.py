"""The reference's drop-in entry points called from C++ (cl_forward_amd/
clf_dropin, a caller like the Gaudi offload handler), checked against the
oracle event by event:

* gpuPixelSearchByTriplet over 16,005 events, so that the chunking into
  launches of at most 16000 events (GpuPixelSearchByTriplet.cpp:44-65)
  runs a second chunk;
* invokeParallelSearch over ranges that start past the first event
  (KernelInvoker.cpp:14-31: output[startingEvent + i]), others left empty;
* gpuPixelSearchByTriplet over the batch1000 workload.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

from cl_forward_amd import event as ev_mod
from cl_forward_amd import generator
from tests.helpers import sub_event

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "cl_forward_amd", "clf_dropin")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(tmp_path, blobs, *args, timeout=300):
    src = tmp_path / "events.frames"
    out = tmp_path / "tracks.out"
    ev_mod.write_frame_stream(src, blobs)
    r = subprocess.run([DROPIN, str(src), "--out", str(out), *map(str, args)], capture_output=True,
                       text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return ev_mod.read_sized_stream(out), json.loads(r.stdout.strip().splitlines()[-1])


@pytest.fixture(scope="module")
def pool():
    """200 small distinct events: 8-sensor slices of the plain events,
    every third hit (about 130 hits each)."""
    out = []
    for i in range(200):
        ev = generator.plain_event(i % 51)
        lo = (i * 7) % 40
        out.append(sub_event(ev, lo, lo + 8, stride=3).pack())
    return out


def test_gaudi_entry_chunks_past_16000(tmp_path, oracle_mod, pool):
    n = 16005
    blobs = [pool[i % len(pool)] for i in range(n)]
    want = oracle_mod.search_batch(pool, n_threads=16, keep_output=True)[0]
    _say(f"gpuPixelSearchByTriplet over {n} events (chunks of 16000 and 5)")
    outs, timing = _run(tmp_path, blobs, "--entry", "gaudi")
    assert len(outs) == n and timing["events"] == n
    bad = [i for i in range(n) if outs[i] != want[i % len(pool)]]
    assert not bad, f"{len(bad)} events differ, first {bad[:5]}"
    assert sum(len(o) for o in outs) > 0


@pytest.mark.parametrize("start,count", [(37, 300), (16000, 5), (199, 1), (5, 0)])
def test_invoke_parallel_search_range(tmp_path, oracle_mod, pool, start, count):
    n = 16005
    blobs = [pool[i % len(pool)] for i in range(n)]
    want = oracle_mod.search_batch(pool, n_threads=16, keep_output=True)[0]
    _say(f"invokeParallelSearch({start}, {count}) over {n} input events")
    outs, _ = _run(tmp_path, blobs, "--entry", "invoke", "--start", start, "--count", count)
    assert len(outs) == n
    for i in range(n):
        if start <= i < start + count:
            assert outs[i] == want[i % len(pool)], i
        else:
            assert outs[i] == b"", i


def test_gaudi_entry_batch1000(tmp_path, oracle_mod):
    blobs = generator.workload_blobs("batch1000")
    want = oracle_mod.search_batch(blobs, n_threads=16, keep_output=True)[0]
    _say("gpuPixelSearchByTriplet over batch1000, 3 timed calls")
    outs, timing = _run(tmp_path, blobs, "--entry", "gaudi", "--reps", 3, "--warmup", 1)
    assert outs == want
    _say(f"drop-in call (host in, host out): median {timing['median_ms']} ms")
