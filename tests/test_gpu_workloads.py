"""The benchmark workloads and the hand-traced events on the GPU, checked
byte for byte against the oracle through the C ABI.

* batch1000 (1,000 events, Poisson(7.6) vertices x Poisson(20) tracks, 1 %
  noise): every event through the tuned launch replayed from its hipGraph,
  as the benchmark runs it, and the 16 costliest events by the work queue's
  cost estimate plus a stride sample through two other launches;
* stress1000 (Poisson(15) x Poisson(25), 5 % noise): the 16 costliest plus a
  stride sample through its tuned launch;
* the automatic launch shape of a context that sets none, on small and on
  large events;
* the events whose 24-slot records overflow, worked out by hand, and the
  hand-worked candidate windows.
"""
from __future__ import annotations

import struct
import sys

import numpy as np
import pytest

from cl_forward_amd import autotune, generator
from cl_forward_amd.search import MODE_BATCHED, MODE_PLAIN, Context
from tests.parity import assert_event_equal
from tests.helpers import sparse_line_event, window_pin_events

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def queue_cost(blob: bytes) -> float:
    """The work queue's cost estimate (N + sum over sensors of hits(s) x
    hits(s-2)), which orders events largest first."""
    S, N = struct.unpack_from("<ii", blob, 0)
    nums = np.frombuffer(blob, "<i4", count=S, offset=8 + 8 * S).astype(np.float64)
    return float(N + np.sum(nums[2:] * nums[:-2]))


def costliest_and_stride(blobs, n_top=16, n_stride=64):
    order = sorted(range(len(blobs)), key=lambda i: -queue_cost(blobs[i]))
    top = order[:n_top]
    rest = [i for i in range(len(blobs)) if i not in set(top)]
    step = max(1, len(rest) // n_stride)
    return sorted(top + rest[::step][:n_stride])


def _check(ctx, oracle, blobs, tag, mode=MODE_BATCHED, reps=1):
    want = oracle.search_batch(blobs, n_threads=16, keep_output=True)
    with ctx.batch(blobs) as b:
        for r in range(reps):
            b.run(mode)
            outs, stats = b.results()
            for i in range(len(blobs)):
                assert_event_equal(outs[i], stats[i], want[0][i], want[1][i], f"{tag}[{i}] run {r}")
    return want


@pytest.fixture(scope="module")
def batch1000():
    return generator.workload_blobs("batch1000")


def test_batch1000_tuned_graph_all_events(oracle_mod, batch1000):
    with Context(0) as ctx:
        cfg = autotune.load_tuned(ctx.num_cus, "batch1000") or {}
        cfg["use_graph"] = 1
        ctx.set_config(**cfg)
        _say(f"batch1000: all 1000 events, tuned launch {ctx.config()}, graph replays x2")
        want = _check(ctx, oracle_mod, batch1000, "batch1000/tuned", reps=2)
    st = want[1]
    # every event of this workload reaches the overflow point of the
    # 24-slot record at least once; each such track is counted
    assert sum(s["n_overflow"] > 0 for s in st) > 900
    assert all((s["flags"] & 1) == (s["n_overflow"] > 0) for s in st)


@pytest.mark.parametrize("block,lds_hits,graph", [(256, 256, 0), (1024, 512, 1)])
def test_batch1000_costliest_and_sample_other_launches(oracle_mod, batch1000, block, lds_hits, graph):
    idx = costliest_and_stride(batch1000)
    blobs = [batch1000[i] for i in idx]
    _say(f"batch1000: {len(blobs)} events (16 costliest + stride), block {block}, lds {lds_hits}, "
         f"graph {graph}")
    with Context(0) as ctx:
        ctx.set_config(block_size=block, lds_hits=lds_hits, use_graph=graph)
        _check(ctx, oracle_mod, blobs, f"batch1000/{block}x{lds_hits}")


@pytest.mark.parametrize("grain", [2, 5, 16])
def test_batch1000_queue_grain(oracle_mod, batch1000, grain):
    """Several consecutive events of the cost-ordered queue per dequeue
    (the last run shorter than the grain when it does not divide the
    count) give the same bytes."""
    idx = costliest_and_stride(batch1000, n_top=16, n_stride=51)
    blobs = [batch1000[i] for i in idx]
    _say(f"batch1000: {len(blobs)} events, queue grain {grain}")
    with Context(0) as ctx:
        cfg = autotune.load_tuned(ctx.num_cus, "batch1000") or {}
        cfg.update(use_graph=1, grain=grain)
        ctx.set_config(**cfg)
        _check(ctx, oracle_mod, blobs, f"batch1000/grain{grain}")


def test_batch1000_sample_plain_mode(oracle_mod, batch1000):
    idx = costliest_and_stride(batch1000, n_top=4, n_stride=12)
    blobs = [batch1000[i] for i in idx]
    _say(f"batch1000: {len(blobs)} events, plain mode (one launch per event)")
    with Context(0) as ctx:
        _check(ctx, oracle_mod, blobs, "batch1000/plain", mode=MODE_PLAIN)


def test_stress1000_costliest_and_sample(oracle_mod):
    blobs_all = generator.workload_blobs("stress1000")
    idx = costliest_and_stride(blobs_all, n_top=16, n_stride=48)
    blobs = [blobs_all[i] for i in idx]
    with Context(0) as ctx:
        cfg = autotune.load_tuned(ctx.num_cus, "stress1000") or {}
        cfg["use_graph"] = 1
        ctx.set_config(**cfg)
        _say(f"stress1000: {len(blobs)} events (16 costliest + stride), tuned {ctx.config()}")
        _check(ctx, oracle_mod, blobs, "stress1000/tuned")


def test_automatic_launch_shape(oracle_mod, batch1000):
    """A context starts with block_size 0: the launch shape is chosen per
    batch from its mean event size (above 4096 hits 1024 threads with
    512-hit slots, else 512 with 256).  A batch of batch1000 events (mean
    2.8k hits) and one of the 16 costliest stress1000 events (mean 12.4k)
    both come out identical to the oracle, twice from their graphs; a set
    block size with lds_hits 0 takes 256-hit slots."""
    small = [batch1000[i] for i in costliest_and_stride(batch1000, n_top=2, n_stride=30)]
    stress = generator.workload_blobs("stress1000")
    big = [stress[i] for i in sorted(range(len(stress)), key=lambda i: -queue_cost(stress[i]))[:16]]
    with Context(0) as ctx:
        c = ctx.config()
        assert c["block_size"] == 0 and c["lds_hits"] == 0, c
        for name, blobs, large in (("small", small, False), ("large", big, True)):
            mean = float(np.mean([struct.unpack_from("<ii", b, 0)[1] for b in blobs]))
            assert (mean > 4096) == large, mean
            _say(f"automatic launch: {len(blobs)} events of {mean:.0f} hits on average")
            _check(ctx, oracle_mod, blobs, f"auto/{name}", reps=2)
        ctx.set_config(block_size=256, lds_hits=0)
        _say("block size 256, lds_hits 0")
        _check(ctx, oracle_mod, small[:8], "block256/lds0")


PLANES48 = [-300 + 22 * k for k in range(48)]


@pytest.mark.parametrize("drop", [(), (21,)])
def test_full_record_overflow_events(oracle_mod, drop):
    """One line through 48 planes: two tracks fill their 24 slots and find a
    25th hit (marked used, then the track stops), with and without a skipped
    plane after the record is full."""
    keep = [k for k in range(48) if k not in drop]
    blob = sparse_line_event([(1.0, 2.0, 0.01, 0.02, keep)], PLANES48)
    _say(f"full-record events, dropped planes {drop}")
    with Context(0) as ctx:
        for block, lds in ((64, 256), (512, 256), (1024, 512)):
            ctx.set_config(block_size=block, lds_hits=lds)
            want = _check(ctx, oracle_mod, [blob], f"overflow{drop}/{block}")
            assert want[1][0]["n_overflow"] == 2


def test_windows_hand_worked_on_gpu(gpu_ctx, oracle_mod):
    cases = window_pin_events()
    got = gpu_ctx.fill_candidates([c[1] for c in cases])
    for (name, blob, want), (g1, g2, flags) in zip(cases, got):
        N = len(g1) // 2
        for h in range(N):
            w1, w2 = want.get(h, ((-1, -1), (-1, -1)))
            assert (int(g1[2 * h]), int(g1[2 * h + 1])) == w1, (name, h)
            assert (int(g2[2 * h]), int(g2[2 * h + 1])) == w2, (name, h)
        assert flags == 0
    # and through the whole search (the windows the seeding uses)
    for name, blob, _ in cases:
        o_out, o_st = oracle_mod.search_event(blob)
        outs, stats = gpu_ctx.search([blob])
        assert_event_equal(outs[0], stats[0], o_out, o_st, name)


def test_event_above_lds_used_flags(oracle_mod):
    """An event of 37,054 hits (a plain event plus 720 banded noise hits per
    sensor): above 32,768 hits the used flags live in HBM bytes instead of
    the LDS bitmask, and every sensor exceeds an LDS slot."""
    from cl_forward_amd import generator
    from tests.helpers import banded_noise_event

    blob = banded_noise_event(generator.plain_event(3), 720, 5).pack()
    n = struct.unpack_from("<i", blob, 4)[0]
    assert n > 32768
    _say(f"event of {n} hits (used flags in HBM)")
    with Context(0) as ctx:
        for block, lds in ((512, 256), (1024, 512), (256, 256)):
            ctx.set_config(block_size=block, lds_hits=lds)
            _check(ctx, oracle_mod, [blob, generator.plain_event(4).pack()], f"big/{block}x{lds}")


@pytest.fixture(scope="module")
def window_events(mcdata_sorted, mcdata_raw, plain51_blobs):
    from cl_forward_amd import generator

    blobs = [c[1] for c in window_pin_events()]
    blobs += [mcdata_sorted[k] for k in ("0", "49", "77")] + [mcdata_raw["15"]]
    blobs += plain51_blobs[:4] + [generator.stress_event(i).pack() for i in (0, 7)]
    blobs += [generator.batch_event(i).pack() for i in (1, 2)]
    return blobs


@pytest.mark.parametrize("block,lds_hits", [(64, 256), (256, 256), (512, 256), (512, 512),
                                            (1024, 256)])
def test_windows_search_path(oracle_mod, window_events, block, lds_hits):
    """The windows as the search computes them (ring slots with the x-bucket
    index, staged by the configuration's own staging code; sensors above
    the slot size from HBM) against the oracle's fillCandidates."""
    _say(f"search-path windows: block {block}, lds {lds_hits}, {len(window_events)} events")
    with Context(0) as ctx:
        ctx.set_config(block_size=block, lds_hits=lds_hits)
        got = ctx.fill_candidates(window_events, search_path=True)
    for i, (b, (g1, g2, flags)) in enumerate(zip(window_events, got)):
        o1, o2 = oracle_mod.fill_candidates(b)
        bad1 = np.nonzero(g1 != o1)[0][:5]
        bad2 = np.nonzero(g2 != o2)[0][:5]
        assert bad1.size == 0 and bad2.size == 0, (
            f"event {i}: h1 {bad1.tolist()} {g1[bad1].tolist()} vs {o1[bad1].tolist()}; "
            f"h2 {bad2.tolist()} {g2[bad2].tolist()} vs {o2[bad2].tolist()}")
        assert flags == oracle_mod.input_flags(b)
