"""Serving loop on the GPU, both engines (the resident kernel and the
batch dispatcher): every served event's tracks must be byte-identical to the
plain path's for the same event, whether it ran alone or among others, and
to the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from cl_forward_amd._lib import SERVE_BATCHES, SERVE_RESIDENT
from cl_forward_amd.search import MODE_PLAIN, ClfError, Server, tracks_digest
from tests.helpers import line_event

pytestmark = pytest.mark.gpu

ENGINES = pytest.mark.parametrize("engine", [SERVE_RESIDENT, SERVE_BATCHES],
                                  ids=["resident", "batches"])


@pytest.fixture(scope="module")
def plain_ref(gpu_ctx, plain51_blobs):
    outs, stats = gpu_ctx.search(plain51_blobs, MODE_PLAIN)
    return outs, stats


@ENGINES
def test_serve_single_matches_plain_and_oracle(gpu_ctx, oracle_mod, single_blob, engine):
    ref, _ = gpu_ctx.search([single_blob], MODE_PLAIN)
    want = oracle_mod.search_event(single_blob)[0]
    with Server(gpu_ctx, engine=engine, slots=2, max_batch_events=16,
                max_batch_hits=1 << 16) as srv:
        t = srv.submit(single_blob, tag=7)
        raw, stats, timing = srv.wait(t, timeout_ms=20000)
    assert raw == ref[0] == want
    assert stats["n_tracks"] == len(raw) // 100
    assert timing["batch_events"] == 1
    assert timing["submit_ns"] <= timing["done_ns"]
    if engine == SERVE_BATCHES:
        assert timing["submit_ns"] <= timing["launch_ns"] <= timing["done_ns"]


@ENGINES
def test_serve_mix_equals_alone(gpu_ctx, plain51_blobs, plain_ref, mcdata_sorted, engine):
    """All events submitted at once (they share batches) and then one at a
    time (each alone) give the plain path's bytes."""
    ref_outs, _ = plain_ref
    mc = [mcdata_sorted[k] for k in list(mcdata_sorted)[:12]]
    mc_ref, _ = gpu_ctx.search(mc, MODE_PLAIN)
    events = plain51_blobs + mc
    want = ref_outs + mc_ref
    with Server(gpu_ctx, engine=engine, slots=3, max_batch_events=32, max_batch_hits=1 << 18,
                max_wait_us=200) as srv:
        tickets = [srv.submit(e, tag=i) for i, e in enumerate(events)]
        got = [srv.wait(t, timeout_ms=30000) for t in tickets]
        if engine == SERVE_BATCHES:
            assert any(g[2]["batch_events"] > 1 for g in got), "nothing was batched"
        for i, (raw, stats, _) in enumerate(got):
            assert raw == want[i], f"event {i} differs in the mix"
        for i in (0, 5, 51, len(events) - 1):
            raw, _, timing = srv.wait(srv.submit(events[i]), timeout_ms=30000)
            assert timing["batch_events"] == 1  # alone
            assert raw == want[i], f"event {i} differs alone"


@ENGINES
def test_serve_submit_many(gpu_ctx, plain51_blobs, plain_ref, mcdata_sorted, engine):
    """clf_server_submit_many queues a whole group in one call with the same
    results as one submit each, and stops at the first event refused: the
    ones before it stay queued and finish exactly."""
    ref_outs, _ = plain_ref
    mc = [mcdata_sorted[k] for k in list(mcdata_sorted)[:12]]
    mc_ref, _ = gpu_ctx.search(mc, MODE_PLAIN)
    events = plain51_blobs + mc
    want = ref_outs + mc_ref
    with Server(gpu_ctx, engine=engine, slots=3, max_batch_events=32, max_batch_hits=1 << 18,
                max_wait_us=200) as srv:
        tickets = srv.submit_many(events[:40], tags=range(40)) + srv.submit_many(events[40:])
        assert len(tickets) == len(events) and 0 not in tickets
        for i, t in enumerate(tickets):
            assert srv.wait(t, timeout_ms=30000)[0] == want[i], f"event {i} differs"
        assert srv.submit_many([]) == []
        with pytest.raises(ClfError) as e:  # malformed second event
            srv.submit_many([events[3], events[4][:100], events[5]])
        assert e.value.status == 1 and len(e.value.tickets) == 1
        assert srv.wait(e.value.tickets[0], timeout_ms=30000)[0] == want[3]
    small = line_event([(1.0, 2.0, 0.01, -0.02), (-3.0, 1.0, -0.03, 0.01)],
                       [-300 + 22 * k for k in range(20)]).pack()
    with Server(gpu_ctx, engine=engine, slots=1, max_batch_events=4, max_batch_hits=1000,
                max_event_hits=1000, max_requests=2) as srv:
        with pytest.raises(ClfError) as e:  # request table full after two
            srv.submit_many([small, small, small])
        assert e.value.status == 5 and len(e.value.tickets) == 2
        a, b = (srv.wait(t, timeout_ms=20000)[0] for t in e.value.tickets)
        assert a == b


@ENGINES
def test_serve_replay_digests(gpu_ctx, plain51_blobs, plain_ref, engine):
    """Open-loop replay of 4000 arrivals drawn from the 51 events; every
    result's digest equals the plain path's."""
    ref_outs, ref_stats = plain_ref
    want = np.array([tracks_digest(o) for o in ref_outs], np.uint64)
    rng = np.random.default_rng(4050)
    n = 4000
    pick = rng.integers(0, len(plain51_blobs), n).astype(np.int32)
    arrival = np.cumsum(rng.exponential(1e9 / 100000.0, n)).astype(np.int64)
    with Server(gpu_ctx, engine=engine, slots=4, max_batch_events=128,
                max_batch_hits=1 << 19) as srv:
        rec, summ = srv.replay(plain51_blobs, pick, arrival, submit_threads=2)
    ok = rec["status"] == 0
    assert summ["rejected"] == int((~ok).sum()) == 0
    assert summ["completed"] == n
    assert np.array_equal(rec["digest"], want[pick])
    assert np.array_equal(rec["n_tracks"], np.array([s["n_tracks"] for s in ref_stats])[pick])
    assert (rec["latency_ns"] > 0).all() and summ["p99_ns"] >= summ["p50_ns"] > 0
    assert summ["batches"] >= 1 and summ["mean_batch_events"] >= 1.0


@ENGINES
def test_serve_small_arena_wraps(gpu_ctx, plain51_blobs, plain_ref, engine):
    """A 1 MiB input arena holds about 25 of these events: the ring wraps
    many times, submits beyond it are refused with CLF_E_CAPACITY, and every
    accepted event is still exact."""
    ref_outs, _ = plain_ref
    want = np.array([tracks_digest(o) for o in ref_outs], np.uint64)
    with Server(gpu_ctx, engine=engine, slots=2, max_batch_events=8, max_batch_hits=1 << 16,
                arena_bytes=1 << 20, out_arena_bytes=8 << 20) as srv:
        for rnd in range(3):
            for i, e in enumerate(plain51_blobs):
                raw, _, _ = srv.wait(srv.submit(e), timeout_ms=30000)
                assert raw == ref_outs[i], (rnd, i)
        rng = np.random.default_rng(5)
        n = 1500
        pick = rng.integers(0, len(plain51_blobs), n).astype(np.int32)
        arrival = np.cumsum(rng.exponential(1e9 / 200000.0, n)).astype(np.int64)
        rec, summ = srv.replay(plain51_blobs, pick, arrival, submit_threads=2)
    ok = rec["status"] == 0
    assert summ["completed"] == int(ok.sum()) and summ["completed"] + summ["rejected"] == n
    assert (rec["status"][~ok] == 5).all()
    assert np.array_equal(rec["digest"][ok], want[pick[ok]])


@ENGINES
def test_serve_admission_errors(gpu_ctx, single_blob, engine):
    with Server(gpu_ctx, engine=engine, slots=1, max_batch_events=4, max_batch_hits=1000,
                max_event_hits=1000, max_requests=2) as srv:
        with pytest.raises(ClfError) as e:
            srv.submit(single_blob[:100])
        assert e.value.status == 1
        with pytest.raises(ClfError) as e:  # 2,500 hits > the per-event limit
            srv.submit(single_blob)
        assert e.value.status == 5
        small = line_event([(1.0, 2.0, 0.01, -0.02), (-3.0, 1.0, -0.03, 0.01)],
                           [-300 + 22 * k for k in range(20)]).pack()
        t1 = srv.submit(small)
        t2 = srv.submit(small)
        with pytest.raises(ClfError) as e:  # request table full until released
            srv.submit(small)
        assert e.value.status == 5
        a = srv.wait(t1, timeout_ms=20000)[0]
        b = srv.wait(t2, timeout_ms=20000)[0]
        assert a == b
        srv.wait(srv.submit(small), timeout_ms=20000)
        with pytest.raises(ClfError):
            srv.release(t1)  # already released


def test_resident_idle_exit_and_relaunch(gpu_ctx, plain51_blobs, plain_ref):
    """The resident kernel leaves after idle_us without work; a later submit
    relaunches it and is served exactly."""
    import time

    ref_outs, _ = plain_ref
    with Server(gpu_ctx, engine=SERVE_RESIDENT, idle_us=2000) as srv:
        for rnd in range(3):
            for i in (0, 7, 20):
                raw, _, _ = srv.wait(srv.submit(plain51_blobs[i]), timeout_ms=30000)
                assert raw == ref_outs[i], (rnd, i)
            time.sleep(0.05)  # 25 x idle_us: the kernel has left


def test_resident_feeder_stall_loses_no_request(gpu_ctx, plain51_blobs, plain_ref):
    """The first launch's feeder stalls 10 x idle_us with no heartbeat
    before its first hand-out (feeder_pause_us, a testing control; the
    stall is bounded, so every wait still ends).  Each waiting worker
    withdraws its idle entry after 4 x idle_us and leaves; the feeder passes
    the withdrawn entries over and, with no worker left, leaves with the
    requests still waiting; the host launches the kernel again, and every
    request is served exactly.  A hand-out to a departed worker would lose
    its request (its wait times out), and a feeder that stayed on alone
    would hold every request until the server closed."""
    ref_outs, _ = plain_ref
    idle_us, pause_us = 2000, 20000
    picks = [i % len(plain51_blobs) for i in range(64)]
    with Server(gpu_ctx, engine=SERVE_RESIDENT, idle_us=idle_us, feeder_pause_us=pause_us) as srv:
        tickets = srv.submit_many([plain51_blobs[i] for i in picks])
        assert len(tickets) == len(picks)
        lat_ms = []
        for t, i in zip(tickets, picks):
            raw, _, tm = srv.wait(t, timeout_ms=20000)
            assert raw == ref_outs[i], i
            lat_ms.append((tm["done_ns"] - tm["submit_ns"]) / 1e6)
        # the stall happened: nothing came back before it ended
        assert min(lat_ms) > 0.75 * pause_us / 1000.0, lat_ms[:4]
        # later requests on the same server see no stall
        for i in (3, 11):
            raw, _, tm = srv.wait(srv.submit(plain51_blobs[i]), timeout_ms=20000)
            assert raw == ref_outs[i]


def test_resident_output_ring_wraps_and_keeps_unreleased(gpu_ctx, plain51_blobs, plain_ref):
    """A 4 MiB device output ring, where each admitted event reserves its
    capacity bound (about 0.6 MB) and uses about 28 kB: after one round of
    the 51 events, three results are held unreleased; two more rounds run
    the ring past its end (allocations straddling it are copied past the
    pinned mirror's end as well) while the held results keep their bytes;
    once they are released, a replay wraps the ring many times, refusing
    what does not fit, every accepted result exact."""
    ref_outs, _ = plain_ref
    want = np.array([tracks_digest(o) for o in ref_outs], np.uint64)
    with Server(gpu_ctx, engine=SERVE_RESIDENT, out_arena_bytes=4 << 20, arena_bytes=4 << 20) as srv:
        for i, e in enumerate(plain51_blobs):
            assert srv.wait(srv.submit(e), timeout_ms=30000)[0] == ref_outs[i], i
        held = []
        for i in (0, 9, 33):
            t = srv.submit(plain51_blobs[i])
            assert srv.wait(t, timeout_ms=30000, release=False)[0] == ref_outs[i]
            held.append((t, i))
        for rnd in range(2):
            for i, e in enumerate(plain51_blobs):
                raw, _, _ = srv.wait(srv.submit(e), timeout_ms=30000)
                assert raw == ref_outs[i], (rnd, i)
        for t, i in held:
            raw, _, _ = srv.wait(t, timeout_ms=1000)  # still intact; now released
            assert raw == ref_outs[i], i
        rng = np.random.default_rng(8)
        n = 2000
        pick = rng.integers(0, len(plain51_blobs), n).astype(np.int32)
        # about 3 in flight at a time, which the ring admits; its 56 MB of
        # tracks wrap it about 14 times
        arrival = np.cumsum(rng.exponential(1e9 / 5000.0, n)).astype(np.int64)
        rec, summ = srv.replay(plain51_blobs, pick, arrival, submit_threads=2)
    ok = rec["status"] == 0
    assert summ["completed"] == int(ok.sum()) > n // 2
    assert (rec["status"][~ok] == 5).all()
    assert np.array_equal(rec["digest"][ok], want[pick[ok]])


def test_resident_serves_stress_events(gpu_ctx, oracle_mod):
    """Events of the stress workload (up to ~25,000 hits, sensors above an
    LDS slot) through the resident engine: identical to the oracle."""
    from cl_forward_amd import generator

    blobs = [generator.stress_event(i).pack() for i in (0, 1, 2, 7)]
    with Server(gpu_ctx, engine=SERVE_RESIDENT) as srv:
        tickets = [srv.submit(b) for b in blobs]
        got = [srv.wait(t, timeout_ms=60000)[0] for t in tickets]
    for b, g in zip(blobs, got):
        assert g == oracle_mod.search_event(b)[0]
