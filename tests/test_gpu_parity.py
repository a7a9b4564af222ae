"""HIP path against the oracle on the same inputs (bit-exact).

Every comparison goes through the C ABI of libclf.so; the oracle is the
checker only."""
from __future__ import annotations

import struct
import sys

import numpy as np
import pytest

from cl_forward_amd import generator
from cl_forward_amd.search import MODE_BATCHED, MODE_PLAIN
from tests.parity import assert_event_equal, canonical, describe_diff

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check_windows(ctx, oracle, blobs, tag):
    got = ctx.fill_candidates(blobs)
    for i, (b, (g1, g2, flags)) in enumerate(zip(blobs, got)):
        o1, o2 = oracle.fill_candidates(b)
        if not (np.array_equal(g1, o1) and np.array_equal(g2, o2)):
            bad1 = np.nonzero(g1 != o1)[0][:5]
            bad2 = np.nonzero(g2 != o2)[0][:5]
            raise AssertionError(f"{tag}[{i}]: windows differ at h1 {bad1.tolist()} "
                                 f"{g1[bad1].tolist()} vs {o1[bad1].tolist()}, h2 {bad2.tolist()} "
                                 f"{g2[bad2].tolist()} vs {o2[bad2].tolist()}")
        assert flags == oracle.input_flags(b), f"{tag}[{i}] input flags"


def _check_search(ctx, oracle, blobs, tag, mode=MODE_BATCHED):
    outs, stats = ctx.search(blobs, mode)
    for i, b in enumerate(blobs):
        o_out, o_st = oracle.search_event(b)
        assert_event_equal(outs[i], stats[i], o_out, o_st, f"{tag}[{i}]")
    return outs, stats


def test_fill_candidates_mcdata_sorted(gpu_ctx, oracle_mod, mcdata_sorted):
    _say("fill_candidates: 52 mcdata events, x-sorted")
    _check_windows(gpu_ctx, oracle_mod, list(mcdata_sorted.values()), "mcdata")


def test_fill_candidates_mcdata_unsorted(gpu_ctx, oracle_mod, mcdata_raw):
    _say("fill_candidates: 52 mcdata events as stored (unsorted: exact scans)")
    _check_windows(gpu_ctx, oracle_mod, list(mcdata_raw.values()), "mcdata-raw")


def test_fill_candidates_synthetic(gpu_ctx, oracle_mod, single_blob, plain51_blobs):
    _say("fill_candidates: single + 51 plain events")
    _check_windows(gpu_ctx, oracle_mod, [single_blob] + plain51_blobs, "synthetic")


def test_search_single(gpu_ctx, oracle_mod, single_blob):
    _say("search: single event (150 tracks, seed 1), batched and plain")
    for mode in (MODE_BATCHED, MODE_PLAIN):
        _check_search(gpu_ctx, oracle_mod, [single_blob], f"single/mode{mode}", mode)


def test_search_mcdata(gpu_ctx, oracle_mod, mcdata_sorted, oracle_digests):
    _say("search: 52 mcdata events (reference-presorted), batched")
    outs, stats = _check_search(gpu_ctx, oracle_mod, list(mcdata_sorted.values()), "mcdata")
    import hashlib

    for (stem, _), out in zip(mcdata_sorted.items(), outs):
        assert hashlib.sha256(out).hexdigest() == oracle_digests["mcdata"][stem]["sha256"], stem


def test_search_mcdata_unsorted(gpu_ctx, oracle_mod, mcdata_raw):
    _say("search: 52 mcdata events unsorted (Gaudi path without presort)")
    _check_search(gpu_ctx, oracle_mod, list(mcdata_raw.values()), "mcdata-raw")


def test_search_plain51_both_modes(gpu_ctx, oracle_mod, plain51_blobs):
    _say("search: 51 plain events, plain mode and batched mode")
    a, _ = _check_search(gpu_ctx, oracle_mod, plain51_blobs, "plain51/plain", MODE_PLAIN)
    b, _ = _check_search(gpu_ctx, oracle_mod, plain51_blobs, "plain51/batched", MODE_BATCHED)
    assert a == b


@pytest.fixture(scope="module")
def config_blobs(plain51_blobs, mcdata_sorted):
    # plain events, two mcdata events, two stress events whose largest
    # sensors (over 256 hits) exceed a 256-hit LDS slot and are read from HBM,
    # and a dense event (1,300 tracks from one vertex, ttf ring span 2,879)
    # whose ttf lists outgrow their LDS part in every configuration
    blobs = plain51_blobs[:6] + [mcdata_sorted["77"], mcdata_sorted["3"]]
    blobs += [generator.stress_event(i).pack() for i in (0, 7)]
    blobs.append(generator.generate_event(77, n_tracks=1300, noise_fraction=0.02).pack())
    return blobs


@pytest.mark.parametrize("block", [64, 128, 256, 512, 1024])
@pytest.mark.parametrize("lds_hits", [256, 512])
def test_launch_configs_identical(oracle_mod, config_blobs, block, lds_hits):
    from cl_forward_amd.search import Context

    _say(f"search: block size {block}, lds slot {lds_hits} hits")
    with Context(0) as ctx:
        ctx.set_config(block_size=block, lds_hits=lds_hits)
        _check_search(ctx, oracle_mod, config_blobs, f"block{block}/lds{lds_hits}")


@pytest.mark.parametrize("block,lds_hits", [(256, 256), (512, 256), (512, 512), (1024, 256), (1024, 512)])
def test_repeated_runs_identical(oracle_mod, config_blobs, block, lds_hits):
    """A resident batch run eight times gives the oracle's bytes every time:
    an unsynchronised read between the waves of a workgroup shows up as
    run-to-run differences before it shows up in one run."""
    from cl_forward_amd.search import Context

    _say(f"search: block size {block}, lds slot {lds_hits} hits, 8 runs")
    want = [oracle_mod.search_event(x)[0] for x in config_blobs]
    with Context(0) as ctx:
        ctx.set_config(block_size=block, lds_hits=lds_hits)
        with ctx.batch(config_blobs * 4) as b:
            for rep in range(8):
                b.run(MODE_BATCHED)
                outs, _ = b.results()
                for i, o in enumerate(outs):
                    assert o == want[i % len(want)], f"run {rep}, event {i}"


@pytest.mark.parametrize("heavy,graph", [(1, 1), (3, 1), (4, 0), (64, 1)])
def test_heavy_lane_identical(oracle_mod, config_blobs, heavy, graph):
    """The costliest events in a second launch of 512-thread workgroups
    (more than the batch holds: all of them), graph-captured or not."""
    from cl_forward_amd.search import Context

    _say(f"search: heavy lane of {heavy} events, graph {graph}")
    with Context(0) as ctx:
        ctx.set_config(block_size=256, lds_hits=512, heavy_events=heavy, heavy_block=512,
                       use_graph=graph)
        _check_search(ctx, oracle_mod, config_blobs, f"heavy{heavy}")
        with ctx.batch(config_blobs) as b:
            for _ in range(3):
                b.run(MODE_BATCHED)
            again, _ = b.results()
        assert again == [oracle_mod.search_event(x)[0] for x in config_blobs]


@pytest.mark.parametrize("grid", [2, 8])
def test_work_slots_and_event_slices(oracle_mod, grid):
    """Work slices per workgroup slot (reused event after event) when the
    slots fit in the batch's per-event slices, else one slice per event:
    one stress event beside seven small ones gives both (2 slots of its
    slice fit, 8 do not), each run three times over the same buffers."""
    from cl_forward_amd.search import Context, track_capacity

    blobs = [generator.stress_event(3).pack()]
    blobs += [generator.generate_event(500 + i, n_tracks=4, noise_fraction=0.0).pack() for i in range(7)]
    slices = [(track_capacity(b) + 3) // 4 * 4 for b in blobs]
    fits = grid * max(slices) <= sum(slices)
    assert fits == (grid == 2), (grid, slices)
    _say(f"search: grid {grid}, work slots {'fit' if fits else 'do not fit'}")
    want = [oracle_mod.search_event(x)[0] for x in blobs]
    with Context(0) as ctx:
        ctx.set_config(block_size=512, lds_hits=256, grid_size=grid)
        _check_search(ctx, oracle_mod, blobs, f"grid{grid}")
        with ctx.batch(blobs) as b:
            for rep in range(3):
                b.run(MODE_BATCHED)
                outs, _ = b.results()
                assert outs == want, f"run {rep}"


def test_determinism_and_graph_replays(gpu_ctx, plain51_blobs):
    _say("search: repeated graph replays give identical bytes")
    with gpu_ctx.batch(plain51_blobs) as b:
        b.run(MODE_BATCHED)
        first, _ = b.results()
        for r in range(4):
            # host work between replays
            _ = sum(len(x) for x in first)
            b.run(MODE_BATCHED)
            again, _ = b.results()
            assert again == first, f"replay {r}"
        b.run(MODE_PLAIN)
        plain, _ = b.results()
        assert plain == first


def test_every_run_leaves_its_queue_words_clear(oracle_mod, config_blobs):
    """A graph holds the search launch alone: each run's last workgroup
    stores the packed total and clears the allocator, the finished-workgroup
    count and the queue counters for the next run.  Replays of one batch's
    graph between every other kind of run on the same buffers (plain
    launches, a phase profile, uncaptured runs, the timing runs, the heavy
    lane, other launch shapes, a coarser queue grain) give the oracle's
    bytes each time, through the packed total and the offsets."""
    from cl_forward_amd.search import Context

    want = [oracle_mod.search_event(x)[0] for x in config_blobs]
    with Context(0) as ctx:
        ctx.set_config(block_size=512, lds_hits=256, use_graph=1)
        with ctx.batch(config_blobs) as b:
            def check(tag, mode=MODE_BATCHED):
                b.run(mode)
                outs, _ = b.results()
                assert outs == want, tag

            check("first replay")
            check("second replay")
            check("plain launches", MODE_PLAIN)
            check("replay after plain launches")
            b.profile(MODE_BATCHED)
            check("replay after a batched profile")
            b.profile(MODE_PLAIN)
            check("replay after a plain profile")
            ctx.set_config(use_graph=0)  # same launch shape: the graph stays
            check("uncaptured run")
            ctx.set_config(use_graph=1)
            check("replay after an uncaptured run")
            b.time(MODE_BATCHED, warmup=1, reps=2)
            check("replay after the timing runs")
            for kw in (dict(heavy_events=3, heavy_block=512), dict(heavy_events=0),
                       dict(block_size=1024, lds_hits=512), dict(grain=2),
                       dict(block_size=512, lds_hits=256, grain=1)):
                ctx.set_config(**kw)
                _say(f"search: queue words after a change to {kw}")
                check(f"new graph {kw}")
                check(f"its replay {kw}")


def test_stress_events(gpu_ctx, oracle_mod):
    _say("search: 6 high-occupancy events (sensors above 128 hits)")
    blobs = [generator.stress_event(i).pack() for i in range(6)]
    _check_search(gpu_ctx, oracle_mod, blobs, "stress")


def _blob(S, N, Z, starts, nums, x, y, z, ids=None):
    ids = np.arange(N, dtype=np.uint32) if ids is None else ids
    return (struct.pack("<ii", S, N) + np.asarray(Z, "<i4").tobytes() +
            np.asarray(starts, "<i4").tobytes() + np.asarray(nums, "<i4").tobytes() +
            np.asarray(ids, "<u4").tobytes() + np.asarray(x, "<f4").tobytes() +
            np.asarray(y, "<f4").tobytes() + np.asarray(z, "<f4").tobytes())


def test_edge_cases(gpu_ctx, oracle_mod, single_blob):
    _say("search: empty, tiny, malformed and sensor-sparse events")
    from cl_forward_amd.event import parse_event

    empty = _blob(0, 0, [], [], [], [], [], [])
    ev = parse_event(single_blob)
    # only the first 4 sensors: the sensor loop never seeds
    n4 = int(ev.hitNums[:4].sum())
    four = _blob(4, n4, ev.sensor_Zs[:4], ev.hitStarts[:4], ev.hitNums[:4], ev.X[:n4], ev.Y[:n4],
                 ev.Z[:n4], ev.hit_IDs[:n4])
    # every other sensor emptied
    nums = ev.hitNums.copy()
    nums[::2] = 0
    sparse = _blob(ev.n_sensors, ev.n_hits, ev.sensor_Zs, ev.hitStarts, nums, ev.X, ev.Y, ev.Z, ev.hit_IDs)
    truncated = single_blob[:-4]
    blobs = [empty, four, sparse, truncated, single_blob]
    outs, stats = gpu_ctx.search(blobs, MODE_BATCHED)
    for i in (0, 1, 2, 4):
        o_out, o_st = oracle_mod.search_event(blobs[i])
        assert_event_equal(outs[i], stats[i], o_out, o_st, f"edge[{i}]")
    assert outs[3] == b"" and stats[3]["flags"] & (1 << 7), "truncated blob must come back INVALID"


def _few_sensor_blobs(single_blob):
    """Events of 5 to 9 sensors (the first k of the synthetic event, and 7
    from its middle): around the ring of six or seven sensor slots, where
    the first steps stage nothing or the last sensor."""
    from cl_forward_amd.event import parse_event

    ev = parse_event(single_blob)
    out = []
    for k in range(5, 10):
        n = int(ev.hitNums[:k].sum())
        out.append(_blob(k, n, ev.sensor_Zs[:k], ev.hitStarts[:k], ev.hitNums[:k], ev.X[:n], ev.Y[:n],
                         ev.Z[:n], ev.hit_IDs[:n]))
    s0, k = 20, 7
    b0 = int(ev.hitStarts[s0])
    nums = ev.hitNums[s0:s0 + k]
    n = int(nums.sum())
    starts = np.asarray(ev.hitStarts[s0:s0 + k], np.int64) - b0
    out.append(_blob(k, n, ev.sensor_Zs[s0:s0 + k], starts, nums, ev.X[b0:b0 + n], ev.Y[b0:b0 + n],
                     ev.Z[b0:b0 + n], ev.hit_IDs[b0:b0 + n]))
    return out


@pytest.mark.parametrize("block,lds_hits", [(64, 256), (256, 256), (512, 256), (512, 512), (1024, 512)])
def test_few_sensor_events(oracle_mod, single_blob, block, lds_hits):
    """5 to 9 sensors, in launch configurations that stage at the top of a
    step into six slots and during the gather into seven: identical to the
    oracle."""
    from cl_forward_amd.search import Context

    _say(f"search: 5-9 sensor events, block size {block}, lds slot {lds_hits}")
    blobs = _few_sensor_blobs(single_blob)
    with Context(0) as ctx:
        ctx.set_config(block_size=block, lds_hits=lds_hits)
        _check_search(ctx, oracle_mod, blobs, f"few-sensors/block{block}")


def _degenerate_blobs(single_blob):
    """Events outside the validated domain, where the kernels take the
    reference's full formulas (no x-range cuts, every division done)."""
    from cl_forward_amd.event import parse_event

    ev = parse_event(single_blob)
    S, N = ev.n_sensors, ev.n_hits
    out = []

    def blob(x, y, z):
        return _blob(S, N, ev.sensor_Zs, ev.hitStarts, ev.hitNums, x, y, z, ev.hit_IDs)

    def sensor(s):
        b = int(ev.hitStarts[s])
        return slice(b, b + int(ev.hitNums[s]))

    # two sensors with the same hit z (z ranges overlap)
    z = ev.Z.copy()
    z[sensor(10)] = z[sensor(12)].mean()
    z[sensor(12)] = z[sensor(12)].mean()
    out.append(blob(ev.X, ev.Y, z))
    # NaN and infinite coordinates
    x, y = ev.X.copy(), ev.Y.copy()
    x[sensor(20)][:3] = np.nan
    y[sensor(30)][:2] = np.inf
    out.append(blob(x, y, ev.Z))
    # a sensor 1e-6 from its neighbour in z (ranges apart, tiny gaps)
    z = ev.Z.copy()
    z[sensor(15)] = np.float32(z[sensor(16)].max()) + np.float32(1e-6) * np.abs(z[sensor(16)].max())
    out.append(blob(ev.X, ev.Y, z))
    # coordinates at the edge of the domain (|x| near 1e5)
    x = ev.X.copy()
    x[sensor(25)] *= np.float32(9.0e4) / max(float(np.abs(x[sensor(25)]).max()), 1.0)
    out.append(blob(x, ev.Y, ev.Z))
    return out


@pytest.mark.parametrize("block,lds_hits", [(256, 256), (512, 256), (1024, 512)])
def test_degenerate_domain_events(oracle_mod, single_blob, block, lds_hits):
    """Overlapping or nearly touching sensor z ranges, NaN/inf and extreme
    coordinates: bytes, counts and flags equal to the oracle."""
    from cl_forward_amd.search import Context

    _say(f"search: degenerate-domain events, block size {block}, lds slot {lds_hits}")
    blobs = _degenerate_blobs(single_blob)
    with Context(0) as ctx:
        ctx.set_config(block_size=block, lds_hits=lds_hits)
        _check_search(ctx, oracle_mod, blobs, f"degenerate/block{block}")


@pytest.mark.parametrize("block,lds_hits", [(256, 256), (512, 256), (1024, 512)])
def test_mcdata_all_launch_configs(oracle_mod, mcdata_sorted, mcdata_raw, block, lds_hits):
    """All 52 real-geometry events (two hit z per sensor, sensors over 128
    hits), presorted and as stored, in the tuned and the extreme workgroup
    sizes: identical to the oracle."""
    from cl_forward_amd.search import Context

    _say(f"search: 104 mcdata events, block size {block}, lds slot {lds_hits}")
    blobs = list(mcdata_sorted.values()) + list(mcdata_raw.values())
    with Context(0) as ctx:
        ctx.set_config(block_size=block, lds_hits=lds_hits)
        _check_search(ctx, oracle_mod, blobs, f"mcdata/block{block}")
