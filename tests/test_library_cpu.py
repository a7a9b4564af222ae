"""Host-side checks of the product library and package (no GPU needed)."""
from __future__ import annotations

import itertools
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from cl_forward_amd import event as ev_mod
from cl_forward_amd import generator
from tests.helpers import raw_blob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "cl_forward_amd", "libclf.so")


@pytest.fixture(scope="module")
def lib():
    from cl_forward_amd import _lib

    return _lib.load()


def _declared_c_symbols():
    names = set()
    for h in ("clf.h",):
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(clf_[a-z_0-9]+)\s*\(", text))
    return names


def _dynamic_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_exports_every_declared_c_symbol(lib):
    from cl_forward_amd import _lib

    declared = _declared_c_symbols()
    assert declared, "no declarations parsed"
    syms = _dynamic_symbols()
    missing = sorted(declared - syms)
    assert not missing, missing
    assert declared == set(_lib.EXPORTED), sorted(declared ^ set(_lib.EXPORTED))
    for name in declared:
        getattr(lib, name)


def test_library_exports_reference_cpp_entry_points():
    demangled = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True,
                               check=True).stdout
    for sig in ("independent_execute(", "independent_post_execute(", "gpuPixelSearchByTriplet(",
                "gpuPixelSearchByTripletInvocation(", "invokeParallelSearch("):
        assert sig in demangled, sig


def test_reference_signature_headers_compile(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "GpuPixelSearchByTriplet.h"\n#include "KernelInvoker.h"\n#include "clf.h"\n'
        "int f(const std::vector<const std::vector<uint8_t>*>& in,\n"
        "      std::vector<std::vector<uint8_t> >& out) {\n"
        "  clf_track t; (void)t; static_assert(sizeof(clf_track) == 100, \"Track is 100 bytes\");\n"
        "  return gpuPixelSearchByTriplet(in, out) + invokeParallelSearch(0, 0, in, out);\n}\n")
    subprocess.run(["g++", "-std=c++11", "-fsyntax-only", f"-I{INCLUDE}", str(src)], check=True)
    csrc = tmp_path / "use.c"
    csrc.write_text('#include "clf.h"\nint g(void) { return (int)sizeof(clf_track) + CLF_OK; }\n')
    subprocess.run(["gcc", "-std=c99", "-fsyntax-only", f"-I{INCLUDE}", str(csrc)], check=True)


def test_reference_caller_links_against_the_library(tmp_path):
    """A caller of the reference's five C++ entry points links against
    libclf.so as the integration notes describe (not run: no GPU here)."""
    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include <cstddef>\n#include "GpuPixelSearchByTriplet.h"\n#include "KernelInvoker.h"\n'
        "int main() {\n"
        "  std::vector<std::vector<uint8_t> > events, out;\n"
        "  std::vector<const std::vector<uint8_t>*> ptrs;\n"
        "  for (size_t i = 0; i < events.size(); ++i) ptrs.push_back(&events[i]);\n"
        "  int r = gpuPixelSearchByTriplet(ptrs, out) + gpuPixelSearchByTripletInvocation(ptrs, out);\n"
        "  r += invokeParallelSearch(0, 0, ptrs, out) + independent_execute(events, out);\n"
        "  independent_post_execute(out);\n"
        "  return r;\n}\n")
    exe = tmp_path / "caller"
    lib_dir = os.path.dirname(LIB)
    subprocess.run(["g++", "-std=c++11", f"-I{INCLUDE}", str(src), "-o", str(exe), f"-L{lib_dir}", "-lclf",
                    f"-Wl,-rpath,{lib_dir}"], check=True)
    assert exe.exists()


def test_dropin_caller_checks_its_input(tmp_path):
    """clf_dropin reads the offload-frame stream before touching the device:
    a missing file or a truncated frame is refused with status 2."""
    exe = os.path.join(os.path.dirname(LIB), "clf_dropin")
    assert os.path.exists(exe), "build clf_dropin first"
    bad = tmp_path / "bad.frames"
    good = ev_mod.frame(generator.single_event().pack())
    bad.write_bytes(good + good[:-3])
    for path in (str(tmp_path / "missing"), str(bad)):
        r = subprocess.run([exe, path], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "clf_dropin" in r.stderr


def test_frame_streams_round_trip(tmp_path):
    blobs = [generator.single_event().pack(), b"", generator.plain_event(2).pack()]
    p = tmp_path / "s.frames"
    ev_mod.write_frame_stream(p, blobs)
    raw = p.read_bytes()
    got, pos = [], 0
    while pos < len(raw):
        name, data = ev_mod.unframe(raw[pos:])
        got.append(data)
        pos += 8 + len(name) + len(data)
    assert got == blobs
    q = tmp_path / "s.sized"
    q.write_bytes(b"".join(len(b).to_bytes(8, "little") + b for b in blobs))
    assert ev_mod.read_sized_stream(q) == blobs


def test_status_strings_and_no_device(lib):
    assert lib.clf_status_string(0) == b"ok"
    assert lib.clf_version().startswith(b"clf")
    import ctypes

    h = ctypes.c_void_p()
    st = lib.clf_context_create(0, ctypes.byref(h))
    if st == 0:
        lib.clf_context_destroy(h)
    else:
        assert st == 4 and b"device" in lib.clf_last_error()
    assert lib.clf_context_create(0, None) == 1


def test_track_capacity(lib):
    from cl_forward_amd import search

    blob = generator.single_event().pack()
    S, N = 48, ev_mod.parse_event(blob).n_hits
    assert search.track_capacity(blob) == N + 64 * S + 64
    assert search.track_capacity(blob[:-1]) == -1
    assert search.track_capacity(b"") == -1
    assert search.track_capacity(raw_blob(1, 1, [0], [0], [2], [0], [0], [0])) == -1


def test_product_presort_matches_reference_golden(mcdata_raw, mcdata_sorted):
    from cl_forward_amd import search

    for k, blob in mcdata_raw.items():
        b = bytearray(blob)
        search.presort_event(b)
        assert bytes(b) == mcdata_sorted[k], k
    with pytest.raises(TypeError):
        search.presort_event(b"immutable")


@pytest.mark.parametrize("values", [[1, 1], [3, 3, 1, 2, 2], [4, 1, 4, 1, 4, 1], [2, 2, 2, 1, 1, 1]])
def test_product_presort_permutations_vs_reference(oracle_mod, values):
    from cl_forward_amd import search

    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    for perm in set(itertools.permutations(values)):
        n = len(perm)
        blob = raw_blob(1, n, [0], [0], [n], np.array(perm, np.float32), np.arange(n) * 0.25,
                        np.zeros(n), np.arange(7, 7 + n))
        b = bytearray(blob)
        search.presort_event(b)
        assert bytes(b) == oracle_mod.ref_presort(blob), perm


def test_product_presort_rejects_nan():
    from cl_forward_amd import search

    b = bytearray(raw_blob(1, 2, [0], [0], [2], [np.nan, 1.0], [0, 0], [0, 0]))
    with pytest.raises(ValueError):
        search.presort_event(b)


# ---------------------------------------------------------------------------
# generator and event I/O
# ---------------------------------------------------------------------------
def test_generator_is_deterministic_and_well_formed():
    a = generator.single_event().pack()
    assert a == generator.single_event().pack()
    ev, truth = generator.single_event(return_truth=True)
    assert 2000 <= ev.n_hits <= 3200 and truth.n_particles == 150
    assert len(np.unique(ev.hit_IDs)) == ev.n_hits
    assert np.array_equal(ev.sensor_Zs, -300 + 22 * np.arange(48))
    sensor = ev.sensor_of_hit()
    assert np.all(ev.Z == ev.sensor_Zs[sensor].astype(np.float32))
    for s in range(48):
        b, n = int(ev.hitStarts[s]), int(ev.hitNums[s])
        x = ev.X[b:b + n]
        assert np.all(np.diff(x) > 0), "x strictly increasing (ties redrawn)"
        assert np.all(np.abs(x) <= 42.1) and np.all(np.abs(ev.Y[b:b + n]) <= 42.1)
    noise = int((truth.particle < 0).sum())
    assert abs(noise - round(0.01 * (ev.n_hits - noise))) <= 1


def test_workload_sizes():
    hits = [generator.plain_event(i).n_hits for i in range(0, 51, 10)]
    assert all(500 <= h <= 7000 for h in hits), hits
    hits = [generator.batch_event(i).n_hits for i in range(0, 1000, 97)]
    assert min(hits) >= 50 and max(hits) <= 14000, hits
    st = generator.stress_event(0)
    assert st.n_hits > 3000


def test_track_record_parsing():
    rec = np.zeros(2, ev_mod.TRACK_DTYPE)
    rec[0]["hitsNum"] = 3
    rec[0]["hits"][:3] = [5, 7, 9]
    rec[1]["hitsNum"] = 4
    rec[1]["hits"][:4] = [1, 2, 3, 4]
    assert ev_mod.track_lists(rec.tobytes()) == [(5, 7, 9), (1, 2, 3, 4)]
    with pytest.raises(ValueError):
        ev_mod.tracks_from_bytes(b"\0" * 99)


def test_bench_rejects_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_c_struct_layouts_match_ctypes(tmp_path):
    """Every struct of clf.h has the same size and field offsets in C as in
    the ctypes binding."""
    import ctypes

    from cl_forward_amd import _lib

    structs = {"clf_event_stats": _lib.EventStats, "clf_launch_config": _lib.LaunchConfig,
               "clf_result": _lib.Result, "clf_timing": _lib.Timing,
               "clf_server_config": _lib.ServerConfig, "clf_served": _lib.Served,
               "clf_replay_record": _lib.ReplayRecord, "clf_replay_summary": _lib.ReplaySummary}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "clf.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{INCLUDE}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_tracks_digest_properties(lib):
    """The digest is a function of the bytes and the count: equal inputs
    agree, a one-bit change or a reordering of records changes it."""
    from cl_forward_amd.search import tracks_digest

    rng = np.random.default_rng(3)
    a = rng.integers(0, 2**32, 25 * 37, dtype=np.uint64).astype(np.uint32).tobytes()
    assert tracks_digest(a) == tracks_digest(bytes(a))
    b = bytearray(a)
    b[1234] ^= 1
    assert tracks_digest(bytes(b)) != tracks_digest(a)
    swapped = a[100:200] + a[:100] + a[200:]
    assert tracks_digest(swapped) != tracks_digest(a)
    assert tracks_digest(a[:100]) != tracks_digest(a[:200])
    assert tracks_digest(b"") == tracks_digest(b"")


def test_serve_arrivals_reproducible():
    from cl_forward_amd import generator

    p1, a1 = generator.serve_arrivals(0.5, 1e5, 1000, 256)
    p2, a2 = generator.serve_arrivals(0.5, 1e5, 1000, 256)
    assert np.array_equal(p1, p2) and np.array_equal(a1, a2)
    assert a1[0] == 0 and (np.diff(a1) >= 0).all()
    assert 0 <= p1.min() and p1.max() < 256
    mean_gap = float(np.diff(a1).mean())
    assert 0.8e4 < mean_gap < 1.2e4  # 1e5 events/s -> 10 us


def _digest_py(data: bytes, n_tracks: int) -> int:
    """Independent restatement of clf_tracks_digest in Python integers."""
    M = (1 << 64) - 1
    P = 0x9E3779B97F4A7C15
    h = [(0x243F6A8885A308D3 * (2 * l + 1)) & M for l in range(8)]

    def step(x, w):
        x = ((x ^ w) * P) & M
        return x ^ (x >> 29)

    nbytes = 100 * n_tracks
    i = 0
    while i + 64 <= nbytes:
        for l in range(8):
            h[l] = step(h[l], int.from_bytes(data[i + 8 * l: i + 8 * l + 8], "little"))
        i += 64
    l = 0
    while i < nbytes:
        h[l] = step(h[l], int.from_bytes(data[i: i + 4], "little"))
        i += 4
        l = (l + 1) & 7
    r = (n_tracks * P) & M
    for l in range(8):
        r = ((r ^ h[l]) * P) & M
        r ^= r >> 32
    return r


def test_tracks_digest_matches_restatement(lib):
    """Every tail length (0..15 trailing words after the 64-byte blocks)."""
    from cl_forward_amd.search import tracks_digest

    rng = np.random.default_rng(11)
    for n in list(range(0, 18)) + [37, 100]:
        data = rng.integers(0, 256, 100 * n, dtype=np.uint8).tobytes()
        assert tracks_digest(data) == _digest_py(data, n), n


def test_autotune_keeps_single_launch_within_margin():
    from cl_forward_amd.autotune import HEAVY_MARGIN, pick_best

    def row(ms, heavy=0, block=256):
        return {"block_size": block, "grid_size": 0, "lds_hits": 256, "heavy_events": heavy,
                "heavy_block": 512, "ms_run": ms}

    single = row(1.00)
    near = row(1.00 * (1.0 - HEAVY_MARGIN / 2), heavy=8)
    far = row(1.00 * (1.0 - 2 * HEAVY_MARGIN), heavy=16)
    assert pick_best([single, near]) is single
    assert pick_best([single, near, far]) is far
    assert pick_best([row(2.0), single, row(1.5, block=512)]) is single


def test_autotune_prefers_resident_grid_within_margin():
    from cl_forward_amd.autotune import GRID_MARGIN, pick_best

    def row(ms, grid, block=1024):
        return {"block_size": block, "grid_size": grid, "lds_hits": 512, "heavy_events": 0,
                "heavy_block": 512, "ms_run": ms}

    big = row(1.000, 1024)
    resident = row(1.000 * (1.0 + GRID_MARGIN / 2), 0)
    slow_resident = row(1.000 * (1.0 + 2 * GRID_MARGIN), 0)
    other_shape = row(1.001, 0, block=512)
    assert pick_best([big, resident]) is resident
    assert pick_best([big, slow_resident]) is big
    # the tie-break stays within the fastest launch shape
    assert pick_best([big, other_shape]) is big


def test_autotune_keeps_grain_one_within_margin():
    from cl_forward_amd.autotune import GRAIN_MARGIN, pick_best

    def row(ms, grain):
        return {"block_size": 512, "grid_size": 0, "lds_hits": 256, "heavy_events": 0,
                "heavy_block": 512, "grain": grain, "ms_run": ms}

    one = row(1.0, 1)
    near = row(1.0 * (1.0 - GRAIN_MARGIN / 2), 4)
    far = row(1.0 * (1.0 - 2 * GRAIN_MARGIN), 8)
    assert pick_best([one, near]) is one
    assert pick_best([one, near, far]) is far
    assert pick_best([one, row(1.2, 2)]) is one


def test_event_arrays_point_at_the_events():
    """The binding hands `bytes` events to the library by their own buffer
    (no copy) and copies other buffers once; every pointer reads back the
    event's bytes, and empty or missing events pass a null pointer and
    size 0."""
    import ctypes

    from cl_forward_amd.search import _event_arrays

    blobs = [generator.single_event().pack(),
             raw_blob(2, 3, [0, 22], [0, 2], [2, 1], [1.0, 2.0, 3.0], [0.0] * 3, [0.0, 0.0, 22.0])]
    events = [blobs[0], bytearray(blobs[1]), memoryview(blobs[0]), b"", None]
    holders, ptrs, sizes = _event_arrays(events)
    want = [blobs[0], blobs[1], blobs[0], b"", b""]
    assert [sizes[i] for i in range(len(events))] == [len(w) for w in want]
    for i, w in enumerate(want):
        if w:
            assert ctypes.string_at(ptrs[i], sizes[i]) == w
        else:
            assert not ptrs[i]
    # the bytes event itself is what the pointer addresses
    assert ctypes.cast(ptrs[0], ctypes.c_void_p).value == \
        ctypes.cast(ctypes.c_char_p(blobs[0]), ctypes.c_void_p).value
    assert holders[0] is blobs[0]
