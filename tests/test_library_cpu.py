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
