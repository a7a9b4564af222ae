"""The reference's result text (printTrack files), its framing and the
track comparator, on the CPU: the C-ABI writer against the restatement in
oracle/results_text.py, which is itself pinned to the C++ stream library and
to hand-derived cases."""
from __future__ import annotations

import io
import os
import struct
import subprocess

import numpy as np
import pytest

from cl_forward_amd import comparetracks
from cl_forward_amd import event as ev_mod
from oracle import results_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STANDALONE = os.path.join(ROOT, "cl_forward_amd", "clf_standalone")


def _blob(sensor_z, hits):
    """Event blob: sensor_z list, hits as (sensor, id, x, y, z) in order."""
    S = len(sensor_z)
    by = [[h for h in hits if h[0] == s] for s in range(S)]
    starts, nums, flat = [], [], []
    for s in range(S):
        starts.append(len(flat))
        nums.append(len(by[s]))
        flat += by[s]
    N = len(flat)
    b = struct.pack("<ii", S, N) + np.array(sensor_z, "<i4").tobytes()
    b += np.array(starts, "<i4").tobytes() + np.array(nums, "<i4").tobytes()
    b += np.array([h[1] for h in flat], "<u4").tobytes()
    for k in (2, 3, 4):
        b += np.array([h[k] for h in flat], "<f4").tobytes()
    return b


def _tracks(*hit_lists):
    out = b""
    for hl in hit_lists:
        r = np.zeros(25, "<u4")
        r[0] = len(hl)
        r[1:1 + len(hl)] = hl
        out += r.tobytes()
    return out


def test_cpp_stream_float_format_pinned(tmp_path):
    """cpp_float + width 6 equals what libstdc++ writes with setw(6) and the
    default float format, for values across the %g regimes."""
    vals = [0.0, -0.0, 1.5, -12.3456789, 0.0001234, 1e-05, 123456.7, 1234567.0, -300.0, 3.14159274,
            41.999, -0.5, 7.0e-7, 99999.95, 2.5e10, 100000.5, -0.25]
    src = tmp_path / "fmt.cpp"
    src.write_text('#include <iostream>\n#include <iomanip>\nint main(){float v[]={'
                   + ",".join(repr(float(np.float32(v))) + "f" for v in vals)
                   + '};for(float x: v) std::cout << "[" << std::setw(6) << x << "]\\n";}\n')
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-std=c++11", str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    want = ["[%6s]" % results_text.cpp_float(v) for v in vals]
    assert got[: len(vals)] == want


@pytest.mark.parametrize("order,modules", [((4.0, 6.0), (0, 0)), ((6.0, 4.0), (1, 1))])
def test_z_to_module_grows_in_hit_order(lib, order, modules):
    """Sensors at z 0 and 10.  A hit at z 4 maps to 0 (nearest), and then a
    hit at z 6 finds the key 4 added by that hit at distance 2 before the
    sensor at 10 (distance 4): module 0.  In the other hit order, 6 maps to
    10's module 1 and 4 then finds the key 6 first: module 1."""
    m = results_text.z_to_module([0, 10], list(order))
    assert (m[int(order[0])], m[int(order[1])]) == modules
    blob = _blob([0, 10], [(0, 11, 1.0, 2.0, order[0]), (1, 12, 3.0, 4.0, order[1])])
    text = results_text.format_tracks(blob, _tracks([0, 1]))
    assert [int(l.split("module")[1].split(",")[0]) for l in text.splitlines()[1:3]] == list(modules)
    from cl_forward_amd.search import format_tracks

    assert format_tracks(blob, _tracks([0, 1])) == text


def test_z_truncates_toward_zero():
    m = results_text.z_to_module([-4, 4], [-3.7, 3.7])
    assert m[-3] == 0 and m[3] == 1


def test_known_answer_text(lib):
    """One event, two tracks, typed out by hand from the C++ format rules.
    Hits are stored sensor by sensor: index 0 is ID 1234567 and index 1 is
    ID 7 (sensor 0), index 2 is ID 42 (sensor 1)."""
    blob = _blob([-300, -278], [(0, 1234567, 1.5, -0.25, -300.0), (1, 42, -12.3456789, 0.0001234, -278.0),
                                (0, 7, 100000.5, 2.0, -300.0)])
    tr = _tracks([0, 2], [1, 2])
    want = ("Track #0, length 2\n"
            "  1234567 (0) module  0, x    1.5, y  -0.25, z   -300\n"
            "       42 (2) module  1, x -12.3457, y 0.0001234, z   -278\n"
            "\n"
            "Track #1, length 2\n"
            "        7 (1) module  0, x 100000, y      2, z   -300\n"
            "       42 (2) module  1, x -12.3457, y 0.0001234, z   -278\n"
            "\n")
    assert results_text.format_tracks(blob, tr) == want
    from cl_forward_amd.search import format_tracks

    assert format_tracks(blob, tr) == want


def test_writer_matches_restatement_on_mcdata(lib, oracle_mod, mcdata_sorted):
    from cl_forward_amd.search import format_tracks

    for k in list(mcdata_sorted)[:10]:
        blob = mcdata_sorted[k]
        tracks = oracle_mod.search_event(blob)[0]
        assert format_tracks(blob, tracks) == results_text.format_tracks(blob, tracks), k


def test_writer_rejects_bad_tracks(lib, single_blob):
    from cl_forward_amd.search import ClfError, format_tracks

    N = struct.unpack_from("<i", single_blob, 4)[0]
    with pytest.raises(ClfError):
        format_tracks(single_blob, _tracks([0, N]))  # hit index out of range
    bad = bytearray(_tracks([0, 1]))
    bad[0] = 25
    with pytest.raises(ClfError):
        format_tracks(single_blob, bytes(bad))
    with pytest.raises(ClfError):
        format_tracks(single_blob[:50], b"")


def test_unframe_matches_python_reader(lib, mcdata_raw):
    from cl_forward_amd.search import ClfError, unframe

    for k in list(mcdata_raw)[:5]:
        fr = ev_mod.frame(mcdata_raw[k]) + b"trailing bytes are ignored"
        assert unframe(fr) == ev_mod.unframe(fr) == (ev_mod.HANDLER_NAME, mcdata_raw[k])
    for bad in (b"", b"\x05\x00\x00\x00ab", struct.pack("<i", -1) + b"x" * 8,
                struct.pack("<i", 2) + b"ab" + struct.pack("<i", 100) + b"short"):
        with pytest.raises(ClfError):
            unframe(bad)


def _text(*tracks):
    out = []
    for n, hits in enumerate(tracks):
        out.append(f"Track #{n}, length {len(hits)}\n")
        for hid in hits:
            out.append(f" {hid:8d} (0) module  1, x      1, y      2, z      3\n")
        out.append("\n")
    return "".join(out)


def test_comparator_legacy_and_strict():
    a = comparetracks.read_gpupixel_text(_text([1, 2, 3], [4, 5, 6], [7, 8, 9, 10]))
    same = comparetracks.read_gpupixel_text(_text([7, 8, 9, 10], [1, 2, 3], [4, 5, 6]))
    buf = io.StringIO()
    assert comparetracks.compare_tracks(a, same, out=buf) == (3, 0)
    assert buf.getvalue().rstrip().endswith("All tracks are equal")
    assert comparetracks.strict_compare(a, same, out=io.StringIO())
    # same first hit and length, one hit different: legacy calls it equal,
    # strict does not
    b = comparetracks.read_gpupixel_text(_text([1, 2, 99], [4, 5, 6], [7, 8, 9, 10]))
    assert comparetracks.compare_tracks(a, b, out=io.StringIO()) == (3, 0)
    assert not comparetracks.strict_compare(a, b, out=io.StringIO())
    # a shorter track and a missing one
    c = comparetracks.read_gpupixel_text(_text([1, 2, 3], [7, 8, 9]))
    buf = io.StringIO()
    assert comparetracks.compare_tracks(a, c, out=buf) == (1, 2)
    text = buf.getvalue()
    assert "Track ID 4 has no corresponding track" in text
    assert "Tracks with ID 7 differ:" in text and " nohits: 4 vs 3" in text
    assert " - 10 module 1 x 1 y 2 z 3" in text
    assert text.rstrip().endswith("0.333333333333 % of tracks are identical ( 1 identical, 2 different)")


def test_comparator_cli(tmp_path):
    pa, pb = tmp_path / "a.out", tmp_path / "b.out"
    pa.write_text(_text([1, 2, 3]))
    pb.write_text(_text([1, 2, 3]))
    assert comparetracks.main([str(pa), str(pb), "--strict"]) == 0
    pb.write_text(_text([1, 2, 4]))
    assert comparetracks.main([str(pa), str(pb), "--strict"]) == 1
    assert comparetracks.main([str(pa)]) == 2


def test_standalone_usage_and_errors(tmp_path, single_blob):
    r = subprocess.run([STANDALONE], capture_output=True, text=True)
    assert r.returncode == 255 and "Usage" in r.stderr
    r = subprocess.run([STANDALONE, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 255 and "either could not be opened, or is empty" in r.stderr
    (tmp_path / "bad.dat").write_bytes(b"\x03\x00")
    r = subprocess.run([STANDALONE, str(tmp_path), "1"], capture_output=True, text=True)
    assert r.returncode == 255 and "bad.dat" in r.stderr
    if os.path.exists("/dev/kfd"):
        return  # the device error below needs a machine without a GPU
    (tmp_path / "bad.dat").write_bytes(ev_mod.frame(single_blob))
    r = subprocess.run([STANDALONE, str(tmp_path), "2"], capture_output=True, text=True)
    assert r.returncode == 255 and "Error in function clf_context_create" in r.stderr
    assert "2 files read" in r.stdout


@pytest.fixture(scope="module")
def lib():
    from cl_forward_amd import _lib

    return _lib.load()
