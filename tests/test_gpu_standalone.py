"""The standalone program end to end on the GPU: framed .dat files of the
reference's mcdata events -> read, x-presort in place, search, result text
files.  Every results/<i>.out must equal the restated writer applied to the
oracle's tracks of the reference-presorted event, byte for byte."""
from __future__ import annotations

import os
import subprocess

import pytest

from cl_forward_amd import event as ev_mod
from cl_forward_amd.search import MODE_PLAIN, Server
from oracle import results_text

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STANDALONE = os.path.join(ROOT, "cl_forward_amd", "clf_standalone")


def test_standalone_mcdata_results_text(tmp_path, oracle_mod, mcdata_raw, mcdata_sorted):
    keys = list(mcdata_raw)[:8]
    data = tmp_path / "data"
    res = tmp_path / "results"
    data.mkdir()
    res.mkdir()
    for i, k in enumerate(keys):
        (data / f"{i:02d}.dat").write_bytes(ev_mod.frame(mcdata_raw[k]) + b"trailer")
    n = 11  # round robin: events 8, 9, 10 are files 0, 1, 2 again
    print(f"[standalone] {len(keys)} files, {n} events", flush=True)
    r = subprocess.run([STANDALONE, str(data), str(n), "--results", str(res), "--sorted"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert f"{n} files read" in out
    assert "Invoking gpuPixelSearchByTriplet with %d events" % n in out
    assert "Time averages:" in out
    lines = [l for l in out.splitlines() if l.startswith(" nthreads (64, ")]
    assert len(lines) == 4, out
    assert "Size of output: %d entries" % n in out
    for i in range(n):
        k = keys[i % len(keys)]
        blob = mcdata_sorted[k]
        want = results_text.format_tracks(blob, oracle_mod.search_event(blob)[0])
        got = (res / f"{i}.out").read_text()
        assert got == want, f"event {i} (mcdata {k})"


def test_server_submit_framed(gpu_ctx, plain51_blobs):
    ref, _ = gpu_ctx.search(plain51_blobs[:5], MODE_PLAIN)
    with Server(gpu_ctx) as srv:
        for i in range(5):
            raw, _, _ = srv.wait(srv.submit_framed(ev_mod.frame(plain51_blobs[i]) + b"tail"),
                                 timeout_ms=30000)
            assert raw == ref[i]
