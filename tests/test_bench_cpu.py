"""The benchmark's traffic figures on the CPU: they come from the committed
rocprofv3 counters with the copy-kernel calibration."""
from __future__ import annotations

import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_traffic_figures_follow_from_the_counters():
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        data = json.load(f)
    assert {"batch1000", "stress1000"} <= set(data)
    for name, t in data.items():
        # the calibration: a copy of C bytes each way reads FETCH_SIZE f KiB
        # and WRITE_SIZE w KiB, so the true-bytes factors are C/(1024 f), C/(1024 w)
        m = re.search(r"over (\d+) bytes: FETCH_SIZE ([\d.]+) KiB, WRITE_SIZE ([\d.]+) KiB", t["calibration"])
        assert m, name
        copy, f_kib, w_kib = int(m.group(1)), float(m.group(2)), float(m.group(3))
        assert abs(t["fetch_scale"] - copy / (1024 * f_kib)) < 1e-3, name
        assert abs(t["write_scale"] - copy / (1024 * w_kib)) < 1e-3, name
        want = 1024 * (t["fetch_kib_raw"] * copy / (1024 * f_kib) + t["write_kib_raw"] * copy / (1024 * w_kib))
        assert abs(t["bytes_per_launch"] - want) <= 2, name
        # and the profile summary of the same workload quotes the same megabytes
        md = os.path.join(ROOT, "profiles", f"search_kernel_{name}.md")
        with open(md) as f:
            assert f"{t['bytes_per_launch'] / 1e6:.2f} MB" in f.read(), md
