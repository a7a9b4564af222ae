"""Writes the committed fixtures under tests/golden/.

Run in a container that has the reference tree at /root/reference and the
oracle built (``make -C oracle all ref``):

    python tests/golden/make_golden.py [--digests-only]

``--digests-only`` rewrites oracle_digests.json alone, from the committed
mcdata fixtures (no reference tree needed).

Fixtures (data only):

* ``mcdata_events.npz`` — the 52 event payloads of the reference's
  ``mcdata/*.dat`` files (the bytes after the ``.dat`` framing,
  IndependentEntrypoint.cpp:89-123), unsorted, keyed by file stem.
* ``mcdata_presort_ref.npz`` — the same events after the REFERENCE's own
  ``preorder_by_x`` (Tools.cpp:46-66, compiled from its source into
  oracle/_ref/), the golden vectors of the presort.
* ``mcdata_layout_ref.json`` — the nine array offsets the reference's
  ``setHPointersFromInput`` (Tools.cpp:68-80) computes for each event.
* ``oracle_digests.json`` — sha256 of the oracle's output bytes and its
  per-event counters, for the mcdata events (presorted by the reference)
  and for the synthetic workloads "single" and "plain51"; a regression
  pin of the oracle itself.
"""
from __future__ import annotations

import glob
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cl_forward_amd import event, generator  # noqa: E402
from oracle import oracle  # noqa: E402

MCDATA = "/root/reference/mcdata"


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def write_digests(srt):
    digests = {"mcdata": {}, "single": {}, "plain51": {}}
    for stem, arr in srt.items():
        out, st = oracle.search_event(arr.tobytes())
        digests["mcdata"][stem] = {"sha256": sha(out), **st}
    out, st = oracle.search_event(generator.single_event().pack())
    digests["single"]["0"] = {"sha256": sha(out), **st}
    for i in range(51):
        out, st = oracle.search_event(generator.plain_event(i).pack())
        digests["plain51"][str(i)] = {"sha256": sha(out), **st}
    with open(os.path.join(HERE, "oracle_digests.json"), "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)


def main():
    oracle.build()
    if "--digests-only" in sys.argv[1:]:
        z = np.load(os.path.join(HERE, "mcdata_presort_ref.npz"))
        write_digests({k: z[k] for k in sorted(z.files, key=int)})
        print("wrote oracle_digests.json")
        return
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref_tools.so is missing (make -C oracle ref)")
    files = sorted(glob.glob(os.path.join(MCDATA, "*.dat")), key=lambda p: int(os.path.basename(p)[:-4]))
    raw, srt, layout = {}, {}, {}
    for p in files:
        stem = os.path.basename(p)[:-4]
        blob = event.read_dat(p)
        raw[stem] = np.frombuffer(blob, np.uint8)
        sorted_blob = oracle.ref_presort(blob)
        srt[stem] = np.frombuffer(sorted_blob, np.uint8)
        layout[stem] = oracle.ref_layout_offsets(blob)
    np.savez_compressed(os.path.join(HERE, "mcdata_events.npz"), **raw)
    np.savez_compressed(os.path.join(HERE, "mcdata_presort_ref.npz"), **srt)
    with open(os.path.join(HERE, "mcdata_layout_ref.json"), "w") as f:
        json.dump(layout, f, indent=0, sort_keys=True)

    write_digests(srt)
    print("wrote fixtures for", len(raw), "mcdata events")


if __name__ == "__main__":
    main()
