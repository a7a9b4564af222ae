"""Compare two track result files in the reference's text format.

    python -m cl_forward_amd.comparetracks <tracks-A> <tracks-B> [--inverse] [--strict]
                                           [--prpixel-first]

A Python 3 restatement of the reference's ``tools/comparetracks.py``:

* the default mode keeps its semantics and messages: each track of the first
  file (keyed by its first hit ID; a repeated first ID is reported and
  skipped) is looked up in the second by first hit ID; a missing track or a
  different hit count is a difference, listed hit by hit; an equal hit
  count counts as identical (the legacy tool compares nothing else); the
  last line is ``All tracks are equal`` or the identical fraction (printed
  with the legacy tool's ``%`` label although it is a fraction);
* both files are read in the text format of the result writer
  (``Track #n, length h`` blocks), as the legacy main() reads them;
  ``--prpixel-first`` reads the first file as a PrPixelTracking debug log
  instead (the legacy tool's unused second reader);
* ``--strict`` compares the two files as multisets of tracks (ordered hit-ID
  lists) and reports what only one side has; exit status 0 only when they
  are equal.  This is the check that bit-exact parity needs.
"""
from __future__ import annotations

import re
import sys
from collections import Counter


def _flt(x) -> str:
    # the legacy tool ran under Python 2, whose print wrote floats as %.12g
    return "%.12g" % x


def read_gpupixel_text(text: str) -> list[dict]:
    tracks = []
    for i in re.finditer(r"Track #(?P<ntrack>\d+), length (?P<nhits>\d+)\n(?P<hits>.*?)\n\n", text,
                         re.DOTALL):
        hits = []
        for j in re.finditer(r" (?P<hitid>\d+) (\(\d+\) )?module *(?P<module>\d+), x *(?P<x>[\d\.\-]+),"
                             r" y *(?P<y>[\d\.\-]+), z *(?P<z>[\d\.\-]+)", i.group(0), re.DOTALL):
            hits.append({k: j.group(k) for k in ("hitid", "module", "x", "y", "z")})
        tracks.append({"ntrack": i.group("ntrack"), "nhits": i.group("nhits"), "hits": hits})
    return tracks


def read_prpixel_text(text: str) -> list[dict]:
    tracks = []
    for i in re.finditer(r"Store track Nb (?P<ntrack>\d+)[^\d]nhits (?P<nhits>\d+).*?PrPixelTracking"
                         r"[ \t]*?(INFO ===|DEBUG)", text, re.DOTALL):
        hits = []
        for j in re.finditer(r"PrPixelTracking.*?(?P<hitid>\d+) *module *(?P<module>\d+) x *(?P<x>[\d\.\-]+)"
                             r" y *(?P<y>[\d\.\-]+) z *(?P<z>[\d\-\.]+) used \d", i.group(0), re.DOTALL):
            hits.append({k: j.group(k) for k in ("hitid", "module", "x", "y", "z")})
        tracks.append({"ntrack": i.group("ntrack"), "nhits": i.group("nhits"), "hits": hits})
    return tracks


def _read(path: str, prpixel: bool = False) -> list[dict]:
    with open(path) as f:
        text = f.read()
    return read_prpixel_text(text) if prpixel else read_gpupixel_text(text)


def _hit_line(prefix: str, h: dict) -> str:
    return f"{prefix} {h['hitid']} module {h['module']} x {h['x']} y {h['y']} z {h['z']}"


def find_track(track: dict, tracks: list[dict]):
    hitid = track["hits"][0]["hitid"]
    for t in tracks:
        if hitid == t["hits"][0]["hitid"]:
            return t
    return None


def compare_track(a: dict, tracks: list[dict], out) -> bool:
    b = find_track(a, tracks)
    first = a["hits"][0]["hitid"]
    if b is None:
        print("Track ID", first, "has no corresponding track", file=out)
        for h in a["hits"]:
            print(_hit_line(" -", h), file=out)
        return False
    if a["nhits"] != b["nhits"]:
        print("Tracks with ID", first, "differ:", file=out)
        print(" nohits:", a["nhits"], "vs", b["nhits"], file=out)
        print(" hit differences:", file=out)
        ida = [h["hitid"] for h in a["hits"]]
        idb = [h["hitid"] for h in b["hits"]]
        for hid in [x for x in ida if x in idb]:
            print(_hit_line("  ", [h for h in a["hits"] if h["hitid"] == hid][0]), file=out)
        for hid in [x for x in ida if x not in idb]:
            print(_hit_line(" -", [h for h in a["hits"] if h["hitid"] == hid][0]), file=out)
        for hid in [x for x in idb if x not in ida]:
            print(_hit_line(" +", [h for h in b["hits"] if h["hitid"] == hid][0]), file=out)
        return False
    print("Track ID", first, "are equal! :)", file=out)
    return True


def compare_tracks(list1: list[dict], list2: list[dict], out=sys.stdout) -> tuple[int, int]:
    """The legacy comparison; returns (identical, different)."""
    seen = []
    identical = different = 0
    for t in list1:
        first = t["hits"][0]["hitid"]
        if first not in seen:
            seen.append(first)
            if compare_track(t, list2, out):
                identical += 1
            else:
                different += 1
            print(file=out)
        else:
            print("Track ", first, "appears repeated in the first list of tracks", file=out)
    if different == 0:
        print("All tracks are equal", file=out)
    else:
        print(_flt(float(identical) / (identical + different)), "% of tracks are identical (",
              identical, "identical,", different, "different)", file=out)
    return identical, different


def strict_compare(list1: list[dict], list2: list[dict], out=sys.stdout) -> bool:
    """Multiset comparison of the tracks as ordered hit-ID tuples."""
    ca = Counter(tuple(h["hitid"] for h in t["hits"]) for t in list1)
    cb = Counter(tuple(h["hitid"] for h in t["hits"]) for t in list2)
    only_a = ca - cb
    only_b = cb - ca
    common = sum((ca & cb).values())
    for t, k in sorted(only_a.items()):
        print(" - only in first (x%d):" % k, " ".join(t), file=out)
    for t, k in sorted(only_b.items()):
        print(" + only in second (x%d):" % k, " ".join(t), file=out)
    ok = not only_a and not only_b
    print("strict: %d tracks in common, %d only in the first, %d only in the second: %s" % (
        common, sum(only_a.values()), sum(only_b.values()), "identical" if ok else "DIFFERENT"),
        file=out)
    return ok


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    flags = {a for a in argv if a.startswith("--")}
    files = [a for a in argv if not a.startswith("--")]
    if len(files) < 2:
        print("Usage: comparetracks.py <prpixel-outputlog> <gpupixel-outputlog> [--inverse] [--strict]"
              " [--prpixel-first]")
        return 2
    first = _read(files[0], prpixel="--prpixel-first" in flags)
    second = _read(files[1])
    if "--inverse" in flags:
        first, second = second, first
    if "--strict" in flags:
        return 0 if strict_compare(first, second) else 1
    compare_tracks(first, second)
    return 0


if __name__ == "__main__":
    sys.exit(main())
