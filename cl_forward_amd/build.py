"""Builds cl_forward_amd/libclf.so for gfx950 with hipcc (in-tree).

    python -m cl_forward_amd.build [--arch gfx950] [--jobs N]

Every translation unit is compiled with -ffp-contract=off (no FMA
contraction: the tracker's cuts are discrete decisions on FP32 values
computed in the reference's operation order) and with correctly rounded
FP32 division.  Packed FP32 instructions are disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libclf.so")
STANDALONE = os.path.join(PKG, "clf_standalone")
DROPIN = os.path.join(PKG, "clf_dropin")
SOURCES = ["clf_kernels.hip", "clf_api.cpp", "clf_serve.cpp", "clf_results.cpp", "clf_presort.cpp",
           "clf_shims.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags(arch: str) -> list[str]:
    return [
        f"--offload-arch={arch}",
        "-O3",
        "-fPIC",
        "-std=c++17",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-command-line-argument",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
    ]


def build(arch: str = "gfx950", jobs: int = 4, verbose: bool = False) -> str:
    cc = hipcc()
    obj_dir = os.path.join(PKG, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    flags = common_flags(arch)

    def compile_one(src):
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={arch}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp,
           "-L/opt/rocm/lib", "-lhsa-runtime64", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the standalone program and the drop-in caller, host code over the library
    for exe, src in ((STANDALONE, "clf_standalone.cpp"), (DROPIN, "clf_dropin.cpp")):
        exe_tmp = exe + ".tmp"
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", os.path.join(CSRC, src),
               f"-L{PKG}", "-lclf", "-Wl,-rpath,$ORIGIN", "-o", exe_tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{src} link failed\n{r.stdout}\n{r.stderr}")
        os.replace(exe_tmp, exe)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default=os.environ.get("PYTORCH_ROCM_ARCH", "gfx950"))
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.arch, a.jobs, a.verbose))


if __name__ == "__main__":
    sys.exit(main())
