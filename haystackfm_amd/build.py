"""Build libhsfm.so (the gfx950 query path) in-tree with hipcc.

    python -m haystackfm_amd.build          # or haystackfm_amd.build.build()

The library has no torch dependency; it links only the HIP runtime.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhsfm.so")
CALIB_LIB = os.path.join(HERE, "libhsfm_calib.so")  # measurement aid (random-load ceiling), not the query path
SOURCES = ["hsfm_runtime.hip", "hsfm_index.hip", "hsfm_locate.hip", "hsfm_mem.hip", "hsfm_lsi.hip",
           "hsfm_construct.hip"]
ARCH = os.environ.get("HSFM_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep the compiler from emitting packed-FP32 VALU forms
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 4) -> str:
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    tmp = CALIB_LIB + ".tmp"
    cmd = [cc, *HIPCC_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "hsfm_calib.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("calibration library build failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, CALIB_LIB)
    return LIB


def _drain(procs):
    while procs:
        src, p = procs.pop(0)
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; that warning is expected
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n" + "\n".join(lines))
        if lines and any(ln.strip() for ln in lines):
            sys.stderr.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
