"""Build the gfx950 HIP library ``libsimuse.so`` (and the CPU oracle).

Run ``python dioma-core_amd/build_lib.py`` from anywhere.  The library is
written next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsimuse.so")

SOURCES = ["net_kernels.hip", "net_api.hip", "snapshot.hip", "kuramoto.hip", "batch.hip", "server.hip", "pool.hip",
           "model.hip"]

# -ffp-contract=off: every float multiply and add rounds on its own, which is
# what makes the LIF/Izhikevich/ring arithmetic bit-identical to the
# reference's non-FMA CPU build.  Packed-FP32 instructions are disabled.
HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fPIC",
    "-shared",
    "-Wl,--no-undefined",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> str:
    """Compile every translation unit in parallel (one hipcc each), then link."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    odir = os.path.join(HERE, "_build")
    os.makedirs(odir, exist_ok=True)
    cflags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-Wl,--no-undefined")]
    jobs = []
    for src in srcs:
        obj = os.path.join(odir, os.path.basename(src) + ".o")
        cmd = [hipcc(), *cflags, "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for obj, p in jobs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("hipcc failed building libsimuse.so")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-fPIC", "-shared", "-Wl,--no-undefined", "-o", tmp,
           *[o for o, _ in jobs]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libsimuse.so")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    odir = os.path.join(REPO, "oracle")
    cmd = ["make", "-s", "-C", odir, "-B", "liboracle.so"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    return os.path.join(odir, "liboracle.so")


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_oracle(verbose=True))
