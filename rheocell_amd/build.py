"""Build recipe of the HIP library (gfx950) -- plain hipcc, in-tree output."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", n) for n in ("rc_kernels.hip", "rc_api.hip", "rc_conv.hip", "rc_net.hip")]
OUT = os.path.join(HERE, "librheocell.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # IEEE evaluation of every expression (no FMA contraction): parity with the reference source
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # leading scalar and pointer kernel arguments arrive preloaded in SGPRs
    # (the site-form iteration puts the pointers of its first loads there)
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
    "-I" + os.path.join(ROOT, "include"),
]


HOST_SOURCES = [os.path.join(HERE, "host", "rc_host.cpp"), os.path.join(HERE, "host", "rheoCell_hip.cpp")]
HOST_EXE = os.path.join(HERE, "rheoCell_hip")


def build_host(verbose: bool = False) -> str:
    """The C++ drop-in (FD_HIPBackend + SimManager + main), linked to librheocell.so."""
    tmp = HOST_EXE + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wno-unused-result", "-pthread", "-I" + os.path.join(ROOT, "include")] + HOST_SOURCES + [
        "-L" + HERE, "-lrheocell", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib",
        "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building rheoCell_hip")
    os.replace(tmp, HOST_EXE)
    if verbose:
        print("built", HOST_EXE)
    return HOST_EXE


def hipcc() -> str:
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.isabs(p) and os.path.exists(p) or not os.path.isabs(p)):
            return p
    return "hipcc"


def build_library(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + SOURCES + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building librheocell.so")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build_library(verbose=True)
    build_host(verbose=True)
