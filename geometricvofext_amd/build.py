"""Builds libsimplic_amd.so in-tree with hipcc for gfx950.

Flags that matter for parity: -ffp-contract=off (no FMA contraction, so every
a*b+c is rounded twice as in an x86-64 build of the reference) and no fast-math
(correctly rounded FP64 division and square root).  Packed FP32 instructions
are disabled for the whole library.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsimplic_amd.so")
SOURCES = ["simplic_kernels.hip", "harness_kernels.hip", "learned_kernels.hip", "simplic_api.cpp", "learned_api.cpp"]
HEADERS = ["simplic_device.hpp", "simplic_types.hpp", "learned_device.hpp", "learned_kernels.hpp",
           "learned_model.hpp", "api_common.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _objects_and_commands(build_dir):
    cmds = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") or src.endswith(".cpp") else []
        cmds.append([HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    return objs, cmds


def build(verbose: bool = False, jobs: int = 5) -> str:
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs, cmds = _objects_and_commands(build_dir)
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(link), flush=True)
    _wait(subprocess.Popen(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    os.replace(tmp, LIB)
    return LIB


def _wait(p):
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
    lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(lines))
    if lines:
        print("\n".join(lines), file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
