"""Builds the gfx950 shared library `fastwarp_amd/lib/libfastwarp_hip.so` in-tree with hipcc.

Every translation unit is compiled in parallel into `fastwarp_amd/_build/`, then linked. HIP code
is compiled with the packed-FP32 instructions disabled (`-packed-fp32-ops`).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "lib" / "libfastwarp_hip.so"
BUILD = HERE / "_build"

SOURCES = [
    CSRC / "hip" / "kernels.hip",
    CSRC / "hip" / "mi_kernels.hip",
    CSRC / "hip" / "backend.hip",
    CSRC / "hip" / "mi_backend.hip",
    CSRC / "hip" / "noalloc.hip",
    CSRC / "hip" / "batch.hip",
    CSRC / "capi.cpp",
    CSRC / "serve.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FW_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-std=c++17",
    "-O3",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{HERE.parent / 'include'}",
]


def _run(cmds: list[list[str]]) -> None:
    procs = [subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for c in cmds]
    failed = []
    for c, p in zip(cmds, procs):
        out, _ = p.communicate()
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and "unknown target feature" not in l]
        if p.returncode != 0:
            failed.append((c, "\n".join(lines)))
    if failed:
        for c, text in failed:
            sys.stderr.write(" ".join(c) + "\n" + text + "\n")
        raise RuntimeError(f"{len(failed)} compile step(s) failed")


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        if not src.exists():
            continue
        obj = BUILD / (src.name + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)])
    _run(cmds)
    tmp = LIB.with_suffix(".so.tmp")
    _run([[HIPCC, *FLAGS, "-shared", *map(str, objs), "-o", str(tmp), "-lpthread"]])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
