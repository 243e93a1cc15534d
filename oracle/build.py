"""ORACLE — test infrastructure only. Builds the f64 CPU restatement `oracle/lib/libfastwarp_oracle.so`
with g++ (no GPU code). The reference itself is Rust and cannot be compiled in this image, so there
is no `oracle/_ref` build; the oracle is pinned by the reference's fixtures and closed forms instead.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "libfastwarp_oracle.so"


def build(verbose: bool = False) -> Path:
    LIB.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [
        os.environ.get("CXX", "g++"),
        "-std=c++17",
        "-O3",
        "-fPIC",
        "-shared",
        str(HERE / "capi.cpp"),
        "-o",
        str(tmp),
        "-lpthread",
    ]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
