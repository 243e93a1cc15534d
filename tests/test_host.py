"""The product's host code (drivers' building blocks) pinned on the CPU: builds tests/cpp/host_checks.cpp
with g++ and runs it (MT19937 published outputs, the ITK shrink rule, Deriche gain and variance, the
convergence-monitor closed form, CG on a quadratic, Jacobians, f32 bin thresholds, taps, windowing)."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent


def test_host_checks(tmp_path):
    exe = tmp_path / "host_checks"
    subprocess.run(["g++", "-std=c++17", "-O2", str(ROOT / "cpp" / "host_checks.cpp"), "-o", str(exe), "-lpthread"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all host checks passed" in out.stdout
