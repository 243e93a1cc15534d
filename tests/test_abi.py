"""The C-ABI boundary on the CPU: the built libraries load and export every declared entry point."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

from fastwarp_amd import backend

ROOT = Path(__file__).resolve().parent.parent


def declared_functions() -> list[str]:
    names = []
    for header in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names += re.findall(r"\b(fw_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_stage_interface():
    names = declared_functions()
    for required in ["fw_cc_pair", "fw_warp", "fw_smooth_field", "fw_invert_pair", "fw_compose", "fw_resample_field",
                     "fw_mi_dense_pair", "fw_mi_level_value_and_derivative", "fw_register"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    handle = C.CDLL(str(backend.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(handle, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    assert sorted(backend.exported_symbols()) == declared_functions()


def test_a_context_on_a_machine_without_a_gpu_fails_cleanly():
    """Without a device the library reports an error instead of falling back to the CPU."""
    import os

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(backend.RegError):
        backend.HipBackend(0)


def test_a_server_on_a_machine_without_a_gpu_fails_cleanly():
    import os

    from fastwarp_amd.serve import Server

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(backend.RegError):
        Server(0, slots=2)


def test_the_header_is_plain_c(tmp_path):
    """The boundary is a C ABI: the header compiles as C99 with every warning on."""
    import subprocess
    from pathlib import Path

    src = tmp_path / "use.c"
    src.write_text('#include "fastwarp_hip.h"\nint main(void) { fw_served_timing t; fw_affine a; (void)t; (void)a; return 0; }\n')
    include = Path(__file__).resolve().parent.parent / "include"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", f"-I{include}", "-fsyntax-only", str(src)],
                   check=True)


def test_oracle_library_loads():
    import oracle

    assert abs(oracle.cubic_bspline(0.0) - 4.0 / 6.0) < 1e-15


def test_served_stream_mix_follows_its_recipe():
    """The served-stream driver's request mix: about 70 % full normalisations, spacings 2, 1.5 and
    1 mm, two channels only on full requests, the same mix for the same seed."""
    from fastwarp_amd.serve_stream import GRIDS, make_mix

    mix = make_mix(2000)
    full = [m for m in mix if m[2]]
    assert 0.66 < len(full) / len(mix) < 0.74
    assert {m[0] for m in mix} == set(GRIDS) and {m[1] for m in full} == {1, 2}
    assert all(m[1] == 1 for m in mix if not m[2])
    assert make_mix(50) == make_mix(50)
