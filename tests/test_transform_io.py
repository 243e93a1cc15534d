"""ANTs affine files (0GenericAffine.mat): the ITK MATLAB Level-4 layout of
crates/fastwarp-reg/src/transform/itk_mat.rs, checked against bytes built here from the format's
definition (five little-endian u32, the NUL-terminated name, f64 or f32 values), and the
reference's own cases (itk_mat.rs:196-246: a 193-byte bit-exact round trip, f32 variables,
bad type codes and truncation refused). Host only."""
from __future__ import annotations

import struct

import numpy as np
import pytest

from fastwarp_amd._abi import Affine
from fastwarp_amd.backend import RegError, read_affine_mat, write_affine_mat


def sample() -> Affine:
    # itk_mat.rs:191-197
    return Affine.make([[0.9, 0.1, -0.05], [0.0, 1.1, 0.2], [0.03, -0.1, 1.0]], (1.0, -2.0, 3.0), (-5.9, 3.1, -46.7))


def variable(name: str, values, code: int = 0) -> bytes:
    fmt = "<d" if code == 0 else "<f"
    head = struct.pack("<5I", code, len(values), 1, 0, len(name) + 1) + name.encode() + b"\0"
    return head + b"".join(struct.pack(fmt, v) for v in values)


def params(a: Affine):
    return list(a.matrix) + list(a.translation), list(a.center)


def test_written_file_is_the_itk_layout_and_193_bytes(tmp_path):
    a = sample()
    path = tmp_path / "0GenericAffine.mat"
    write_affine_mat(path, a)
    got = path.read_bytes()
    p, fixed = params(a)
    assert len(got) == 193
    assert got == variable("AffineTransform_double_3_3", p) + variable("fixed", fixed)


def test_round_trip_is_bit_exact(tmp_path):
    a = sample()
    path = tmp_path / "a.mat"
    write_affine_mat(path, a)
    b = read_affine_mat(path)
    assert params(b) == params(a)


def test_float_variables_are_read(tmp_path):
    a = sample()
    p, fixed = params(a)
    path = tmp_path / "f.mat"
    path.write_bytes(variable("AffineTransform_float_3_3", p, 10) + variable("fixed", fixed, 10))
    b = read_affine_mat(path)
    assert np.max(np.abs(np.array(params(b)[0]) - p)) < 1e-6
    assert np.max(np.abs(np.array(params(b)[1]) - fixed)) < 1e-5


def test_translation_and_matrix_offset_forms(tmp_path):
    path = tmp_path / "t.mat"
    path.write_bytes(variable("TranslationTransform_double_3_3", [1.5, -2.0, 0.25]) + variable("fixed", []))
    b = read_affine_mat(path)
    assert list(b.translation) == [1.5, -2.0, 0.25] and list(b.matrix) == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    p, fixed = params(sample())
    path.write_bytes(variable("MatrixOffsetTransformBase_double_3_3", p) + variable("fixed", fixed))
    assert params(read_affine_mat(path)) == (p, fixed)


@pytest.mark.parametrize("case", ["bad type", "truncated", "empty", "no fixed", "wrong second name", "field class"])
def test_malformed_files_are_refused(tmp_path, case):
    p, fixed = params(sample())
    good = variable("AffineTransform_double_3_3", p) + variable("fixed", fixed)
    data = {
        "bad type": b"\x01" + good[1:],
        "truncated": good[:-1],
        "empty": b"",
        "no fixed": variable("AffineTransform_double_3_3", p),
        "wrong second name": variable("AffineTransform_double_3_3", p) + variable("center", fixed),
        "field class": variable("DisplacementFieldTransform_double_3_3", p) + variable("fixed", fixed),
    }[case]
    path = tmp_path / "bad.mat"
    path.write_bytes(data)
    with pytest.raises(RegError):
        read_affine_mat(path)


def test_an_unwritable_path_is_refused(tmp_path):
    with pytest.raises(RegError, match="cannot open"):
        write_affine_mat(tmp_path / "missing" / "dir" / "a.mat", sample())
