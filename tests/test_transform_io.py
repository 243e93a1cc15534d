"""ANTs transform files. The affine (0GenericAffine.mat): the ITK MATLAB Level-4 layout of
crates/fastwarp-reg/src/transform/itk_mat.rs, checked against bytes built here from the format's
definition (five little-endian u32, the NUL-terminated name, f64 or f32 values), and the
reference's own cases (itk_mat.rs:196-246: a 193-byte bit-exact round trip, f32 variables,
bad type codes and truncation refused). The composites (Composite.h5 / InverseComposite.h5,
itk_h5.rs and ants_io.rs): the ITK file of the reference's tests, written by libhdf5 (its values
as the reference's test states them), bit-exact round trips, the pair's chains, and damaged files
refused. Host only."""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np
import pytest

from fastwarp_amd._abi import Affine, Grid
from fastwarp_amd.backend import (RegError, read_affine_mat, read_composite_h5, write_affine_mat, write_composite_h5,
                                  write_composite_pair)


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


# ---- ITK HDF5 transform files (Composite.h5 / InverseComposite.h5): itk_h5.rs, ants_io.rs ----


GOLDEN_H5 = Path(__file__).resolve().parent / "golden" / "itk" / "ctest.h5"


def field_grid() -> Grid:
    # an oblique grid: axes permuted and flipped
    return Grid.make((5, 6, 7), (2.0, 2.5, 3.0), (-4.0, -5.0, -6.0), [[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])


def field_values(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 2.0, (5 * 6 * 7, 3)).astype(np.float32).astype(np.float64)


def grid_tuple(g: Grid):
    return tuple(g.size), tuple(g.spacing), tuple(g.origin), tuple(g.direction)


def test_itk_file_written_by_libhdf5_is_read():
    # itk_h5.rs:205-237: ITK 5.4.7 / HDF5 1.12.1 wrote an affine diag(2, 1, 1) then a 6x5x4 field of (1, 0, 0)
    chain = read_composite_h5(GOLDEN_H5)
    assert len(chain) == 2 and isinstance(chain[0], Affine)
    a = chain[0]
    assert list(a.matrix) == [2.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    assert list(a.translation) == [0.0] * 3 and list(a.center) == [0.0] * 3
    g, v = chain[1]
    assert tuple(g.size) == (6, 5, 4) and tuple(g.spacing) == (1.0, 1.0, 1.0) and tuple(g.origin) == (0.0, 0.0, 0.0)
    assert tuple(g.direction) == (1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0)
    assert v.shape == (120, 3) and np.all(v == [1.0, 0.0, 0.0])


def test_composite_round_trip_is_bit_exact(tmp_path):
    a = sample()
    g = field_grid()
    v = field_values(3)
    path = tmp_path / "Composite.h5"
    write_composite_h5(path, [a, (g, v)])
    assert path.read_bytes()[:8] == b"\x89HDF\r\n\x1a\n"
    back = read_composite_h5(path)
    assert params(back[0]) == params(a)
    gb, vb = back[1]
    assert grid_tuple(gb) == grid_tuple(g)
    assert np.array_equal(vb.view(np.uint64), v.view(np.uint64))


def test_the_libhdf5_file_rewritten_reads_back_the_same(tmp_path):
    chain = read_composite_h5(GOLDEN_H5)
    path = tmp_path / "again.h5"
    write_composite_h5(path, chain)
    again = read_composite_h5(path)
    assert params(again[0]) == params(chain[0])
    assert grid_tuple(again[1][0]) == grid_tuple(chain[1][0]) and np.array_equal(again[1][1], chain[1][1])


def test_pair_files_hold_the_chains_of_ants(tmp_path):
    # ants_io.rs:100-141: Composite = [affine, field], InverseComposite = [inverse field, affine^-1]
    a = sample()
    g = field_grid()
    fwd, inv = field_values(4), field_values(5)
    prefix = tmp_path / "sub-01_"
    write_composite_pair(prefix, a, g, fwd, inv)
    forward = read_composite_h5(tmp_path / "sub-01_Composite.h5")
    inverse = read_composite_h5(tmp_path / "sub-01_InverseComposite.h5")
    assert params(forward[0]) == params(a) and np.array_equal(forward[1][1], fwd)
    assert np.array_equal(inverse[0][1], inv)
    ai = inverse[1]
    m = np.array(a.matrix).reshape(3, 3)
    mi = np.array(ai.matrix).reshape(3, 3)
    assert np.allclose(mi @ m, np.eye(3), atol=1e-13)
    # same centre, translation -M^-1 t: the inverse maps a's image of a point back to it
    c = np.array(a.center)
    p = np.array([5.0, -9.0, 4.0])
    y = m @ (p - c) + c + np.array(a.translation)
    back = mi @ (y - np.array(ai.center)) + np.array(ai.center) + np.array(ai.translation)
    assert np.allclose(back, p, atol=1e-12) and list(ai.center) == list(a.center)


def test_linear_only_pair_and_refusals(tmp_path):
    a = sample()
    write_composite_pair(tmp_path / "lin_", a)
    assert len(read_composite_h5(tmp_path / "lin_Composite.h5")) == 1
    assert len(read_composite_h5(tmp_path / "lin_InverseComposite.h5")) == 1
    singular = Affine.make([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [0.0, 0.0, 1.0]])
    with pytest.raises(RegError, match="singular"):
        write_composite_pair(tmp_path / "bad_", singular)
    assert not (tmp_path / "bad_Composite.h5").exists()
    with pytest.raises(RegError, match="empty chain"):
        write_composite_h5(tmp_path / "empty.h5", [])
    with pytest.raises(ValueError):
        write_composite_h5(tmp_path / "short.h5", [(field_grid(), np.zeros(10))])
    damaged = bytearray(GOLDEN_H5.read_bytes())
    damaged[60] ^= 0x40
    (tmp_path / "damaged.h5").write_bytes(bytes(damaged))
    with pytest.raises(RegError, match="checksum"):
        read_composite_h5(tmp_path / "damaged.h5")
    (tmp_path / "notes.h5").write_bytes(b"not an hdf5 file at all, just text" * 4)
    with pytest.raises(RegError, match="signature"):
        read_composite_h5(tmp_path / "notes.h5")


def test_a_1mm_field_pair_streams_through_the_writer(tmp_path):
    """The 1 mm grid of the plain workload (182 x 218 x 182, 7.2 M voxels): each file holds 174 MB
    of f64 parameters, widened from the field's f32 values in blocks as they are written; both
    files read back with every value, and the metadata's lengths take 8-byte fields."""
    g = Grid.make((182, 218, 182), (1.0, 1.0, 1.0), (-90.0, -126.0, -72.0))
    n = 182 * 218 * 182
    rng = np.random.default_rng(17)
    fwd = rng.normal(0.0, 2.0, (n, 3)).astype(np.float32).astype(np.float64)
    inv = -fwd
    write_composite_pair(tmp_path / "big_", sample(), g, fwd, inv)
    size = (tmp_path / "big_Composite.h5").stat().st_size
    assert size >= 3 * n * 8 and size < 3 * n * 8 + (1 << 16)
    forward = read_composite_h5(tmp_path / "big_Composite.h5")
    assert np.array_equal(forward[1][1], fwd)
    del forward
    inverse = read_composite_h5(tmp_path / "big_InverseComposite.h5")
    assert np.array_equal(inverse[0][1], inv)


def test_composite_files_are_written_whole_or_not_at_all(tmp_path):
    # fs_write.rs write_complete: a temporary file beside the target, renamed once complete
    a = sample()
    with pytest.raises(RegError, match="cannot open"):
        write_composite_h5(tmp_path / "missing" / "x.h5", [a])
    assert not (tmp_path / "missing").exists()
    path = tmp_path / "c.h5"
    path.write_bytes(b"old")
    write_composite_h5(path, [a, (field_grid(), field_values(6))])
    assert len(read_composite_h5(path)) == 2
    assert sorted(p.name for p in tmp_path.iterdir()) == ["c.h5"]
