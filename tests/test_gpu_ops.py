"""GPU parity of every SyN-stage op against the f64 oracle, through the C ABI.

Inputs are rounded to f32 before both sides see them, so the comparison measures the kernels'
arithmetic. Errors are the largest absolute error over the largest reference magnitude, as the
reference's known-answer test reports them (crates/fastwarp-gpu/src/parity.rs:25-54, 258-285).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from fastwarp_amd._abi import Affine, Grid, InvertParams
from tests import cases

pytestmark = pytest.mark.gpu

GRIDS = {
    "odd": cases.grid((11, 9, 8), (1.5, 2.0, 1.25), (-8.0, 5.0, -3.5)),
    "oblique": cases.grid((23, 19, 17), (1.1, 0.9, 1.3), (-12.5, 7.25, -4.0), oblique=0.35),
    "thin": cases.grid((40, 3, 5), (1.0, 1.0, 1.0)),
    "medium": cases.grid((67, 71, 45), (1.0, 1.0, 1.0), (-30.0, -20.0, -10.0)),
    "flipped": Grid.make((30, 26, 22), (1.0, 1.0, 2.0), (12.0, 9.0, -5.0), np.diag([-1.0, -1.0, 1.0])),
}


def other_grid(g: Grid) -> Grid:
    nx, ny, nz = g.shape_xyz
    return cases.grid((nx + 3, max(ny - 1, 2), nz + 2), tuple(s * f for s, f in zip(g.spacing, (0.8, 1.15, 0.9))),
                      tuple(o + d for o, d in zip(g.origin, (1.3, -0.9, 0.7))), oblique=0.17)


@pytest.mark.parametrize("name", list(GRIDS))
def test_field_round_trip(backend, name):
    g = GRIDS[name]
    f = cases.field(g, 2.0, 3)
    back = backend.download_field(backend.upload_field(g, f))
    assert cases.rel_err(back, f) <= 2e-7


@pytest.mark.parametrize("name", ["odd", "oblique", "medium", "flipped"])
@pytest.mark.parametrize("mode", ["plain", "affine", "pre"])
def test_warp_matches_oracle(backend, name, mode):
    g = GRIDS[name]
    src_grid = other_grid(g)
    src = cases.image(src_grid, 5)
    fld = cases.field(g, 1.4 * g.spacing[0], 7)
    aff = Affine.make([[1.02, 0.03, -0.01], [-0.02, 0.98, 0.02], [0.01, -0.03, 1.01]], (0.7, -0.4, 0.3), tuple(g.index_to_point(np.array(g.shape_xyz) / 2)))
    pre_g = cases.grid(tuple(max(s - 2, 2) for s in g.shape_xyz), g.spacing, g.origin) if mode == "pre" else None
    pre = cases.field(pre_g, 1.0, 9) if pre_g else None
    want = oracle.warp(src_grid, src, g, fld, pre_grid=pre_g, pre=pre, affine=aff if mode != "plain" else None)
    d_src = backend.upload_image(src_grid, src)
    d_f = backend.upload_field(g, fld)
    d_pre = backend.upload_field(pre_g, pre) if pre_g else None
    got = backend.download_image(backend.warp_through_field(d_src, d_f, d_pre, aff if mode != "plain" else None))
    # voxels whose source point sits on the buffer edge may flip inside/outside in f32
    flips = np.count_nonzero((got == 0) != (want == 0))
    keep = (got != 0) & (want != 0)
    assert flips <= 1 + g.n // 500
    assert cases.rel_err(got[keep], want[keep]) <= 1e-5


@pytest.mark.parametrize("name", ["odd", "oblique", "medium"])
def test_warp_mask_matches_oracle(backend, name):
    g = GRIDS[name]
    mg = other_grid(g)
    i, j, k = cases.coords(mg)
    mask = ((i + 2 * j + k) % 5 != 0).astype(float)
    fld = cases.field(g, 1.4 * g.spacing[0], 11)
    prod = (cases.image(g, 2) > 0.4).astype(float)
    want = oracle.warp(mg, mask, g, fld, nearest=True, product=prod)
    got = backend.download_image(
        backend.warp_mask_through_field(backend.upload_image(mg, mask), backend.upload_field(g, fld), None, None,
                                        backend.upload_image(g, prod))
    )
    assert set(np.unique(got)) <= {0.0, 1.0}
    assert np.count_nonzero(got != want) <= 1 + g.n // 500


def cc_case(g: Grid, seed: int):
    f = cases.image(g, seed)
    # the moving image: the fixed one seen through a smooth warp, plus its own texture
    m = oracle.warp(g, f, g, cases.field(g, 1.2, seed + 1))
    m = (0.7 * m + 0.3 * cases.image(g, seed + 2)).astype(np.float32).astype(np.float64)
    return f, m


@pytest.mark.parametrize("name", list(GRIDS))
@pytest.mark.parametrize("radius", [2, 4])
def test_cc_pair_matches_oracle(backend, name, radius):
    """single-kernel CC gradient within 1e-5 of the FP64 evaluation (range-relative)."""
    g = GRIDS[name]
    f, m = cc_case(g, 3)
    a = oracle.cc(g, f, m, radius)
    b = oracle.cc(g, m, f, radius)
    (v0, d0), (v1, d1) = backend.cc_value_derivative_pair(backend.upload_image(g, f), backend.upload_image(g, m), radius)
    assert abs(v0 - a["value"]) <= 1e-5 * abs(a["value"])
    assert abs(v1 - b["value"]) <= 1e-5 * abs(b["value"])
    assert cases.rel_err(backend.download_field(d0), a["derivative"]) <= 1e-5
    assert cases.rel_err(backend.download_field(d1), b["derivative"]) <= 1e-5


@pytest.mark.parametrize("name", ["odd", "oblique", "medium"])
def test_cc_pair_masked_matches_oracle(backend, name):
    g = GRIDS[name]
    f, m = cc_case(g, 5)
    i, j, k = cases.coords(g)
    c = np.array(g.shape_xyz) / 2
    valid = (((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) > 4.0).astype(float)
    a = oracle.cc(g, f, m, 2, mask=valid)
    b = oracle.cc(g, m, f, 2, mask=valid)
    ((v0, d0), (v1, d1)), norms = backend.cc_value_derivative_pair_masked_with_norms(
        backend.upload_image(g, f), backend.upload_image(g, m), backend.upload_image(g, valid), 2
    )
    assert abs(v0 - a["value"]) <= 1e-5 * abs(a["value"])
    assert abs(v1 - b["value"]) <= 1e-5 * abs(b["value"])
    g0, g1 = backend.download_field(d0), backend.download_field(d1)
    assert cases.rel_err(g0, a["derivative"]) <= 1e-5
    assert cases.rel_err(g1, b["derivative"]) <= 1e-5
    assert np.all(g0[valid == 0] == 0)
    assert abs(norms[0] - np.linalg.norm(a["derivative"])) <= 1e-5 * np.linalg.norm(a["derivative"])


def test_cc_of_an_image_with_itself_is_minus_one_at_full_size(backend):
    """Size-independent property at the 1 mm grid: CC(I, I) = -1 and a zero derivative."""
    g = cases.grid((182, 218, 182))
    f = cases.image(g, 1)
    d = backend.upload_image(g, f)
    (v0, d0), (v1, d1) = backend.cc_value_derivative_pair(d, d, 4)
    assert abs(v0 + 1.0) < 1e-6 and abs(v1 + 1.0) < 1e-6
    assert np.max(np.abs(backend.download_field(d0))) < 1e-5


def test_cc_rejects_bad_radius(backend):
    g = GRIDS["odd"]
    d = backend.upload_image(g, cases.image(g, 1))
    from fastwarp_amd.backend import RegError

    with pytest.raises(RegError, match="outside 1..=8"):
        backend.cc_value_derivative_pair(d, d, 9)


@pytest.mark.parametrize("name", list(GRIDS))
@pytest.mark.parametrize("variance", [3.0, 0.25, 20.0])
def test_smooth_matches_oracle(backend, name, variance):
    # variance 20 needs radii above 8 on the larger axes: the three-pass path; the others run fused
    g = GRIDS[name]
    f = cases.field(g, 2.0, 13)
    want = oracle.smooth(g, f, variance)
    got = backend.download_field(backend.smooth_field(backend.upload_field(g, f), variance))
    assert cases.rel_err(got, want) <= 2e-6


@pytest.mark.parametrize("variance", [3.0, 0.25])
def test_smooth_matches_oracle_on_a_large_grid(backend, variance):
    """Grids of four million voxels and more take the one-pass three-axis kernel."""
    g = cases.grid((171, 160, 150), (1.0, 1.0, 1.0), (-85.0, -80.0, -75.0))
    f = cases.field(g, 2.0, 17)
    want = oracle.smooth(g, f, variance)
    got = backend.download_field(backend.smooth_field(backend.upload_field(g, f), variance))
    assert cases.rel_err(got, want) <= 2e-6


@pytest.mark.parametrize("name", ["odd", "oblique", "medium"])
def test_norms_scale_axpy(backend, name):
    g = GRIDS[name]
    a, b = cases.field(g, 2.0, 1), cases.field(g, 0.5, 2)
    da, db = backend.upload_field(g, a), backend.upload_field(g, b)
    ma, la = oracle.norms(g, a)
    mb, _ = oracle.norms(g, b)
    ga, gb = backend.max_scaled_norm_pair(da, db)
    assert abs(ga - ma) <= 1e-6 * ma and abs(gb - mb) <= 1e-6 * mb
    assert abs(backend.l2_norm(da) - la) <= 1e-6 * la
    backend.scale(da, 0.37)
    backend.axpy(da, -1.5, db)
    assert cases.rel_err(backend.download_field(da), 0.37 * a - 1.5 * b) <= 1e-6


@pytest.mark.parametrize("name", list(GRIDS))
def test_compose_matches_oracle(backend, name):
    g = GRIDS[name]
    u, w = cases.field(g, 1.5 * g.spacing[0], 3), cases.field(g, 1.2 * g.spacing[0], 4)
    want = oracle.compose(g, u, w)
    got = backend.download_field(backend.compose(backend.upload_field(g, u), backend.upload_field(g, w)))
    assert cases.rel_err(got, want) <= 1e-5


@pytest.mark.parametrize("name", ["odd", "oblique", "medium", "flipped"])
def test_invert_matches_oracle(backend, name):
    g = GRIDS[name]
    u = cases.field(g, 0.6 * g.spacing[0], 5)
    init = cases.field(g, 0.1, 6)
    want, (it, mean, mx) = oracle.invert(g, u, init)
    got, st = backend.invert(backend.upload_field(g, u), backend.upload_field(g, init))
    assert st.iterations == it
    assert cases.rel_err(backend.download_field(got), want) <= 1e-4
    assert abs(st.max_error_norm - mx) <= 1e-4 * mx and abs(st.mean_error_norm - mean) <= 1e-4 * mean


def test_invert_honours_zero_iterations_and_the_limit(backend):
    g = GRIDS["medium"]
    u = cases.field(g, 0.8, 5)
    p0 = InvertParams(0, 0.001, 0.1, 1)
    got, st = backend.invert(backend.upload_field(g, u), backend.zero_field(g), p0)
    assert st.iterations == 0 and st.max_error_norm == 0.0
    assert np.all(backend.download_field(got) == 0)
    p3 = InvertParams(3, 0.0, 0.0, 1)
    want, (it, _, _) = oracle.invert(g, u, None, p3)
    got, st = backend.invert(backend.upload_field(g, u), backend.zero_field(g), p3)
    assert st.iterations == it == 3
    assert cases.rel_err(backend.download_field(got), want) <= 1e-4


@pytest.mark.parametrize("name", ["odd", "oblique", "medium"])
def test_resample_matches_oracle(backend, name):
    g = GRIDS[name]
    f = cases.field(g, 2.0, 8)
    target = other_grid(g)
    want = oracle.resample(g, f, target)
    got = backend.download_field(backend.resample_field(backend.upload_field(g, f), target))
    flips = np.count_nonzero(np.all(got == 0, axis=1) != np.all(want == 0, axis=1))
    keep = ~(np.all(got == 0, axis=1) | np.all(want == 0, axis=1))
    assert flips <= 1 + target.n // 500
    assert cases.rel_err(got[keep], want[keep]) <= 1e-5


@pytest.mark.parametrize("size,spacing,mm", [((37, 29, 23), (1.5, 0.8, 2.0), 1.0), ((40, 44, 36), (0.7, 0.7, 1.3), 1.2),
                                             ((121, 140, 97), (1.2, 1.2, 1.2), 1.0), ((20, 21, 22), (1.0, 1.0, 1.0), 1.0)])
def test_isotropic_resample_matches_oracle(backend, size, spacing, mm):
    """register.rs:112-142 on the device against the oracle's f64 rule: the same grid, values within
    1e-6 of the image's range (f32 weights against f64), and a copy when already isotropic."""
    g = cases.grid(size, spacing, (-11.0, 4.5, 2.25), oblique=0.2)
    v = cases.image(g, 5)
    og, want = oracle.resample_isotropic(g, v, mm)
    got_img = backend.resample_isotropic(backend.upload_image(g, v), mm)
    got = backend.download_image(got_img)
    assert got.shape == want.shape, (got.shape, want.shape)
    gg = got_img.grid
    assert list(gg.size) == list(og.size)
    assert np.allclose(list(gg.spacing), list(og.spacing), rtol=0, atol=0)
    assert np.array_equal(np.array(list(gg.origin)), np.array(list(og.origin)))
    assert np.array_equal(np.array(list(gg.direction)), np.array(list(og.direction)))
    assert np.max(np.abs(got - want)) <= 1e-6 * np.max(np.abs(want))
    if spacing == (mm, mm, mm):
        assert np.array_equal(got, v.astype(np.float32).astype(np.float64))
