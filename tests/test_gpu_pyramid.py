"""GPU parity of the level-image preparation that the HIP engine runs on the device in place of the
reference's host code: the recursive Gaussian of the pyramid (pyramid.rs:184-465, against the f64
host filter) and the fixed-image values at linear-stage sample points (mi.rs:161-207,
linear.rs:251-279, against the oracle's nearest-voxel and linear rules).

Errors are the largest absolute error over the largest reference magnitude (parity.rs `compare`).
The device runs the recursions in f64 and stores each axis pass in f32, so the result carries a
few f32 roundings (about 1e-7 of the range) whatever sigma is.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from fastwarp_amd._abi import Grid
from tests import cases

pytestmark = pytest.mark.gpu

GRIDS = {
    "odd": cases.grid((11, 9, 8), (1.5, 2.0, 1.25), (-8.0, 5.0, -3.5)),
    "oblique": cases.grid((23, 19, 17), (1.1, 0.9, 1.3), (-12.5, 7.25, -4.0), oblique=0.35),
    "medium": cases.grid((67, 71, 45), (1.0, 1.0, 1.0), (-30.0, -20.0, -10.0)),
    "flipped": Grid.make((30, 26, 22), (1.0, 1.0, 2.0), (12.0, 9.0, -5.0), np.diag([-1.0, -1.0, 1.0])),
}


@pytest.mark.parametrize("sigma_vox", [0.7, 2.0, 3.0, 6.0])
@pytest.mark.parametrize("name", list(GRIDS))
def test_smooth_image_matches_the_f64_filter(backend, name, sigma_vox):
    g = GRIDS[name]
    img = cases.image(g, 5)
    sigma_mm = np.array(list(g.spacing)) * sigma_vox
    want = oracle.smooth_image(g, img, sigma_mm)
    got = backend.download_image(backend.smooth_image(backend.upload_image(g, img), sigma_mm))
    assert cases.rel_err(got, want) <= 5e-7, (name, sigma_vox)


@pytest.mark.parametrize("shape", [(11, 9, 8), (182, 40, 9), (67, 71, 45), (1600, 5, 4)])
def test_lines_staged_in_lds_give_the_bits_of_the_direct_form(backend, shape):
    """The x lines of the recursive Gaussian go through LDS tiles (a partial last tile on most of
    these shapes; the 1600-voxel lines exceed the tile budget and take the direct form): the same
    bits as the direct form (tuning knob deriche.lds = 0)."""
    g = cases.grid(shape, (1.0, 1.2, 0.9), (-5.0, 3.0, 1.0))
    img = cases.image(g, 7)
    out = []
    for lds in (1, 0):
        backend.set_tuning("deriche.lds", lds)
        try:
            out.append(backend.download_image(backend.smooth_image(backend.upload_image(g, img), (2.0, 1.5, 1.0))))
        finally:
            backend.set_tuning("deriche.lds", -1)
    assert np.array_equal(out[0], out[1])


def test_smooth_image_keeps_constants_and_copies_at_zero(backend):
    g = GRIDS["medium"]
    c = np.full(g.n, 0.625)
    got = backend.download_image(backend.smooth_image(backend.upload_image(g, c), (3.0, 3.0, 3.0)))
    assert np.max(np.abs(got - 0.625)) <= 2e-6
    img = cases.image(g, 2)
    same = backend.download_image(backend.smooth_image(backend.upload_image(g, img), (0.0, 0.0, 0.0)))
    assert np.array_equal(same, img)


def test_smooth_image_rejects_short_axes(backend):
    from fastwarp_amd.backend import RegError

    g = cases.grid((40, 3, 5))
    with pytest.raises(RegError, match="4 voxels"):
        backend.smooth_image(backend.upload_image(g, cases.image(g, 1)), (1.0, 1.0, 1.0))


@pytest.mark.parametrize("name", list(GRIDS))
def test_sample_values_match_the_oracle(backend, name):
    g = GRIDS[name]
    img = cases.image(g, 9)
    rng = np.random.default_rng(4)
    # continuous indices from a voxel outside each face to a voxel past it, so both rules' edges
    # ([-0.5, n - 0.5) for linear, rounding to [0, n) for nearest) are crossed
    size = np.array(g.shape_xyz, dtype=float)
    idx = rng.uniform(-1.2, size + 0.2, size=(4000, 3))
    pts = np.array([g.index_to_point(c) for c in idx])
    want_n, want_l = oracle.sample_values(g, img, pts)
    got_n, got_l = backend.sample_values(backend.upload_image(g, img), pts)
    assert np.array_equal(np.isnan(got_n), np.isnan(want_n))
    assert np.array_equal(np.isnan(got_l), np.isnan(want_l))
    ok_n, ok_l = ~np.isnan(want_n), ~np.isnan(want_l)
    assert np.array_equal(got_n[ok_n], want_n[ok_n])
    assert np.max(np.abs(got_l[ok_l] - want_l[ok_l])) <= 2e-6
