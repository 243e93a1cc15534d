"""The reference's edge cases run through the HIP path:

* the ITK 5.4.5 Mattes fixtures through the dense MI pair in both directions, with masks
  (crates/fastwarp-gpu/tests/mi_dense_parity.rs:452-495: value within 1e-6 relative, derivative
  within 1e-3 of its range off the z faces, no zero/non-zero mismatch there);
* the sampled MI histogram cells against the oracle's f64 PDF (north star: 1e-5);
* the nearest-voxel rule's round-half-up ties (mask.rs:99-111);
* a warp through a prior field against its closed form (cpu/mod.rs:60-86);
* the device centre of gravity of a blob (init.rs:107-140);
* the wait timeout (device.rs:919-928), forced from the host.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from fastwarp_amd._abi import Affine, Grid, LinearTransform, MiBinning
from fastwarp_amd.backend import MiLevel
from tests import cases
from tests.test_gpu_mi import fixed_bins, smooth_vol
from tests.test_oracle import blob, fixture_grid, load_fixture

pytestmark = pytest.mark.gpu


def off_z_faces(g: Grid) -> np.ndarray:
    _, _, k = cases.coords(g)
    return (k > 0) & (k < g.shape_xyz[2] - 1)


def compare_fixture(value, field, fx, g, label):
    want = fx["derivative"].reshape(-1, 3)
    keep = off_z_faces(g)
    assert abs(value - fx["value"][0]) <= 1e-6 * abs(fx["value"][0]), (label, value, fx["value"][0])
    rng = np.max(np.abs(want))
    err = np.max(np.abs(field[keep] - want[keep])) / rng
    assert err <= 1e-3, (label, err)
    zero_got = np.all(field[keep] == 0, axis=1)
    zero_want = np.all(want[keep] == 0, axis=1)
    assert np.array_equal(zero_got, zero_want), label


def test_dense_mattes_pair_matches_the_itk_fixtures_in_both_directions(backend):
    plain = load_fixture("virtual_b32")
    g = fixture_grid(plain, "fixed")
    assert g.shape_xyz == fixture_grid(plain, "moving").shape_xyz and not np.any(plain["field"])
    up = lambda v: backend.upload_image(g, v)
    a, b = up(plain["fixed"]), up(plain["moving"])
    (v1, d1), _ = backend.mi_value_derivative_pair(a, b, 32)
    compare_fixture(v1, backend.download_field(d1), plain, g, "virtual_b32")
    masked = load_fixture("virtual_b32_masked")
    swapped = load_fixture("virtual_b32_masked_swapped")
    (w1, e1), (w2, e2) = backend.mi_value_derivative_pair_masked(a, b, up(masked["fixed_mask"]), up(masked["moving_mask"]), 32)
    compare_fixture(w1, backend.download_field(e1), masked, g, "virtual_b32_masked")
    compare_fixture(w2, backend.download_field(e2), swapped, g, "virtual_b32_masked second direction")


@pytest.mark.parametrize("kind", [0, 1])
def test_sampled_mi_histogram_cells_match_the_oracle(backend, kind):
    """The joint PDF from the device histogram (2^-32 fixed-point cells) against the oracle's f64
    PDF, cell by cell, within 1e-5 of the largest cell; the same valid count."""
    g = cases.grid((40, 36, 30), (1.2, 1.0, 1.4), (-20.0, -15.0, -18.0), oblique=0.3)
    f = smooth_vol(g, (0, 0, 0))
    m = smooth_vol(g, (1.1, -0.8, 0.6), 0.3)
    rng = np.random.default_rng(3)
    pts = np.array([g.index_to_point(p) for p in rng.uniform([4, 4, 4], np.array(g.shape_xyz) - 5, size=(20000, 3))])
    b = MiBinning.from_ranges(32, (f.min(), f.max()), (m.min(), m.max()))
    fb = fixed_bins(g, f, pts, b)
    centre = tuple(g.index_to_point(np.array(g.shape_xyz) / 2))
    params = (np.array([0.02, -0.01, 0.03, 0.8, -0.5, 0.4]) if kind == 0 else
              np.array([1.02, 0.01, 0.0, -0.01, 0.97, 0.02, 0.0, 0.01, 1.01, 0.5, -0.3, 0.2]))
    t = LinearTransform.make(kind, params, centre, Affine.make(translation=(0.1, 0.05, -0.1)))
    want = oracle.mi_sampled(b, g, m, pts, fb, t, derivative=False)
    level = MiLevel(backend, backend.upload_image(g, m), pts, fb, b)
    joint, marginal, valid = level.histogram(t, 32)
    assert valid == want["num_valid"]
    assert marginal.sum() == valid
    pdf = joint.reshape(-1) / joint.sum()
    ref = want["joint_pdf"]
    assert np.max(np.abs(pdf - ref)) <= 1e-5 * np.max(ref), np.max(np.abs(pdf - ref)) / np.max(ref)


def test_nearest_voxel_ties_round_half_up_on_the_device(backend):
    src = Grid.make((6, 5, 4), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    dst = Grid.make((6, 5, 4), (1.0, 1.0, 1.0), (0.5, 0.0, 0.0))
    i, j, k = cases.coords(src)
    mask = ((i + 2 * j + 3 * k) % 4 != 1).astype(float)
    zero = np.zeros((dst.n, 3))
    want = oracle.warp(src, mask, dst, zero, nearest=True)
    got = backend.download_image(backend.warp_mask_through_field(backend.upload_image(src, mask), backend.upload_field(dst, zero)))
    di, dj, dk = cases.coords(dst)
    # the upper voxel of every tie, and nothing past the last one
    upper = np.where(di < 5, ((di + 1 + 2 * dj + 3 * dk) % 4 != 1).astype(float), 0.0)
    assert np.array_equal(want, upper)
    assert np.array_equal(got, want)


def test_warp_through_a_prior_field_matches_its_closed_form_on_the_device(backend):
    g = cases.grid((10, 9, 8), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    pre_g = cases.grid((12, 11, 10), (1.0, 1.0, 1.0), (-1.0, -1.0, -1.0))
    i, j, k = cases.coords(g)
    image = 3.0 * i - 2.0 * j + 0.5 * k + 7.0
    u = np.tile([0.75, -0.5, 0.25], (g.n, 1))
    pre = np.tile([1.0, 0.25, -0.5], (pre_g.n, 1))
    got = backend.download_image(backend.warp_through_field(backend.upload_image(g, image), backend.upload_field(g, u),
                                                            backend.upload_field(pre_g, pre)))
    x, y, z = i + 1.75, j - 0.25, k - 0.25
    inside = (x >= -0.5) & (x < 9.5) & (y >= -0.5) & (y < 8.5) & (z >= -0.5) & (z < 7.5)
    interior = inside & (x <= 9) & (y >= 0) & (z >= 0)
    want = 3.0 * x - 2.0 * y + 0.5 * z + 7.0
    assert np.max(np.abs(got[interior] - want[interior])) <= 1e-5 * np.max(np.abs(want)), "prior-field warp"
    assert np.all(got[~inside] == 0.0)


def test_device_centre_of_gravity_of_a_blob_is_its_centre(backend):
    g = Grid.make((40, 36, 32), (1.0, 1.5, 2.0), (-20.0, -30.0, -10.0))
    c = backend.centre_of_gravity(backend.upload_image(g, blob(g, (3.0, -4.0, 5.0))))
    assert np.max(np.abs(np.asarray(c) - np.array([3.0, -4.0, 5.0]))) < 1e-3, c


def test_a_wait_past_the_wait_timeout_fails_the_call():
    """device.rs:919-928: FASTWARP_GPU_WAIT_TIMEOUT_S bounds every wait on the device; with a
    nanosecond the first wait of a 1 mm CC pair gives up (its kernels still finish normally and
    the context's destruction waits for them), and the error names the setting."""
    import os

    from fastwarp_amd.backend import HipBackend, RegError

    g = cases.grid((182, 218, 182))
    f, m = cases.image(g, 1), cases.image(g, 2)
    os.environ["FASTWARP_GPU_WAIT_TIMEOUT_S"] = "1e-9"
    try:
        b = HipBackend(0)
    finally:
        del os.environ["FASTWARP_GPU_WAIT_TIMEOUT_S"]
    try:
        # the uploads wait for their copies too: whichever wait comes first gives up
        with pytest.raises(RegError, match="FASTWARP_GPU_WAIT_TIMEOUT_S"):
            b.cc_value_derivative_pair(b.upload_image(g, f), b.upload_image(g, m), 4)
    finally:
        b.close()
