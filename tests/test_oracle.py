"""Pins of the f64 oracle (CPU only): the reference's ITK fixtures, brute force and closed forms.

Each check targets a mistake the restatement could plausibly make (a dropped term, a wrong sign
or index, a transposed operand); none re-types the oracle's own formula.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle
from fastwarp_amd._abi import Affine, Grid, InvertParams, LinearTransform, MiBinning
from tests import cases

GOLDEN = Path(__file__).resolve().parent / "golden" / "mattes"


# --------------------------------------------------------------- ITK Mattes fixtures -------

def load_fixture(name: str) -> dict:
    out = {}
    for line in (GOLDEN / f"{name}.txt").read_text().splitlines():
        if not line or line.startswith("#"):
            continue
        key, vals = line.split(":", 1)
        out[key] = np.array([float(v) for v in vals.split()])
    return out


def fixture_grid(fx: dict, role: str) -> Grid:
    return Grid.make(fx[f"{role}_size"].astype(int), fx[f"{role}_spacing"], fx[f"{role}_origin"], fx[f"{role}_direction"])


def check_itk(name: str):
    """ITK 5.4.5 fixture (crates/fastwarp-reg/tests/mattes_itk.rs:95-132): value within 1e-12
    relative, valid count exact, derivative within 1e-10 of its range, identical zero pattern."""
    fx = load_fixture(name)
    fg, mg = fixture_grid(fx, "fixed"), fixture_grid(fx, "moving")
    r = oracle.mi_dense(
        fg, fx["fixed"], mg, fx["moving"], int(fx["bins"][0]), field_grid=fg, field=fx["field"],
        fixed_mask=fx.get("fixed_mask"), moving_mask=fx.get("moving_mask"),
    )
    want = fx["derivative"].reshape(-1, 3)
    assert r["num_valid"] == int(fx["valid"][0])
    assert abs(r["value"] - fx["value"][0]) <= 1e-12 * abs(fx["value"][0])
    rng = np.max(np.abs(want))
    assert rng > 0
    assert np.max(np.abs(r["derivative"] - want)) <= 1e-10 * rng
    zero_got = np.all(r["derivative"] == 0, axis=1)
    zero_want = np.all(want == 0, axis=1)
    assert np.array_equal(zero_got, zero_want)


@pytest.mark.parametrize(
    "name", ["virtual_b32", "virtual_b32_masked", "virtual_b32_masked_swapped", "oblique_b32", "oblique_b12", "oblique_b32_masked"]
)
def test_dense_mattes_matches_itk_fixtures(name):
    check_itk(name)


def test_a_perturbed_oracle_fails_the_fixture():
    """The fixture discriminates: a one-percent nonlinear change of the moving intensities must move
    the derivative far beyond the 1e-10 bound."""
    fx = load_fixture("oblique_b32")
    fg, mg = fixture_grid(fx, "fixed"), fixture_grid(fx, "moving")
    r = oracle.mi_dense(fg, fx["fixed"], mg, fx["moving"] + 0.01 * fx["moving"] ** 2, 32, field_grid=fg, field=fx["field"])
    want = fx["derivative"].reshape(-1, 3)
    assert np.max(np.abs(r["derivative"] - want)) > 1e-6 * np.max(np.abs(want))


# ------------------------------------------------------------------------------ CC ---------

def noise_image(g: Grid, seed: int) -> np.ndarray:
    n = g.n
    return (((np.arange(n) + seed) * 2654435761) % 1000) / 1000.0


def brute_cc(g: Grid, f: np.ndarray, m: np.ndarray, valid: np.ndarray, r: int, idx: int) -> float:
    nx, ny, nz = g.shape_xyz
    F = f.reshape(nz, ny, nx)
    M = m.reshape(nz, ny, nx)
    V = valid.reshape(nz, ny, nx)
    k, rem = divmod(idx, nx * ny)
    j, i = divmod(rem, nx)
    sl = (slice(max(k - r, 0), k + r + 1), slice(max(j - r, 0), j + r + 1), slice(max(i - r, 0), i + r + 1))
    a, b, v = F[sl][V[sl]], M[sl][V[sl]], len(F[sl][V[sl]])
    ffv = np.sum(a * a) - np.sum(a) ** 2 / v
    mmv = np.sum(b * b) - np.sum(b) ** 2 / v
    fmv = np.sum(a * b) - np.sum(a) * np.sum(b) / v
    return fmv * fmv / (ffv * mmv) if abs(ffv * mmv) > np.finfo(float).eps else 1.0


def test_cc_of_an_image_with_itself_is_minus_one_with_zero_derivative():
    g = cases.grid((10, 10, 10))
    v = noise_image(g, 1)
    r = oracle.cc(g, v, v, 4)
    assert r["num_valid"] == 1000
    assert abs(r["value"] + 1.0) < 1e-12
    assert np.max(np.abs(r["derivative"])) < 1e-9


def test_cc_window_statistics_match_brute_force_with_invalid_voxels():
    g = cases.grid((9, 9, 9))
    f, m = noise_image(g, 3), noise_image(g, 11)
    shift = Affine.make(translation=(2.3, -1.2, 0.4))
    r = oracle.cc(g, f, m, 2, moving_affine=shift, derivative=False)
    assert 100 < r["num_valid"] < 729
    # the moving image as the metric sees it: interpolated at the shifted points
    ident = np.zeros((g.n, 3))
    warped = oracle.warp(g, m, g, ident, affine=shift)
    total = 0.0
    for idx in np.flatnonzero(r["valid"]):
        want = brute_cc(g, f, warped, r["valid"], 2, int(idx))
        assert abs(r["local_cc"][idx] - want) < 1e-9
        total += want
    assert abs(r["value"] + total / r["num_valid"]) < 1e-9


def test_cc_mask_restricts_sums_and_zeroes_outside():
    g = cases.grid((9, 9, 9))
    f, m = noise_image(g, 3), noise_image(g, 11)
    i, j, k = cases.coords(g)
    inside = ~((np.sqrt((i - 4) ** 2 + (j - 3.5) ** 2 + (k - 5) ** 2) < 2.6) | (i == 0) | ((j == 8) & (k > 4)))
    r = oracle.cc(g, f, m, 2, mask=np.where(inside, 3.0, 0.0))
    assert np.array_equal(r["valid"], inside)
    for idx in np.flatnonzero(inside)[::7]:
        assert abs(r["local_cc"][idx] - brute_cc(g, f, m, inside, 2, int(idx))) < 1e-9
    assert np.all(r["derivative"][~inside] == 0)
    with pytest.raises(oracle.OracleError, match="inside both"):
        oracle.cc(g, f, m, 2, mask=np.zeros(g.n))


def test_cc_coefficient_is_the_derivative_of_local_cc_wrt_the_centre():
    g = cases.grid((9, 9, 9))
    f, m = noise_image(g, 2), noise_image(g, 9)
    base = oracle.cc(g, f, m, 2)
    h = 1e-6
    for idx in [g.n // 2, 0, 8 + 9 * 3 + 81 * 5, 1 + 9 * 7 + 81 * 2]:
        up, dn = m.copy(), m.copy()
        up[idx] += h
        dn[idx] -= h
        fd = (oracle.cc(g, f, up, 2)["local_cc"][idx] - oracle.cc(g, f, dn, 2)["local_cc"][idx]) / (2 * h)
        assert abs(fd - base["coefficient"][idx]) < 1e-5 * (1 + abs(fd))


def test_cc_is_invariant_to_affine_intensity_changes():
    g = cases.grid((10, 10, 10))
    f = noise_image(g, 5)
    m = cases.image(g, 3, f32=False)
    base = oracle.cc(g, f, m, 3)
    a = oracle.cc(g, f, -2.5 * m + 7.0, 3)
    b = oracle.cc(g, 4.0 * f - 1.0, m, 3)
    for other in (a, b):
        assert abs(base["value"] - other["value"]) < 1e-9
    assert np.max(np.abs(base["derivative"] - a["derivative"])) < 1e-8


def test_cc_derivative_sign_decreases_the_value():
    """Stepping the moving image along the returned derivative lowers the value (ITK sign)."""
    g = cases.grid((16, 16, 16))
    f = cases.image(g, 1, noise=0.0, f32=False)
    m = cases.image(g, 2, noise=0.0, f32=False)
    r = oracle.cc(g, f, m, 3)
    # the derivative wrt a uniform translation t of the moving transform is -sum(d)/N
    d = r["derivative"].sum(axis=0) / r["num_valid"]
    step = 0.05 / np.linalg.norm(d)
    moved = oracle.cc(g, f, m, 3, moving_affine=Affine.make(translation=tuple(step * d)), derivative=False)
    assert moved["value"] < r["value"]


# ------------------------------------------------------------------------- Gaussian -------

def series_i(n: int, t: float) -> float:
    term = (t / 2.0) ** n
    for k in range(n):
        term /= k + 1
    s = 0.0
    for k in range(60):
        s += term
        term *= (t / 2.0) ** 2 / ((k + 1) * (k + 1 + n))
    return s


@pytest.mark.parametrize("variance", [0.5, 1.0, 3.0, 8.0])
def test_gaussian_taps_match_the_bessel_series(variance):
    taps = oracle.gaussian_taps(variance)
    r = len(taps) // 2
    assert abs(taps.sum() - 1.0) < 1e-12
    assert np.allclose(taps, taps[::-1], atol=1e-15)
    raw = np.array([np.exp(-variance) * series_i(n, variance) for n in range(r + 1)])
    one_sided = raw[0] + 2 * raw[1:].sum()
    assert one_sided >= 0.999 - 1e-5
    assert raw[0] + 2 * raw[1:r].sum() < 0.999 + 1e-5
    assert np.max(np.abs(taps[r:] - raw / one_sided)) < 5e-6


def test_variance_three_has_radius_six_and_the_itk_centre_tap():
    taps = oracle.gaussian_taps(3.0)
    assert len(taps) == 13
    assert abs(taps[6] - 0.2433) < 5e-4
    short = oracle.gaussian_taps(3.0, max_width=3)
    assert len(short) == 7 and abs(short.sum() - 1) < 1e-12


def test_bessel_polynomials_match_the_series():
    for y in [0.3, 1.0, 3.0, 3.74, 3.76, 6.0, 12.0]:
        assert abs(oracle.bessel(0, y) - series_i(0, y)) < 2e-7 * max(series_i(0, y), 1)
        assert abs(oracle.bessel(1, y) - series_i(1, y)) < 2e-7 * max(series_i(1, y), 1)
        assert abs(oracle.bessel(3, y) - series_i(3, y)) < 2e-6 * max(series_i(3, y), 1e-3)


def test_smoothing_a_delta_returns_the_kernel_and_pins_the_faces():
    g = cases.grid((21, 21, 21))
    f = np.zeros((g.n, 3))
    c = 10 + 21 * (10 + 21 * 10)
    f[c, 0] = 1.0
    s = oracle.smooth(g, f, 3.0)
    taps = oracle.gaussian_taps(3.0, max_width=21)
    r = len(taps) // 2
    for t in range(len(taps)):
        i = 10 + t - r
        assert abs(s[i + 21 * (10 + 21 * 10), 0] - taps[t] * taps[r] * taps[r]) < 1e-14
    face = ~cases.interior(g)
    assert np.all(s[face] == 0)


def test_smoothing_preserves_constants_and_blends_below_half_variance():
    g = cases.grid((15, 15, 15))
    f = np.tile([2.0, -1.0, 0.5], (g.n, 1))
    mid = 7 + 15 * (7 + 15 * 7)
    assert np.allclose(oracle.smooth(g, f, 3.0)[mid], [2.0, -1.0, 0.5], atol=1e-12)
    assert np.allclose(oracle.smooth(g, f, 0.25)[mid], [2.0, -1.0, 0.5], atol=1e-12)
    assert np.array_equal(oracle.smooth(g, f, 0.0), f)
    # below variance 0.5 the result is w1*smooth + (1-w1)*original: a delta keeps half its peak
    d = np.zeros((g.n, 3))
    d[mid, 0] = 1.0
    taps = oracle.gaussian_taps(0.25, max_width=15)
    r = len(taps) // 2
    assert abs(oracle.smooth(g, d, 0.25)[mid, 0] - (0.5 * taps[r] ** 3 + 0.5)) < 1e-14


# ------------------------------------------------------- compose, invert, resample --------

def test_compose_adds_constants_and_is_zero_outside():
    g = cases.grid((8, 8, 8), (2.0, 2.0, 2.0))
    u = np.tile([1.0, 2.0, 3.0], (g.n, 1))
    w = np.tile([0.5, 0.0, -0.5], (g.n, 1))
    assert np.allclose(oracle.compose(g, u, w), [1.5, 2.0, 2.5])
    far = np.tile([100.0, 0.0, 0.0], (g.n, 1))
    assert np.allclose(oracle.compose(g, u, far), [100.0, 0, 0])


def test_compose_follows_the_warping_field_for_a_linear_displacement():
    g = cases.grid((10, 10, 10))
    i, j, k = cases.coords(g)
    u = np.stack([0.1 * i, 0 * i, 0 * i], axis=1)
    w = np.tile([1.0, 0, 0], (g.n, 1))
    c = oracle.compose(g, u, w)
    idx = 3 + 10 * (4 + 10 * 4)
    assert abs(c[idx, 0] - (1.0 + 0.1 * 4.0)) < 1e-12


def sine_field(g: Grid, amp: float) -> np.ndarray:
    nx, ny, nz = g.shape_xyz
    i, j, k = cases.coords(g)
    wv = np.sin(np.pi * i / (nx - 1)) * np.sin(np.pi * j / (ny - 1)) * np.sin(np.pi * k / (nz - 1))
    return np.stack([amp * wv, -0.7 * amp * wv, 0.5 * amp * wv], axis=1)


def test_inverse_composes_to_the_identity_within_the_filter_tolerance():
    g = cases.grid((24, 24, 24), (1.5, 1.5, 1.5))
    u = sine_field(g, 2.0)
    inv, (it, mean, mx) = oracle.invert(g, u)
    assert 2 <= it <= 20
    res = oracle.compose(g, u, inv)
    inner = cases.interior(g)
    assert np.max(np.linalg.norm(res[inner] / 1.5, axis=1)) < 0.1
    assert np.all(inv[~inner] == 0)
    again, (it2, _, mx2) = oracle.invert(g, u, inv)
    assert it2 == 1 and mx2 < 0.1


def test_inversion_stops_at_the_iteration_limit_and_reports_the_pre_update_error():
    g = cases.grid((12, 12, 12))
    u = sine_field(g, 3.0)
    p = InvertParams(3, 0.0, 0.0, 1)
    _, (it, _, mx3) = oracle.invert(g, u, params=p)
    _, (it1, _, mx1) = oracle.invert(g, u, params=InvertParams(1, 0.0, 0.0, 1))
    assert it == 3 and it1 == 1 and mx1 > mx3
    z, (itz, _, _) = oracle.invert(g, np.zeros((g.n, 3)))
    assert itz == 1 and np.all(z == 0)


def test_resampling_reproduces_a_linear_field_and_is_zero_outside():
    coarse = cases.grid((8, 8, 8), (2.0, 2.0, 2.0))
    i, j, k = cases.coords(coarse)
    f = np.stack([0.1 * 2 * i + 1.0, 0.2 * 2 * j, -0.05 * 2 * k], axis=1)
    fine = cases.grid((13, 13, 13))
    r = oracle.resample(coarse, f, fine)
    idx = 5 + 13 * (6 + 13 * 7)
    assert abs(r[idx, 0] - (0.1 * 5 + 1.0)) < 1e-12 and abs(r[idx, 1] - 0.2 * 6) < 1e-12
    wide = cases.grid((8, 8, 8), (2.0, 2.0, 2.0), (-20.0, 0.0, 0.0))
    assert np.all(oracle.resample(coarse, f, wide)[0 + 8 * (3 + 8 * 3)] == 0)


def test_norms():
    g = cases.grid((5, 4, 3))
    i, _, _ = cases.coords(g)
    f = np.stack([i, 0 * i, 0 * i], axis=1)
    mx, l2 = oracle.norms(g, f)
    assert mx == 4.0 and abs(l2 - np.sqrt(np.sum(i * i))) < 1e-12
    f[17, 0] = np.nan
    assert np.isnan(oracle.norms(g, f)[0])


# ---------------------------------------------------------------------------- warp --------

def test_warp_is_exact_on_an_affine_ramp_and_zero_outside():
    g = cases.grid((8, 7, 6), (1.5, 1.5, 2.0), (-3.0, 2.0, 1.0))
    i, j, k = cases.coords(g)
    ramp = 2 * i + 5 * j - k
    zero = np.zeros((g.n, 3))
    assert np.allclose(oracle.warp(g, ramp, g, zero), ramp, atol=1e-12)
    moved = oracle.warp(g, ramp, g, zero, affine=Affine.make(translation=(3.0, 0, 0)))
    idx = 2 + 8 * (3 + 7 * 2)
    assert abs(moved[idx] - (2 * 4 + 15 - 2)) < 1e-12
    assert moved[7 + 8 * (3 + 7 * 2)] == 0.0


def test_nearest_voxel_rounds_half_up_and_is_zero_outside():
    """mask.rs:99-111 (ITK NearestNeighborInterpolateImageFunction): a point exactly half way
    between voxels takes the upper one; a point whose rounded index leaves the buffer is outside.
    The field grid sits half a voxel along x from the mask's, so every sample is a tie."""
    src = Grid.make((6, 5, 4), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    dst = Grid.make((6, 5, 4), (1.0, 1.0, 1.0), (0.5, 0.0, 0.0))
    i, j, k = cases.coords(src)
    values = 1.0 + i + 10 * j + 100 * k
    zero = np.zeros((dst.n, 3))
    got = oracle.warp(src, values, dst, zero, nearest=True)
    di, dj, dk = cases.coords(dst)
    want = np.where(di < 5, 1.0 + (di + 1) + 10 * dj + 100 * dk, 0.0)
    assert set(np.unique(got)) <= {0.0, 1.0}
    # the warp of a mask reports membership: {0, 1} of the nearest voxel's non-zero value
    assert np.array_equal(got, (want != 0).astype(float))
    # a displacement of just under half a voxel back keeps the lower voxel everywhere
    back = np.tile([-1e-9, 0.0, 0.0], (dst.n, 1))
    below = oracle.warp(src, (values % 2 == 0).astype(float), dst, back, nearest=True)
    assert np.array_equal(below, ((1.0 + di + 10 * dj + 100 * dk) % 2 == 0).astype(float))


def test_warp_through_a_prior_field_follows_both_displacements():
    """cpu/mod.rs:60-86: out(x) = I(q + pre(q)), q = x + u(x), with the prior field on its own grid;
    constant displacements on a linear image give out(x) = I(x + u + pre) exactly where the
    point is inside, and zero where it leaves the image."""
    g = cases.grid((10, 9, 8), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    pre_g = cases.grid((12, 11, 10), (1.0, 1.0, 1.0), (-1.0, -1.0, -1.0))
    i, j, k = cases.coords(g)
    image = 3.0 * i - 2.0 * j + 0.5 * k + 7.0
    u = np.tile([0.75, -0.5, 0.25], (g.n, 1))
    pre = np.tile([1.0, 0.25, -0.5], (pre_g.n, 1))
    got = oracle.warp(g, image, g, u, pre_grid=pre_g, pre=pre)
    x, y, z = i + 1.75, j - 0.25, k - 0.25
    inside = (x >= -0.5) & (x < 9.5) & (y >= -0.5) & (y < 8.5) & (z >= -0.5) & (z < 7.5)
    # the linear interpolation of a linear image is exact inside the grid's voxel centres
    interior = inside & (x <= 9) & (y >= 0) & (z >= 0)
    want = 3.0 * x - 2.0 * y + 0.5 * z + 7.0
    assert np.allclose(got[interior], want[interior], atol=1e-12)
    assert np.all(got[~inside] == 0.0)
    # without the prior the same point moves by u alone
    alone = oracle.warp(g, image, g, u)
    x1 = i + 0.75
    ok = (x1 <= 9) & (j - 0.5 >= 0) & (k + 0.25 <= 7)
    assert np.allclose(alone[ok], (3.0 * x1 - 2.0 * (j - 0.5) + 0.5 * (k + 0.25) + 7.0)[ok], atol=1e-12)


def blob(g: Grid, centre) -> np.ndarray:
    """init.rs:92-105: exp(-|p - c|^2 / 18) over the physical points of the grid."""
    i, j, k = cases.coords(g)
    pts = np.stack([i, j, k], 1) * np.array(list(g.spacing)) + np.array(list(g.origin))
    return np.exp(-np.sum((pts - np.array(centre)) ** 2, axis=1) / 18.0)


def test_centre_of_gravity_of_a_blob_is_its_centre():
    """init.rs:107-140: the intensity centre of a Gaussian blob on an anisotropic grid is the blob's
    centre (1e-3 mm); an image without mass is an error."""
    g = Grid.make((40, 36, 32), (1.0, 1.5, 2.0), (-20.0, -30.0, -10.0))
    c = oracle.centre_of_gravity(g, blob(g, (3.0, -4.0, 5.0)))
    assert np.max(np.abs(c - np.array([3.0, -4.0, 5.0]))) < 1e-3, c
    with pytest.raises(oracle.OracleError):
        oracle.centre_of_gravity(Grid.make((4, 4, 4), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)), np.zeros(64))


# ------------------------------------------------------------------------ sampled MI ------

def smooth_vol(g: Grid, shift, phase=0.0) -> np.ndarray:
    i, j, k = cases.coords(g)
    x, y, z = i - shift[0], j - shift[1], k - shift[2]
    return 1.0 + np.sin(0.35 * x + phase) * np.cos(0.3 * y) + 0.6 * np.sin(0.25 * z + 0.15 * x)


def test_cubic_bspline_partition_of_unity_and_derivative():
    for t in [0.0, 0.1, 0.37, 0.5, 0.99]:
        assert abs(sum(oracle.cubic_bspline(k - t) for k in range(-1, 3)) - 1.0) < 1e-12
        assert abs(sum(oracle.cubic_bspline(k - t, True) for k in range(-1, 3))) < 1e-12
    for u in [-1.7, -1.2, -0.6, -0.1, 0.2, 0.9, 1.4, 1.9]:
        h = 1e-6
        fd = (oracle.cubic_bspline(u + h) - oracle.cubic_bspline(u - h)) / (2 * h)
        assert abs(fd - oracle.cubic_bspline(u, True)) < 1e-6


def fixed_bins(g, f, pts, b):
    from scipy.ndimage import map_coordinates

    vol = f.reshape(g.shape_xyz[::-1])
    vals = map_coordinates(vol, [pts[:, 2], pts[:, 1], pts[:, 0]], order=1)
    t = vals / b.fixed_bin_size - b.fixed_normalized_min
    return np.clip(t.astype(np.int64), 2, b.bins - 3).astype(np.uint32)


@pytest.mark.parametrize("kind", [0, 1])
def test_sampled_mi_derivative_matches_central_differences(kind):
    """metric/mi.rs:973-1065: the analytic parameter derivative against -(v(+h) - v(-h)) / 2h."""
    g = cases.grid((26, 26, 26))
    f = smooth_vol(g, (0, 0, 0))
    m = smooth_vol(g, (1.1, -0.8, 0.6))
    rng = np.random.default_rng(3)
    pts = rng.uniform(6.0, 19.0, size=(3000, 3))
    b = MiBinning.from_ranges(32, (f.min(), f.max()), (m.min(), m.max()))
    fb = fixed_bins(g, f, pts, b)
    if kind == 0:
        params = np.array([0.01, -0.02, 0.015, 0.4, -0.3, 0.2])
        outer = Affine.make()
    else:
        params = np.array([1.0, 0.01, 0.0, -0.01, 1.0, 0.0, 0.0, 0.0, 1.0, 0.3, -0.2, 0.1])
        outer = Affine.make(translation=(0.1, 0.05, -0.1))
    t = LinearTransform.make(kind, params, (13.0, 13.0, 13.0), outer)
    base = oracle.mi_sampled(b, g, m, pts, fb, t)
    fd = np.zeros(len(params))
    for q in range(len(params)):
        h = 1e-4 if (kind == 0 and q < 3) else 1e-3
        up, dn = params.copy(), params.copy()
        up[q] += h
        dn[q] -= h
        vu = oracle.mi_sampled(b, g, m, pts, fb, LinearTransform.make(kind, up, (13.0, 13.0, 13.0), outer), False)
        vd = oracle.mi_sampled(b, g, m, pts, fb, LinearTransform.make(kind, dn, (13.0, 13.0, 13.0), outer), False)
        fd[q] = -(vu["value"] - vd["value"]) / (2 * h)
    assert np.max(np.abs(fd)) > 1e-4
    assert np.max(np.abs(base["derivative"] - fd)) <= 0.08 * np.max(np.abs(fd))


def test_sampled_mi_pdf_is_normalised_and_value_is_minus_mi():
    g = cases.grid((14, 14, 14))
    f = smooth_vol(g, (0, 0, 0))
    m = smooth_vol(g, (0.7, -0.4, 0.2), 0.2)
    rng = np.random.default_rng(5)
    pts = rng.uniform(0.0, 13.0, size=(700, 3))
    b = MiBinning.from_ranges(32, (f.min(), f.max()), (m.min(), m.max()))
    fb = fixed_bins(g, f, pts, b)
    r = oracle.mi_sampled(b, g, m, pts, fb, LinearTransform.make(0, np.zeros(6)))
    p = r["joint_pdf"].reshape(32, 32)
    assert abs(p.sum() - 1) < 1e-12
    pf, pm = p.sum(axis=1), p.sum(axis=0)
    nz = p > 1e-12
    mi = np.sum(p[nz] * np.log(p[nz] / np.outer(pf, pm)[nz]))
    assert mi > 0 and abs(r["value"] + mi) < 1e-9


# ------------------------------------------------------------------ isotropic resample (register.rs) --


def test_isotropic_resample_follows_the_pipeline_grid_rule():
    """register.rs:447-459: 21 voxels at 0.5 mm become 11 at 1 mm, same origin, voxel (3, 4, 5)
    takes source voxel (6, 8, 10), and a grid already at 1 mm is returned as it is."""
    g = Grid.make((21, 21, 21), (0.5, 0.5, 0.5), (-3.0, 2.0, 1.5))
    v = np.random.default_rng(4).random(g.n)
    og, r = oracle.resample_isotropic(g, v, 1.0)
    assert list(og.size) == [11, 11, 11] and list(og.spacing) == [1.0, 1.0, 1.0]
    assert list(og.origin) == list(g.origin)
    assert abs(r[3 + 11 * (4 + 11 * 5)] - v[6 + 21 * (8 + 21 * 10)]) < 1e-6
    og2, same = oracle.resample_isotropic(og, r, 1.0)
    assert list(og2.size) == [11, 11, 11] and np.array_equal(same, r)


def test_isotropic_resample_interpolates_between_source_voxels():
    """register.rs:462-471: [0, 3, 6, 9] at 1.5 mm resampled to 1 mm is [0, 2, 4, 6, 8]."""
    g = Grid.make((4, 1, 1), (1.5, 1.0, 1.0))
    og, r = oracle.resample_isotropic(g, np.array([0.0, 3.0, 6.0, 9.0]), 1.0)
    assert list(og.size) == [5, 1, 1]
    assert np.max(np.abs(r - [0.0, 2.0, 4.0, 6.0, 8.0])) < 1e-6


@pytest.mark.parametrize("spacing,mm", [((1.5, 0.8, 2.0), 1.0), ((0.7, 0.7, 1.3), 1.2), ((1.0, 2.0, 3.0), 2.5)])
def test_isotropic_resample_reproduces_a_linear_function(spacing, mm):
    """Trilinear interpolation is exact on a linear function: the resampled volume is the function
    at the new voxels' source indices (clamped at the last voxel), rounded to f32."""
    size = (13, 9, 7)
    g = Grid.make(size, spacing, (4.0, -2.0, 1.0))
    i, j, k = np.meshgrid(*[np.arange(n, dtype=float) for n in size], indexing="ij")
    f = lambda a, b, c: 0.3 * a - 1.7 * b + 0.45 * c + 2.0
    v = f(i, j, k).transpose(2, 1, 0).reshape(-1)  # x-fastest
    og, r = oracle.resample_isotropic(g, v, mm)
    idx = [np.minimum(np.arange(og.size[a]) * (mm / spacing[a]), size[a] - 1) for a in range(3)]
    for a in range(3):
        assert og.size[a] == int(np.floor((size[a] - 1) * spacing[a] / mm + 1e-9)) + 1
    ci, cj, ck = np.meshgrid(*idx, indexing="ij")
    want = f(ci, cj, ck).transpose(2, 1, 0).reshape(-1).astype(np.float32).astype(np.float64)
    assert np.max(np.abs(r - want)) <= 2e-6 * np.max(np.abs(want))


def test_isotropic_resample_rejects_a_non_positive_spacing():
    g = Grid.make((4, 4, 4))
    with pytest.raises(oracle.OracleError, match="positive"):
        oracle.resample_isotropic(g, np.zeros(g.n), 0.0)
