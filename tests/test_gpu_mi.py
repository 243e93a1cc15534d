"""GPU parity of the Mattes MI kernels against the f64 oracle: the linear-stage sampled metric
(value and parameter derivative) and the dense SyN pair (values and per-voxel derivatives)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from fastwarp_amd._abi import Affine, Grid, LinearTransform, MiBinning
from fastwarp_amd.backend import MiLevel
from tests import cases

pytestmark = pytest.mark.gpu


def smooth_vol(g: Grid, shift, phase=0.0) -> np.ndarray:
    i, j, k = cases.coords(g)
    x, y, z = i - shift[0], j - shift[1], k - shift[2]
    v = 1.0 + np.sin(0.35 * x + phase) * np.cos(0.3 * y) + 0.6 * np.sin(0.25 * z + 0.15 * x)
    return v.astype(np.float32).astype(np.float64)


def fixed_bins(g, f, pts, b):
    """prepare_samples' fixed bin (linear.rs:251-279) from the oracle's own interpolation."""
    n = pts.shape[0]
    # a 1-D carrier grid whose voxel i sits at (i, 0, 0): displacing it by pts - (i, 0, 0) samples f at pts
    carrier = np.stack([np.arange(n, dtype=float), np.zeros(n), np.zeros(n)], axis=1)
    vals = oracle.warp(g, f, Grid.make((n, 1, 1)), pts - carrier)
    t = vals / b.fixed_bin_size - b.fixed_normalized_min
    return np.clip(t.astype(np.int64), 2, b.bins - 3).astype(np.uint32)


CASES = [
    ("rigid-identity", 0, np.zeros(6), Affine.make()),
    ("rigid-moved", 0, np.array([0.03, -0.02, 0.05, 1.4, -0.9, 0.6]), Affine.make(translation=(0.3, -0.2, 0.1))),
    ("affine", 1, np.array([1.02, 0.01, 0.0, -0.01, 0.97, 0.02, 0.0, 0.01, 1.01, 0.5, -0.3, 0.2]), Affine.make(translation=(0.1, 0.05, -0.1))),
]


@pytest.mark.parametrize("name,kind,params,outer", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("oblique", [0.0, 0.3])
def test_sampled_mi_matches_oracle(backend, name, kind, params, outer, oblique):
    g = cases.grid((40, 36, 30), (1.2, 1.0, 1.4), (-20.0, -15.0, -18.0), oblique=oblique)
    f = smooth_vol(g, (0, 0, 0))
    m = smooth_vol(g, (1.1, -0.8, 0.6), 0.3)
    rng = np.random.default_rng(7)
    idx = rng.uniform([4, 4, 4], np.array(g.shape_xyz) - 5, size=(6000, 3))
    pts = np.array([g.index_to_point(p) for p in idx])
    b = MiBinning.from_ranges(32, (f.min(), f.max()), (m.min(), m.max()))
    fb = fixed_bins(g, f, pts, b)
    centre = tuple(g.index_to_point(np.array(g.shape_xyz) / 2))
    t = LinearTransform.make(kind, params, centre, outer)
    want = oracle.mi_sampled(b, g, m, pts, fb, t)
    level = MiLevel(backend, backend.upload_image(g, m), pts, fb, b)
    v = level.value(t)
    v2, d = level.value_and_derivative(t)
    assert v == v2
    assert abs(v - want["value"]) <= 1e-5 * abs(want["value"])
    rel = np.linalg.norm(d - want["derivative"]) / np.linalg.norm(want["derivative"])
    assert rel <= 5e-4, (d, want["derivative"])


@pytest.mark.parametrize("bins", [32, 12, 64])
def test_batched_values_equal_single_values(backend, bins):
    """One pass over the samples for up to four transforms (the line search's speculative probes):
    the integer histograms make every batched value bit-identical to its single evaluation."""
    g = cases.grid((40, 36, 30), (1.2, 1.0, 1.4), (-20.0, -15.0, -18.0), oblique=0.3)
    f = smooth_vol(g, (0, 0, 0))
    m = smooth_vol(g, (1.1, -0.8, 0.6), 0.3)
    rng = np.random.default_rng(11)
    pts = np.array([g.index_to_point(p) for p in rng.uniform([4, 4, 4], np.array(g.shape_xyz) - 5, size=(9000, 3))])
    b = MiBinning.from_ranges(bins, (f.min(), f.max()), (m.min(), m.max()))
    fb = fixed_bins(g, f, pts, b)
    level = MiLevel(backend, backend.upload_image(g, m), pts, fb, b)
    centre = tuple(g.index_to_point(np.array(g.shape_xyz) / 2))
    ts = []
    for q in range(7):
        if q % 3 == 2:
            p = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]) + rng.normal(0, 0.02, 12)
            ts.append(LinearTransform.make(1, p, centre, Affine.make(translation=(0.1, 0.0, -0.1))))
        else:
            ts.append(LinearTransform.make(0, np.concatenate([rng.normal(0, 0.04, 3), rng.normal(0, 1.5, 3)]), centre))
    single = np.array([level.value(t) for t in ts])
    for n in range(1, 8):
        assert np.array_equal(level.values(ts[:n]), single[:n]), n
    from fastwarp_amd.backend import RegError

    away = LinearTransform.make(0, [0, 0, 0, 500.0, 0, 0], centre)
    with pytest.raises(RegError, match="outside the moving image buffer"):
        level.values([ts[0], away, ts[1]])


@pytest.mark.parametrize("bins", [32, 12, 64])
def test_device_finish_matches_the_f64_finish_of_the_same_cells(backend, bins):
    """The linear stage finishes its histograms on the device (k_mi_finish). On the integer cells
    of the same pass, the CPU engine's finish (mi.rs:702-769, f64) gives the same value to the
    rounding of another summation order (1e-12 relative); the value with the derivative, whose pass
    reads the device's ratio weights, is the same value."""
    g = cases.grid((40, 36, 30), (1.2, 1.0, 1.4), (-20.0, -15.0, -18.0), oblique=0.3)
    f = smooth_vol(g, (0, 0, 0))
    m = smooth_vol(g, (1.1, -0.8, 0.6), 0.3)
    rng = np.random.default_rng(5)
    pts = np.array([g.index_to_point(p) for p in rng.uniform([4, 4, 4], np.array(g.shape_xyz) - 5, size=(12000, 3))])
    b = MiBinning.from_ranges(bins, (f.min(), f.max()), (m.min(), m.max()))
    level = MiLevel(backend, backend.upload_image(g, m), pts, fixed_bins(g, f, pts, b), b)
    centre = tuple(g.index_to_point(np.array(g.shape_xyz) / 2))
    for q in range(4):
        t = LinearTransform.make(0, np.concatenate([rng.normal(0, 0.04, 3), rng.normal(0, 1.5, 3)]), centre)
        joint, marginal, valid = level.histogram(t, bins)
        want, _ = oracle.mi_finish(bins, b.moving_bin_size, valid, joint, marginal)
        got = level.value(t)
        assert abs(got - want) <= 1e-12 * abs(want), (q, got, want)
        assert level.value_and_derivative(t)[0] == got, q


@pytest.mark.parametrize("bins", [32, 12])
def test_fused_slot_sum_and_finish_give_the_bits_of_the_two_launches(backend, bins):
    """Passes of few blocks add their slots and finish them in one launch (k_mi_sum_finish); the
    cells are the same integers as k_mi_sum_slots adds, so values (single and batched) and
    derivatives are bit-identical with the fusion switched off (tuning knob mi.fused_slots = 0)."""
    g = cases.grid((40, 36, 30), (1.2, 1.0, 1.4), (-20.0, -15.0, -18.0), oblique=0.3)
    f = smooth_vol(g, (0, 0, 0))
    m = smooth_vol(g, (1.1, -0.8, 0.6), 0.3)
    rng = np.random.default_rng(9)
    pts = np.array([g.index_to_point(p) for p in rng.uniform([4, 4, 4], np.array(g.shape_xyz) - 5, size=(5000, 3))])
    b = MiBinning.from_ranges(bins, (f.min(), f.max()), (m.min(), m.max()))
    level = MiLevel(backend, backend.upload_image(g, m), pts, fixed_bins(g, f, pts, b), b)
    centre = tuple(g.index_to_point(np.array(g.shape_xyz) / 2))
    ts = [LinearTransform.make(0, np.concatenate([rng.normal(0, 0.04, 3), rng.normal(0, 1.5, 3)]), centre) for _ in range(5)]
    ts.append(LinearTransform.make(1, np.array([1.01, 0.02, 0.0, -0.01, 0.98, 0.01, 0.0, 0.02, 1.0, 0.4, -0.2, 0.3]), centre))
    runs = []
    for fused in (64, 0):
        backend.set_tuning("mi.fused_slots", fused)
        try:
            runs.append(([level.value(t) for t in ts], list(level.values(ts[:4])), [level.value_and_derivative(t) for t in ts]))
        finally:
            backend.set_tuning("mi.fused_slots", -1)
    (v1, b1, d1), (v0, b0, d0) = runs
    assert v1 == v0 and b1 == b0
    for (x, dx), (y, dy) in zip(d1, d0):
        assert x == y and np.array_equal(dx, dy)


def test_sampled_mi_is_deterministic_and_errors_without_overlap(backend):
    g = cases.grid((24, 24, 24))
    f = smooth_vol(g, (0, 0, 0))
    pts = np.array([g.index_to_point(p) for p in np.random.default_rng(1).uniform(2, 21, size=(3000, 3))])
    b = MiBinning.from_ranges(32, (f.min(), f.max()), (f.min(), f.max()))
    fb = fixed_bins(g, f, pts, b)
    level = MiLevel(backend, backend.upload_image(g, f), pts, fb, b)
    t = LinearTransform.make(0, [0.01, 0, 0, 0.5, 0, 0], (12, 12, 12))
    a = level.value_and_derivative(t)
    c = level.value_and_derivative(t)
    assert a[0] == c[0] and np.array_equal(a[1], c[1])
    from fastwarp_amd.backend import RegError

    with pytest.raises(RegError, match="outside the moving image buffer"):
        level.value(LinearTransform.make(0, [0, 0, 0, 500.0, 0, 0], (12, 12, 12)))


@pytest.mark.parametrize("shape,oblique,masks", [((23, 19, 17), 0.35, False), ((48, 52, 40), 0.0, False), ((30, 28, 26), 0.2, True)])
@pytest.mark.parametrize("bins", [32, 12])
def test_dense_mi_pair_matches_oracle(backend, shape, oblique, masks, bins):
    g = cases.grid(shape, (1.1, 0.9, 1.3), (-12.5, 7.25, -4.0), oblique=oblique)
    f = cases.image(g, 3)
    m = oracle.warp(g, f, g, cases.field(g, 1.0, 4))
    m = (0.6 * m + 0.4 * cases.image(g, 9)) ** 1.3
    m = m.astype(np.float32).astype(np.float64)
    fm = mm = None
    if masks:
        i, j, k = cases.coords(g)
        fm = (i + j > 6).astype(float)
        mm = (k + i < shape[0] + shape[2] - 8).astype(float)
    a = oracle.mi_dense(g, f, g, m, bins, fixed_mask=fm, moving_mask=mm)
    b = oracle.mi_dense(g, m, g, f, bins, fixed_mask=mm, moving_mask=fm)
    up = backend.upload_image
    (v0, d0), (v1, d1) = backend.mi_value_derivative_pair_masked(
        up(g, f), up(g, m), up(g, fm) if fm is not None else None, up(g, mm) if mm is not None else None, bins
    )
    assert abs(v0 - a["value"]) <= 1e-5 * abs(a["value"])
    assert abs(v1 - b["value"]) <= 1e-5 * abs(b["value"])
    inner = cases.interior(g)
    assert cases.rel_err(backend.download_field(d0)[inner], a["derivative"][inner]) <= 1e-4
    assert cases.rel_err(backend.download_field(d1)[inner], b["derivative"][inner]) <= 1e-4
