"""CC pair precision on the inputs where fixed-point or f32 window sums fail: the pyramid levels of
the phantom registration (whose zero-padded corners leave windows of tiny values with the largest
coefficients of the image), the 1 mm grid, and images on an intensity offset
(crates/fastwarp-gpu/tests/cc_precision.rs:115-170, backgrounds 100 and 1000). North-star bound:
the single-kernel CC gradient within 1e-5 of the FP64 evaluation, as the largest absolute error
over the largest reference magnitude; values within 1e-5 relative."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from fastwarp_amd._abi import Grid
from fastwarp_amd.phantom import make_pair
from tests import cases

pytestmark = pytest.mark.gpu


def level_images(shape, spacing, seed, shrink, sigma):
    """The preprocessed phantom pair smoothed by `sigma` voxels and sampled on the shrunken level
    grid (the images a SyN level's first iteration compares), rounded to f32."""
    pair = make_pair(shape, spacing, seed=seed)
    g = pair.grid
    f = oracle.preprocess(g, pair.fixed[0])
    m = oracle.preprocess(g, pair.moving[0])
    if shrink == 1:
        return g, f.astype(np.float32).astype(np.float64), m.astype(np.float32).astype(np.float64)
    nx, ny, nz = g.shape_xyz
    lg = Grid.make((nx // shrink, ny // shrink, nz // shrink), tuple(s * shrink for s in g.spacing),
                   tuple(o + (shrink - 1) * s / 2 for o, s in zip(g.origin, g.spacing)))
    sig = [sigma * s for s in g.spacing]
    zero = np.zeros((lg.n, 3))
    fs = oracle.warp(g, oracle.smooth_image(g, f, sig), lg, zero)
    ms = oracle.warp(g, oracle.smooth_image(g, m, sig), lg, zero)
    return lg, fs.astype(np.float32).astype(np.float64), ms.astype(np.float32).astype(np.float64)


def check_pair(backend, g, f, m, radius, valid=None, bound=1e-5):
    a = oracle.cc(g, f, m, radius, mask=valid)
    b = oracle.cc(g, m, f, radius, mask=valid)
    df, dm = backend.upload_image(g, f), backend.upload_image(g, m)
    if valid is None:
        (v0, d0), (v1, d1) = backend.cc_value_derivative_pair(df, dm, radius)
    else:
        (v0, d0), (v1, d1) = backend.cc_value_derivative_pair_masked(df, dm, backend.upload_image(g, valid), radius)
    g0, g1 = backend.download_field(d0), backend.download_field(d1)
    report = {"value": (abs(v0 - a["value"]) / abs(a["value"]), abs(v1 - b["value"]) / abs(b["value"])),
              "derivative": (cases.rel_err(g0, a["derivative"]), cases.rel_err(g1, b["derivative"]))}
    assert max(report["value"]) <= 1e-5, report
    assert max(report["derivative"]) <= bound, report
    return report


@pytest.mark.parametrize("shrink,sigma", [(4, 2.0), (2, 1.0), (1, 0.0)])
def test_cc_pair_on_the_2mm_phantom_levels(backend, shrink, sigma):
    g, f, m = level_images((96, 112, 96), 2.0, 1, shrink, sigma)
    check_pair(backend, g, f, m, 4)


def test_cc_pair_on_the_1mm_phantom(backend):
    g, f, m = level_images((182, 218, 182), 1.0, 2, 1, 0.0)
    check_pair(backend, g, f, m, 4)


def offset_image(g: Grid, seed: int, shift: float, background: float) -> np.ndarray:
    """cc_precision.rs:36-60: a textured ball on a flat background, intensities background + [0.05, 1]."""
    i, j, k = cases.coords(g)
    nx, ny, nz = g.shape_xyz
    r2 = (((i - 0.5 * (nx - 1) - shift) / (0.34 * nx)) ** 2 + ((j - 0.5 * (ny - 1)) / (0.34 * ny)) ** 2
          + ((k - 0.5 * (nz - 1)) / (0.34 * nz)) ** 2)
    base = cases.image(g, seed, noise=0.2, f32=False)
    v = np.where(r2 < 1.0, background + 0.05 + 0.95 * base, background)
    return v.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("background,moving_background,radius", [(100.0, 30.0, 2), (1000.0, 300.0, 4)])
def test_cc_pair_with_an_intensity_offset(backend, background, moving_background, radius):
    g = cases.grid((32, 32, 32))
    f = offset_image(g, 3, 0.0, background)
    m = offset_image(g, 4, 1.3, moving_background)
    check_pair(backend, g, f, m, radius)
    i, _, _ = cases.coords(g)
    check_pair(backend, g, f, m, radius, valid=(i >= 5).astype(float))
