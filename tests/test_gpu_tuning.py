"""The tuning table on the device: fw_set_tuning rejects unknown knobs and out-of-range values, and
every launch shape the autotuner may pick gives the default shape's results, exactly where the
arithmetic is integer (the MI histogram cells) and to rounding elsewhere (the CC z runs, the
inversion's block reductions, the one-pass or separable Gaussian).
"""
from __future__ import annotations

import numpy as np
import pytest

from fastwarp_amd._abi import InvertParams, LinearTransform, MiBinning
from fastwarp_amd.autotune import CANDIDATES, size_class
from fastwarp_amd.backend import HipBackend, MiLevel, RegError
from tests import cases
from tests.test_gpu_mi import fixed_bins, smooth_vol

pytestmark = pytest.mark.gpu


@pytest.fixture()
def tuned():
    b = HipBackend(0)
    # start from the library's defaults whatever table the context loaded
    for cands in CANDIDATES.values():
        for cand in cands:
            for q in range(0, len(cand), 2):
                b.set_tuning(cand[q], -1)
    yield b
    b.close()


def knobs(b: HipBackend, cls: int, cand, on: bool = True):
    for q in range(0, len(cand), 2):
        b.set_tuning(f"{cand[q]}@{cls}", cand[q + 1] if on else -1)


def test_unknown_knobs_and_bad_values_are_rejected(tuned):
    for key, value in (("cc.tile_z", 0), ("mi.threads", 384), ("gauss.one_pass", 2), ("no.such_knob", 1)):
        with pytest.raises(RegError, match="tuning"):
            tuned.set_tuning(key, value)
    tuned.set_tuning("no.such_knob", -1)  # removing what is not there is no error


def test_cc_z_runs_change_results_only_by_rounding(tuned):
    g = cases.grid((91, 109, 91), (2.0, 2.0, 2.0))
    f = tuned.upload_image(g, cases.image(g, 1))
    m = tuned.upload_image(g, cases.image(g, 2))
    (v0, d0), (w0, e0) = tuned.cc_value_derivative_pair(f, m, 4)
    a0, b0 = tuned.download_field(d0), tuned.download_field(e0)
    rng = np.max(np.abs(a0))
    for cand in CANDIDATES["cc"]:
        knobs(tuned, size_class(g.n), cand)
        (v, d), (w, e) = tuned.cc_value_derivative_pair(f, m, 4)
        knobs(tuned, size_class(g.n), cand, False)
        assert abs(v - v0) <= 1e-6 * abs(v0) and abs(w - w0) <= 1e-6 * abs(w0), (cand, v, v0)
        assert np.max(np.abs(tuned.download_field(d) - a0)) <= 1e-6 * rng, cand
        assert np.max(np.abs(tuned.download_field(e) - b0)) <= 1e-6 * rng, cand


def test_mi_launch_shapes_give_the_same_histogram_cells(tuned):
    g = cases.grid((40, 36, 30), (1.2, 1.0, 1.4), (-20.0, -15.0, -18.0), oblique=0.3)
    f = smooth_vol(g, (0, 0, 0))
    m = smooth_vol(g, (1.1, -0.8, 0.6), 0.3)
    pts = np.array([g.index_to_point(p) for p in np.random.default_rng(4).uniform([4, 4, 4], np.array(g.shape_xyz) - 5, size=(150000, 3))])
    binning = MiBinning.from_ranges(32, (f.min(), f.max()), (m.min(), m.max()))
    level = MiLevel(tuned, tuned.upload_image(g, m), pts, fixed_bins(g, f, pts, binning), binning)
    t = LinearTransform.make(0, [0.02, -0.01, 0.03, 0.8, -0.5, 0.4], tuple(g.index_to_point(np.array(g.shape_xyz) / 2)))
    j0, m0, n0 = level.histogram(t, 32)
    v0, d0 = level.value_and_derivative(t)
    cls = size_class(len(pts))
    for cand in CANDIDATES["mi"]:
        knobs(tuned, cls, cand)
        j, mm, n = level.histogram(t, 32)
        knobs(tuned, cls, cand, False)
        assert n == n0 and np.array_equal(j, j0) and np.array_equal(mm, m0), cand
    for cand in CANDIDATES["mi_deriv"]:
        knobs(tuned, cls, cand)
        v, d = level.value_and_derivative(t)
        knobs(tuned, cls, cand, False)
        assert v == v0, cand
        assert np.max(np.abs(np.asarray(d) - np.asarray(d0))) <= 1e-9 * np.max(np.abs(d0)), cand
    # the batched passes' block cap: the same integer cells, so the same values as single passes
    centre = tuple(g.index_to_point(np.array(g.shape_xyz) / 2))
    probes = [LinearTransform.make(0, [0.02 * q, -0.01, 0.03, 0.8, -0.5 * q, 0.4], centre) for q in range(4)]
    single = [level.value(p) for p in probes]
    for cand in CANDIDATES["mi_batch"]:
        knobs(tuned, cls, cand)
        batched = list(level.values(probes))
        knobs(tuned, cls, cand, False)
        assert batched == single, cand


def test_inversion_blocks_and_gaussian_passes_change_results_only_by_rounding(tuned):
    g = cases.grid((64, 72, 60), (2.0, 2.0, 2.0))
    i, j, k = cases.coords(g)
    u = np.stack([np.sin(0.11 * i + 0.05 * j), np.cos(0.07 * j - 0.04 * k), np.sin(0.09 * k + 0.03 * i)], 1) * 1.5
    du = tuned.upload_field(g, u)
    params = InvertParams(5, 0.0, 0.0, 1)
    want, _ = tuned.invert(du, tuned.zero_field(g), params)
    want = tuned.download_field(want)
    cls = size_class(g.n)
    for cand in CANDIDATES["inv"]:
        knobs(tuned, cls, cand)
        got, _ = tuned.invert(du, tuned.zero_field(g), params)
        knobs(tuned, cls, cand, False)
        assert np.max(np.abs(tuned.download_field(got) - want)) <= 1e-5, cand
    smoothed = {}
    for cand in CANDIDATES["gauss"]:
        knobs(tuned, cls, cand)
        smoothed[cand] = tuned.download_field(tuned.smooth_field(tuned.copy_field(du), 3.0))
        knobs(tuned, cls, cand, False)
    a, b = smoothed.values()
    assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(a))
