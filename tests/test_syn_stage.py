"""One SyN stage with masks and the fields of earlier stages (crates/fastwarp-reg/src/syn.rs
run_syn_from with SynStart::prior and the recipe's masks; fw_syn_stage):

* masks of ones leave a run unchanged, bit for bit, and half masks change a mixed CC + Mattes run
  that still follows the shift (syn.rs:1240-1298, the reference's own test, on the CPU engine and
  on the HIP path);
* a stage with no iterations returns the earlier stages' fields exactly (the collapse of
  syn.rs:746-774 composes them with zero fields), and earlier fields off the last level's grid are
  refused (syn.rs:804-817);
* the CPU engine's own SyN loop (oracle/syn.hpp) and the product's generic loop over the same ops
  give the same bits;
* the HIP path against the CPU engine with masks and mixed metrics, and over two stages, to the
  accuracy bounds of whole registrations (fields mean 0.1, p99 0.5 voxel; energies 1e-4).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from fastwarp_amd._abi import Affine, Grid
from fastwarp_amd.register import cohort_plan


def blob(grid: Grid, shift_mm: float) -> np.ndarray:
    """A smooth ellipsoidal blob centred `shift_mm` along x from the grid centre."""
    nx, ny, nz = grid.shape_xyz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    sp = grid.spacing
    x = (i - (nx - 1) / 2) * sp[0] - shift_mm
    y = (j - (ny - 1) / 2) * sp[1]
    z = (k - (nz - 1) / 2) * sp[2]
    return np.exp(-(x**2 / 40.0 + y**2 / 30.0 + z**2 / 25.0)).reshape(-1)


def inverted(v: np.ndarray) -> np.ndarray:
    return 2.0 - 1.5 * v


def stage_plan(metrics, iterations=(20, 10)):
    """syn.rs:1219-1229 `mattes_recipe`: step 0.2, 20/10 iterations at shrink 2/1, sigmas 1/0."""
    p = cohort_plan(len(metrics))
    levels = len(iterations)
    p.set_syn(list(iterations), [2, 1][-levels:], [1.0, 0.0][-levels:])
    p.syn_gradient_step = 0.2
    for c, m in enumerate(metrics):
        p.syn_channel_metrics[c] = m
        p.channel_weights[c] = 1.0 / len(metrics)
    return p


GRID = Grid.make((12, 12, 11), (2.0, 2.0, 2.0), (0.0, 0.0, 0.0))
CC, MATTES = 0, 1


def channels():
    f = blob(GRID, 0.0)
    return [f, f], [blob(GRID, 1.5), inverted(blob(GRID, 1.5))]


def half_mask() -> np.ndarray:
    nx, ny, nz = GRID.shape_xyz
    k = np.repeat(np.arange(nz), nx * ny)
    return (k < 8).astype(float)


def centre_shift(forward: np.ndarray) -> float:
    """x displacement at index (5.5, 5.5, 5.0), trilinear (the reference's centre_shift)."""
    nx, ny, _ = GRID.shape_xyz
    v = forward.reshape(GRID.shape_xyz[2], ny, nx, 3)
    return float(v[5, 5:7, 5:7, 0].mean())


def test_masks_of_ones_leave_a_run_unchanged_on_the_cpu_engine():
    fixed, moving = channels()
    ones = (GRID, np.ones(GRID.n))
    for metric, c in ((MATTES, 1), (CC, 0)):
        plan = stage_plan([metric])
        _, pf, pi = oracle.syn_stage(plan, GRID, [fixed[c]], GRID, [moving[c]])
        for fm, mm in ((ones, None), (None, ones), (ones, ones)):
            _, f, i = oracle.syn_stage(plan, GRID, [fixed[c]], GRID, [moving[c]], fixed_mask=fm, moving_mask=mm)
            assert np.array_equal(f, pf) and np.array_equal(i, pi), (metric, fm is None, mm is None)


def test_half_masks_change_a_mixed_run_that_still_follows_the_shift():
    fixed, moving = channels()
    plan = stage_plan([CC, MATTES])
    _, open_f, _ = oracle.syn_stage(plan, GRID, fixed, GRID, moving)
    half = (GRID, half_mask())
    _, closed_f, _ = oracle.syn_stage(plan, GRID, fixed, GRID, moving, fixed_mask=half, moving_mask=half)
    for f in (open_f, closed_f):
        assert np.all(np.isfinite(f))
        assert abs(centre_shift(f) - 1.5) < 1.0, centre_shift(f)
    assert not np.array_equal(open_f, closed_f)


def test_a_stage_without_iterations_returns_the_earlier_fields():
    fixed, moving = channels()
    _, f1, i1 = oracle.syn_stage(stage_plan([CC]), GRID, [fixed[0]], GRID, [moving[0]])
    _, f2, i2 = oracle.syn_stage(stage_plan([CC], iterations=(0, 0)), GRID, [fixed[0]], GRID, [moving[0]],
                                 prior=(GRID, f1, i1))
    assert np.array_equal(f2, f1) and np.array_equal(i2, i1)


def test_earlier_fields_off_the_last_level_grid_are_refused():
    fixed, moving = channels()
    coarse = Grid.make((6, 6, 5), (4.0, 4.0, 4.0), (1.0, 1.0, 1.0))
    zero = np.zeros((coarse.n, 3))
    with pytest.raises(RuntimeError, match="same shrink factor"):
        oracle.syn_stage(stage_plan([CC]), GRID, [fixed[0]], GRID, [moving[0]], prior=(coarse, zero, zero))


def test_the_cpu_engines_own_syn_loop_gives_the_bits_of_the_generic_loop():
    """The CPU engine runs its own restatement of syn.rs (oracle/syn.hpp); the product's generic
    SyN loop (which defers read-backs and pairs operations for the device) over the same f64 ops
    must give the same bits: one CC channel, CC + Mattes with half masks, and a second stage
    after a first."""
    fixed, moving = channels()
    half = (GRID, half_mask())
    _, f1, i1 = oracle.syn_stage(stage_plan([CC]), GRID, [fixed[0]], GRID, [moving[0]])
    cases = [
        (stage_plan([CC]), [fixed[0]], [moving[0]], {}),
        (stage_plan([CC, MATTES]), fixed, moving, {"fixed_mask": half, "moving_mask": half}),
        (stage_plan([CC]), [fixed[0]], [moving[0]], {"prior": (GRID, f1, i1)}),
    ]
    for k, (plan, f, m, kw) in enumerate(cases):
        own = oracle.syn_stage(plan, GRID, f, GRID, m, **kw)
        own_trace = [(t.value, tuple(t.inversion_steps)) for t in oracle.last_trace()]
        oracle.generic_syn(True)
        try:
            generic = oracle.syn_stage(plan, GRID, f, GRID, m, **kw)
            generic_trace = [(t.value, tuple(t.inversion_steps)) for t in oracle.last_trace()]
        finally:
            oracle.generic_syn(False)
        assert own_trace == generic_trace, k
        assert np.array_equal(own[1], generic[1]) and np.array_equal(own[2], generic[2]), k


# ---------------------------------------------------------------------------- the HIP path --


def hip_stage(backend, plan, fixed, moving, fixed_mask=None, moving_mask=None, prior=None):
    from fastwarp_amd.register import syn_stage

    up = lambda v: backend.upload_image(GRID, v)
    r = syn_stage(backend, plan, [up(v) for v in fixed], [up(v) for v in moving],
                  fixed_mask=up(fixed_mask) if fixed_mask is not None else None,
                  moving_mask=up(moving_mask) if moving_mask is not None else None,
                  prior=prior[0] if prior else None, prior_inverse=prior[1] if prior else None)
    return r


def fields_report(got_f, got_i, want_f, want_i):
    sp = GRID.spacing[0]
    out = {}
    for name, g, w in (("forward", got_f, want_f), ("inverse", got_i, want_i)):
        d = np.linalg.norm(g - w, axis=1) / sp
        out[name] = (float(d.mean()), float(np.percentile(d, 99)))
    return out


@pytest.mark.gpu
def test_masks_of_ones_leave_a_run_unchanged_on_the_device(backend):
    fixed, moving = channels()
    ones = np.ones(GRID.n)
    for metric, c in ((MATTES, 1), (CC, 0)):
        plan = stage_plan([metric])
        plain = hip_stage(backend, plan, [fixed[c]], [moving[c]])
        pf, pi = backend.download_field(plain.forward), backend.download_field(plain.inverse)
        for fm, mm in ((ones, None), (None, ones), (ones, ones)):
            r = hip_stage(backend, plan, [fixed[c]], [moving[c]], fm, mm)
            f, i = backend.download_field(r.forward), backend.download_field(r.inverse)
            assert np.array_equal(f, pf) and np.array_equal(i, pi), (metric, fm is None, mm is None)


@pytest.mark.gpu
def test_masked_mixed_metric_stage_matches_the_cpu_engine(backend):
    fixed, moving = channels()
    plan = stage_plan([CC, MATTES])
    half = half_mask()
    want, wf, wi = oracle.syn_stage(plan, GRID, fixed, GRID, moving, fixed_mask=(GRID, half), moving_mask=(GRID, half))
    got = hip_stage(backend, plan, fixed, moving, half, half)
    report = fields_report(backend.download_field(got.forward), backend.download_field(got.inverse), wf, wi)
    energies = [(got.result.syn_last_energy[l], want.syn_last_energy[l]) for l in range(want.syn_levels)]
    for name in ("forward", "inverse"):
        assert report[name][0] <= 0.1 and report[name][1] <= 0.5, (report, energies)
    for g, w in energies:
        assert abs(g - w) <= 1e-4 * abs(w), (report, energies)


@pytest.mark.gpu
def test_two_stages_collapse_as_on_the_cpu_engine(backend):
    fixed, moving = channels()
    plan = stage_plan([CC])
    _, f1, i1 = oracle.syn_stage(plan, GRID, [fixed[0]], GRID, [moving[0]])
    want, wf, wi = oracle.syn_stage(plan, GRID, [fixed[0]], GRID, [moving[0]], prior=(GRID, f1, i1))
    first = hip_stage(backend, plan, [fixed[0]], [moving[0]])
    got = hip_stage(backend, plan, [fixed[0]], [moving[0]], prior=(first.forward, first.inverse))
    report = fields_report(backend.download_field(got.forward), backend.download_field(got.inverse), wf, wi)
    for name in ("forward", "inverse"):
        assert report[name][0] <= 0.1 and report[name][1] <= 0.5, report
    # a stage without iterations hands the earlier fields back exactly
    empty = hip_stage(backend, stage_plan([CC], iterations=(0, 0)), [fixed[0]], [moving[0]],
                      prior=(first.forward, first.inverse))
    assert np.array_equal(backend.download_field(empty.forward), backend.download_field(first.forward))
    assert np.array_equal(backend.download_field(empty.inverse), backend.download_field(first.inverse))
