"""Whole registrations: the f64 CPU engine recovers a known transform (CPU), and the HIP engine
matches it on the same phantom and plan within the accuracy bounds (GPU): affine parameters
within 1e-3 (translations within 0.05 mm), displacement fields mean |d| <= 0.1 voxel and
p99 <= 0.5 voxel, per-level final SyN energies within 1e-4 relative."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from fastwarp_amd.phantom import make_pair
from fastwarp_amd.register import cohort_plan


def small_plan():
    p = cohort_plan(1)
    p.set_linear([40, 20], [2, 1], [1.0, 0.0])
    p.set_syn([15, 8], [2, 1], [1.0, 0.0])
    return p


def affine_matrix(a) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    return np.array(list(a.matrix)).reshape(3, 3), np.array(list(a.translation)), np.array(list(a.center))


def test_cpu_engine_recovers_a_known_affine():
    pair = make_pair((40, 44, 36), 4.0, seed=3, affine_only=True)
    res, _, _ = oracle.register(small_plan(), pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)
    m, t, c = affine_matrix(res.affine)
    am, at, ac = affine_matrix(pair.affine)
    rng = np.random.default_rng(0)
    pts = rng.uniform(-30, 30, size=(50, 3))
    # the registration maps fixed points to moving points: it inverts the phantom's A
    mapped = (am @ (pts - ac).T).T + ac + at
    back = (m @ (mapped - c).T).T + c + t
    err = np.linalg.norm(back - pts, axis=1)
    assert np.max(err) < 1.0, err.max()


def test_cpu_engine_rejects_bad_plans():
    pair = make_pair((24, 24, 24), 4.0, seed=1, affine_only=True)
    p = small_plan()
    p.syn_levels = 0
    with pytest.raises(oracle.OracleError, match="levels"):
        oracle.register(p, pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)
    p = small_plan()
    bad = [pair.fixed[0].copy()]
    bad[0][5] = np.nan
    with pytest.raises(oracle.OracleError, match="non-finite"):
        oracle.register(p, pair.grid, bad, pair.grid, pair.moving, fields=False)


def linear_plan():
    """Linear stages only, run to convergence (window 10, threshold 1e-6) at two levels."""
    p = cohort_plan(1)
    p.set_linear([400, 200], [2, 1], [1.0, 0.0])
    p.run_syn = 0
    return p


def syn_plan(channels: int):
    """SyN only, from the centre-of-mass start, three levels."""
    p = cohort_plan(channels)
    p.set_syn([30, 20, 10], [4, 2, 1], [2.0, 1.0, 0.0])
    p.run_linear = 0
    return p


@pytest.mark.gpu
def test_hip_linear_stages_match_the_cpu_engine(backend):
    from fastwarp_amd.register import register

    pair = make_pair((48, 56, 48), 4.0, seed=5, disp_max_mm=6.0)
    plan = linear_plan()
    want, _, _ = oracle.register(plan, pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)
    got = register(backend, plan, pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)
    gm, gt, gc = affine_matrix(got.affine)
    wm, wt, wc = affine_matrix(want.affine)
    # the same map about the same centre
    gt_at_wc = gt + gc - gm @ gc - (wc - gm @ wc)
    # Euler parameters are three angles (radians) then three translations (mm)
    rigid = np.abs(np.array(list(got.result.rigid_parameters)) - np.array(list(want.rigid_parameters)))
    report = {"matrix": float(np.max(np.abs(gm - wm))), "translation_mm": float(np.max(np.abs(gt_at_wc - wt))),
              "rigid_angles": float(rigid[:3].max()), "rigid_translation_mm": float(rigid[3:].max())}
    assert report["matrix"] <= 1e-3 and report["translation_mm"] <= 0.05, report
    assert report["rigid_angles"] <= 1e-3 and report["rigid_translation_mm"] <= 0.05, report


@pytest.mark.gpu
@pytest.mark.parametrize("channels", [1, 2])
def test_hip_syn_matches_the_cpu_engine(backend, channels):
    from fastwarp_amd.register import register

    pair = make_pair((48, 56, 48), 4.0, seed=5, disp_max_mm=6.0, channels=channels)
    plan = syn_plan(channels)
    want, wf, wi = oracle.register(plan, pair.grid, pair.fixed, pair.grid, pair.moving)
    got = register(backend, plan, pair.grid, pair.fixed, pair.grid, pair.moving)
    spacing = pair.grid.spacing[0]
    report = {}
    for name, g_field, w_field in (("forward", got.forward, wf), ("inverse", got.inverse, wi)):
        d = np.linalg.norm(g_field - w_field, axis=1) / spacing
        report[name] = (float(d.mean()), float(np.percentile(d, 99)))
    energies = [(got.result.syn_last_energy[l], want.syn_last_energy[l]) for l in range(want.syn_levels)]
    iters = [(got.result.syn_iterations[l], want.syn_iterations[l]) for l in range(want.syn_levels)]
    for name in ("forward", "inverse"):
        assert report[name][0] <= 0.1 and report[name][1] <= 0.5, (report, energies, iters)
    for g, w in energies:
        assert abs(g - w) <= 1e-4 * abs(w), (report, energies, iters)
