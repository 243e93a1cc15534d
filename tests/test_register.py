"""Whole registrations: the f64 CPU engine recovers a known transform (CPU), and the HIP engine
matches it on the same phantom and plan within the north-star bounds (GPU): affine parameters
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


@pytest.mark.gpu
@pytest.mark.parametrize("channels", [1, 2])
def test_hip_registration_matches_the_cpu_engine(backend, channels):
    from fastwarp_amd.register import register

    pair = make_pair((40, 44, 36), 4.0, seed=5, disp_max_mm=6.0, channels=channels)
    plan = small_plan()
    plan.channels = channels
    for c in range(channels):
        plan.channel_weights[c] = 1.0 / channels
    want, wf, wi = oracle.register(plan, pair.grid, pair.fixed, pair.grid, pair.moving)
    got = register(backend, plan, pair.grid, pair.fixed, pair.grid, pair.moving)
    gm, gt, gc = affine_matrix(got.affine)
    wm, wt, wc = affine_matrix(want.affine)
    assert np.max(np.abs(gm - wm)) <= 1e-3
    # the same map: compare translations about the same centre
    gt_at_wc = gt + gc - gm @ gc - (wc - gm @ wc)
    assert np.max(np.abs(gt_at_wc - wt)) <= 0.05
    assert np.max(np.abs(np.array(list(got.result.rigid_parameters)) - np.array(list(want.rigid_parameters)))) <= 1e-3
    spacing = pair.grid.spacing[0]
    for g_field, w_field in ((got.forward, wf), (got.inverse, wi)):
        d = np.linalg.norm(g_field - w_field, axis=1) / spacing
        assert d.mean() <= 0.1 and np.percentile(d, 99) <= 0.5, (d.mean(), np.percentile(d, 99))
    for level in range(want.syn_levels):
        assert got.result.syn_iterations[level] == want.syn_iterations[level]
        g, w = got.result.syn_last_energy[level], want.syn_last_energy[level]
        assert abs(g - w) <= 1e-4 * abs(w), (level, g, w)
