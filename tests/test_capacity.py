"""Admission model (the role of crates/fastwarp-gpu/src/capacity.rs and fits_gpu): the planned
device bytes of one registration, checked on the CPU for its structure and on the GPU against the
engine's own buffer-pool high-water mark (the plan must bound the measured peak, and not by more
than a quarter, as the reference's capacity tests bound its pool peak)."""
from __future__ import annotations

import pytest

from fastwarp_amd.backend import plan_capacity
from fastwarp_amd.phantom import make_pair
from fastwarp_amd.register import cohort_plan
from tests import cases

G1 = cases.grid((182, 218, 182))


def test_the_syn_stage_dominates_and_grows_with_channels():
    one = plan_capacity(cohort_plan(1), G1)
    two = plan_capacity(cohort_plan(2), G1)
    n = one.fixed_voxels
    assert n == 182 * 218 * 182 and one.moving_voxels == n
    # four distinct sigmas (3, 2, 1 and the upload itself), two images of 4 bytes each
    assert one.smoothing_cache_bytes == 4 * 8 * n
    assert two.smoothing_cache_bytes == 8 * 8 * n
    assert one.syn_bytes > one.linear_bytes > 0
    assert 137 * n < one.syn_bytes < 150 * n
    assert two.syn_bytes - one.syn_bytes > 40 * n
    assert one.peak_bytes >= one.syn_bytes and one.largest_buffer_bytes == 12 * n


def test_stages_that_do_not_run_need_nothing():
    p = cohort_plan(1)
    p.run_syn = 0
    lin = plan_capacity(p, G1)
    assert lin.syn_bytes == 0 and lin.linear_bytes > 0
    p = cohort_plan(1)
    p.run_linear = 0
    syn = plan_capacity(p, G1)
    assert syn.linear_bytes == 0 and syn.syn_bytes > 0


def test_a_larger_moving_grid_costs_its_cached_images():
    big = cases.grid((200, 230, 200))
    same = plan_capacity(cohort_plan(1), G1)
    more = plan_capacity(cohort_plan(1), G1, big)
    assert more.smoothing_cache_bytes - same.smoothing_cache_bytes == 4 * 4 * (big.n - G1.n)


@pytest.mark.gpu
@pytest.mark.parametrize("channels", [1, 2])
@pytest.mark.parametrize("stages", ["linear", "syn", "both"])
def test_the_plan_bounds_the_measured_pool_peak(channels, stages):
    from fastwarp_amd.backend import HipBackend
    from fastwarp_amd.register import register

    pair = make_pair((64, 72, 60), 2.0, seed=4, disp_max_mm=4.0, channels=channels)
    plan = cohort_plan(channels)
    plan.set_linear([20, 10, 5, 3], [8, 4, 2, 1], [3.0, 2.0, 1.0, 0.0])
    plan.set_syn([6, 4, 3, 2], [8, 4, 2, 1], [3.0, 2.0, 1.0, 0.0])
    if stages == "linear":
        plan.run_syn = 0
    if stages == "syn":
        plan.run_linear = 0
    b = HipBackend(0)
    try:
        register(b, plan, pair.grid, pair.fixed, pair.grid, pair.moving)
        _, peak = b.memory_stats()
        cap = plan_capacity(plan, pair.grid)
        assert b.fits(plan, pair.grid)
        assert b.registrations_at_once(plan, pair.grid, max=4) == 4
    finally:
        b.close()
    assert peak <= cap.peak_bytes, (peak, cap.peak_bytes)
    assert cap.peak_bytes - (16 << 20) <= 1.25 * peak, (peak, cap.peak_bytes)
