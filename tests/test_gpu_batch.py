"""Batched registration of several subjects against one template (fw_register_batch): the SyN
iterations of all subjects run in lockstep, each step one launch over every subject, and the
template's smoothed levels are built once; and many independent linear fits of uneven grids side
by side on worker streams (fw_linear_batch). Each subject's or fit's result must be the bits of its
own fw_register_resident run (crates/fastwarp-reg/src/syn.rs:776-986 and linear.rs:1148-1212 per
subject): the same affine, the same CG values and SyN energies, iteration and inversion step counts
per iteration, and the same fields. Those single runs are themselves held to the CPU engine by
tests/test_register.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from fastwarp_amd.phantom import make_pair
from fastwarp_amd.register import (last_batch_trace, last_trace, linear_batch, plan_2mm, register_batch,
                                   register_resident)

pytestmark = pytest.mark.gpu


def affine_array(a) -> np.ndarray:
    return np.array(list(a.matrix) + list(a.translation) + list(a.center))


def steps(trace):
    return [(t.stage, t.level, t.iteration, t.value, tuple(t.inversion_steps), t.update_scale[0], t.update_scale[1])
            for t in trace]


def cohort(channels: int, subjects: int, shape=(96, 112, 96), spacing=2.0):
    """One template (the fixed channels of the first phantom) and the moving channels of
    `subjects` phantoms of other seeds."""
    template = make_pair(shape, spacing, seed=11, channels=channels)
    moving = [make_pair(shape, spacing, seed=20 + s, channels=channels).moving for s in range(subjects)]
    return template.grid, template.fixed, moving


@pytest.mark.parametrize("channels,subjects", [(1, 3), (2, 2)])
def test_batched_subjects_give_the_bits_of_their_single_runs(backend, channels, subjects):
    grid, fixed, moving = cohort(channels, subjects)
    plan = plan_2mm(channels)
    plan.set_linear([30, 15, 8], [4, 2, 1], [2.0, 1.0, 0.0])
    up = lambda vols: [backend.upload_image(grid, v) for v in vols]
    tmpl = up(fixed)
    subj = [up(m) for m in moving]
    batch = register_batch(backend, plan, tmpl, subj, workers=2)
    traces = [steps(last_batch_trace(s)) for s in range(subjects)]
    for s in range(subjects):
        alone = register_resident(backend, plan, tmpl, subj[s])
        want = steps(last_trace())
        got = batch[s]
        label = f"subject {s} of {subjects}, {channels} channel(s)"
        assert np.array_equal(affine_array(got.affine), affine_array(alone.affine)), label
        assert list(got.result.syn_iterations)[: plan.syn_levels] == list(alone.result.syn_iterations)[: plan.syn_levels], label
        assert traces[s] == want, label
        for name in ("forward", "inverse"):
            a = backend.download_field(getattr(got, name))
            b = backend.download_field(getattr(alone, name))
            assert np.array_equal(a, b), (label, name, float(np.max(np.abs(a - b))))


def test_batched_two_channel_subjects_at_1mm_give_the_bits_of_their_single_runs(backend):
    """The batched multichannel workload's shape (1 mm, 182 x 218 x 182, two channels, CC radius 4
    on both, the cohort pyramid 8/4/2/1) with two subjects and a shortened plan: every subject's
    affine, trace and fields are the bits of its own run."""
    from fastwarp_amd.register import cohort_plan

    grid, fixed, moving = cohort(2, 2, shape=(182, 218, 182), spacing=1.0)
    plan = cohort_plan(2)
    plan.set_linear([20, 10, 5, 2], [8, 4, 2, 1], [3.0, 2.0, 1.0, 0.0])
    plan.set_syn([10, 6, 3, 2], [8, 4, 2, 1], [3.0, 2.0, 1.0, 0.0])
    up = lambda vols: [backend.upload_image(grid, v) for v in vols]
    tmpl = up(fixed)
    subj = [up(m) for m in moving]
    batch = register_batch(backend, plan, tmpl, subj)
    traces = [steps(last_batch_trace(s)) for s in range(2)]
    for s in range(2):
        alone = register_resident(backend, plan, tmpl, subj[s])
        assert np.array_equal(affine_array(batch[s].affine), affine_array(alone.affine)), s
        assert traces[s] == steps(last_trace()), s
        for name in ("forward", "inverse"):
            a = backend.download_field(getattr(batch[s], name))
            b = backend.download_field(getattr(alone, name))
            assert np.array_equal(a, b), (s, name, float(np.max(np.abs(a - b))))


def test_batch_without_subjects_is_an_error(backend):
    from fastwarp_amd.backend import RegError

    grid, fixed, _ = cohort(1, 0, shape=(24, 24, 24), spacing=4.0)
    with pytest.raises(RegError, match="no subjects"):
        register_batch(backend, plan_2mm(1), [backend.upload_image(grid, fixed[0])], [])


def test_linear_fits_side_by_side_give_the_bits_of_their_single_runs(backend):
    """Six rigid+affine fits of uneven grids and spacings (the batched linear workload's generator,
    smaller), on three worker streams, against each fit run alone."""
    shapes = [((40, 44, 36), 4.0), ((64, 64, 64), 2.0), ((56, 72, 48), 3.0), ((48, 40, 52), 2.5), ((72, 80, 64), 2.0),
              ((44, 48, 40), 3.5)]
    plan = plan_2mm(1)
    plan.set_linear([30, 15, 8], [4, 2, 1], [2.0, 1.0, 0.0])
    plan.run_syn = 0
    fixed, moving = [], []
    for i, (shape, spacing) in enumerate(shapes):
        pair = make_pair(shape, spacing, seed=40 + i, affine_only=True)
        fixed.append(backend.upload_image(pair.grid, pair.fixed[0]))
        moving.append(backend.upload_image(pair.grid, pair.moving[0]))
    results = linear_batch(backend, plan, fixed, moving, workers=3)
    traces = [steps(last_batch_trace(i)) for i in range(len(shapes))]
    for i in range(len(shapes)):
        alone = register_resident(backend, plan, [fixed[i]], [moving[i]], fields=False)
        label = f"fit {i}, grid {shapes[i]}"
        assert np.array_equal(affine_array(results[i].affine), affine_array(alone.affine)), label
        assert np.array_equal(np.array(results[i].linear_iterations), np.array(alone.result.linear_iterations)), label
        assert traces[i] == steps(last_trace()), label


def test_graph_replays_give_the_bits_of_direct_launches(backend):
    """The batched SyN iterations replay a HIP graph of each iteration's update launches (captured
    once per active set and buffer parity); with graphs switched off (the tuning knob
    batch.graphs = 0) the same launches are issued one by one. Three subjects (whose convergence
    may shrink the active set within a level): the same traces and fields."""
    grid, fixed, moving = cohort(1, 3)
    plan = plan_2mm(1)
    plan.set_linear([30, 15, 8], [4, 2, 1], [2.0, 1.0, 0.0])
    up = lambda vols: [backend.upload_image(grid, v) for v in vols]
    tmpl = up(fixed)
    subj = [up(m) for m in moving]
    runs = []
    for graphs in (1, 0):
        backend.set_tuning("batch.graphs", graphs)
        try:
            batch = register_batch(backend, plan, tmpl, subj)
        finally:
            backend.set_tuning("batch.graphs", -1)
        runs.append(([steps(last_batch_trace(s)) for s in range(3)],
                     [[backend.download_field(getattr(r, name)) for name in ("forward", "inverse")] for r in batch]))
    (t_graph, f_graph), (t_direct, f_direct) = runs
    # SyN iterations per level of each subject (stage 2), for the messages
    iterations = [[sum(1 for t in tr if t[0] == 2 and t[1] == lv) for lv in range(plan.syn_levels)] for tr in t_graph]
    for s in range(3):
        assert t_graph[s] == t_direct[s], (s, iterations)
        for a, b in zip(f_graph[s], f_direct[s]):
            assert np.array_equal(a, b), (s, iterations)
