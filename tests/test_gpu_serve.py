"""The registration server (fw_server_*): requests of different sizes, channel counts and stage
plans run together on one device give the same bits as each run alone, the same bits as the
resident path on a single context (upload_image + fw_register_resident), and the server's
admission, cancellation and error paths behave as documented.

Determinism is the property under test: every slot owns its stream, buffers and scratch, the
histograms are integer sums and every reduction has a fixed order, so a request's result cannot
depend on what runs beside it."""
from __future__ import annotations

import time

import numpy as np
import pytest

from fastwarp_amd.backend import FW_ERR_CANCELLED, RegError
from fastwarp_amd.phantom import make_pair
from fastwarp_amd.register import cohort_plan, register_resident
from fastwarp_amd.serve import Server, Ticket

pytestmark = pytest.mark.gpu


def short_plan(channels: int = 1, syn: bool = True):
    p = cohort_plan(channels)
    p.set_linear([30, 15, 8], [4, 2, 1], [2.0, 1.0, 0.0])
    p.set_syn([10, 6, 4], [4, 2, 1], [2.0, 1.0, 0.0])
    p.run_syn = 1 if syn else 0
    return p


def requests():
    """Four requests: two sizes, one and two channels, full and rigid+affine only."""
    out = []
    for shape, spacing, seed, channels, syn in (((40, 44, 36), 2.0, 5, 1, True), ((48, 52, 44), 2.0, 9, 1, True),
                                                ((36, 40, 32), 2.5, 13, 2, True), ((44, 48, 40), 2.0, 17, 1, False)):
        pair = make_pair(shape, spacing, seed=seed, channels=channels, disp_max_mm=3.0)
        f32 = [np.asarray(v, dtype=np.float32) for v in pair.fixed], [np.asarray(v, dtype=np.float32) for v in pair.moving]
        out.append((short_plan(channels, syn), pair.grid, f32[0], f32[1]))
    return out


def outputs(served):
    a = served.affine
    return (np.array(list(a.matrix) + list(a.translation) + list(a.center)), served.forward, served.inverse)


def same(x, y) -> bool:
    return all((p is None and q is None) or np.array_equal(p, q) for p, q in zip(x, y))


def test_results_are_the_same_alone_and_in_the_mix():
    reqs = requests()
    with Server(slots=1) as alone:
        solo = [outputs(alone.wait(alone.submit(r[0], r[1], r[2], r[1], r[3]))) for r in reqs]
    with Server(slots=3) as mixed:
        # twice each, interleaved, so that every request overlaps others of another size
        tickets = [mixed.submit(r[0], r[1], r[2], r[1], r[3]) for _ in range(2) for r in reqs]
        got = [mixed.wait(t) for t in tickets]
        assert mixed.counts() == (0, 0, 0)
    slots = {g.timing.slot for g in got}
    assert len(slots) > 1, slots
    for i, g in enumerate(got):
        assert g.result.seconds_total > 0.0
        assert same(outputs(g), solo[i % len(reqs)]), f"request {i % len(reqs)} differs in the mix"


def test_served_equals_the_resident_path(backend):
    plan, grid, fixed, moving = requests()[1]
    imgs_f = [backend.upload_image(grid, v.astype(np.float64)) for v in fixed]
    imgs_m = [backend.upload_image(grid, v.astype(np.float64)) for v in moving]
    dev = register_resident(backend, plan, imgs_f, imgs_m)
    with Server(slots=2) as s:
        got = s.wait(s.submit(plan, grid, fixed, grid, moving))
    assert np.array_equal(np.array(list(got.affine.matrix)), np.array(list(dev.affine.matrix)))
    assert np.array_equal(got.forward, backend.download_field(dev.forward))
    assert np.array_equal(got.inverse, backend.download_field(dev.inverse))


def test_timing_and_admission_order():
    reqs = requests()
    with Server(slots=1) as s:
        t0 = time.perf_counter()
        tickets = [s.submit(r[0], r[1], r[2], r[1], r[3], fields=False) for r in reqs]
        got = [s.wait(t) for t in tickets]
        wall = time.perf_counter() - t0
    # one slot: requests run one after another in arrival order
    for a, b in zip(got, got[1:]):
        # (less the gap between the two submissions, well under 50 ms)
        assert b.timing.queued_s >= a.timing.queued_s + a.timing.run_s - 0.05
    for g in got:
        assert g.timing.slot == 0 and g.timing.footprint_bytes > (256 << 20)
        assert abs(g.timing.total_s - (g.timing.queued_s + g.timing.run_s)) < 1e-6
        assert g.forward is None and g.inverse is None
    assert got[-1].timing.total_s <= wall


def test_cancel_a_waiting_request():
    reqs = requests()
    with Server(slots=1) as s:
        first = s.submit(reqs[1][0], reqs[1][1], reqs[1][2], reqs[1][1], reqs[1][3])
        last = s.submit(reqs[0][0], reqs[0][1], reqs[0][2], reqs[0][1], reqs[0][3])
        s.cancel(last)
        with pytest.raises(RegError) as e:
            s.wait(last)
        assert e.value.status == FW_ERR_CANCELLED and "before it started" in str(e.value)
        assert s.wait(first).timing.slot == 0


def test_a_failing_request_leaves_the_others_alone():
    reqs = requests()
    bad = [v.copy() for v in reqs[0][2]]
    bad[0][7] = np.nan
    with Server(slots=2) as s:
        t_bad = s.submit(reqs[0][0], reqs[0][1], bad, reqs[0][1], reqs[0][3])
        t_ok = s.submit(reqs[0][0], reqs[0][1], reqs[0][2], reqs[0][1], reqs[0][3])
        with pytest.raises(RegError, match="non-finite"):
            s.wait(t_bad)
        assert s.wait(t_ok).forward is not None


def test_footprint_beyond_the_budget_is_refused():
    plan, grid, fixed, moving = requests()[0]
    # every request carries the 256 MiB instance overhead beyond its plan
    with Server(slots=1, budget_bytes=200 << 20) as s:
        with pytest.raises(RegError, match="beyond the budget"):
            s.submit(plan, grid, fixed, grid, moving)
        with pytest.raises(RegError, match="unknown ticket"):
            s.wait(Ticket(12345, None, None))


def test_dropped_tickets_and_a_second_wait():
    """Tickets dropped without a wait keep their arrays alive inside the server until it stops
    (the slots still read the inputs and write the fields), and a ticket is waited for once."""
    import gc

    reqs = requests()
    s = Server(slots=2)
    for r in reqs:
        s.submit(r[0], r[1], [v.copy() for v in r[2]], r[1], [v.copy() for v in r[3]])
    gc.collect()
    s.close()
    with Server(slots=2) as s:
        t = s.submit(reqs[1][0], reqs[1][1], reqs[1][2], reqs[1][1], reqs[1][3])
        got = s.wait(t)
        assert got.forward is not None and np.all(np.isfinite(got.forward))
        with pytest.raises(RegError, match="unknown ticket"):
            s.wait(t)
