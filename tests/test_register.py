"""Whole registrations: the f64 CPU engine recovers a known transform (CPU), and the HIP engine
matches it on the same phantom and plan (GPU).

Free running, the HIP result meets the north-star bounds against the CPU engine: affine
parameters within 1e-3 (translations within 0.05 mm), displacement fields mean |d| <= 0.1 voxel
and p99 <= 0.5 voxel. Two discrete effects keep free-running trajectories from agreeing iteration
by iteration: golden-section comparisons of MI values that agree to about 1e-7 (f32 level images)
send the linear stages' CG along slightly different paths (affines about 1e-4 apart), and the
first inversion of SyN's first iteration measures an update whose largest step is exactly the
0.1-voxel tolerance, a tie each engine breaks by its own rounding. So the SyN stage is compared in
lockstep: the CPU engine replays the HIP run's SyN decisions (iteration and inversion step counts)
from the HIP run's affine, every iteration's energy agrees to 1e-6 relative, and every decision
the CPU rules would have taken otherwise is shown to be a tie."""
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


def test_the_cpu_engines_own_linear_chain_gives_the_bits_of_the_generic_chain():
    """The CPU engine runs its own restatement of linear.rs and optimizer.rs (oracle/linear.hpp:
    no prefetch, no level reuse, the plain golden section); the product's generic chain over the
    same f64 ops must give the same bits: the affine, every level's CG values and counts. Three
    pairs: two shifted phantoms and one with a level skipped for want of samples."""
    cases = [(make_pair((40, 44, 36), 4.0, seed=3, affine_only=True), small_plan()),
             (make_pair((48, 40, 44), 3.0, seed=7, affine_only=True), small_plan())]
    p = cohort_plan(1)
    p.set_linear([30, 15, 8], [8, 2, 1], [2.0, 1.0, 0.0])
    cases.append((make_pair((24, 24, 24), 4.0, seed=1, affine_only=True), p))
    for k, (pair, plan) in enumerate(cases):
        plan.run_syn = 0
        runs = []
        for generic in (False, True):
            oracle.generic_linear(generic)
            try:
                res, _, _ = oracle.register(plan, pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)
                runs.append((res, [(t.stage, t.level, t.iteration, t.value) for t in oracle.last_trace()]))
            finally:
                oracle.generic_linear(False)
        (own, own_trace), (gen, gen_trace) = runs
        assert own_trace and own_trace == gen_trace, k
        assert [list(x) for x in own.linear_iterations] == [list(x) for x in gen.linear_iterations], k
        for a, b in zip(affine_matrix(own.affine), affine_matrix(gen.affine)):
            assert np.array_equal(a, b), k


def test_a_given_initial_transform_starts_the_chain_where_the_centre_of_mass_would():
    """stages.rs:74-86, 319-325: the chain starts from the centre of mass, the identity or a given
    affine; the centre-of-mass affine given explicitly reproduces the default run bit for bit,
    and a non-finite or singular start is rejected (stages.rs:139-155)."""
    pair = make_pair((40, 44, 36), 4.0, seed=3)
    p = small_plan()
    a, fa, _ = oracle.register(p, pair.grid, pair.fixed, pair.grid, pair.moving)
    p0 = small_plan()
    p0.run_linear = 0
    com, _, _ = oracle.register(p0, pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)
    p2 = small_plan()
    p2.initial_kind = 2
    p2.initial_affine = com.affine
    b, fb, _ = oracle.register(p2, pair.grid, pair.fixed, pair.grid, pair.moving)
    assert np.array_equal(fa, fb)
    assert list(a.affine.matrix) == list(b.affine.matrix)
    p1 = small_plan()
    p1.initial_kind = 1
    p1.run_linear = 0
    ident, _, _ = oracle.register(p1, pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)
    assert list(ident.affine.matrix) == [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    assert list(ident.affine.translation) == [0.0, 0.0, 0.0] and list(ident.affine.center) == [0.0, 0.0, 0.0]
    bad = small_plan()
    bad.initial_kind = 2
    with pytest.raises(oracle.OracleError, match="singular"):
        oracle.register(bad, pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)
    bad.initial_affine = com.affine
    bad.initial_affine.translation[0] = float("nan")
    with pytest.raises(oracle.OracleError, match="non-finite"):
        oracle.register(bad, pair.grid, pair.fixed, pair.grid, pair.moving, fields=False)


def test_a_lockstep_replay_of_a_run_reproduces_it():
    """The oracle replaying its own run's SyN decisions (iteration and inversion step counts) from
    its own collapsed affine gives the same fields bit for bit and no disagreement; forcing one
    extra inversion step is reported as a disagreement of the oracle's own rule."""
    pair = make_pair((40, 44, 36), 4.0, seed=3)
    p = small_plan()
    res, f, i = oracle.register(p, pair.grid, pair.fixed, pair.grid, pair.moving)
    trace = oracle.last_trace()
    assert sum(1 for t in trace if t.stage == 2) == sum(res.syn_iterations[:res.syn_levels])
    _, f2, i2, dis = oracle.register_lockstep(p, pair.grid, pair.fixed, pair.grid, pair.moving, trace, start=res.affine)
    assert np.array_equal(f, f2) and np.array_equal(i, i2) and dis == []
    first = next(t for t in trace if t.stage == 2)
    first.inversion_steps[0] += 1
    _, _, _, dis = oracle.register_lockstep(p, pair.grid, pair.fixed, pair.grid, pair.moving, trace, start=res.affine)
    assert dis and dis[0][0] == 0 and dis[0][2] == 0


def lockstep_report(backend, pair, plan, channels):
    """The HIP registration against the CPU engine: the free-running results, and the CPU engine's
    SyN stage replaying the HIP run's discrete decisions from the HIP run's affine."""
    from fastwarp_amd.register import last_trace, register

    want, wf, wi = oracle.register(plan, pair.grid, pair.fixed, pair.grid, pair.moving)
    got = register(backend, plan, pair.grid, pair.fixed, pair.grid, pair.moving)
    trace = last_trace()
    ls, lf, li, dis = oracle.register_lockstep(plan, pair.grid, pair.fixed, pair.grid, pair.moving, trace, start=got.affine)
    lt = oracle.last_trace()
    spacing = pair.grid.spacing[0]
    gm, gt, gc = affine_matrix(got.affine)
    wm, wt, wc = affine_matrix(want.affine)
    gt_at_wc = gt + gc - gm @ gc - (wc - gm @ wc)
    r = {"matrix": float(np.max(np.abs(gm - wm))), "translation_mm": float(np.max(np.abs(gt_at_wc - wt)))}
    for name, g_field, w_field, l_field in (("forward", got.forward, wf, lf), ("inverse", got.inverse, wi, li)):
        d = np.linalg.norm(g_field - w_field, axis=1) / spacing
        e = np.linalg.norm(g_field - l_field, axis=1) / spacing
        r[name] = (float(d.mean()), float(np.percentile(d, 99)))
        r["lockstep_" + name] = (float(e.mean()), float(np.percentile(e, 99)))
    hs = [t for t in trace if t.stage == 2]
    os_ = [t for t in lt if t.stage == 2]
    r["syn_iterations"] = (list(got.result.syn_iterations[:plan.syn_levels]), list(want.syn_iterations[:plan.syn_levels]))
    r["linear_iterations"] = ([list(got.result.linear_iterations[s][:plan.linear_levels]) for s in range(2)],
                              [list(want.linear_iterations[s][:plan.linear_levels]) for s in range(2)])
    r["free_energy_rel"] = [abs(got.result.syn_last_energy[l] - want.syn_last_energy[l]) / abs(want.syn_last_energy[l])
                            for l in range(plan.syn_levels)]
    r["lockstep_energy_rel"] = max(abs(a.value - b.value) / abs(b.value) for a, b in zip(hs, os_))
    r["lockstep_convergence_abs"] = max(abs(a.convergence_value - b.convergence_value) for a, b in zip(hs, os_))
    r["disagreements"] = dis
    # the oracle's own convergence rule at the HIP run's stopping points
    stops = []
    for level in range(plan.syn_levels):
        lv = [t for t in os_ if t.level == level]
        stopped_early = len(lv) < plan.syn_iterations[level]
        for k, t in enumerate(lv):
            rule_stops = t.convergence_value < plan.syn_convergence_threshold
            hip_stops = stopped_early and k == len(lv) - 1
            if rule_stops != hip_stops:
                stops.append((level, k, t.convergence_value))
    r["convergence_disagreements"] = stops
    return r, hs, os_


def check_lockstep(r, plan, energy=1e-6, tie=1e-9, convergence=1e-8):
    # north-star bounds on the free-running results
    assert r["matrix"] <= 1e-3 and r["translation_mm"] <= 0.05, r
    for name in ("forward", "inverse"):
        assert r[name][0] <= 0.1 and r[name][1] <= 0.5, r
        assert r["lockstep_" + name][0] <= 1e-4 and r["lockstep_" + name][1] <= 1e-3, r
    # the SyN stage from the same affine, with the same decisions, follows the oracle's trajectory
    assert r["lockstep_energy_rel"] <= energy, r
    assert r["lockstep_convergence_abs"] <= convergence, r
    # every decision the oracle's own rules would have taken otherwise is an exact tie: the first
    # inversion of a level's first iteration measures the update itself, whose largest step is the
    # gradient step (0.1 voxel), the inversion's max-error tolerance
    for call, body, natural, mx, mn in r["disagreements"]:
        assert abs(mx - 0.1) <= tie or abs(mn - 0.001) <= tie, r
    for level, k, cv in r["convergence_disagreements"]:
        assert abs(cv - plan.syn_convergence_threshold) <= convergence, r


@pytest.mark.gpu
@pytest.mark.parametrize("seed,channels", [(1, 1), (1, 2), (2, 1), (3, 1)])
def test_hip_registration_at_2mm_matches_the_cpu_engine(backend, seed, channels):
    """The 2 mm plain registration (96 x 112 x 96, centre of mass, Rigid, Affine, SyN) against
    the f64 CPU engine on the same phantom: the north-star bounds on the free-running affine
    (1e-3, translations 0.05 mm) and fields (mean 0.1, p99 0.5 voxel); the SyN stage replayed by
    the CPU engine from the HIP run's affine and decisions agrees at every iteration (energy 1e-6
    relative, convergence value 1e-8, fields 1e-4 voxel mean), and every decision the CPU rules
    would have taken otherwise is a tie."""
    from fastwarp_amd.register import plan_2mm

    pair = make_pair((96, 112, 96), 2.0, seed=seed, channels=channels)
    plan = plan_2mm(channels)
    r, _, _ = lockstep_report(backend, pair, plan, channels)
    print(r)
    check_lockstep(r, plan)


@pytest.mark.gpu
def test_hip_registration_at_1mm_matches_the_cpu_engine(backend):
    """The 1 mm grid (182 x 218 x 182, displacement up to 6 mm) through all four levels of both
    stages with a shortened plan, against the f64 CPU engine as above."""
    pair = make_pair((182, 218, 182), 1.0, seed=2, disp_max_mm=6.0)
    plan = cohort_plan(1)
    plan.set_linear([40, 20, 10, 4], [8, 4, 2, 1], [3.0, 2.0, 1.0, 0.0])
    plan.set_syn([20, 10, 4, 2], [8, 4, 2, 1], [3.0, 2.0, 1.0, 0.0])
    r, _, _ = lockstep_report(backend, pair, plan, 1)
    print(r)
    check_lockstep(r, plan, energy=1e-5)
