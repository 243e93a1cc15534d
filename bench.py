"""Benchmark of the plain HIP registration path on one MI355X.

A step is one full 1 mm normalisation (centre of mass, Rigid and Affine Mattes MI, SyN CC radius 4
with the cohort schedule 8x4x2x1) of the seeded 182x218x182 phantom pair. W untimed warmup steps,
then K timed steps; prints one JSON line with registrations/hour, per-step latency, the roofline of
the dominant kernel (HIP events carried by every one of its launches in the timed region), the
per-kernel breakdown of the last warmup step, the served stream (256 mixed requests with Poisson
arrivals at 80 % load against serial plain execution of the same mix), and one complete
registration on the f64 CPU engine as a baseline (the 2 mm one by default).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "1mm SyN normalizations/hour batched+served vs plain; p50/p99 latency; HBM GB/s"


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def dominant(profile: dict) -> str:
    """The kernel family with the largest total time."""
    return max(profile.items(), key=lambda kv: kv[1][1])[0]


def roofline(name: str, timed: dict, steps: int, breakdown: dict) -> dict:
    """Kernel `name` over the timed region, every launch carrying HIP events on the context's
    stream: the algorithmic bytes of all its launches over the sum of their durations, the same
    quantity a rocprofv3 kernel trace of the run gives (total bytes over total duration)."""
    launches, ms, nbytes, n_timed = timed[name]
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {
        "kernel": name,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "algorithmic_bytes_per_launch": round(nbytes / launches),
        "launches_per_step": round(launches / steps, 1),
        "ms_per_step": round(ms / steps, 3),
        "mean_launch_us": round(1e3 * ms / launches, 2),
        "share_of_gpu_time": round(breakdown[name][1] / sum(v[1] for v in breakdown.values()), 4),
    }


def cpu_info() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(config: str) -> dict:
    """The f64 CPU engine (the oracle's restatement of engine="cpu") running one complete plain
    registration of `config` (every stage and level, to convergence as planned), timed end to end
    on this host's cores."""
    import oracle
    from fastwarp_amd.phantom import CONFIGS, make_pair
    from fastwarp_amd.register import cohort_plan, plan_2mm

    cfg = CONFIGS[config]
    pair = make_pair(**cfg)
    plan = cohort_plan(1) if config == "1mm" else plan_2mm(1)
    t0 = time.perf_counter()
    res, _, _ = oracle.register(plan, pair.grid, pair.fixed, pair.grid, pair.moving, fields=True)
    wall = time.perf_counter() - t0
    threads = int(os.environ.get("FW_THREADS") or os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {
        "value": round(3600.0 / wall, 3),
        "unit": "registrations/hour",
        "cores": threads,
        "kind": "port",
        "sample": f"one complete {config} plain registration ({cfg['shape_xyz']} at {cfg['spacing']} mm, seed {cfg['seed']}: "
        f"centre of mass, Rigid, Affine, SyN) on the f64 CPU engine, {wall:.1f} s end to end "
        f"(linear {res.seconds_linear:.1f} s, SyN {res.seconds_syn:.1f} s) on {threads} threads of {os.cpu_count()} "
        f"logical CPUs ({cpu_info()})",
        "seconds": round(wall, 2),
        "config": config,
    }


def served(slots: int, requests: int, load: float) -> dict:
    """The served-stream workload: `requests` registrations (70 % full normalisations at 2, 1.5 or
    1 mm with one or two channels, 30 % rigid+affine only; f32 volumes from host memory, fields back
    to host memory) arriving as a Poisson stream at `load` times the rate the server saturates at
    on the same mix, through `slots` slots; against serial plain execution of the same mix (one
    slot, one request at a time, queued back to back) on a prefix of it."""
    from fastwarp_amd.serve import Server
    from fastwarp_amd.serve_stream import make_mix, pairs_for, run_stream

    mix = make_mix(requests)
    pairs = pairs_for(mix)
    prefix = mix[: max(16, requests // 8)]
    with Server(slots=1) as one:
        run_stream(one, sorted(set(mix)), pairs, None)  # warm every kind of request
        serial_wall, _, serial_fail = run_stream(one, prefix, pairs, None)
    serial = len(prefix) / serial_wall
    with Server(slots=slots) as s:
        run_stream(s, sorted(set(mix)) * slots, pairs, None)
        sat_wall, _, sat_fail = run_stream(s, prefix, pairs, None)
        saturated = len(prefix) / sat_wall
        wall, lat, fail = run_stream(s, mix, pairs, load * saturated)
    lat = np.array(lat)
    value = requests / wall
    failures = serial_fail + sat_fail + fail
    log(f"served: {slots} slots, {requests} Poisson requests at {load:.0%} load in {wall:.1f} s "
        f"({3600 * value:.0f}/h; serial plain {3600 * serial:.0f}/h; saturated {3600 * saturated:.0f}/h)")
    kinds = {}
    for spacing, channels, full in mix:
        key = f"{'full' if full else 'linear'} {spacing} mm x{channels}"
        kinds[key] = kinds.get(key, 0) + 1
    return {"value": round(3600.0 * value, 3), "unit": "registrations/hour", "slots": slots, "requests": requests,
            "load": load, "arrival_per_hour": round(3600.0 * load * saturated, 1),
            "saturated_per_hour": round(3600.0 * saturated, 1),
            "serial_plain_per_hour": round(3600.0 * serial, 1), "vs_serial_plain": round(value / serial, 3),
            "latency_s": {q: round(float(np.percentile(lat, int(q[1:]))), 3) for q in ("p50", "p95", "p99")},
            "mix": dict(sorted(kinds.items())), "failed": len(failures),
            "path": "fw_server: Poisson arrivals, admitted by HBM footprint to slots with their own HIP streams, "
                    "f32 volumes uploaded per request, fields downloaded to the host per full request"}


def batched(backend, subjects: int) -> dict:
    """The batched multichannel workload: `subjects` two-channel subjects (T1-like and T2-like
    phantom channels, 1 mm, 182 x 218 x 182; template seed 4000, subjects 4001 on) registered to
    one shared template resident in HBM,
    full normalisations (centre of mass, Rigid+Affine MI, SyN CC radius 4 on both channels with
    equal weights, 100/70/50/20) through fw_register_batch, against the same subjects registered
    one at a time (fw_register_resident)."""
    from fastwarp_amd.phantom import make_pair
    from fastwarp_amd.register import cohort_plan, register_batch, register_resident

    t = time.perf_counter()
    tmpl_pair = make_pair((182, 218, 182), 1.0, seed=4000, disp_max_mm=6.0, channels=2)
    grid = tmpl_pair.grid
    tmpl = [backend.upload_image(grid, v) for v in tmpl_pair.fixed]
    del tmpl_pair
    subj = []
    for s in range(subjects):
        p = make_pair((182, 218, 182), 1.0, seed=4001 + s, disp_max_mm=6.0, channels=2)
        subj.append([backend.upload_image(grid, v) for v in p.moving])
    log(f"batched: template and {subjects} two-channel subjects ready in {time.perf_counter() - t:.1f} s")
    plan = cohort_plan(2)

    def free(rs):
        for r in rs:
            for f in (r.forward, r.inverse):
                if f is not None:
                    f.free()

    free(register_batch(backend, plan, tmpl, subj))  # warm the pool at this size
    t = time.perf_counter()
    rs = register_batch(backend, plan, tmpl, subj)
    batch_wall = time.perf_counter() - t
    syn_s = rs[0].result.seconds_syn
    linear_s = sum(r.result.seconds_linear for r in rs)
    free(rs)
    # the SyN stage's kernels (on this context; the linear chains run on the workers' own streams),
    # every launch timed, in one more run (no profiling focus left from an earlier measurement)
    backend.set_profiling_focus(None)
    backend.set_profiling(True)
    free(register_batch(backend, plan, tmpl, subj))
    prof = {k: v for k, v in backend.take_profile().items() if not k.startswith("host:")}
    backend.set_profiling(False)
    t = time.perf_counter()
    plain_syn = 0.0
    for s in range(subjects):
        r = register_resident(backend, plan, tmpl, subj[s])
        plain_syn += r.result.seconds_syn
        free([r])
    plain_wall = time.perf_counter() - t
    log(f"batched: {subjects} subjects in {batch_wall:.2f} s (linear {linear_s:.2f} s, batched SyN {syn_s:.2f} s); "
        f"one at a time {plain_wall:.2f} s (SyN {plain_syn:.2f} s)")
    return {"value": round(3600.0 * subjects / batch_wall, 3), "unit": "registrations/hour", "subjects": subjects,
            "channels": 2, "grid": [182, 218, 182], "spacing_mm": 1.0,
            "plain_per_hour": round(3600.0 * subjects / plain_wall, 3), "vs_plain": round(plain_wall / batch_wall, 3),
            "syn_s": {"batched": round(syn_s, 3), "one_at_a_time": round(plain_syn, 3),
                      "vs_one_at_a_time": round(plain_syn / max(syn_s, 1e-9), 3)},
            "linear_s": round(linear_s, 3),
            "syn_kernels": {k: {"launches": v[0], "ms": round(v[1], 3), "gbs": round(v[2] / max(v[1], 1e-9) / 1e6, 1)}
                            for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])},
            "path": "fw_register_batch: shared template resident in HBM, linear chains side by side on worker streams, "
                    "SyN in lockstep with one launch per step over all subjects; fields left in HBM"}


def fit_sizes(fits: int, seed: int = 3):
    """The batched linear workload's pairs: grids drawn uniformly per axis from 64^3 to
    192 x 224 x 192, isotropic spacing drawn from {1, 1.5, 2} mm, phantom seed 1000 + i for pair i,
    random affine only."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(fits):
        shape = (int(rng.integers(64, 193)), int(rng.integers(64, 225)), int(rng.integers(64, 193)))
        out.append((shape, float(rng.choice([1.0, 1.5, 2.0])), 1000 + i))
    return out


def make_fit(args):
    from fastwarp_amd.phantom import make_pair

    shape, spacing, seed = args
    p = make_pair(shape, spacing, seed=seed, affine_only=True)
    g = p.grid
    return (shape, tuple(g.spacing), tuple(g.origin), p.fixed[0].astype(np.float32), p.moving[0].astype(np.float32))


def linear_fits(backend, data, workers: int) -> dict:
    """The batched linear workload: len(data) independent rigid+affine Mattes MI fits of uneven
    sizes (the cohort's linear recipe: 32 bins, 25 % regular sampling, shrink 8/4/2/1), raw volumes
    resident in HBM, side by side on `workers` worker streams (fw_linear_batch), against the same
    fits one at a time (fw_register_resident without SyN)."""
    from fastwarp_amd._abi import Grid
    from fastwarp_amd.register import cohort_plan, linear_batch, register_resident

    plan = cohort_plan(1)
    plan.run_syn = 0
    fixed, moving = [], []
    voxels = 0
    for shape, spacing, origin, f, m in data:
        g = Grid.make(shape, spacing, origin)
        fixed.append(backend.upload_image(g, f))
        moving.append(backend.upload_image(g, m))
        voxels += int(np.prod(shape))
    linear_batch(backend, plan, fixed[:workers], moving[:workers], workers=workers)  # warm the workers
    t = time.perf_counter()
    linear_batch(backend, plan, fixed, moving, workers=workers)
    batch_wall = time.perf_counter() - t
    t = time.perf_counter()
    for f, m in zip(fixed, moving):
        register_resident(backend, plan, [f], [m], fields=False)
    plain_wall = time.perf_counter() - t
    n = len(data)
    log(f"linear fits: {n} fits ({voxels / n / 1e6:.2f} M voxels on average) in {batch_wall:.2f} s on {workers} workers; "
        f"one at a time {plain_wall:.2f} s")
    return {"value": round(3600.0 * n / batch_wall, 3), "unit": "fits/hour", "fits": n, "workers": workers,
            "mean_voxels": int(voxels / n), "plain_per_hour": round(3600.0 * n / plain_wall, 3),
            "vs_plain": round(plain_wall / batch_wall, 3),
            "path": "fw_linear_batch: raw volumes resident in HBM, fits taken in turn by worker contexts with their own "
                    "HIP streams"}


def kernels(repeats: int) -> int:
    """Each hot-path op on the 1 mm pair at full resolution (182x218x182), `repeats` launches
    after one warm-up: one JSON line per kernel family with its mean launch time (HIP events on
    the context's stream) and algorithmic bytes per launch against the HBM peak."""
    from fastwarp_amd._abi import InvertParams, LinearTransform, MiBinning
    from fastwarp_amd.backend import HipBackend, MiLevel
    from fastwarp_amd.phantom import CONFIGS, make_pair

    pair = make_pair(**CONFIGS["1mm"])
    g = pair.grid
    b = HipBackend(0)
    f = b.upload_image(g, pair.fixed[0])
    m = b.upload_image(g, pair.moving[0])
    u = b.upload_field(g, pair.displacement)
    # a second, different field, so that the composition reads two fields' worth of data
    u2 = b.upload_field(g, 0.7 * pair.displacement[::-1].copy())
    n = pair.fixed[0].size
    # samples of the finest linear level: every 4th voxel, jittered by a third of a voxel
    rng = np.random.default_rng(0)
    idx = np.arange(0, n, 4)
    size, origin, spacing = np.array(list(g.size)), np.array(list(g.origin)), np.array(list(g.spacing))
    ijk = np.stack([idx % size[0], (idx // size[0]) % size[1], idx // (size[0] * size[1])], 1).astype(float)
    pts = origin + (ijk + rng.normal(0, 1 / 3, ijk.shape)) * spacing
    fv, mv = pair.fixed[0], pair.moving[0]
    binning = MiBinning.from_ranges(32, (fv.min(), fv.max()), (mv.min(), mv.max()))
    fb = np.clip(((fv[idx] - binning.fixed_min) / binning.fixed_bin_size + 2).astype(np.uint32), 2, 29)
    level = MiLevel(b, m, pts, fb, binning)
    centre = tuple(origin + (size - 1) / 2 * spacing)
    t_rigid = LinearTransform.make(0, [0.01, -0.02, 0.015, 0.5, -0.3, 0.2], centre)
    no_stop = InvertParams(20, 0.0, 0.0, 1)

    ops = {
        "cc_pair r4": lambda: b.cc_value_derivative_pair(f, m, 4),
        "smooth_field var3": lambda: b.smooth_field(b.copy_field(u), 3.0),
        "smooth_image sigma 3 vox": lambda: b.smooth_image(f, (3.0, 3.0, 3.0)),
        "compose": lambda: b.compose(u, u2),
        "warp": lambda: b.warp_through_field(m, u),
        "invert (20 iterations)": lambda: b.invert(u, b.zero_field(g), no_stop),
        "max_scaled_norm_pair": lambda: b.max_scaled_norm_pair(u, u),
        "scale": lambda: b.scale(u, 1.0),
        "mi value (1.8M samples)": lambda: level.value(t_rigid),
        "mi value+derivative": lambda: level.value_and_derivative(t_rigid),
    }
    for name, op in ops.items():
        op()
        b.set_profiling(True)
        for _ in range(repeats):
            op()
        prof = b.take_profile()
        b.set_profiling(False)
        for k, (launches, ms, nbytes, _) in sorted(prof.items()):
            if k.startswith("host:") or launches == 0:
                continue
            mean_us = 1e3 * ms / launches
            gbs = nbytes / launches / (mean_us * 1e-6) / 1e9 if nbytes else 0.0
            print(json.dumps({"op": name, "kernel": k, "grid": size.tolist(), "launches": launches,
                              "mean_us": round(mean_us, 2), "bytes_per_launch": round(nbytes / launches),
                              "achieved_gbs": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-engine baseline")
    ap.add_argument("--cpu-config", default="2mm", choices=["1mm", "2mm"],
                    help="the plain registration the CPU engine runs for the baseline (the 2 mm one by default)")
    ap.add_argument("--config", default="1mm", choices=["1mm", "2mm"])
    ap.add_argument("--kernels", action="store_true", help="time each hot-path kernel at 1 mm instead")
    ap.add_argument("--slots", type=int, default=8, help="slots of the served stream (0 skips it)")
    ap.add_argument("--requests", type=int, default=256, help="requests of the served stream")
    ap.add_argument("--load", type=float, default=0.8, help="arrival rate of the served stream as a share of saturation")
    ap.add_argument("--batch", type=int, default=8, help="subjects of the batched two-channel workload (0 skips it)")
    ap.add_argument("--fits", type=int, default=64, help="fits of the batched linear workload (0 skips it)")
    ap.add_argument("--workers", type=int, default=8, help="worker streams of the batched linear fits")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for {args.gpus}): registrations are independent and would shard across "
              "GPUs by request, one server per GPU", file=sys.stderr)
        return 2
    if args.kernels:
        return kernels(max(args.steps, 1) * 10)

    fit_data = []
    if args.fits > 0:
        # generated in worker processes before anything touches the GPU
        from concurrent.futures import ProcessPoolExecutor

        t = time.perf_counter()
        with ProcessPoolExecutor(max_workers=8) as pool:
            fit_data = list(pool.map(make_fit, fit_sizes(args.fits)))
        log(f"linear fits: {args.fits} phantom pairs ready in {time.perf_counter() - t:.1f} s")

    from fastwarp_amd.backend import HipBackend
    from fastwarp_amd.phantom import CONFIGS, make_pair
    from fastwarp_amd.register import cohort_plan, plan_2mm, register, register_resident

    cfg = CONFIGS[args.config]
    t = time.perf_counter()
    pair = make_pair(**cfg)
    log(f"phantom {cfg['shape_xyz']} at {cfg['spacing']} mm ready in {time.perf_counter() - t:.1f} s")
    plan = cohort_plan(1) if args.config == "1mm" else plan_2mm(1)
    backend = HipBackend(0)
    # The raw volumes are in HBM before the timed region starts (fw_upload_image), and the output
    # fields stay there (fw_register_resident); the host-buffer path (fw_register: PCIe both ways
    # and host preprocessing) is timed after it and reported beside it.
    fixed_dev = [backend.upload_image(pair.grid, pair.fixed[0])]
    moving_dev = [backend.upload_image(pair.grid, pair.moving[0])]

    def step():
        r = register_resident(backend, plan, fixed_dev, moving_dev)
        for f in (r.forward, r.inverse):
            if f is not None:
                f.free()
        return r

    # The per-kernel breakdown comes from the last warmup step, every launch timed; the timed
    # steps time every launch of the dominant kernel family only.
    breakdown = None
    for w in range(args.warmup):
        if w == args.warmup - 1:
            backend.set_profiling(True)
        t = time.perf_counter()
        r = step()
        log(f"warmup {w + 1}/{args.warmup}: {time.perf_counter() - t:.2f} s (linear {r.result.seconds_linear:.2f}, syn {r.result.seconds_syn:.2f})")
    if args.warmup:
        breakdown = {k: v for k, v in backend.take_profile().items() if not k.startswith("host:")}
        backend.set_profiling(False)
        focus = dominant(breakdown)
        backend.set_profiling_focus(focus, 1)
    backend.set_profiling(True)
    lat = []
    last = None
    t_all = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        last = step()
        lat.append(time.perf_counter() - t)
        log(f"step {k + 1}/{args.steps}: {lat[-1]:.2f} s (linear {last.result.seconds_linear:.2f}, syn {last.result.seconds_syn:.2f})")
    total = time.perf_counter() - t_all
    profile = backend.take_profile()
    host = {k[5:]: round(v[1] / 1e3 / args.steps, 4) for k, v in profile.items() if k.startswith("host:")}
    profile = {k: v for k, v in profile.items() if not k.startswith("host:")}
    backend.set_profiling(False)
    if breakdown is None:
        breakdown = profile
        focus = dominant(profile)
    held, peak = backend.memory_stats()
    # the host-buffer path on the same pair: one untimed run, then as many timed as above (at most 3)
    register(backend, plan, pair.grid, pair.fixed, pair.grid, pair.moving)
    host_steps = max(1, min(args.steps, 3))
    t = time.perf_counter()
    for _ in range(host_steps):
        register(backend, plan, pair.grid, pair.fixed, pair.grid, pair.moving)
    host_s = (time.perf_counter() - t) / host_steps
    log(f"host-buffer path: {host_s:.3f} s per registration")
    line = {
        "metric": METRIC,
        "value": round(3600.0 * args.steps / total, 3),
        "unit": "registrations/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * total / args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain full normalization {args.config}: centre of mass, Rigid+Affine Mattes MI (32 bins, 25% regular), "
            f"SyN CC radius 4, grid {cfg['shape_xyz']} at {cfg['spacing']} mm, seeded phantom seed {cfg['seed']}",
            "grid": list(cfg["shape_xyz"]),
            "spacing_mm": cfg["spacing"],
            "channels": 1,
            "path": "plain one-at-a-time HIP, raw volumes resident in HBM, fields left in HBM",
        },
        "host_buffers": {"value": round(3600.0 / host_s, 3), "unit": "registrations/hour", "ms_per_step": round(1000.0 * host_s, 1),
                         "path": "fw_register: f64 host volumes in, host preprocessing, PCIe uploads, fields downloaded to the host"},
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "phases_s": {"linear": round(last.result.seconds_linear, 3), "syn": round(last.result.seconds_syn, 3),
                     "preprocess": round(last.result.seconds_preprocess, 3)},
        "linear_levels": [[{"iterations": last.result.linear_iterations[st][l], "evaluations": last.result.linear_evaluations[st][l],
                            "seconds": round(last.result.linear_level_seconds[st][l], 4)} for l in range(4)] for st in range(2)],
        "syn_levels": [{"iterations": last.result.syn_iterations[l],
                        "mean_inversion_steps": round(last.result.syn_inversion_steps[l] / max(4 * last.result.syn_iterations[l], 1), 2),
                        "seconds": round(last.result.syn_level_seconds[l], 4)}
                       for l in range(last.result.syn_levels)],
        "host_s_per_step": dict(sorted(host.items(), key=lambda kv: -kv[1])),
        "device_peak_bytes": int(peak),
        "roofline": roofline(focus, profile, args.steps, breakdown),
        "kernels_from": "the last warmup step, every launch timed" if args.warmup else f"the {args.steps} timed steps",
        "kernels": {k: {"launches": v[0], "ms": round(v[1], 3), "gbs": round(v[2] / max(v[1], 1e-9) / 1e6, 1)}
                    for k, v in sorted(breakdown.items(), key=lambda kv: -kv[1][1])},
    }
    if fit_data:
        backend.release_idle_buffers()
        try:
            line["linear_fits"] = linear_fits(backend, fit_data, args.workers)
        except Exception as e:  # the plain line is still reported
            log(f"batched linear fits failed: {e}")
            line["linear_fits"] = {"error": str(e)}
        fit_data = None
    if args.batch > 0:
        backend.release_idle_buffers()
        try:
            line["batched"] = batched(backend, args.batch)
        except Exception as e:  # the plain line is still reported
            log(f"batched workload failed: {e}")
            line["batched"] = {"error": str(e)}
    if args.slots > 0:
        backend.release_idle_buffers()
        try:
            line["served"] = served(args.slots, args.requests, args.load)
        except Exception as e:  # the plain line is still reported
            log(f"served stream failed: {e}")
            line["served"] = {"error": str(e)}
    if not args.no_cpu:
        log(f"CPU-engine baseline: one complete {args.cpu_config} registration")
        line["cpu_baseline"] = cpu_baseline(args.cpu_config)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
